"""Pins for the FP64 oracle (``oracle/ref_np.py``), CPU only.

Each test checks an oracle function against something that does not share its
code: a known-answer value, an exact rational, a brute-force evaluation with a
different loop structure, a scipy-assembled matrix, or a closed form that a
plausible mistake (wrong sign, dropped term, shifted index, transposed axis)
would break.  The reference's own property tests (tests/test_pfflip.py,
tests/test_poisson.py in the reference repository) are restated at the end.
"""

import math
from fractions import Fraction

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import ref_np as O


# --------------------------------------------------------------- helpers
def assemble(res, bx, by, bz):
    """Assembled ``A p_i = sum_f beta_f (p_i - p_nbr)`` (reference
    tests/test_poisson.py:42-78), built here from explicit loops over faces."""
    nx, ny, nz = res
    n = nx * ny * nz
    idx = np.arange(n).reshape(res)
    rows, cols, vals = [], [], []
    diag = np.zeros(n)
    for axis, bf in enumerate((bx, by, bz)):
        for f in range(1, res[axis]):
            sl_f = [slice(None)] * 3
            sl_f[axis] = f
            sl_a = [slice(None)] * 3
            sl_a[axis] = f - 1
            sl_b = [slice(None)] * 3
            sl_b[axis] = f
            b = np.asarray(bf[tuple(sl_f)], np.float64).ravel()
            a_id = idx[tuple(sl_a)].ravel()
            b_id = idx[tuple(sl_b)].ravel()
            rows += [a_id, b_id]
            cols += [b_id, a_id]
            vals += [-b, -b]
            np.add.at(diag, a_id, b)
            np.add.at(diag, b_id, b)
    rows.append(np.arange(n))
    cols.append(np.arange(n))
    vals.append(diag)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, n))


def direct_solve(A, b):
    """Exact solve of the singular Neumann system: pin p_0 = 0 (drop row and
    column 0), then remove the mean.  ``b`` must have zero mean."""
    A = A.tocsc()
    x = np.zeros(A.shape[0])
    x[1:] = spla.spsolve(A[1:, 1:], np.asarray(b).ravel()[1:])
    return x - x.mean()


def random_beta(res, rng, contrast=1.0):
    """Reference tests/test_poisson.py:23-39."""
    nx, ny, nz = res
    bx = rng.random((nx + 1, ny, nz))
    by = rng.random((nx, ny + 1, nz))
    bz = rng.random((nx, ny, nz + 1))
    if contrast != 1.0:
        for arr in (bx, by, bz):
            arr[:, : arr.shape[1] // 2, :] /= contrast
    bx[0] = bx[-1] = 0.0
    by[:, 0] = by[:, -1] = 0.0
    bz[:, :, 0] = bz[:, :, -1] = 0.0
    return bx, by, bz


# --------------------------------------------------------------- seeding
def test_pcg64_known_answer():
    """numpy's PCG64 stream for seed 42 (documented values)."""
    rng = np.random.default_rng(42)
    assert rng.random(3).tolist() == [0.7739560485559633, 0.4388784397520523, 0.8585979199113825]


@pytest.mark.parametrize("ppc,jitter", [(8, 0.8), (27, 0.5)])
def test_lattice_sites_and_jitter_bound(ppc, jitter):
    res, dx = (3, 4, 5), 0.25
    pos = O.lattice(res, dx, ppc, np.random.default_rng(7), jitter)
    pa = round(ppc ** (1 / 3))
    step = dx / pa
    assert pos.shape == (np.prod(res) * ppc, 3)
    # row r of the 'ij' meshgrid is site (r // (Y*Z), (r // Z) % Y, r % Z)
    Y, Z = res[1] * pa, res[2] * pa
    r = np.arange(len(pos))
    site = np.stack([r // (Y * Z), (r // Z) % Y, r % Z], 1)
    off = pos / step - (site + 0.5)
    assert np.abs(off).max() <= 0.5 * jitter + 1e-12
    assert np.abs(off).max() > 0.45 * jitter  # the jitter is actually applied


def test_lattice_rejects_non_cube():
    with pytest.raises(AssertionError):
        O.lattice((2, 2, 2), 0.5, 9, np.random.default_rng(0))


# --------------------------------------------------------------- rho~_0
def _exact_rho0(pa):
    """Exact rational sum over *all* lattice sites inside the support sphere."""
    acc = Fraction(0)
    for a in range(-2 * pa, 2 * pa):
        for b in range(-2 * pa, 2 * pa):
            for c in range(-2 * pa, 2 * pa):
                s = Fraction((2 * a + 1) ** 2 + (2 * b + 1) ** 2 + (2 * c + 1) ** 2, 4 * pa * pa)
                if s < 1:
                    acc += (1 - s) ** 3
    return acc


@pytest.mark.parametrize("ppc,exact", [(8, Fraction(643, 128)), (27, Fraction(33491, 1944))])
def test_rho_tilde_0_exact(ppc, exact):
    pa = round(ppc ** (1 / 3))
    assert _exact_rho0(pa) == exact
    got = O.rho_tilde_0(1.0 / 64, 1.0 / 64, ppc)
    assert abs(got - float(exact)) <= 1e-13 * float(exact)


def test_rho_tilde_0_closed_form():
    """Reference tests/test_pfflip.py:35-41 and :156-163."""
    dx = 1.0 / 8
    n_p = 8 / dx**3
    analytic = n_p * 64.0 * np.pi * dx**3 / 315.0
    assert abs(O.rho_tilde_0(dx, dx, 8) - analytic) / analytic < 0.05
    for ppc in (8, 27):
        assert abs(O.rho_tilde_0(1 / 16, 1 / 16, ppc) - 0.638 * ppc) / (0.638 * ppc) < 0.10


# --------------------------------------------------------------- P2G
def _brute_splat(px, val, pmass, prad, pstate, shape, o, dx):
    """Every particle against every grid point: no candidate index ranges."""
    I, J, K = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    cen = np.stack([(I + o[0]) * dx, (J + o[1]) * dx, (K + o[2]) * dx], -1)
    m = np.zeros(shape)
    mu = np.zeros(shape)
    for p in range(len(px)):
        if pstate[p] != 0:
            continue
        d2 = ((cen - px[p].astype(np.float64)) ** 2).sum(-1)
        t = np.maximum(1.0 - d2 / float(prad[p]) ** 2, 0.0)
        w = t**3
        m += w * float(pmass[p])
        if val is not None:
            mu += w * float(pmass[p]) * float(val[p])
    return m, mu


def _tiny_particles(res, dx, n, seed):
    rng = np.random.default_rng(seed)
    L = np.array(res) * dx
    # include particles outside the box by up to 1.2 dx and exactly on faces
    px = (rng.random((n, 3)) * (L + 2.4 * dx) - 1.2 * dx).astype(np.float32)
    px[:4] = np.float32(dx) * np.array([[1, 2, 1], [0, 0, 0], [2.5, 1.5, 0.5], [3, 3, 3]], np.float32)
    pv = rng.standard_normal((n, 3)).astype(np.float32)
    ptype = (rng.random(n) < 0.5).astype(np.int32)
    pmass = np.where(ptype == 1, 1000.0, 1.0).astype(np.float32)
    prad = np.full(n, dx, np.float32)
    pstate = np.where(rng.random(n) < 0.1, 1, 0).astype(np.int32)
    return px, pv, ptype, pmass, prad, pstate


def test_p2g_matches_brute_force():
    res, dx = (4, 3, 5), 0.25
    px, pv, ptype, pmass, prad, pstate = _tiny_particles(res, dx, 300, 3)
    for ax, shape in enumerate(O.face_shapes(res)):
        m, mu = O.p2g_face(px, pv, pmass, prad, pstate, shape, ax, dx)
        bm, bmu = _brute_splat(px, pv[:, ax], pmass, prad, pstate, shape, O.face_offset(ax), dx)
        assert np.allclose(m, bm, rtol=1e-12, atol=1e-12)
        assert np.allclose(mu, bmu, rtol=1e-12, atol=1e-9)
    cm, cw = O.p2g_cell(px, pmass, prad, pstate, res, dx)
    bm, _ = _brute_splat(px, None, pmass, prad, pstate, res, np.full(3, 0.5), dx)
    bw, _ = _brute_splat(px, None, np.ones(len(px)), prad, pstate, res, np.full(3, 0.5), dx)
    assert np.allclose(cm, bm, rtol=1e-12, atol=1e-12)
    assert np.allclose(cw, bw, rtol=1e-12, atol=1e-12)


def test_p2g_droplets_excluded():
    """Reference tests/test_pfflip.py:170-189: pstate != 0 does not splat."""
    res, dx = (4, 4, 4), 0.25
    px, pv, ptype, pmass, prad, pstate = _tiny_particles(res, dx, 200, 5)
    all_on = np.zeros_like(pstate)
    m_all, _ = O.p2g_cell(px, pmass, prad, all_on, res, dx)
    m_some, _ = O.p2g_cell(px, pmass, prad, pstate, res, dx)
    m_off, _ = O.p2g_cell(px[pstate != 0], pmass[pstate != 0], prad[pstate != 0],
                          all_on[pstate != 0], res, dx)
    assert np.allclose(m_all, m_some + m_off, rtol=1e-12)


def test_face_velocity_threshold():
    m = np.array([0.0, 1e-13, 2.0])
    v, ok = O.face_velocity_from_momentum(m, np.array([1.0, 1.0, 3.0]), 1e-12)
    assert v.tolist() == [0.0, 0.0, 1.5] and ok.tolist() == [0, 0, 1]


# --------------------------------------------------------------- binning
def test_bin_keys_hand_cases():
    res, dx, T = (16, 8, 8), 1.0 / 16, 8
    px = np.array([[0.0, 0.0, 0.0],          # cell (0,0,0)
                   [15.99 / 16, 0.1, 0.2],   # cell (15,1,3) -> tile (1,0,0)
                   [-0.5, 2.0, 0.5],         # clamped to (0,7,8->7)
                   [9 / 16, 3 / 16, 7 / 16]], np.float32)
    keys = O.bin_keys(px, np.zeros(4, np.int32), res, dx, T)
    # tile grid (2,1,1): tile id = tx; local = (li*8 + lj)*8 + lk
    exp = [0, 1 * 512 + (7 * 8 + 1) * 8 + 3, (0 * 8 + 7) * 8 + 7, 512 + (1 * 8 + 3) * 8 + 7]
    assert keys.tolist() == exp
    dead = O.bin_keys(px, np.array([0, 1, 2, 0], np.int32), res, dx, T)
    assert dead.tolist() == [exp[0], 0xFFFFFFFF, 0xFFFFFFFF, exp[3]]


def test_sort_is_stable_counting_sort():
    keys = np.array([5, 1, 5, 0xFFFFFFFF, 1, 0, 5], np.uint32)
    perm, start = O.sort_by_key(keys, 8)
    assert perm.tolist() == [5, 1, 4, 0, 2, 6]
    assert start.tolist() == [0, 1, 3, 3, 3, 3, 6, 6, 6]


# --------------------------------------------------------------- phase field
def test_phase_field_perfect_lattice_closed_form():
    """Zero-jitter lattice: interior raw density is exactly rho~_0 * rho_phase."""
    cfg = O.OracleConfig(res=(6, 6, 6), dx=1 / 6, ppc=8)
    rt0, rho_0l, rho_min = O.derived_densities(cfg)
    pos = O.lattice(cfg.res, cfg.dx, 8, np.random.default_rng(0), jitter=0.0).astype(np.float32)
    n = len(pos)
    for rho, expect in ((cfg.rho_l, math.sqrt((rho_0l - rho_min) / rho_0l)), (cfg.rho_g, 0.0)):
        cm, _ = O.p2g_cell(pos, np.full(n, rho, np.float32), np.full(n, cfg.dx, np.float32),
                           np.zeros(n, np.int32), cfg.res, cfg.dx)
        inner = cm[1:-1, 1:-1, 1:-1]
        assert np.allclose(inner, rt0 * rho, rtol=1e-6)
        phi, tag = O.cell_phase(cm, rho_min, cfg.alpha_phi, rho_0l)
        if expect == 0.0:
            assert phi.max() == 0.0 and tag.max() == 0
        else:
            assert np.allclose(phi[1:-1, 1:-1, 1:-1], expect, rtol=1e-6)
            assert tag[1:-1, 1:-1, 1:-1].min() == 1


def test_face_beta_cases():
    """Liquid-liquid faces get phi=1, air-air phi=0, mixed use eq. 7; walls
    copy the inner tag (reference kernels.py:283-311)."""
    tag = np.array([[[1]], [[0]], [[0]]], np.int32).reshape(3, 1, 1)
    mass = np.array([9.0, 5.0, 5.0, 5.0]).reshape(4, 1, 1)
    f, beta = O.face_beta(mass, tag, 0, rho_min=1.0, alpha_phi=1.0, rho_0l=16.0,
                          rho_l=1000.0, rho_g=1.0)
    assert f.ravel().tolist() == [1.0, 0.5, 0.0, 0.0]
    assert np.allclose(beta.ravel(), [1e-3, 1 / 500.5, 1.0, 1.0])


# --------------------------------------------------------------- operator
@pytest.mark.parametrize("contrast", [1.0, 1000.0])
def test_operator_matches_assembled(contrast):
    res = (5, 4, 6)
    rng = np.random.default_rng(0)
    bx, by, bz = random_beta(res, rng, contrast)
    A = assemble(res, bx, by, bz)
    assert abs(A - A.T).max() < 1e-12
    assert np.abs(np.asarray(A.sum(axis=1)).ravel()).max() < 1e-12
    p = rng.random(res)
    diag = O.compute_diag(bx, by, bz)
    assert np.allclose(diag.ravel(), A.diagonal(), rtol=1e-14)
    got = O.apply_A(p, bx, by, bz, diag)
    assert np.allclose(got.ravel(), A @ p.ravel(), rtol=1e-12, atol=1e-14)


def test_galerkin_equals_ptap():
    res = (8, 4, 6)
    rng = np.random.default_rng(2)
    bx, by, bz = random_beta(res, rng, 1000.0)
    A = assemble(res, bx, by, bz)
    cres = (4, 2, 3)
    agg = np.zeros(res, np.int64)
    I, J, K = np.meshgrid(*[np.arange(n) for n in res], indexing="ij")
    agg = ((I // 2) * cres[1] + J // 2) * cres[2] + K // 2
    P = sp.csr_matrix((np.ones(agg.size), (np.arange(agg.size), agg.ravel())),
                      shape=(agg.size, int(np.prod(cres))))
    fs = O.face_shapes(cres)
    cb = [O.galerkin_coarsen_faces(b, ax, fs[ax]) for ax, b in enumerate((bx, by, bz))]
    Ac = assemble(cres, *cb)
    assert abs(P.T @ A @ P - Ac).max() < 1e-12
    # R = P^T and P as used by the V-cycle
    r = rng.random(res)
    assert np.allclose(O.restrict_sum(r).ravel(), P.T @ r.ravel())
    c = rng.random(cres)
    assert np.allclose(O.prolong(c).ravel(), P @ c.ravel())


def test_level_dims():
    assert O.level_dims((64, 64, 64), 4) == [(64,) * 3, (32,) * 3, (16,) * 3, (8,) * 3, (4,) * 3]
    assert len(O.level_dims((128,) * 3, 4)) == 6 and len(O.level_dims((256,) * 3, 4)) == 7
    assert O.level_dims((12, 8, 8), 4) == [(12, 8, 8), (6, 4, 4)]
    assert O.level_dims((10, 10, 10), 4) == [(10,) * 3, (5,) * 3]


def test_colour_is_checkerboard():
    m0 = O.colour_mask((3, 3, 2), 0)
    assert m0[0, 0, 0] and not m0[1, 0, 0] and not m0[0, 0, 1] and m0[1, 1, 0]
    assert (m0 ^ O.colour_mask((3, 3, 2), 1)).all()


def test_rbgs_half_sweep_equals_scipy_rows():
    res = (4, 5, 6)
    rng = np.random.default_rng(4)
    bx, by, bz = random_beta(res, rng, 1000.0)
    A = assemble(res, bx, by, bz)
    diag = O.compute_diag(bx, by, bz)
    p = rng.standard_normal(res)
    b = rng.standard_normal(res)
    omega = 6 / 7
    for colour in (0, 1):
        ref = p.ravel().copy()
        res_v = b.ravel() - A @ ref
        m = O.colour_mask(res, colour).ravel() & (A.diagonal() > 0)
        ref[m] += omega * res_v[m] / A.diagonal()[m]
        got = p.copy()
        O.rbgs_sweep(got, b, bx, by, bz, diag, colour, omega)
        assert np.allclose(got.ravel(), ref, rtol=1e-13, atol=1e-13)
        p = got


@pytest.mark.parametrize("contrast", [1.0, 1000.0])
def test_mgpcg_matches_direct_solve(contrast):
    """Reference tests/test_poisson.py:113-145, tightened for FP64."""
    res = (16, 16, 16)
    rng = np.random.default_rng(1)
    bx, by, bz = random_beta(res, rng, contrast)
    A = assemble(res, bx, by, bz)
    b = rng.standard_normal(res)
    b -= b.mean()
    ref = direct_solve(A, b)
    s = O.MGPCGOracle(res)
    s.build_hierarchy(bx, by, bz)
    x, it, rel = s.solve(b, tol=1e-10, max_iter=200)
    assert rel < 1e-10 and it < 60
    err = np.abs(x.ravel() - ref).max() / np.abs(ref).max()
    assert err < 1e-7


# --------------------------------------------------------------- G2P / advect
def test_sample_face_reproduces_linear_fields():
    res, dx = (6, 5, 7), 0.1
    rng = np.random.default_rng(9)
    coef = rng.standard_normal(4)
    x = (rng.random((500, 3)) * (np.array(res) - 2) + 1) * dx  # keep away from clamps
    for ax, shape in enumerate(O.face_shapes(res)):
        o = O.face_offset(ax)
        I, J, K = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
        f = coef[0] * (I + o[0]) * dx + coef[1] * (J + o[1]) * dx + coef[2] * (K + o[2]) * dx + coef[3]
        got = O.sample_face(f, x, dx, ax)
        assert np.allclose(got, x @ coef[:3] + coef[3], rtol=1e-12, atol=1e-12)
    # clamping: a constant field stays constant everywhere, even outside
    f = np.full(O.face_shapes(res)[1], 2.5)
    assert np.allclose(O.sample_face(f, x * 3 - 1, dx, 1), 2.5)


def test_rk3_is_third_order_taylor_for_linear_fields():
    res, dx, dt = (10, 10, 10), 0.1, 0.05
    A = np.array([[0.1, -0.7, 0.2], [0.5, -0.2, 0.1], [0.0, 0.3, 0.1]])
    vel = []
    for ax, shape in enumerate(O.face_shapes(res)):
        o = O.face_offset(ax)
        I, J, K = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
        pts = np.stack([(I + o[0]) * dx, (J + o[1]) * dx, (K + o[2]) * dx], -1)
        vel.append(pts @ A[ax])
    rng = np.random.default_rng(3)
    x = 0.3 + 0.4 * rng.random((50, 3))
    n = len(x)
    new = O.advect_rk3(x, np.ones(n, np.int32), np.zeros(n, np.int32), vel, vel, dt, dx,
                       np.zeros(3), np.ones(3))
    M = dt * A
    taylor = np.eye(3) + M + M @ M / 2 + M @ M @ M / 6
    assert np.allclose(new, x @ taylor.T, rtol=1e-12, atol=1e-13)


def test_g2p_flip_pic_blend():
    res, dx = (4, 4, 4), 0.25
    shapes = O.face_shapes(res)
    c = np.array([0.3, -0.2, 0.7])
    vel = [np.full(s, 1.0 + c[a]) for a, s in enumerate(shapes)]
    dvel = [np.full(s, c[a]) for a, s in enumerate(shapes)]
    x = np.array([[0.5, 0.5, 0.5], [0.2, 0.7, 0.9]])
    pv = np.array([[1.0, 2.0, 3.0], [-1.0, 0.0, 1.0]])
    ptype = np.array([1, 0], np.int32)
    new = O.g2p(x, pv, ptype, np.zeros(2, np.int32), vel, dvel, dx, 0.985, 0.96)
    for p, a in ((0, 0.985), (1, 0.96)):
        assert np.allclose(new[p], a * (pv[p] + c) + (1 - a) * (1 + c))


def test_reflect_at_walls():
    lo, hi = np.full(3, 0.1), np.full(3, 0.9)
    x = np.array([[0.05, 0.5, 0.95], [0.5, 0.5, 0.5]])
    v = np.array([[-1.0, 2.0, 3.0], [-1.0, -1.0, -1.0]])
    nx, nv = O.reflect_at_walls(x, v, np.zeros(2, np.int32), lo, hi)
    assert nx[0].tolist() == [0.1, 0.5, 0.9] and nv[0].tolist() == [0.0, 2.0, 0.0]
    assert nv[1].tolist() == [-1.0, -1.0, -1.0]


# --------------------------------------------------------------- projection
def test_projection_of_exact_solve_is_divergence_free():
    res, dx, dt = (6, 5, 4), 0.2, 0.01
    rng = np.random.default_rng(11)
    beta = list(random_beta(res, rng, 1000.0))
    for b in beta:
        b += 0.0
    vel = [rng.standard_normal(s) for s in O.face_shapes(res)]
    for ax in range(3):
        O.zero_solid_faces(beta[ax], vel[ax], ax)
    rhs = O.build_rhs(vel, None, dx, dt, 1.0, 0.0)
    A = assemble(res, *beta)
    p = direct_solve(A, rhs - rhs.mean()).reshape(res)
    before = O.max_divergence(vel, dx)
    for ax in range(3):
        O.apply_pressure_gradient(p, beta[ax], vel[ax], ax, dx, dt)
    # closed box: divergence sums to 0, so removing the mean of rhs is exact
    assert O.max_divergence(vel, dx) < 1e-8 * before


# --------------------------------------------------------------- extrapolation
def test_extrapolation_single_liquid_cell():
    res = (5, 5, 5)
    liq = np.zeros(res, np.int32)
    liq[2, 2, 2] = 1
    u = np.arange(6 * 5 * 5, dtype=np.float64).reshape(6, 5, 5) + 1.0
    vl, kn = O.seed_liquid_velocity(u, liq, 0)
    assert kn.sum() == 2 and kn[2, 2, 2] == 1 and kn[3, 2, 2] == 1
    assert vl[2, 2, 2] == u[2, 2, 2] and vl[0, 0, 0] == 0.0
    v1, k1 = O.extrapolate_sweep(vl, kn)
    # face (1,2,2) sees only (2,2,2); face (2,3,2) sees only (2,2,2)
    assert v1[1, 2, 2] == u[2, 2, 2] and v1[2, 3, 2] == u[2, 2, 2]
    # face (4,2,2) sees (3,2,2); face (2,2,2) kept
    assert v1[4, 2, 2] == u[3, 2, 2] and v1[2, 2, 2] == u[2, 2, 2]
    assert k1.sum() == 2 + 2 + 2 * 4   # seeds + 2 axial + 4 lateral per seed
    assert v1[3, 3, 2] == u[3, 2, 2]
    v2, k2 = O.extrapolate_sweep(v1, k1)
    # face (2,3,3) now sees (2,3,2) and (2,2,3): the mean of two equal values
    assert k2[2, 3, 3] == 1 and v2[2, 3, 3] == u[2, 2, 2]
    # face (1,3,2) sees (1,2,2) and (2,3,2), both carry u[2,2,2]
    assert v2[1, 3, 2] == u[2, 2, 2]


# --------------------------------------------------------------- whole step
@pytest.fixture(scope="module")
def pool():
    cfg = O.OracleConfig(res=(16, 16, 16), dx=1.0 / 16)
    sim = O.OracleSim(cfg)
    sim.set_particles(*O.still_pool(cfg, fill=0.5))
    return sim


def test_step_phase_field_saturates(pool):
    """Reference tests/test_pfflip.py:44-62."""
    pool.step()
    phi = pool.phi_cell
    ny = phi.shape[1]
    assert phi[:, 3 * ny // 4:, :].max() == 0.0
    liq = phi[:, : ny // 4, :]
    assert liq.mean() > 0.95 and liq.min() > 0.6


def test_step_hydrostatic_and_conservation(pool):
    """Reference tests/test_pfflip.py:77-103 (shortened)."""
    liq = pool.ptype == 1
    m0 = pool.pmass.sum()
    y0 = pool.px[liq, 1].mean()
    speeds = []
    for _ in range(4):
        pool.step()
        speeds.append(np.linalg.norm(pool.pv[liq], axis=1).max())
    assert pool.pmass.sum() == m0 and (pool.ptype == 1).sum() == liq.sum()
    g_dt = 9.81 * pool.stats[-1]["dt"]
    assert max(speeds) < 0.5 * g_dt
    assert abs(pool.px[liq, 1].mean() - y0) < 0.01


def test_projection_reduces_divergence_on_dam_break():
    """Reference tests/test_pfflip.py:106-120."""
    cfg = O.OracleConfig(res=(16, 16, 16), dx=1.0 / 16, pcg_tol=1e-5)
    sim = O.OracleSim(cfg)
    sim.set_particles(*O.dam_break(cfg))
    sim._p2g()
    sim._phase_field_and_coefficients()
    sim.vel_star = [v.copy() for v in sim.vel]
    dt = 1e-3
    sim._add_forces(dt)
    before = sim._max_divergence()
    sim._pressure_solve(dt)
    sim._project(dt)
    assert sim._max_divergence() < 0.05 * before


# --------------------------------------------------------------- droplets
def test_rng_uniform_and_normal_statistics():
    """The stream is uniform on [0, 1) in 24-bit steps and its Box-Muller
    normals have the moments of N(0, 1); neighbouring particles' streams
    are uncorrelated."""
    n = 1 << 18
    st = O.rng_init(7, np.arange(n))
    st1, u = O.rng_uniform(st)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert np.all(u * 16777216.0 == np.floor(u * 16777216.0))
    assert abs(u.mean() - 0.5) < 4 * math.sqrt(1 / 12 / n)
    assert abs(u.var() - 1 / 12) < 2e-3
    hist = np.bincount((u * 64).astype(int), minlength=64)
    chi2 = float(((hist - n / 64) ** 2 / (n / 64)).sum())
    assert chi2 < 63 + 6 * math.sqrt(2 * 63)
    assert abs(np.corrcoef(u[:-1], u[1:])[0, 1]) < 0.01
    _, u2 = O.rng_uniform(st1)
    assert abs(np.corrcoef(u, u2)[0, 1]) < 0.01
    _, z = O.rng_normal(O.rng_init(11, np.arange(n)))
    assert abs(z.mean()) < 4 / math.sqrt(n)
    assert abs(z.std() - 1.0) < 0.01
    assert abs((z ** 4).mean() - 3.0) < 0.1
    # distinct seeds give distinct streams
    _, ua = O.rng_uniform(O.rng_init(1, np.arange(1000)))
    _, ub = O.rng_uniform(O.rng_init(2, np.arange(1000)))
    assert not np.array_equal(ua, ub)


def test_lognormal_radii_have_the_requested_mean():
    """E[exp(mu + sigma Z)] = exp(mu + sigma^2 / 2) = mean_radius."""
    mean_r = 0.25 / 64
    mu, sigma = O.lognormal_params(mean_r, 0.5)
    n = 1 << 18
    _, r = O._draw_radius(O.rng_init(3, np.arange(n)), mu, sigma, 0.0, 1.0)
    assert abs(r.mean() / mean_r - 1.0) < 4 * math.sqrt(math.exp(sigma ** 2) - 1) / math.sqrt(n)
    assert abs(np.log(r).std() - sigma) < 0.01


def test_sample_cell_linear_field_and_clamped_edges():
    """Trilinear sampling reproduces a linear field at any interior point and
    holds the edge cell's value within half a cell of the walls."""
    res, dx = (6, 5, 7), 0.1
    c = np.stack(np.meshgrid(*[(np.arange(n) + 0.5) * dx for n in res], indexing="ij"), -1)
    f = 1.0 + 2.0 * c[..., 0] - 3.0 * c[..., 1] + 0.5 * c[..., 2]
    rng = np.random.default_rng(0)
    x = 0.5 * dx + rng.random((500, 3)) * (np.array(res) - 1) * dx
    got = O.sample_cell(f, x, dx)
    assert np.abs(got - (1.0 + 2.0 * x[:, 0] - 3.0 * x[:, 1] + 0.5 * x[:, 2])).max() < 1e-12
    corner = np.array([[0.2 * dx, 0.3 * dx, 0.1 * dx]])
    assert O.sample_cell(f, corner, dx)[0] == f[0, 0, 0]


def _lone_particle_phi(ppc, ptype_lone, ptype_bulk):
    """Eq. (7) phi at a particle placed on a cell centre of a perfect lattice
    of another phase, through the oracle's P2G, phase field and sampling."""
    cfg = O.OracleConfig(res=(8, 8, 8), dx=1 / 8, ppc=ppc)
    rt0, rho_0l, rho_min = O.derived_densities(cfg)
    pos = O.lattice(cfg.res, cfg.dx, ppc, np.random.default_rng(0), jitter=0.0)
    centre = np.array([[4.5, 4.5, 4.5]]) * cfg.dx
    pos = np.concatenate([pos, centre])
    pt = np.full(len(pos), ptype_bulk, np.int32)
    pt[-1] = ptype_lone
    mass = np.where(pt == 1, cfg.rho_l, cfg.rho_g).astype(np.float32)
    cm, _ = O.p2g_cell(pos, mass, np.full(len(pos), cfg.dx, np.float32), np.zeros(len(pos), np.int32),
                       cfg.res, cfg.dx)
    phi, _ = O.cell_phase(cm, rho_min, cfg.alpha_phi, rho_0l)
    return cfg, pos, pt, phi, rt0, rho_0l, rho_min


@pytest.mark.parametrize("ppc,converts", [(8, False), (27, True)])
def test_lone_liquid_particle_converts_only_above_17_ppc(ppc, converts):
    """Reference tests/test_pfflip.py:140-167 and the eq. (7) closed form: a
    liquid particle alone in air sees phi = sqrt((rho_l + rho~_0 rho_g -
    rho_min) / rho_0l) at its own cell centre, which is below 1/2 - delta_phi
    at 27 ppc (0.228) and above it at 8 ppc (0.440)."""
    cfg, pos, pt, phi, rt0, rho_0l, rho_min = _lone_particle_phi(ppc, 1, 0)
    expect = math.sqrt((cfg.rho_l + rt0 * cfg.rho_g - rho_min) / rho_0l)
    # 1e-5: the 27 ppc lattice sites (thirds of a cell) are rounded to FP32
    assert abs(phi[4, 4, 4] - expect) < 1e-5
    mu, sigma = O.lognormal_params(0.25 * cfg.dx)
    st, r, c, f = O.classify_escaped(pos, pt, np.zeros(len(pos), np.int32), np.zeros(len(pos)), phi,
                                     cfg.dx, 0.2, mu, sigma, 0.02 * cfg.dx, 2 * cfg.dx, 42)
    assert abs(f[-1] - expect) < 1e-5
    assert (st[-1] == O.STATE_DROPLET) == converts
    assert c[0] == int(converts) and c[1] == 0 and (st[:-1] == 0).all()
    if converts:
        assert 0.02 * cfg.dx <= r[-1] <= 2 * cfg.dx


def test_air_in_liquid_is_deleted_and_droplet_merges_back():
    """An air particle inside the liquid body sees phi =
    sqrt((rho~_0 rho_l + rho_g - rho_min) / rho_0l) = 0.9966 (> 0.7): a
    bubble, deleted.  A droplet there merges back into FLIP."""
    cfg, pos, pt, phi, rt0, rho_0l, rho_min = _lone_particle_phi(8, 0, 1)
    expect = math.sqrt((rt0 * cfg.rho_l + cfg.rho_g - rho_min) / rho_0l)
    st0 = np.zeros(len(pos), np.int32)
    st0[0] = O.STATE_DROPLET   # a liquid droplet at a lattice site in the bulk
    st0[1] = O.STATE_DEAD
    st, _, c, f = O.classify_escaped(pos, pt, st0, np.zeros(len(pos)), phi, cfg.dx, 0.2, -5.0, 0.5,
                                     0.0, 1.0, 42)
    assert abs(f[-1] - expect) < 1e-12 and st[-1] == O.STATE_DEAD
    assert st[0] == O.STATE_FLIP and st[1] == O.STATE_DEAD
    assert c == [0, 1, 1]


def test_velocity_gradient_of_a_linear_field():
    """u_a = sum_b G_ab x_b: interior cells see sqrt of the sum of G_ab^2 over
    the seven terms the reference keeps (all but dv/dz and dw/dy)."""
    res, dx = (6, 7, 5), 0.1
    G = np.array([[0.3, -1.2, 0.7], [2.0, -0.4, 1.5], [-0.9, 0.8, 0.1]])
    vel = []
    for a, shape in enumerate(O.face_shapes(res)):
        o = [0.5, 0.5, 0.5]
        o[a] = 0.0
        x = np.stack(np.meshgrid(*[(np.arange(n) + o[b]) * dx for b, n in enumerate(shape)],
                                 indexing="ij"), -1)
        vel.append(x @ G[a])
    got = O.velocity_gradient_magnitude(vel, dx)
    keep = np.ones((3, 3), bool)
    keep[1, 2] = keep[2, 1] = False
    expect = math.sqrt(float((G[keep] ** 2).sum()))
    assert np.abs(got[1:-1, 1:-1, 1:-1] - expect).max() < 1e-12
    # at a wall cell the clamped centred difference halves the cross term
    assert got[0, 3, 2] != pytest.approx(expect)


def _one_droplet(x, v, r=0.004):
    return (np.array([x], np.float64), np.array([v], np.float64), np.array([1], np.int32),
            np.array([r]))


def test_ballistic_droplet_matches_the_discrete_closed_form():
    """No drag: N semi-implicit Euler sub-steps of h = dt/N give v = v0 + N h g
    and x = x0 + N h v0 + h^2 g N (N+1)/2, with N = floor(|v0| dt / (dx/2)) + 1."""
    res, dx, dt = (16, 16, 16), 1 / 16, 0.02
    vel = [np.zeros(s) for s in O.face_shapes(res)]
    g = np.array([0.0, 0.0, -9.81])
    for v0, n_expect in (((0.5, 0.0, 0.0), 1), ((2.0, 0.0, 1.0), 2), ((3.5, 3.5, 0.0), 4)):
        x0 = np.array([0.5, 0.5, 0.5])
        v0 = np.array(v0)
        x, v, st, r = O.advance_droplets(*_one_droplet(x0, v0), vel, dt, dx, g, 1.0, 1000.0, 0.0, 0.0,
                                         0.0, np.full(3, 1e-4 * dx), np.full(3, 1 - 1e-4 * dx), 8)
        n = n_expect
        h = dt / n
        assert np.allclose(v[0], v0 + n * h * g, atol=1e-14)
        assert np.allclose(x[0], x0 + n * h * v0 + h * h * g * n * (n + 1) / 2, atol=1e-14)
        assert st[0] == O.STATE_DROPLET


def test_droplet_at_terminal_velocity_stays_there():
    """Gravity balances sphere drag, 1/2 rho_g Cd pi r^2 v^2 = rho_l 4/3 pi r^3 g,
    at v_t = sqrt(8 rho_l r g / (3 rho_g Cd)): a droplet falling through still
    air at v_t keeps its velocity."""
    res, dx, dt = (16, 16, 16), 1 / 16, 1e-3
    vel = [np.zeros(s) for s in O.face_shapes(res)]
    r, rho_l, rho_g, cd, g = 0.004, 1000.0, 1.0, 0.5, 9.81
    vt = math.sqrt(8 * rho_l * r * g / (3 * rho_g * cd))
    x, v, st, _ = O.advance_droplets(*_one_droplet([0.5, 0.5, 0.9], [0.0, 0.0, -vt], r), vel, dt, dx,
                                     (0.0, 0.0, -g), rho_g, rho_l, cd, 0.0, 0.0,
                                     np.full(3, 1e-6), np.full(3, 1 - 1e-6), 8)
    assert abs(v[0, 2] + vt) < 1e-9 * vt and st[0] == O.STATE_DROPLET
    # and half the terminal speed accelerates downwards
    _, v2, _, _ = O.advance_droplets(*_one_droplet([0.5, 0.5, 0.9], [0.0, 0.0, -0.5 * vt], r), vel, dt,
                                     dx, (0.0, 0.0, -g), rho_g, rho_l, cd, 0.0, 0.0,
                                     np.full(3, 1e-6), np.full(3, 1 - 1e-6), 8)
    assert v2[0, 2] < -0.5 * vt


def test_evaporation_follows_r_squared_law_and_kills_small_droplets():
    """dr/dt = -k/r integrates to r^2 = r0^2 - 2 k t; midpoint RK2 over a
    step stays within O(h^3) of it.  A droplet that shrinks below r_min dies
    and keeps its pre-advance state."""
    res, dx = (16, 16, 16), 1 / 16
    vel = [np.zeros(s) for s in O.face_shapes(res)]
    r0, k, dt = 0.004, 1e-7, 0.01
    x, v, st, r = O.advance_droplets(*_one_droplet([0.5, 0.5, 0.5], [0.0, 0.0, 0.0], r0), vel, dt, dx,
                                     (0.0, 0.0, 0.0), 1.0, 1000.0, 0.5, k, 1e-4, np.full(3, 1e-6),
                                     np.full(3, 1 - 1e-6), 8)
    exact = math.sqrt(r0 * r0 - 2 * k * dt)
    assert abs(r[0] - exact) < 1e-3 * (r0 - exact) and r[0] < r0
    x0 = np.array([[0.5, 0.5, 0.5]])
    x, v, st, r = O.advance_droplets(x0.copy(), np.array([[1.0, 0.0, 0.0]]), np.array([1], np.int32),
                                     np.array([0.001]), vel, dt, dx, (0.0, 0.0, 0.0), 1.0, 1000.0, 0.5,
                                     1e-4, 5e-4, np.full(3, 1e-6), np.full(3, 1 - 1e-6), 8)
    assert st[0] == O.STATE_DEAD and np.array_equal(x, x0) and r[0] == 0.001


def test_droplet_leaving_the_domain_dies():
    res, dx, dt = (16, 16, 16), 1 / 16, 0.02
    vel = [np.zeros(s) for s in O.face_shapes(res)]
    x0 = np.array([[0.99, 0.5, 0.5]])
    x, v, st, _ = O.advance_droplets(x0.copy(), np.array([[2.0, 0.0, 0.0]]), np.array([1], np.int32),
                                     np.array([0.004]), vel, dt, dx, (0.0, 0.0, 0.0), 1.0, 1000.0, 0.5,
                                     0.0, 0.0, np.full(3, 1e-4 * dx), np.full(3, 1 - 1e-4 * dx), 8)
    assert st[0] == O.STATE_DEAD and np.array_equal(x, x0)


def test_rasterize_spray_conserves_volume():
    """Trilinear weights sum to one: sum(density) * cell volume equals the
    droplets' total volume when no weight falls outside the field; a droplet
    on a cell centre puts all of it in that cell."""
    shape, dxs = (12, 12, 12), 0.05
    rng = np.random.default_rng(1)
    x = (1.0 + rng.random((200, 3)) * 10.0) * dxs
    r = rng.random(200) * 0.01
    st = np.ones(200, np.int32)
    st[::7] = 0
    dens = O.rasterize_spray(x, st, r, shape, dxs)
    vol = (4 / 3) * math.pi * (r[st == 1] ** 3).sum()
    assert abs(dens.sum() * dxs ** 3 - vol) < 1e-12 * vol
    one = O.rasterize_spray(np.array([[3.5, 4.5, 5.5]]) * dxs, np.array([1]), np.array([0.01]), shape, dxs)
    assert one[3, 4, 5] == pytest.approx((4 / 3) * math.pi * 1e-6 / dxs ** 3, rel=1e-12)
    assert np.count_nonzero(one) == 1
