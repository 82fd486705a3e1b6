"""FP64 numpy restatement of the adaptive PF-FLIP step on an MSBG
(pfflip/mr_sim.py:203-310, pfflip/mr_kernels.py, pfflip/particles.py).

TEST INFRASTRUCTURE ONLY, like ``ref_np.py`` and ``ref_msbg.py``: imported by
``tests/`` and nowhere in the product.  Each function cites the reference
file:line it restates (paths relative to the reference repository root).
Grids are ``ref_msbg.Grid``; cell addressing is vectorised here
(``locate_many`` / ``origin_many``, pinned against the scalar ``Grid.locate``
/ ``Grid.origin``).

P2G is an all-pairs sum: every cell centre and low face against every
particle, weight ``max(1 - d^2/r^2, 0)^3``.  The reference walks the blocks
the support box touches and the cells in an index range widened by one; a
centre or face within ``r`` of the particle always lies in that walk, and
outside ``r`` the weight is zero, so the sums agree.

Randomness (deletion in classify, split jitter in emit) uses the PCG hash
stream of ``ref_np`` (rng_init(seed, index), 24-bit uniforms), the stream the
HIP kernels draw from.

Pinning (``tests/test_oracle_mr.py``): vectorised addressing against the
scalar restatement; the level-0 lattice P2G against the independent lattice
sum of ``calibrate``; the same lattice at level l on a level-l grid (phi = 1,
the per-level normalisation); single-particle face weights against
hand-placed distances; on a uniform grid the projected velocity of an exact
solve is divergence free (this ties the rhs, the operator and the projection
signs together); trilinear face sampling reproduces linear fields; split
children keep the parent's mass and stay inside its cell; deletion happens at
rate 7/8 once the counter fires; the distance field is scipy's exact EDT (the
routine the reference calls).
"""

from __future__ import annotations

import math

import numpy as np

from . import ref_msbg as M
from .ref_np import rng_init, rng_uniform

LIQUID = 1


# ---------------------------------------------------------------- addressing
def locate_many(g: M.Grid, F):
    """msbg.py:81-108 over an (m, 3) int array -> (addr, level), -1 outside."""
    F = np.asarray(F, np.int64).reshape(-1, 3)
    b = F >> g.sh0
    ok = (F >= 0).all(axis=1) & (b[:, 0] < g.nb[0]) & (b[:, 1] < g.nb[1]) & (b[:, 2] < g.nb[2])
    bidx = np.where(ok, (b[:, 0] * g.nb[1] + b[:, 1]) * g.nb[2] + b[:, 2], 0)
    s = np.where(ok, g.slot[bidx], -1)
    ok &= s >= 0
    s = np.where(ok, s, 0)
    lv = g.blev[s]
    n = g.n0 >> lv
    m = g.n0 - 1
    ijk = (F & m) >> lv[:, None]
    addr = g.coff[s] + (ijk[:, 0] * n + ijk[:, 1]) * n + ijk[:, 2]
    return np.where(ok, addr, -1), np.where(ok, lv, -1)


def origin_many(g: M.Grid, addr=None):
    """msbg.py:111-141 -> (finest low corner (m, 3), level)."""
    a = np.arange(g.ncells) if addr is None else np.asarray(addr, np.int64)
    s = g.cslot[a]
    lv = g.blev[s]
    n = g.n0 >> lv
    r = a - g.coff[s]
    i, r = r // (n * n), r % (n * n)
    j, k = r // n, r % n
    b = g.blk[s]
    bz = b % g.nb[2]
    by = (b // g.nb[2]) % g.nb[1]
    bx = b // (g.nb[2] * g.nb[1])
    o = np.stack([(bx << g.sh0) + (i << lv), (by << g.sh0) + (j << lv), (bz << g.sh0) + (k << lv)], axis=1)
    return o, lv


def locate_at(g, x):
    """msbg.py:154-161: world positions -> (addr, level)."""
    return locate_many(g, np.floor(np.asarray(x, np.float64) / g.h0).astype(np.int64))


# ---------------------------------------------------------------- constants
def calibrate(ppc, dx, r):
    """mr_sim.py:104-118: rho~_0, the weight sum of one cell of a level-0
    lattice, summed over lattice offsets (a direct sum, independent of p2g)."""
    per_axis = round(ppc ** (1.0 / 3.0))
    step = dx / per_axis
    R = int(math.ceil(r / step)) + 1
    acc = 0.0
    for a in range(-R, R + 1):
        for b in range(-R, R + 1):
            for c in range(-R, R + 1):
                d2 = ((a + 0.5) ** 2 + (b + 0.5) ** 2 + (c + 0.5) ** 2) * step * step
                t = max(1.0 - d2 / (r * r), 0.0)
                acc += t * t * t
    return acc


class Consts:
    """The scalars of one step (mr_sim.py:60-64, 203-268)."""

    def __init__(self, cfg, rho_tilde_0, dt, domain, max_level):
        self.r0, self.rho_l, self.rho_g = cfg.r_p, cfg.rho_l, cfg.rho_g
        self.rho_tilde_0 = rho_tilde_0
        self.rho_0l = rho_tilde_0 * cfg.rho_l
        self.rho_min = cfg.eta_phi * cfg.rho_g * rho_tilde_0
        self.alpha_phi, self.k_div = cfg.alpha_phi, cfg.k_div
        self.gravity = np.asarray(cfg.gravity, np.float64)
        self.dt = dt
        self.nu_l = (1.0 - cfg.alpha_flip_liquid) * cfg.dx ** 2 / (6.0 * dt)
        self.nu_g = (1.0 - cfg.alpha_flip_air) * cfg.dx ** 2 / (6.0 * dt)
        self.alpha_cap = 0.999
        eps = 1e-4 * cfg.dx
        self.lo = np.full(3, eps)
        self.hi = np.asarray(domain, np.float64) - eps
        self.max_level = max_level


# ---------------------------------------------------------------- P2G
def p2g(g, x, v, ptype, plev, r0, rho_l, rho_g, chunk=64):
    """mr_kernels.py:48-133 with derive_mass_radius (particles.py:234-248):
    -> cmass, cwsum (ncells), fmass, fmom (3 ncells, face = addr * 3 + axis)."""
    x = np.asarray(x, np.float64)
    v = np.asarray(v, np.float64)
    lv = np.asarray(plev, np.int64)
    mp = np.where(np.asarray(ptype) == LIQUID, rho_l, rho_g) * 8.0 ** lv
    r2 = (r0 * 2.0 ** lv) ** 2
    o, cl = origin_many(g)
    h = g.h0 * 2.0 ** cl
    ctr = (o + 0.5 * (1 << cl)[:, None]) * g.h0
    cmass, cwsum = np.zeros(g.ncells), np.zeros(g.ncells)
    fmass, fmom = np.zeros(g.ncells * 3), np.zeros(g.ncells * 3)
    rmax = float(np.sqrt(r2.max())) if len(x) else 0.0
    for a0 in range(0, g.ncells, chunk):
        sl = slice(a0, min(a0 + chunk, g.ncells))
        c = ctr[sl]
        # particles farther than r from the chunk's box carry zero weight
        lo = c.min(axis=0) - h[sl].max() - rmax
        hi = c.max(axis=0) + rmax
        near = np.flatnonzero(((x >= lo) & (x <= hi)).all(axis=1))
        xs, rs, ms, vs = x[near], r2[near], mp[near], v[near]
        d = c[:, None, :] - xs[None, :, :]  # (cells, particles, 3)
        w = np.maximum(1.0 - (d * d).sum(-1) / rs[None, :], 0.0) ** 3
        cmass[sl] = w @ ms
        cwsum[sl] = w.sum(axis=1)
        for ax in range(3):
            df = d.copy()
            df[:, :, ax] -= 0.5 * h[sl][:, None]
            wf = np.maximum(1.0 - (df * df).sum(-1) / rs[None, :], 0.0) ** 3
            fmass[np.arange(sl.start, sl.stop) * 3 + ax] = wf @ ms
            fmom[np.arange(sl.start, sl.stop) * 3 + ax] = wf @ (ms * vs[:, ax])
    return cmass, cwsum, fmass, fmom


def phase_field(rho, rho_min, alpha_phi, rho_0l):
    """mr_kernels.py:141-146."""
    rho = np.asarray(rho, np.float64)
    f = np.minimum(np.sqrt(np.maximum(rho - rho_min, 0.0) / (alpha_phi * rho_0l)), 1.0)
    return np.where(rho < rho_min, 0.0, f)


def finish(g, cmass, fmass, fmom, c: Consts):
    """face_velocity + cell_phase (mr_kernels.py:136-165) -> phi, is_liquid, fvel."""
    _, lv = origin_many(g)
    phi = phase_field(cmass / 8.0 ** lv, c.rho_min, c.alpha_phi, c.rho_0l)
    eps = 1e-12 * c.rho_0l
    fvel = np.where(fmass > eps, fmom / np.where(fmass > eps, fmass, 1.0), 0.0)
    return phi, (phi >= 0.5).astype(np.int32), fvel


def _low_neighbour(g, o, ax):
    e = np.zeros(3, np.int64)
    e[ax] = 1
    return locate_many(g, o - e)


def face_prepare(g, is_liquid, fmass, fvel, c: Consts):
    """face_beta + zero_wall_velocity + u* copy + add_gravity
    (mr_kernels.py:168-227, mr_sim.py:233-245) -> fbeta, fvel, fvel_star."""
    o, lv = origin_many(g)
    norm = 8.0 ** lv
    fbeta = np.zeros(g.ncells * 3)
    u = np.asarray(fvel, np.float64).copy()
    for ax in range(3):
        nb, _ = _low_neighbour(g, o, ax)
        face = np.arange(g.ncells) * 3 + ax
        lo = np.where(nb >= 0, is_liquid[np.maximum(nb, 0)], 0)
        hi = is_liquid
        f = phase_field(fmass[face] / norm, c.rho_min, c.alpha_phi, c.rho_0l)
        f = np.where((lo == 1) & (hi == 1), 1.0, np.where((lo == 0) & (hi == 0), 0.0, f))
        fbeta[face] = np.where(nb >= 0, 1.0 / ((1.0 - f) * c.rho_g + f * c.rho_l), 0.0)
    u = np.where(fbeta == 0.0, 0.0, u)
    star = u.copy()
    gv = np.tile(c.gravity, g.ncells)
    u = np.where(fbeta != 0.0, u + gv * c.dt, u)
    return fbeta, u, star


def face_velocity_high(g, fvel, ax):
    """mr_kernels.py:230-283 for every cell and one axis."""
    o, la = origin_many(g)
    ext = 1 << la
    e = np.zeros(3, np.int64)
    e[ax] = 1
    F = o + e[None, :] * ext[:, None]
    nb, nl = locate_many(g, F)
    out = np.zeros(g.ncells)
    same = (nb >= 0) & (nl >= la)
    out[same] = fvel[nb[same] * 3 + ax]
    finer = (nb >= 0) & (nl < la)
    if finer.any():
        idx = np.flatnonzero(finer)
        sub = 1 << nl[idx]
        a1, a2 = (ax + 1) % 3, (ax + 2) % 3
        acc = np.zeros(len(idx))
        cnt = np.zeros(len(idx))
        for u in range(2):
            for w in range(2):
                q = F[idx].copy()
                q[:, a1] += u * sub
                q[:, a2] += w * sub
                nc, _ = locate_many(g, q)
                ok = nc >= 0
                acc[ok] += fvel[nc[ok] * 3 + ax]
                cnt[ok] += 1
        out[idx] = np.where(cnt > 0, acc / np.maximum(cnt, 1), 0.0)
    return out


def rhs(g, fvel, cwsum, c: Consts):
    """build_rhs (mr_kernels.py:286-313)."""
    o, lv = origin_many(g)
    h = g.h0 * 2.0 ** lv
    flux = sum((face_velocity_high(g, fvel, ax) - fvel[np.arange(g.ncells) * 3 + ax]) * h * h for ax in range(3))
    r = -flux / c.dt
    if c.k_div > 0.0:
        excess = (cwsum / 8.0 ** lv - c.rho_tilde_0) / c.rho_tilde_0
        r = r - c.k_div * np.maximum(excess, 0.0) * h ** 3 / c.dt
    return r


def project(g, p, fbeta, fvel, fvel_star, dt):
    """project + field_difference (mr_kernels.py:316-343, 415-418) -> fvel, fdvel."""
    o, la = origin_many(g)
    hA = g.h0 * 2.0 ** la
    u = np.asarray(fvel, np.float64).copy()
    for ax in range(3):
        nb, nl = _low_neighbour(g, o, ax)
        face = np.arange(g.ncells) * 3 + ax
        ok = (fbeta[face] != 0.0) & (nb >= 0)
        dist = 0.5 * (hA + g.h0 * 2.0 ** np.maximum(nl, 0))
        grad = (p - p[np.maximum(nb, 0)]) / dist
        u[face] = np.where(ok, u[face] - dt * fbeta[face] * grad, u[face])
    return u, u - fvel_star


def max_divergence(g, fvel):
    """mr_kernels.py:346-353."""
    _, lv = origin_many(g)
    h = g.h0 * 2.0 ** lv
    d = sum(face_velocity_high(g, fvel, ax) - fvel[np.arange(g.ncells) * 3 + ax] for ax in range(3))
    return float(np.abs(d / h).max())


# ---------------------------------------------------------------- G2P / advection
def sample_face(g, f, x, ax):
    """mr_kernels.py:359-405 for an (m, 3) array of positions."""
    x = np.asarray(x, np.float64).reshape(-1, 3)
    addr, lv = locate_at(g, x)
    lvc = np.maximum(lv, 0)
    ext = 1 << lvc
    h = g.h0 * ext
    s = x / h[:, None] - np.where(np.arange(3) == ax, 0.0, 0.5)[None, :]
    i0 = np.floor(s).astype(np.int64)
    t = s - i0
    acc = np.zeros(len(x))
    wsum = np.zeros(len(x))
    for di in range(2):
        wi = t[:, 0] if di else 1.0 - t[:, 0]
        for dj in range(2):
            wj = t[:, 1] if dj else 1.0 - t[:, 1]
            for dk in range(2):
                wk = t[:, 2] if dk else 1.0 - t[:, 2]
                w = wi * wj * wk
                nc, _ = locate_many(g, (i0 + np.array([di, dj, dk])) * ext[:, None])
                ok = (w > 0.0) & (nc >= 0) & (addr >= 0)
                acc[ok] += w[ok] * f[nc[ok] * 3 + ax]
                wsum[ok] += w[ok]
    return np.where(wsum > 0.0, acc / np.where(wsum > 0, wsum, 1.0), 0.0)


def sample_velocity(g, f, x):
    return np.stack([sample_face(g, f, x, ax) for ax in range(3)], axis=1)


def alpha_flip(g, x, ptype, c: Consts):
    """per_particle_alpha_flip (particles.py:251-276)."""
    addr, lv = locate_at(g, x)
    dx = g.h0 * 2.0 ** np.where(addr >= 0, lv, 0)
    nu = np.where(np.asarray(ptype) == LIQUID, c.nu_l, c.nu_g)
    return np.clip(1.0 - 6.0 * nu * c.dt / (dx * dx), 0.0, c.alpha_cap)


def g2p(g, x, v, ptype, fvel, fdvel, c: Consts):
    """g2p, eq. (12) (mr_kernels.py:413-430)."""
    a = alpha_flip(g, x, ptype, c)[:, None]
    return a * (np.asarray(v, np.float64) + sample_velocity(g, fdvel, x)) + (1.0 - a) * sample_velocity(g, fvel, x)


def advect(g, x, v, fvel, c: Consts):
    """advect_rk3 + reflect_at_walls (mr_kernels.py:421-461)."""
    x = np.asarray(x, np.float64)
    k1 = sample_velocity(g, fvel, x)
    k2 = sample_velocity(g, fvel, x + 0.5 * c.dt * k1)
    k3 = sample_velocity(g, fvel, x + 0.75 * c.dt * k2)
    xn = np.clip(x + c.dt * ((2.0 / 9.0) * k1 + (3.0 / 9.0) * k2 + (4.0 / 9.0) * k3), c.lo, c.hi)
    vn = np.asarray(v, np.float64).copy()
    vn = np.where(xn <= c.lo, np.maximum(vn, 0.0), np.where(xn >= c.hi, np.minimum(vn, 0.0), vn))
    return xn, vn


# ---------------------------------------------------------------- refinement map
def refine_map(g, x, plev, cfl_eps, max_level):
    """atomic_min_refinement_map (particles.py:49-85): min(l_p) over the
    blocks each particle's box of half width cfl_eps touches."""
    bext = g.n0 * g.h0
    x = np.asarray(x, np.float64)
    plev = np.asarray(plev, np.int64)
    nb = np.array(g.nb)
    lo = np.maximum(np.floor((x - cfl_eps) / bext).astype(np.int64), 0)
    hi = np.minimum(np.floor((x + cfl_eps) / bext).astype(np.int64), nb - 1)
    M_ = np.full(g.nblk, max_level, np.int64)
    # every box spans at most `span` blocks per axis; visit offsets within it
    span = int(max((hi - lo).max(initial=0), 0)) + 1
    for di in range(span):
        for dj in range(span):
            for dk in range(span):
                q = lo + np.array([di, dj, dk])
                ok = (q <= hi).all(axis=1)
                idx = (q[ok, 0] * nb[1] + q[ok, 1]) * nb[2] + q[ok, 2]
                np.minimum.at(M_, idx, plev[ok])
    return M_


# ---------------------------------------------------------------- resample
def block_mixed(g, is_liquid):
    """interface_flag_per_block (mr_kernels.py:464-477): 1 where a block holds both phases."""
    liq = np.zeros(g.nblk, bool)
    air = np.zeros(g.nblk, bool)
    blk = g.blk[g.cslot]
    np.logical_or.at(liq, blk, is_liquid == 1)
    np.logical_or.at(air, blk, is_liquid == 0)
    return (liq & air).astype(np.int32)


def block_distance(mixed, nb, max_dist=64.0):
    """block_distance_field (particles.py:387-419), the scipy path."""
    from scipy.ndimage import distance_transform_edt

    mask = np.asarray(mixed, bool).reshape(nb)
    if not mask.any():
        return np.full(nb, max_dist, np.float32)
    d = distance_transform_edt(~mask).astype(np.float32)
    return np.minimum(d, max_dist).astype(np.float32)


def classify(x, plev, pdefer, dist, bext, beta_bw, max_level, threshold, seed):
    """classify_particles (particles.py:88-168) -> lnew, nout, defer (new)."""
    x = np.asarray(x, np.float64)
    nb = dist.shape
    gpos = x / bext - 0.5
    i0 = np.floor(gpos).astype(np.int64)
    t = gpos - i0
    d = np.zeros(len(x))
    for a in range(2):
        wa = t[:, 0] if a else 1.0 - t[:, 0]
        ii = np.clip(i0[:, 0] + a, 0, nb[0] - 1)
        for b in range(2):
            wb = t[:, 1] if b else 1.0 - t[:, 1]
            jj = np.clip(i0[:, 1] + b, 0, nb[1] - 1)
            for c in range(2):
                wc = t[:, 2] if c else 1.0 - t[:, 2]
                kk = np.clip(i0[:, 2] + c, 0, nb[2] - 1)
                d += wa * wb * wc * dist[ii, jj, kk]
    lp = np.asarray(plev, np.int64)
    ln = np.clip(np.floor(np.abs(beta_bw * d)).astype(np.int64), 0, max_level)
    ln = np.clip(ln, lp - 1, lp + 1)
    defer = np.asarray(pdefer, np.int64).copy()
    nout = np.ones(len(x), np.int64)
    split = ln < lp
    nout[split] = 8
    defer[split] = 0
    coarse = ln > lp
    defer[coarse] += 1
    fire = coarse & (defer >= threshold)
    _, u = rng_uniform(rng_init(seed, np.arange(len(x))))
    prob_del = 1.0 - 2.0 ** (-3.0 * (ln - lp))
    nout[fire & (u < prob_del)] = 0
    defer[~split & ~coarse] = 0
    return ln, nout, defer


def emit(x, v, ptype, plev, pdefer, lnew, nout, threshold, r0, seed):
    """emit_particles (particles.py:171-231) -> x, v, type, level, defer."""
    x = np.asarray(x, np.float64)
    outs = {k: [] for k in ("x", "v", "type", "level", "defer")}
    for p in range(len(x)):
        if nout[p] == 0:
            continue
        if nout[p] == 1:
            promote = lnew[p] > plev[p] and pdefer[p] >= threshold
            outs["x"].append(x[p])
            outs["v"].append(v[p])
            outs["type"].append(ptype[p])
            outs["level"].append(lnew[p] if promote else plev[p])
            outs["defer"].append(0 if promote else pdefer[p])
            continue
        half = 0.5 * r0 * 2.0 ** plev[p]
        st = rng_init(seed, p)
        for ch in range(8):
            s = np.array([(ch & 1) - 0.5, ((ch >> 1) & 1) - 0.5, ((ch >> 2) & 1) - 0.5])
            j = np.zeros(3)
            for a in range(3):
                st, u = rng_uniform(st)
                j[a] = (float(u) - 0.5) * half
            outs["x"].append(x[p] + s * half + j)
            outs["v"].append(v[p])
            outs["type"].append(ptype[p])
            outs["level"].append(plev[p] - 1)
            outs["defer"].append(0)
    return (np.array(outs["x"]).reshape(-1, 3), np.array(outs["v"]).reshape(-1, 3), np.array(outs["type"]),
            np.array(outs["level"]), np.array(outs["defer"]))


# ---------------------------------------------------------------- one step
class State:
    """The adaptive simulator's state (mr_sim.py:40-71): particles as FP64
    arrays, the grid, time and step index."""

    def __init__(self, cfg, nblocks, n0, max_level, beta_bw, x, v, ptype, level=None):
        self.cfg, self.nb, self.n0, self.max_level, self.beta_bw = cfg, tuple(nblocks), n0, max_level, beta_bw
        self.x = np.asarray(x, np.float64).copy()
        self.v = np.asarray(v, np.float64).copy()
        self.type = np.asarray(ptype).astype(np.int64)
        self.level = np.zeros(len(self.x), np.int64) if level is None else np.asarray(level, np.int64).copy()
        self.defer = np.zeros(len(self.x), np.int64)
        self.rho_tilde_0 = calibrate(cfg.ppc, cfg.dx, cfg.r_p)
        self.bext = n0 * cfg.dx
        self.domain = tuple(n * self.bext for n in self.nb)
        self.time, self.step_index = 0.0, 0


def dt_from_cfl(st):
    """mr_sim.py:193-201."""
    cfg = st.cfg
    vmax = float(np.linalg.norm(st.v, axis=1).max()) if len(st.v) else 0.0
    g = float(np.linalg.norm(cfg.gravity))
    vmax = max(vmax, math.sqrt(2.0 * g * cfg.dx) if g > 0 else 0.0, 1e-8)
    return float(np.clip(cfg.cfl * cfg.dx / vmax, cfg.dt_min, cfg.dt_max))


def step(st, dt=None, defer_threshold=3):
    """mr_sim.py:203-310 on the restated stages -> the step record."""
    cfg = st.cfg
    dt = dt_from_cfl(st) if dt is None else dt
    g = M.Grid(st.nb, st.n0, cfg.dx, st.max_level)
    g.build(refine_map(g, st.x, st.level, max(cfg.cfl, 1.0) * cfg.dx, st.max_level))
    c = Consts(cfg, st.rho_tilde_0, dt, st.domain, st.max_level)
    cm, cw, fm, fp = p2g(g, st.x, st.v, st.type, st.level, c.r0, c.rho_l, c.rho_g)
    phi, tag, u = finish(g, cm, fm, fp, c)
    beta, u, star = face_prepare(g, tag, fm, u, c)
    b = rhs(g, u, cw, c)
    p, it, rel = M.solve(g, beta, b, tol=cfg.pcg_tol, max_iter=cfg.pcg_max_iter, preconditioner="mg")
    u, du = project(g, p, beta, u, star, dt)
    div = max_divergence(g, u)
    st.v = g2p(g, st.x, st.v, st.type, u, du, c)
    n_before = len(st.x)
    dist = block_distance(block_mixed(g, tag), g.nb)
    ln, nout, defer = classify(st.x, st.level, st.defer, dist, st.bext, st.beta_bw, st.max_level, defer_threshold,
                               st.step_index)
    st.x, st.v, st.type, st.level, st.defer = emit(st.x, st.v, st.type, st.level, defer, ln, nout, defer_threshold,
                                                   cfg.r_p, st.step_index + 7919)
    st.x, st.v = advect(g, st.x, st.v, u, c)
    st.time += dt
    st.step_index += 1
    return {"step": st.step_index, "dt": dt, "cells": g.ncells, "particles": len(st.x),
            "dn": len(st.x) - n_before, "cg_iter": it, "cg_res": rel, "max_div": div, "grid": g}
