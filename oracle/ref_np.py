"""FP64 numpy restatement of the uniform-grid PF-FLIP step.

TEST INFRASTRUCTURE ONLY.  This module is the checker that the HIP path is
compared against; it is imported by ``tests/``, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py`` and nowhere else.  The product
(``implementation-of-adaptive-phase-field-flip_amd``) never imports it and has
no CPU fallback.

Every function cites the reference file:line it restates (paths relative to
the reference repository root).  Arithmetic is FP64 throughout; the reference
runs FP32 with float atomics, so comparisons against it are tolerance based
except for the integer parts (lattice indices, candidate index ranges,
binning keys, aggregation maps, red-black colouring, level dimensions), which
are exact.

Pinning (what checks each function independently of this file) lives in
``tests/test_oracle.py``:

* ``lattice``               known-answer PCG64 stream values + lattice bounds
* ``rho_tilde_0``           exact rationals 643/128 (8 ppc), 33491/1944 (27 ppc)
                            and the closed form n_p*64*pi*r^3/315 (5 %)
* ``p2g_*``                 brute force over *all* faces (no candidate ranges)
* ``phase_field``/``face_beta``  perfect-lattice closed form, exact 0 in air
* ``apply_A``/``compute_diag``   scipy-assembled matrix (symmetric, zero rows)
* ``galerkin_coarsen_faces``     P^T A P with piecewise-constant P
* ``rbgs_sweep``            scipy rows: red update == weighted Jacobi on red
* ``MGPCGOracle.solve``     scipy ``spsolve`` modulo the constant null space
* ``sample_face``           exact reproduction of linear fields
* ``advect_rk3``            Ralston RK3 == 3rd-order Taylor for linear fields
* ``apply_pressure_gradient`` + ``build_rhs``  div(u) == 0 after an exact solve
* ``extrapolate_sweep``     hand-derived single-cell cases
* whole step                mass / liquid-count conservation, phi == 0 in air,
                            hydrostatic stability (reference tests)
* droplet stage             ``sample_cell`` linear fields + clamped edges;
                            ``classify_escaped`` eq. (7) closed form of a lone
                            particle (converts at 27 ppc, not at 8) and of an
                            air particle in liquid; ``advance_droplets`` discrete
                            ballistic closed form, terminal velocity, r^2 law of
                            evaporation, death on exit; ``velocity_gradient_
                            magnitude`` linear fields; ``rasterize_spray``
                            volume conservation; the random stream by its
                            uniform / normal / log-normal statistics.  Whether
                            the stream reproduces Warp's numbers is parity
                            unpinned (warp-lang is absent).
"""

from __future__ import annotations

import math

import numpy as np

AIR, LIQUID = 0, 1
STATE_FLIP, STATE_DROPLET, STATE_DEAD = 0, 1, 2


# ---------------------------------------------------------------------------
# config (pfflip/config.py:14-135)
# ---------------------------------------------------------------------------
class OracleConfig:
    """Plain restatement of ``Config`` defaults (pfflip/config.py:17-114)."""

    def __init__(self, **kw):
        self.res = (64, 64, 64)
        self.dx = 1.0
        self.ppc = 8
        self.rho_l = 1000.0
        self.density_ratio = 1000.0
        self.gravity = (0.0, -9.81, 0.0)
        self.alpha_phi = 1.0
        self.alpha_flip_liquid = 0.985
        self.alpha_flip_air = 0.96
        self.cfl = 1.0
        self.dt_max = 1.0 / 24.0
        self.dt_min = 1e-6
        self.k_div = 0.0
        self.pcg_tol = 1e-4
        self.pcg_max_iter = 200
        self.mg_pre_sweeps = 3
        self.mg_post_sweeps = 3
        self.mg_coarsest_sweeps = 60
        self.mg_omega = 6.0 / 7.0
        self.mg_min_dim = 4
        self.extrap_cells = 3
        self.seed = 42
        # escaped particles / spray (pfflip/config.py:82-110)
        self.droplets = True
        self.delta_phi = 0.2
        self.droplet_radius_mean = 0.25
        self.droplet_radius_spread = 0.5
        self.droplet_drag_cd = 0.5
        self.evaporation_k = 0.0
        self.spray_shear_threshold = 0.0
        self.spray_emission_rate = 0.02
        self.spray_resolution_factor = 2
        self.droplet_max_substeps = 8
        for k, v in kw.items():
            setattr(self, k, v)

    # pfflip/config.py:117-135
    @property
    def rho_g(self):
        return self.rho_l / self.density_ratio

    @property
    def r_p(self):
        return self.dx

    @property
    def eta_phi(self):
        return math.log(self.density_ratio)

    @property
    def domain_size(self):
        return tuple(n * self.dx for n in self.res)


def face_shapes(res):
    """pfflip/sim.py:47."""
    nx, ny, nz = res
    return ((nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1))


def face_offset(axis):
    """pfflip/kernels.py:25-32: face centre = (index + o_a) * dx."""
    o = np.full(3, 0.5)
    o[axis] = 0.0
    return o


# ---------------------------------------------------------------------------
# scenes (pfflip/scenes.py:17-70)
# ---------------------------------------------------------------------------
def lattice(res, dx, ppc, rng, jitter=0.8):
    """pfflip/scenes.py:17-31.  ``jitter`` = 0.8 is the reference's value."""
    per_axis = round(ppc ** (1.0 / 3.0))
    assert per_axis**3 == ppc, "ppc must be a perfect cube (8, 27, ...)"
    nx, ny, nz = res
    step = dx / per_axis
    idx = np.stack(
        np.meshgrid(np.arange(nx * per_axis), np.arange(ny * per_axis),
                    np.arange(nz * per_axis), indexing="ij"),
        axis=-1,
    ).reshape(-1, 3)
    pos = (idx + 0.5) * step
    pos += (rng.random(pos.shape) - 0.5) * step * jitter
    return pos


def seed_particles(cfg, liquid_mask_fn, velocity_fn=None, jitter=0.8):
    """pfflip/scenes.py:34-45."""
    rng = np.random.default_rng(cfg.seed)
    pos = lattice(cfg.res, cfg.dx, cfg.ppc, rng, jitter)
    liq = np.asarray(liquid_mask_fn(pos), dtype=bool)
    ptype = np.where(liq, LIQUID, AIR).astype(np.int32)
    vel = np.zeros_like(pos) if velocity_fn is None else np.asarray(velocity_fn(pos, liq), float)
    return pos.astype(np.float32), vel.astype(np.float32), ptype


def dam_break(cfg, frac=(0.4, 0.6, 1.0)):
    """pfflip/scenes.py:49-60."""
    sx, sy, sz = cfg.domain_size
    return seed_particles(cfg, lambda p: (p[:, 0] < frac[0] * sx) & (p[:, 1] < frac[1] * sy)
                          & (p[:, 2] < frac[2] * sz))


def still_pool(cfg, fill=0.5):
    """pfflip/scenes.py:63-70."""
    sy = cfg.domain_size[1]
    return seed_particles(cfg, lambda p: p[:, 1] < fill * sy)


# ---------------------------------------------------------------------------
# rho~_0 calibration (pfflip/sim.py:119-140, :104-106)
# ---------------------------------------------------------------------------
def rho_tilde_0(dx, r, ppc):
    """pfflip/sim.py:119-140, verbatim loop order."""
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


def derived_densities(cfg):
    """pfflip/sim.py:104-106: (rho_tilde_0, rho_0l, rho_min)."""
    rt0 = rho_tilde_0(cfg.dx, cfg.r_p, cfg.ppc)
    return rt0, rt0 * cfg.rho_l, cfg.eta_phi * cfg.rho_g * rt0


# ---------------------------------------------------------------------------
# P2G (pfflip/kernels.py:40-49, 100-228)
# ---------------------------------------------------------------------------
def splat_weight(d2, r2):
    """pfflip/kernels.py:40-49, eq. (6)."""
    t = np.maximum(1.0 - d2 / r2, 0.0)
    return t * t * t


def _splat(px, pv_comp, pmass, prad, pstate, shape, o, dx, with_wsum=False):
    """Shared body of p2g_face (kernels.py:100-158) and p2g_cell (:161-209).

    Candidate index range per axis: ``floor(c - R) + 1 .. floor(c + R)`` with
    ``c = x/dx - o`` and ``R = r_p/dx``, clipped to the array (kernels.py:137-146).
    """
    act = pstate == 0
    x = px[act].astype(np.float64)
    m = pmass[act].astype(np.float64)
    rp = prad[act].astype(np.float64)
    u = None if pv_comp is None else pv_comp[act].astype(np.float64)
    r2 = rp * rp
    R = rp / dx
    c = x / dx - o
    shp = np.array(shape)
    lo = np.maximum(np.floor(c - R[:, None]).astype(np.int64) + 1, 0)
    hi = np.minimum(np.floor(c + R[:, None]).astype(np.int64), shp - 1)
    span = int(np.ceil(2.0 * R.max())) + 1 if len(R) else 0
    n = int(np.prod(shape))
    acc_m = np.zeros(n)
    acc_u = np.zeros(n)
    acc_w = np.zeros(n)
    for di in range(span):
        i = lo[:, 0] + di
        for dj in range(span):
            j = lo[:, 1] + dj
            for dk in range(span):
                k = lo[:, 2] + dk
                ok = (i <= hi[:, 0]) & (j <= hi[:, 1]) & (k <= hi[:, 2])
                fx = (i + o[0]) * dx - x[:, 0]
                fy = (j + o[1]) * dx - x[:, 1]
                fz = (k + o[2]) * dx - x[:, 2]
                w = splat_weight(fx * fx + fy * fy + fz * fz, r2)
                ok &= w > 0.0
                flat = (i[ok] * shape[1] + j[ok]) * shape[2] + k[ok]
                acc_m += np.bincount(flat, weights=(w * m)[ok], minlength=n)
                if u is not None:
                    acc_u += np.bincount(flat, weights=(w * m * u)[ok], minlength=n)
                if with_wsum:
                    acc_w += np.bincount(flat, weights=w[ok], minlength=n)
    return acc_m.reshape(shape), acc_u.reshape(shape), acc_w.reshape(shape)


def p2g_face(px, pv, pmass, prad, pstate, shape, axis, dx):
    """pfflip/kernels.py:100-158 -> (mass, mom) on the faces of ``axis``."""
    m, mu, _ = _splat(px, pv[:, axis], pmass, prad, pstate, shape, face_offset(axis), dx)
    return m, mu


def p2g_cell(px, pmass, prad, pstate, res, dx):
    """pfflip/kernels.py:161-209 -> (cell_mass, cell_wsum)."""
    m, _, w = _splat(px, None, pmass, prad, pstate, res, np.full(3, 0.5), dx, with_wsum=True)
    return m, w


def face_velocity_from_momentum(mass, mom, mass_eps):
    """pfflip/kernels.py:212-228 -> (vel, valid)."""
    ok = mass > mass_eps
    vel = np.where(ok, mom / np.where(ok, mass, 1.0), 0.0)
    return vel, ok.astype(np.int32)


# ---------------------------------------------------------------------------
# binning / sort keys (build-defined; no reference counterpart)
# ---------------------------------------------------------------------------
def bin_keys(px, pstate, res, dx, tile):
    """Tile-major cell key of every FLIP particle, ``0xFFFFFFFF`` otherwise.

    cell = clamp(floor(x / dx), 0, n-1) per axis, with ``x / dx`` evaluated
    in FP32 exactly as the device does (the reference evaluates ``x/dx`` in
    FP32 too, pfflip/kernels.py:138-146).  key = tile_id * T^3 + local_id
    with tile_id = (tx*nty + ty)*ntz + tz and local_id = (li*T + lj)*T + lk.
    """
    x = np.asarray(px, np.float32)
    q = np.floor(x / np.float32(dx)).astype(np.int64)
    n = np.array(res, np.int64)
    cell = np.clip(q, 0, n - 1)
    T = tile
    nt = (n + T - 1) // T
    t = cell // T
    loc = cell % T
    tid = (t[:, 0] * nt[1] + t[:, 1]) * nt[2] + t[:, 2]
    lid = (loc[:, 0] * T + loc[:, 1]) * T + loc[:, 2]
    key = tid * T**3 + lid
    key = np.where(np.asarray(pstate) == 0, key, 0xFFFFFFFF)
    return key.astype(np.uint32)


def sort_by_key(keys, nkeys):
    """Stable counting sort: (perm, start) with ties broken by particle index."""
    perm = np.argsort(keys, kind="stable").astype(np.int32)
    valid = keys[perm] != 0xFFFFFFFF
    perm = perm[valid]
    counts = np.bincount(keys[keys != 0xFFFFFFFF].astype(np.int64), minlength=nkeys)
    start = np.zeros(nkeys + 1, np.int64)
    np.cumsum(counts, out=start[1:])
    return perm, start.astype(np.int32)


# ---------------------------------------------------------------------------
# phase field and coefficients (pfflip/kernels.py:234-326)
# ---------------------------------------------------------------------------
def phase_field(rho_raw, rho_min, alpha_phi, rho_0l):
    """pfflip/kernels.py:234-239, eq. (7)."""
    rho_raw = np.asarray(rho_raw, np.float64)
    f = np.minimum(np.sqrt(np.maximum(rho_raw - rho_min, 0.0) / (alpha_phi * rho_0l)), 1.0)
    return np.where(rho_raw < rho_min, 0.0, f)


def cell_phase(cell_mass, rho_min, alpha_phi, rho_0l):
    """pfflip/kernels.py:242-255 -> (phi_cell, is_liquid)."""
    f = phase_field(cell_mass, rho_min, alpha_phi, rho_0l)
    return f, (f >= 0.5).astype(np.int32)


def _adjacent_tags(is_liquid, axis):
    """Tags (a, b) of the low/high cell of every face of ``axis``; a missing
    neighbour copies the other one (kernels.py:283-300)."""
    pad = [(0, 0)] * 3
    pad[axis] = (1, 0)
    lo = np.pad(is_liquid, pad, mode="edge")        # lo[f] = tag[f-1], f=0 -> tag[0]
    pad[axis] = (0, 1)
    hi = np.pad(is_liquid, pad, mode="edge")        # hi[f] = tag[f], f=n -> tag[n-1]
    return lo, hi


def face_beta(mass, is_liquid, axis, rho_min, alpha_phi, rho_0l, rho_l, rho_g):
    """pfflip/kernels.py:258-311 -> (phi_face, beta)."""
    a, b = _adjacent_tags(is_liquid, axis)
    f = phase_field(mass, rho_min, alpha_phi, rho_0l)
    f = np.where((a == 1) & (b == 1), 1.0, np.where((a == 0) & (b == 0), 0.0, f))
    rho = (1.0 - f) * rho_g + f * rho_l
    return f, 1.0 / rho


def wall_mask(shape, axis):
    """Faces with index 0 or n along ``axis`` (kernels.py:314-326)."""
    m = np.zeros(shape, bool)
    sl = [slice(None)] * 3
    sl[axis] = 0
    m[tuple(sl)] = True
    sl[axis] = -1
    m[tuple(sl)] = True
    return m


def solid_face_mask(solid, axis):
    """Faces touching a solid cell (kernels.py:601-627)."""
    a, b = _adjacent_tags_zero(solid, axis)
    return (a == 1) | (b == 1)


def _adjacent_tags_zero(cells, axis):
    """Like ``_adjacent_tags`` but a missing neighbour counts as 0."""
    pad = [(0, 0)] * 3
    pad[axis] = (1, 0)
    lo = np.pad(cells, pad, mode="constant")
    pad[axis] = (0, 1)
    hi = np.pad(cells, pad, mode="constant")
    return lo, hi


def zero_solid_faces(beta, vel, axis):
    """pfflip/kernels.py:314-326 (in place)."""
    m = wall_mask(beta.shape, axis)
    beta[m] = 0.0
    vel[m] = 0.0


def zero_solid_obstacle_faces(solid, beta, vel, axis):
    """pfflip/kernels.py:601-627 (in place)."""
    m = solid_face_mask(solid, axis)
    beta[m] = 0.0
    vel[m] = 0.0


# ---------------------------------------------------------------------------
# forces, divergence, projection (pfflip/kernels.py:332-400)
# ---------------------------------------------------------------------------
def divergence(vel):
    """sum_a (u_{a,i+1} - u_{a,i}) per cell (kernels.py:358, :399)."""
    u, v, w = vel
    return (u[1:] - u[:-1]) + (v[:, 1:] - v[:, :-1]) + (w[:, :, 1:] - w[:, :, :-1])


def build_rhs(vel, cell_wsum, dx, dt, rho_0, k_div):
    """pfflip/kernels.py:338-363."""
    b = -(dx / dt) * divergence(vel)
    if k_div > 0.0:
        excess = (cell_wsum - rho_0) / rho_0
        b = b - dx * dx * k_div * np.maximum(excess, 0.0) / dt
    return b


def apply_pressure_gradient(p, beta, vel, axis, dx, dt):
    """pfflip/kernels.py:366-387 (in place on ``vel``)."""
    sl_f = [slice(None)] * 3
    sl_f[axis] = slice(1, p.shape[axis])
    sl_hi = [slice(None)] * 3
    sl_hi[axis] = slice(1, None)
    sl_lo = [slice(None)] * 3
    sl_lo[axis] = slice(0, -1)
    bf = beta[tuple(sl_f)]
    grad = p[tuple(sl_hi)] - p[tuple(sl_lo)]
    upd = np.where(bf != 0.0, dt * bf * grad / dx, 0.0)
    vel[tuple(sl_f)] -= upd


def max_divergence(vel, dx):
    """pfflip/kernels.py:390-400."""
    return float(np.abs(divergence(vel) / dx).max())


# ---------------------------------------------------------------------------
# extrapolation (pfflip/kernels.py:406-462, sim.py:329-341)
# ---------------------------------------------------------------------------
def seed_liquid_velocity(vel, is_liquid, axis):
    """pfflip/kernels.py:406-425 -> (vel_l, known)."""
    a, b = _adjacent_tags_zero(is_liquid, axis)
    known = np.maximum(a, b).astype(np.int32)
    return np.where(known == 1, vel, 0.0), known


def extrapolate_sweep(vel_l, known):
    """pfflip/kernels.py:428-462 (one Jacobi sweep)."""
    acc = np.zeros_like(vel_l)
    cnt = np.zeros(vel_l.shape, np.int64)
    kv = np.where(known == 1, vel_l, 0.0)
    for ax in range(3):
        for s in (-1, 1):
            src = [slice(None)] * 3
            dst = [slice(None)] * 3
            if s == -1:   # neighbour at index - 1
                src[ax] = slice(0, -1)
                dst[ax] = slice(1, None)
            else:
                src[ax] = slice(1, None)
                dst[ax] = slice(0, -1)
            acc[tuple(dst)] += kv[tuple(src)]
            cnt[tuple(dst)] += known[tuple(src)]
    out = np.where(known == 1, vel_l,
                   np.where(cnt > 0, acc / np.maximum(cnt, 1), vel_l))
    kout = np.where((known == 1) | (cnt > 0), 1, 0).astype(np.int32)
    return out, kout


def extrapolate_liquid_velocity(vel, is_liquid, sweeps):
    """pfflip/sim.py:329-341."""
    out = []
    for ax in range(3):
        vl, kn = seed_liquid_velocity(vel[ax], is_liquid, ax)
        for _ in range(sweeps):
            vl, kn = extrapolate_sweep(vl, kn)
        out.append(vl)
    return out


# ---------------------------------------------------------------------------
# G2P / advection (pfflip/kernels.py:52-94, 468-595)
# ---------------------------------------------------------------------------
def sample_face(f, x, dx, axis):
    """pfflip/kernels.py:52-79: trilinear, indices clamped per axis."""
    o = face_offset(axis)
    g = x / dx - o
    i0 = np.floor(g).astype(np.int64)
    t = g - i0
    acc = np.zeros(len(x))
    n = f.shape
    for di in range(2):
        wi = 1.0 - t[:, 0] if di == 0 else t[:, 0]
        ii = np.clip(i0[:, 0] + di, 0, n[0] - 1)
        for dj in range(2):
            wj = 1.0 - t[:, 1] if dj == 0 else t[:, 1]
            jj = np.clip(i0[:, 1] + dj, 0, n[1] - 1)
            for dk in range(2):
                wk = 1.0 - t[:, 2] if dk == 0 else t[:, 2]
                kk = np.clip(i0[:, 2] + dk, 0, n[2] - 1)
                acc += wi * wj * wk * f[ii, jj, kk]
    return acc


def sample_velocity(vel, x, dx):
    """pfflip/kernels.py:82-94."""
    return np.stack([sample_face(vel[a], x, dx, a) for a in range(3)], axis=1)


def g2p(px, pv, ptype, pstate, vel, dvel, dx, alpha_l, alpha_a):
    """pfflip/kernels.py:468-496, eq. (12).  Returns the new ``pv``."""
    x = px.astype(np.float64)
    act = pstate == 0
    alpha = np.where(ptype == LIQUID, alpha_l, alpha_a)[:, None]
    delta = sample_velocity(dvel, x, dx)
    pic = sample_velocity(vel, x, dx)
    new = alpha * (pv + delta) + (1.0 - alpha) * pic
    return np.where(act[:, None], new, pv)


def advect_rk3(px, ptype, pstate, vel, vel_l, dt, dx, lo, hi):
    """pfflip/kernels.py:508-549 (Ralston RK3).  Returns the new ``px``."""
    x = px.astype(np.float64)
    liq = (ptype == LIQUID)[:, None]

    def vs(xx):
        return np.where(liq, sample_velocity(vel_l, xx, dx), sample_velocity(vel, xx, dx))

    k1 = vs(x)
    k2 = vs(x + 0.5 * dt * k1)
    k3 = vs(x + 0.75 * dt * k2)
    xn = x + dt * ((2.0 / 9.0) * k1 + (3.0 / 9.0) * k2 + (4.0 / 9.0) * k3)
    xn = np.clip(xn, lo, hi)
    return np.where((pstate == 0)[:, None], xn, x)


def reflect_at_walls(px, pv, pstate, lo, hi):
    """pfflip/kernels.py:567-595 -> (px, pv)."""
    x = px.astype(np.float64).copy()
    v = pv.astype(np.float64).copy()
    lo = np.broadcast_to(np.asarray(lo, np.float64), x.shape)
    hi = np.broadcast_to(np.asarray(hi, np.float64), x.shape)
    below = x <= lo
    above = (~below) & (x >= hi)
    x = np.where(below, lo, np.where(above, hi, x))
    v = np.where(below, np.maximum(v, 0.0), np.where(above, np.minimum(v, 0.0), v))
    keep = (pstate == 2)[:, None]
    return np.where(keep, px, x), np.where(keep, pv, v)


def push_out_of_solids(px, pv, pstate, solid, dx):
    """pfflip/kernels.py:630-679 -> (px, pv)."""
    x = px.astype(np.float64).copy()
    v = pv.astype(np.float64).copy()
    nx, ny, nz = solid.shape
    for p in range(len(x)):
        if pstate[p] == 2:
            continue
        i = min(max(int(x[p, 0] / dx), 0), nx - 1)
        j = min(max(int(x[p, 1] / dx), 0), ny - 1)
        k = min(max(int(x[p, 2] / dx), 0), nz - 1)
        if solid[i, j, k] == 0:
            continue
        best, bax, bsg = 1.0e9, 0, 1
        for d in range(6):
            ax, sg = d // 2, (-1 if d % 2 == 0 else 1)
            for t in range(1, 8):
                c = [i, j, k]
                c[ax] += sg * t
                if min(c) < 0 or c[0] >= nx or c[1] >= ny or c[2] >= nz:
                    break
                if solid[c[0], c[1], c[2]] == 0:
                    if float(t) < best:
                        best, bax, bsg = float(t), ax, sg
                    break
        if best > 1.0e8:
            continue
        x[p, bax] += (best + 0.5) * dx * float(bsg)
        if float(bsg) * v[p, bax] < 0.0:
            v[p, bax] = 0.0
    return x, v


def particle_max_speed(pv, pstate):
    """pfflip/kernels.py:552-564."""
    act = pstate == 0
    if not act.any():
        return 0.0
    return float(np.linalg.norm(pv[act].astype(np.float64), axis=1).max())


def dt_from_cfl(cfg, vmax):
    """pfflip/sim.py:158-166."""
    g = float(np.linalg.norm(cfg.gravity))
    vmax = max(vmax, math.sqrt(2.0 * g * cfg.dx) if g > 0 else 0.0, 1e-8)
    return float(np.clip(cfg.cfl * cfg.dx / vmax, cfg.dt_min, cfg.dt_max))


# ---------------------------------------------------------------------------
# MG-PCG (pfflip/poisson.py)
# ---------------------------------------------------------------------------
def level_dims(res, min_dim):
    """pfflip/poisson.py:256-263."""
    out = []
    r = tuple(res)
    while True:
        out.append(r)
        if min(r) // 2 < min_dim or any(n % 2 for n in r):
            break
        r = tuple(n // 2 for n in r)
    return out


def compute_diag(bx, by, bz):
    """pfflip/poisson.py:40-51."""
    return bx[:-1] + bx[1:] + by[:, :-1] + by[:, 1:] + bz[:, :, :-1] + bz[:, :, 1:]


def _shift(p, axis, s):
    """p at index clamp(i + s) along ``axis`` (poisson.py:69-74)."""
    n = p.shape[axis]
    idx = np.clip(np.arange(n) + s, 0, n - 1)
    return np.take(p, idx, axis=axis)


def apply_A(p, bx, by, bz, diag):
    """pfflip/poisson.py:54-88."""
    s = bx[:-1] * _shift(p, 0, -1)
    s = s + bx[1:] * _shift(p, 0, 1)
    s = s + by[:, :-1] * _shift(p, 1, -1)
    s = s + by[:, 1:] * _shift(p, 1, 1)
    s = s + bz[:, :, :-1] * _shift(p, 2, -1)
    s = s + bz[:, :, 1:] * _shift(p, 2, 1)
    return diag * p - s


def colour_mask(res, colour):
    """pfflip/poisson.py:120: cells with ((i+j+k) & 1) == colour."""
    i, j, k = np.meshgrid(*[np.arange(n) for n in res], indexing="ij")
    return ((i + j + k) & 1) == colour


def rbgs_sweep(p, b, bx, by, bz, diag, colour, omega):
    """pfflip/poisson.py:105-126 (in place).  Same-colour cells never neighbour
    each other, so the vectorised update equals the sequential one."""
    m = colour_mask(p.shape, colour) & (diag > 0.0)
    res = b - apply_A(p, bx, by, bz, diag)
    p[m] += omega * res[m] / diag[m]


def galerkin_coarsen_faces(fine, axis, coarse_shape):
    """pfflip/poisson.py:130-163: coarse face = sum of the 4 fine faces."""
    out = np.zeros(coarse_shape)
    nf = fine.shape
    for a in range(2):
        for b in range(2):
            I, J, K = np.meshgrid(*[np.arange(n) for n in coarse_shape], indexing="ij")
            if axis == 0:
                ii, jj, kk = 2 * I, 2 * J + a, 2 * K + b
            elif axis == 1:
                ii, jj, kk = 2 * I + a, 2 * J, 2 * K + b
            else:
                ii, jj, kk = 2 * I + a, 2 * J + b, 2 * K
            ok = (ii < nf[0]) & (jj < nf[1]) & (kk < nf[2])
            out[ok] += fine[ii[ok], jj[ok], kk[ok]]
    return out


def restrict_sum(fine):
    """pfflip/poisson.py:166-175: R = P^T."""
    nx, ny, nz = fine.shape
    return fine.reshape(nx // 2, 2, ny // 2, 2, nz // 2, 2).sum(axis=(1, 3, 5))


def prolong(coarse):
    """pfflip/poisson.py:178-182: piecewise-constant injection."""
    return coarse.repeat(2, 0).repeat(2, 1).repeat(2, 2)


class _OLevel:
    def __init__(self, res):
        self.res = res
        self.bx = np.zeros(face_shapes(res)[0])
        self.by = np.zeros(face_shapes(res)[1])
        self.bz = np.zeros(face_shapes(res)[2])
        self.diag = np.zeros(res)
        self.p = np.zeros(res)
        self.b = np.zeros(res)
        self.r = np.zeros(res)


class MGPCGOracle:
    """pfflip/poisson.py:240-409 in FP64."""

    def __init__(self, res, *, omega=6.0 / 7.0, pre=2, post=2, coarsest=60, min_dim=4):
        self.omega, self.pre, self.post, self.coarsest = omega, pre, post, coarsest
        self.levels = [_OLevel(r) for r in level_dims(res, min_dim)]
        self.nlev = len(self.levels)
        self.ncells = int(np.prod(res))

    def build_hierarchy(self, bx, by, bz):
        """pfflip/poisson.py:295-309."""
        L0 = self.levels[0]
        L0.bx, L0.by, L0.bz = (np.asarray(a, np.float64).copy() for a in (bx, by, bz))
        for lv, L in enumerate(self.levels):
            L.diag = compute_diag(L.bx, L.by, L.bz)
            if lv + 1 < self.nlev:
                C = self.levels[lv + 1]
                fs = face_shapes(C.res)
                C.bx = galerkin_coarsen_faces(L.bx, 0, fs[0])
                C.by = galerkin_coarsen_faces(L.by, 1, fs[1])
                C.bz = galerkin_coarsen_faces(L.bz, 2, fs[2])

    def _smooth(self, L, n, reverse=False):
        """pfflip/poisson.py:312-321."""
        for _ in range(n):
            for c in ((1, 0) if reverse else (0, 1)):
                rbgs_sweep(L.p, L.b, L.bx, L.by, L.bz, L.diag, c, self.omega)

    def _vcycle(self, lv=0):
        """pfflip/poisson.py:323-343."""
        L = self.levels[lv]
        if lv == self.nlev - 1:
            self._smooth(L, self.coarsest)
            L.p -= L.p.mean()
            return
        self._smooth(L, self.pre)
        L.r = L.b - apply_A(L.p, L.bx, L.by, L.bz, L.diag)
        C = self.levels[lv + 1]
        C.b = restrict_sum(L.r)
        C.p = np.zeros(C.res)
        self._vcycle(lv + 1)
        L.p += prolong(C.p)
        self._smooth(L, self.post, reverse=True)

    def precondition(self, r):
        """pfflip/poisson.py:345-352."""
        L0 = self.levels[0]
        L0.b = np.array(r, np.float64)
        L0.p = np.zeros(L0.res)
        self._vcycle(0)
        z = L0.p.copy()
        return z - z.mean()

    def solve(self, b, *, tol=1e-4, max_iter=200):
        """pfflip/poisson.py:355-409 -> (x, it, rel)."""
        L0 = self.levels[0]
        rhs = np.array(b, np.float64)
        rhs -= rhs.mean()
        b_norm = np.abs(rhs).max()
        if b_norm == 0.0:
            return np.zeros(L0.res), 0, 0.0
        x = np.zeros(L0.res)
        r = rhs.copy()
        z = self.precondition(r)
        q = z.copy()
        rz = float((r * z).sum())
        it, rel = 0, 1.0
        for it in range(1, max_iter + 1):
            Ap = apply_A(q, L0.bx, L0.by, L0.bz, L0.diag)
            qAq = float((q * Ap).sum())
            if qAq <= 0.0:
                break
            alpha = rz / qAq
            x += alpha * q
            r_prev = r.copy()
            r -= alpha * Ap
            rel = np.abs(r).max() / b_norm
            if rel < tol:
                break
            z = self.precondition(r)
            rz_new = float((r * z).sum())
            num = rz_new - float((r_prev * z).sum())
            beta = max(num / rz, 0.0) if rz != 0.0 else 0.0
            q = z + beta * q
            rz = rz_new
        x -= x.mean()
        return x, it, float(rel)


# ---------------------------------------------------------------------------
# escaped particles, droplets, spray (pfflip/droplets.py, sim.py:343-401)
# ---------------------------------------------------------------------------
# Random stream.  The reference draws radii from Warp's rand_init/randn
# (warp-lang, version not pinned by the reference), an absent dependency.
# Restated here: the PCG hash of Jarzynski & Olano (JCGT 2020), a stream
# state = hash(seed + hash(offset)), uniforms from the top 24 bits of each
# new state, normals by Box-Muller.  The product uses the same stream, so
# radii agree to the ulps of exp/log/cos; whether these are Warp's numbers
# is parity unpinned (only their distribution is checked).
_M32 = np.uint64(0xFFFFFFFF)


def pcg_hash(v):
    """PCG RXS-M-XS hash of uint32 values (Jarzynski & Olano 2020)."""
    v = np.asarray(v, np.uint64) & _M32
    st = (v * np.uint64(747796405) + np.uint64(2891336453)) & _M32
    w = (((st >> ((st >> np.uint64(28)) + np.uint64(4))) ^ st) * np.uint64(277803737)) & _M32
    return ((w >> np.uint64(22)) ^ w) & _M32


def rng_init(seed, offset):
    """Stream state of particle ``offset`` (droplets.py:91, 149):
    hash(seed + hash(offset))."""
    seed_u = np.asarray(seed, np.int64).astype(np.uint64) & _M32
    return pcg_hash((seed_u + pcg_hash(np.asarray(offset, np.int64).astype(np.uint64) & _M32)) & _M32)


def rng_uniform(state):
    """-> (new state, float32 uniform in [0, 1) from the top 24 bits)."""
    state = pcg_hash(state)
    return state, ((state >> np.uint64(8)).astype(np.float64) / 16777216.0).astype(np.float32)


def rng_normal(state):
    """-> (new state, standard normal), Box-Muller; u1 is floored at 2^-25."""
    state, u1 = rng_uniform(state)
    state, u2 = rng_uniform(state)
    u1 = np.maximum(u1.astype(np.float64), 2.0 ** -25)
    return state, np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2.astype(np.float64))


def lognormal_params(mean_radius, spread=0.5):
    """droplets.py:355-360: (mu, sigma) of a log-normal with the given mean."""
    sigma = spread
    return math.log(mean_radius) - 0.5 * sigma * sigma, sigma


def sample_cell(f, x, dx):
    """droplets.py:42-61: trilinear at cell centres, indices clamped."""
    g = x / dx - 0.5
    i0 = np.floor(g).astype(np.int64)
    t = g - i0
    acc = np.zeros(len(x))
    n = f.shape
    for a in range(2):
        wa = 1.0 - t[:, 0] if a == 0 else t[:, 0]
        ii = np.clip(i0[:, 0] + a, 0, n[0] - 1)
        for b in range(2):
            wb = 1.0 - t[:, 1] if b == 0 else t[:, 1]
            jj = np.clip(i0[:, 1] + b, 0, n[1] - 1)
            for c in range(2):
                wc = 1.0 - t[:, 2] if c == 0 else t[:, 2]
                kk = np.clip(i0[:, 2] + c, 0, n[2] - 1)
                acc += wa * wb * wc * f[ii, jj, kk]
    return acc


def _draw_radius(state, mu, sigma, r_min, r_max):
    state, z = rng_normal(state)
    return state, np.clip(np.exp(mu + sigma * z), r_min, r_max)


def classify_escaped(px, ptype, pstate, pdrop_r, phi, dx, delta_phi, mu, sigma, r_min, r_max, seed):
    """droplets.py:67-106 -> (pstate, pdrop_r, [emitted, bubbles, merged], f)."""
    st = np.asarray(pstate, np.int32).copy()
    r = np.asarray(pdrop_r, np.float64).copy()
    f = sample_cell(phi, np.asarray(px, np.float64), dx)
    idx = np.arange(len(st))
    flip, drop = st == STATE_FLIP, st == STATE_DROPLET
    emit = flip & (ptype == LIQUID) & (f < 0.5 - delta_phi)
    bubble = flip & (ptype != LIQUID) & (f > 0.5 + delta_phi)
    merge = drop & (f > 0.5 + delta_phi)
    if emit.any():
        _, r[emit] = _draw_radius(rng_init(seed, idx[emit]), mu, sigma, r_min, r_max)
    st[emit] = STATE_DROPLET
    st[bubble] = STATE_DEAD
    st[merge] = STATE_FLIP
    return st, r, [int(emit.sum()), int(bubble.sum()), int(merge.sum())], f


def velocity_gradient_magnitude(vel, dx):
    """droplets.py:156-196: |grad u| at cells from the seven terms the
    reference sums (dv/dz and dw/dy are not among them)."""
    u, v, w = (np.asarray(a, np.float64) for a in vel)
    nx, ny, nz = w.shape[0], u.shape[1], u.shape[2]
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    im, ip = np.maximum(i - 1, 0), np.minimum(i + 1, nx - 1)
    jm, jp = np.maximum(j - 1, 0), np.minimum(j + 1, ny - 1)
    km, kp = np.maximum(k - 1, 0), np.minimum(k + 1, nz - 1)
    dudx = (u[i + 1, j, k] - u[i, j, k]) / dx
    dvdy = (v[i, j + 1, k] - v[i, j, k]) / dx
    dwdz = (w[i, j, k + 1] - w[i, j, k]) / dx
    dudy = (0.5 * (u[i, jp, k] + u[i + 1, jp, k]) - 0.5 * (u[i, jm, k] + u[i + 1, jm, k])) / (2 * dx)
    dudz = (0.5 * (u[i, j, kp] + u[i + 1, j, kp]) - 0.5 * (u[i, j, km] + u[i + 1, j, km])) / (2 * dx)
    dvdx = (0.5 * (v[ip, j, k] + v[ip, j + 1, k]) - 0.5 * (v[im, j, k] + v[im, j + 1, k])) / (2 * dx)
    dwdx = (0.5 * (w[ip, j, k] + w[ip, j, k + 1]) - 0.5 * (w[im, j, k] + w[im, j, k + 1])) / (2 * dx)
    return np.sqrt(dudx ** 2 + dvdy ** 2 + dwdz ** 2 + dudy ** 2 + dudz ** 2 + dvdx ** 2 + dwdx ** 2)


def emit_spray_in_shear(px, ptype, pstate, pdrop_r, phi, shear, dx, threshold, rate, mu, sigma,
                        r_min, r_max, seed):
    """droplets.py:112-153 -> (pstate, pdrop_r, n_spray)."""
    st = np.asarray(pstate, np.int32).copy()
    r = np.asarray(pdrop_r, np.float64).copy()
    x = np.asarray(px, np.float64)
    f = sample_cell(phi, x, dx)
    sh = sample_cell(shear, x, dx)
    cand = (st == STATE_FLIP) & (ptype == LIQUID) & ~((f > 0.9) | (f < 0.1)) & ~(sh < threshold)
    idx = np.flatnonzero(cand)
    state, u = rng_uniform(rng_init(seed + 104729, idx))
    go = ~(u.astype(np.float64) > rate)
    idx, state = idx[go], state[go]
    if len(idx):
        _, r[idx] = _draw_radius(state, mu, sigma, r_min, r_max)
    st[idx] = STATE_DROPLET
    return st, r, len(idx)


def drag_acceleration(vrel, r, rho_g, rho_l, cd):
    """droplets.py:200-208: 3 rho_g Cd |w| w / (8 rho_l r), 0 if |w| or r is 0."""
    s = np.linalg.norm(vrel, axis=-1)
    ok = (s > 0.0) & (r > 0.0)
    c = np.where(ok, 3.0 * rho_g * cd * s / (8.0 * rho_l * np.where(ok, r, 1.0)), 0.0)
    return c[:, None] * vrel


def advance_droplets(px, pv, pstate, pdrop_r, vel, dt, dx, gravity, rho_g, rho_l, cd, evap_k,
                     r_min, lo, hi, max_substeps):
    """droplets.py:211-278 -> (px, pv, pstate, pdrop_r).  Semi-implicit
    Euler sub-steps, count from the droplet's own CFL; evaporation
    dr/dt = -k/r by midpoint RK2; a droplet that leaves [lo, hi] dies (and,
    like one that evaporates, keeps its pre-advance x, v and r)."""
    x = np.asarray(px, np.float64).copy()
    v = np.asarray(pv, np.float64).copy()
    st = np.asarray(pstate, np.int32).copy()
    r = np.asarray(pdrop_r, np.float64).copy()
    d = np.flatnonzero(st == STATE_DROPLET)
    if len(d) == 0:
        return x, v, st, r
    xd, vd, rd = x[d].copy(), v[d].copy(), r[d].copy()
    speed = np.linalg.norm(vd, axis=1) + 1e-8
    nsub = np.clip(np.floor(np.minimum(speed * dt / (0.5 * dx), 1e9)).astype(np.int64) + 1, 1,
                   max_substeps)
    h = dt / nsub
    alive = np.ones(len(d), bool)
    g = np.asarray(gravity, np.float64)
    for s in range(int(nsub.max())):
        m = alive & (s < nsub)
        if not m.any():
            break
        ua = sample_velocity(vel, xd[m], dx)
        a = g[None, :] + drag_acceleration(ua - vd[m], rd[m], rho_g, rho_l, cd)
        vd[m] = vd[m] + h[m][:, None] * a
        xd[m] = xd[m] + h[m][:, None] * vd[m]
        if evap_k > 0.0:
            e = np.flatnonzero(m)
            e = e[rd[e] > 0.0]
            k1 = -evap_k / rd[e]
            rm = np.maximum(rd[e] + 0.5 * h[e] * k1, 1e-9)
            rd[e] = rd[e] + h[e] * (-evap_k / rm)
            alive[e[rd[e] <= r_min]] = False
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    out = alive & ((xd <= lo).any(axis=1) | (xd >= hi).any(axis=1))
    keep = alive & ~out
    st[d[~keep]] = STATE_DEAD
    x[d[keep]], v[d[keep]], r[d[keep]] = xd[keep], vd[keep], rd[keep]
    return x, v, st, r


def rasterize_spray(px, pstate, pdrop_r, shape, dx_spray):
    """droplets.py:308-351 in FP64 (the reference adds FP16 atomics)."""
    dens = np.zeros(shape)
    d = np.flatnonzero(pstate == STATE_DROPLET)
    x = np.asarray(px, np.float64)[d]
    r = np.asarray(pdrop_r, np.float64)[d]
    vol = (4.0 / 3.0) * math.pi * r ** 3
    g = x / dx_spray - 0.5
    i0 = np.floor(g).astype(np.int64)
    t = g - i0
    cell_vol = dx_spray ** 3
    for a in range(2):
        wa = 1.0 - t[:, 0] if a == 0 else t[:, 0]
        ii = i0[:, 0] + a
        for b in range(2):
            wb = 1.0 - t[:, 1] if b == 0 else t[:, 1]
            jj = i0[:, 1] + b
            for c in range(2):
                wc = 1.0 - t[:, 2] if c == 0 else t[:, 2]
                kk = i0[:, 2] + c
                ok = ((ii >= 0) & (ii < shape[0]) & (jj >= 0) & (jj < shape[1])
                      & (kk >= 0) & (kk < shape[2]))
                np.add.at(dens, (ii[ok], jj[ok], kk[ok]), (wa * wb * wc * vol / cell_vol)[ok])
    return dens


def droplet_stage(px, pv, ptype, pstate, pdrop_r, phi, vel, cfg, dt, seed):
    """pfflip/sim.py:343-390 in the reference's order: classify, [shear +
    spray], advance.  -> (px, pv, pstate, pdrop_r, counters[4])."""
    dx = cfg.dx
    mu, sigma = lognormal_params(cfg.droplet_radius_mean * dx, cfg.droplet_radius_spread)
    r_min, r_max = 0.02 * dx, 2.0 * dx
    st, r, c, _ = classify_escaped(px, ptype, pstate, pdrop_r, phi, dx, cfg.delta_phi, mu, sigma,
                                   r_min, r_max, seed)
    n_spray = 0
    if cfg.spray_shear_threshold > 0.0:
        shear = velocity_gradient_magnitude(vel, dx)
        st, r, n_spray = emit_spray_in_shear(px, ptype, st, r, phi, shear, dx,
                                             cfg.spray_shear_threshold, cfg.spray_emission_rate,
                                             mu, sigma, r_min, r_max, seed)
    eps = 1e-4 * dx
    lo = np.full(3, eps)
    hi = np.array([s - eps for s in cfg.domain_size])
    x, v, st, r = advance_droplets(px, pv, st, r, vel, dt, dx, cfg.gravity, cfg.rho_g, cfg.rho_l,
                                   cfg.droplet_drag_cd, cfg.evaporation_k * dx * dx, r_min, lo, hi,
                                   cfg.droplet_max_substeps)
    return x, v, st, r, c + [n_spray]


# ---------------------------------------------------------------------------
# the whole step (pfflip/sim.py:168-424)
# ---------------------------------------------------------------------------
class OracleSim:
    """Stage-by-stage restatement of ``Simulator`` (pfflip/sim.py:37-424).

    With ``cfg.droplets`` the droplet stage (sim.py:343-390) runs between
    extrapolation and advection, as in the reference.
    """

    def __init__(self, cfg):
        self.cfg = cfg
        self.res = tuple(cfg.res)
        self.face_shapes = face_shapes(self.res)
        self.rho_tilde_0, self.rho_0l, self.rho_min = derived_densities(cfg)
        self.solver = MGPCGOracle(self.res, omega=cfg.mg_omega, pre=cfg.mg_pre_sweeps,
                                  post=cfg.mg_post_sweeps, coarsest=cfg.mg_coarsest_sweeps,
                                  min_dim=cfg.mg_min_dim)
        self.solid = np.zeros(self.res, np.int32)
        self.has_solids = False
        self.time = 0.0
        self.step_index = 0
        self.stats = []

    def set_particles(self, pos, vel, ptype):
        """pfflip/sim.py:142-155."""
        cfg = self.cfg
        n = len(pos)
        self.n_particles = n
        self.px = np.asarray(pos, np.float32).astype(np.float64)
        self.pv = np.asarray(vel, np.float32).astype(np.float64)
        self.ptype = np.asarray(ptype, np.int32)
        self.pmass = np.where(self.ptype == 1, cfg.rho_l, cfg.rho_g).astype(np.float32)
        self.prad = np.full(n, cfg.r_p, np.float32)
        self.pstate = np.zeros(n, np.int32)
        self.pdrop_r = np.zeros(n)
        self.droplet_stats = {"emitted": 0, "bubbles": 0, "merged": 0, "spray": 0}

    def set_solids(self, mask):
        """pfflip/sim.py:113-116."""
        self.solid = np.ascontiguousarray(mask, np.int32)
        self.has_solids = bool(np.asarray(mask).any())

    def dt_from_cfl(self):
        return dt_from_cfl(self.cfg, particle_max_speed(self.pv, self.pstate))

    def _p2g(self):
        """pfflip/sim.py:205-236."""
        dx = self.cfg.dx
        self.mass, self.mom, self.vel = [], [], []
        for ax in range(3):
            m, mu = p2g_face(self.px, self.pv, self.pmass, self.prad, self.pstate,
                             self.face_shapes[ax], ax, dx)
            self.mass.append(m)
            self.mom.append(mu)
            self.vel.append(face_velocity_from_momentum(m, mu, 1e-12 * self.rho_0l)[0])
        self.cell_mass, self.cell_wsum = p2g_cell(self.px, self.pmass, self.prad,
                                                  self.pstate, self.res, dx)

    def _phase_field_and_coefficients(self):
        """pfflip/sim.py:238-268."""
        cfg = self.cfg
        self.phi_cell, self.is_liquid = cell_phase(self.cell_mass, self.rho_min,
                                                   cfg.alpha_phi, self.rho_0l)
        self.phi_face, self.beta = [], []
        for ax in range(3):
            f, b = face_beta(self.mass[ax], self.is_liquid, ax, self.rho_min, cfg.alpha_phi,
                             self.rho_0l, cfg.rho_l, cfg.rho_g)
            zero_solid_faces(b, self.vel[ax], ax)
            if self.has_solids:
                zero_solid_obstacle_faces(self.solid, b, self.vel[ax], ax)
            self.phi_face.append(f)
            self.beta.append(b)

    def _add_forces(self, dt):
        """pfflip/sim.py:270-285."""
        for ax in range(3):
            g = self.cfg.gravity[ax]
            if g != 0.0:
                self.vel[ax] += g * dt
            zero_solid_faces(self.beta[ax], self.vel[ax], ax)
            if self.has_solids:
                zero_solid_obstacle_faces(self.solid, self.beta[ax], self.vel[ax], ax)

    def _pressure_solve(self, dt):
        """pfflip/sim.py:287-299."""
        cfg = self.cfg
        self.rhs = build_rhs(self.vel, self.cell_wsum, cfg.dx, dt, self.rho_tilde_0, cfg.k_div)
        self.solver.build_hierarchy(*self.beta)
        self.pressure, it, rel = self.solver.solve(self.rhs, tol=cfg.pcg_tol,
                                                   max_iter=cfg.pcg_max_iter)
        return it, rel

    def _project(self, dt):
        """pfflip/sim.py:301-308."""
        for ax in range(3):
            apply_pressure_gradient(self.pressure, self.beta[ax], self.vel[ax], ax,
                                    self.cfg.dx, dt)

    def _max_divergence(self):
        return max_divergence(self.vel, self.cfg.dx)

    def _g2p(self):
        """pfflip/sim.py:316-327."""
        cfg = self.cfg
        self.dvel = [self.vel[a] - self.vel_star[a] for a in range(3)]
        self.pv = g2p(self.px, self.pv, self.ptype, self.pstate, self.vel, self.dvel,
                      cfg.dx, cfg.alpha_flip_liquid, cfg.alpha_flip_air)

    def _extrapolate_liquid_velocity(self):
        self.vel_liq = extrapolate_liquid_velocity(self.vel, self.is_liquid, self.cfg.extrap_cells)

    def _advect(self, dt):
        """pfflip/sim.py:408-424."""
        cfg = self.cfg
        eps = 1e-4 * cfg.dx
        lo = np.full(3, eps)
        hi = np.array([s - eps for s in cfg.domain_size])
        self.px = advect_rk3(self.px, self.ptype, self.pstate, self.vel, self.vel_liq,
                             dt, cfg.dx, lo, hi)
        self.px, self.pv = reflect_at_walls(self.px, self.pv, self.pstate, lo, hi)
        if self.has_solids:
            self.px, self.pv = push_out_of_solids(self.px, self.pv, self.pstate,
                                                  self.solid, cfg.dx)

    def _droplets(self, dt):
        """pfflip/sim.py:343-390."""
        if not self.cfg.droplets:
            return
        self.px, self.pv, self.pstate, self.pdrop_r, c = droplet_stage(
            self.px, self.pv, self.ptype, self.pstate, self.pdrop_r, self.phi_cell, self.vel,
            self.cfg, dt, self.cfg.seed + self.step_index)
        for k, i in (("emitted", 0), ("bubbles", 1), ("merged", 2), ("spray", 3)):
            self.droplet_stats[k] += c[i]

    def step(self, dt=None):
        """pfflip/sim.py:168-202."""
        if dt is None:
            dt = self.dt_from_cfl()
        self._p2g()
        self._phase_field_and_coefficients()
        self.vel_star = [v.copy() for v in self.vel]
        self._add_forces(dt)
        it, rel = self._pressure_solve(dt)
        self._project(dt)
        div_after = self._max_divergence()
        self._g2p()
        self._extrapolate_liquid_velocity()
        self._droplets(dt)
        self._advect(dt)
        self.time += dt
        self.step_index += 1
        rec = {"step": self.step_index, "t": self.time, "dt": dt, "pcg_iter": it,
               "pcg_res": rel, "max_div": div_after,
               "droplets": int((self.pstate == STATE_DROPLET).sum()) if self.cfg.droplets else 0}
        self.stats.append(rec)
        return rec
