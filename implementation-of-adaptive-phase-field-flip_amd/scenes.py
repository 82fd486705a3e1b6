"""Scene set-ups (drop-in for pfflip/scenes.py).

Both phases are seeded everywhere: the air is a fluid with its own
particles.  ``_lattice`` / ``seed`` reproduce the reference's numpy stream
(``default_rng(cfg.seed).random``) and formula bit for bit
(pfflip/scenes.py:17-45); the lattice is generated in row chunks, each
drawing from the PCG64 stream advanced to its first row, which gives the
numbers of one large call while chunks run on several threads and memory
stays bounded at the 256^3 x 27 ppc scale.

The reference scenes are y-up.  The scenes added for the benchmark
configurations (z-up dam break, multi-mode Rayleigh-Taylor, slab surge on a
solid column) take their vertical axis from ``up_axis``.
"""

from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

AIR, LIQUID = 0, 1
_CHUNK = 1 << 22  # lattice rows per chunk
_SEED_THREADS = min(8, os.cpu_count() or 1)


def _lattice_geometry(res, dx, ppc):
    per_axis = round(ppc ** (1.0 / 3.0))
    assert per_axis**3 == ppc, "ppc must be a perfect cube (8, 27, ...)"
    nx, ny, nz = res
    return dx / per_axis, (nx * per_axis, ny * per_axis, nz * per_axis)


def _lattice_rows(r0, r1, step, dims, rng, jitter):
    """Rows [r0, r1) of the jittered lattice; ``rng`` must stand at draw 3 r0
    of the stream."""
    X, Y, Z = dims
    r = np.arange(r0, r1, dtype=np.int64)
    # row r of the 'ij' meshgrid is lattice site (r // (Y Z), (r // Z) % Y, r % Z)
    idx = np.stack([r // (Y * Z), (r // Z) % Y, r % Z], axis=-1)
    pos = (idx + 0.5) * step
    pos += (rng.random(pos.shape) - 0.5) * step * jitter
    return pos


def _lattice_chunks(res, dx, ppc, rng, jitter=0.8):
    """Yield float64 position chunks of the jittered lattice (scenes.py:17-31)."""
    step, dims = _lattice_geometry(res, dx, ppc)
    n = dims[0] * dims[1] * dims[2]
    for r0 in range(0, n, _CHUNK):
        yield _lattice_rows(r0, min(n, r0 + _CHUNK), step, dims, rng, jitter)


def _lattice(res, dx, ppc, rng, jitter=0.8):
    """Jittered ``ppc``-per-cell lattice covering the whole domain."""
    return np.concatenate(list(_lattice_chunks(res, dx, ppc, rng, jitter)), axis=0)


def seed(cfg, liquid_mask_fn, velocity_fn=None, jitter=0.8):
    """Particles from a predicate on world position (scenes.py:34-45).

    ``liquid_mask_fn(pos) -> bool``; ``velocity_fn(pos, is_liquid) -> (n,3)``.
    Both see float64 positions, as in the reference.  ``jitter`` is the
    peak-to-peak jitter in sub-cell units (0.8 in the reference).
    """
    step, dims = _lattice_geometry(cfg.res, cfg.dx, cfg.ppc)
    n = dims[0] * dims[1] * dims[2]
    pos_out = np.empty((n, 3), np.float32)
    vel_out = np.zeros((n, 3), np.float32)
    ptype = np.empty(n, np.int32)

    def chunk(r0):
        # each chunk draws from its own copy of the stream, advanced to the
        # chunk's first draw (one 64-bit draw per coordinate): the numbers are
        # those of one sequential call, and numpy releases the GIL while
        # filling, so chunks run in parallel threads
        bg = np.random.PCG64(cfg.seed)
        bg.advance(3 * r0)
        r1 = min(n, r0 + _CHUNK)
        pos = _lattice_rows(r0, r1, step, dims, np.random.Generator(bg), jitter)
        liq = np.asarray(liquid_mask_fn(pos), dtype=bool)
        ptype[r0:r1] = np.where(liq, LIQUID, AIR)
        if velocity_fn is not None:
            vel_out[r0:r1] = np.asarray(velocity_fn(pos, liq), float)
        pos_out[r0:r1] = pos

    starts = range(0, n, _CHUNK)
    workers = min(len(starts), _SEED_THREADS)
    if workers <= 1:
        for r0 in starts:
            chunk(r0)
    else:
        with ThreadPoolExecutor(max_workers=workers) as ex:
            for _ in ex.map(chunk, starts):
                pass
    return pos_out, vel_out, ptype


# --------------------------------------------------------------- reference scenes
def dam_break(cfg, frac=(0.4, 0.6, 1.0), jitter=0.8):
    """Liquid block ``x < frac0, y < frac1, z < frac2`` (scenes.py:49-60).

    The z-up benchmark dam break (liquid ``x < 0.4, z < 0.6``) is
    ``frac=(0.4, 1.0, 0.6)`` with gravity along -z.
    """
    sx, sy, sz = cfg.domain_size

    def mask(p):
        return (p[:, 0] < frac[0] * sx) & (p[:, 1] < frac[1] * sy) & (p[:, 2] < frac[2] * sz)

    return seed(cfg, mask, jitter=jitter)


def still_pool(cfg, fill=0.5, up_axis=1, jitter=0.8):
    """Hydrostatic rest state (scenes.py:63-70)."""
    h = cfg.domain_size[up_axis]

    def mask(p):
        return p[:, up_axis] < fill * h

    return seed(cfg, mask, jitter=jitter)


def asteroid_impact(cfg, *, water_level=0.45, radius_frac=0.08, speed=None):
    """A dense sphere fired into a deep pool (scenes.py:73-96)."""
    sx, sy, sz = cfg.domain_size
    r = radius_frac * sx
    centre = np.array([0.5 * sx, water_level * sy + 0.35 * sy, 0.5 * sz])
    v_impact = speed if speed is not None else -0.9 * np.sqrt(2 * 9.81 * 0.35 * sy)

    def mask(p):
        return (p[:, 1] < water_level * sy) | (np.linalg.norm(p - centre, axis=1) < r)

    def vel(p, liq):
        v = np.zeros_like(p)
        v[np.linalg.norm(p - centre, axis=1) < r, 1] = v_impact
        return v

    return seed(cfg, mask, vel)


def rayleigh_taylor(cfg, *, ratio=None):
    """Heavy over light with two cosine modes, y-up (scenes.py:99-113)."""
    sx, sy, sz = cfg.domain_size

    def mask(p):
        h = 0.5 * sy + 0.02 * sy * (np.cos(2 * np.pi * p[:, 0] / sx) + np.cos(2 * np.pi * p[:, 2] / sz))
        return p[:, 1] > h

    return seed(cfg, mask)


def wave_vs_tower(cfg, *, sea_level=0.22, column=0.26, tower_r=0.075, tower_at=(0.62, 0.5),
                  surge=None):
    """Surge onto a cylindrical tower, y-up (scenes.py:125-166).

    Returns ``(pos, vel, ptype, solid_mask)``.
    """
    sx, sy, sz = cfg.domain_size
    nx, ny, nz = cfg.res
    v0 = surge if surge is not None else 0.55 * np.sqrt(9.81 * sy)

    def mask(p):
        return (p[:, 1] < sea_level * sy) | ((p[:, 0] < column * sx) & (p[:, 1] < 0.72 * sy))

    def vel(p, liq):
        v = np.zeros_like(p)
        v[(p[:, 0] < column * sx) & (p[:, 1] < 0.72 * sy) & liq, 0] = v0
        return v

    pos, pvel, ptype = seed(cfg, mask, vel)
    ci = (np.arange(nx) + 0.5) / nx
    ck = (np.arange(nz) + 0.5) / nz
    I, K = np.meshgrid(ci, ck, indexing="ij")
    disc = (I - tower_at[0]) ** 2 + (K - tower_at[1]) ** 2 < tower_r**2
    solid = np.broadcast_to(disc[:, None, :], (nx, ny, nz)).copy()
    solid[:, -2:, :] = False
    keep = ~(((pos[:, 0] / sx - tower_at[0]) ** 2 + (pos[:, 2] / sz - tower_at[1]) ** 2) < tower_r**2)
    return pos[keep], pvel[keep], ptype[keep], solid


# --------------------------------------------------------------- benchmark scenes
def rayleigh_taylor_modes(cfg, *, n_modes=8, amplitude=0.02, kmax=4, mode_seed=None,
                          jitter=0.5, up_axis=2):
    """Heavy (liquid, rho_l) over light phase split at half height.

    Interface ``h(x, y) = 0.5 + (amplitude / n_modes) * sum_m cos(2 pi (kx_m x
    + ky_m y) + theta_m)`` in units of the domain size, so |h - 0.5| <=
    amplitude.  ``kx_m, ky_m`` are integers in [1, kmax] and ``theta_m`` is
    uniform in [0, 2 pi), drawn from ``default_rng(mode_seed)`` (default
    ``cfg.seed + 1``).  The two horizontal axes are the ones other than
    ``up_axis``.
    """
    L = cfg.domain_size
    h_ax = [a for a in range(3) if a != up_axis]
    mrng = np.random.default_rng(cfg.seed + 1 if mode_seed is None else mode_seed)
    kx = mrng.integers(1, kmax + 1, n_modes)
    ky = mrng.integers(1, kmax + 1, n_modes)
    th = mrng.random(n_modes) * 2.0 * np.pi

    def mask(p):
        u = p[:, h_ax[0]] / L[h_ax[0]]
        v = p[:, h_ax[1]] / L[h_ax[1]]
        h = np.zeros(len(p))
        for m in range(n_modes):
            h += np.cos(2.0 * np.pi * (kx[m] * u + ky[m] * v) + th[m])
        h = 0.5 + amplitude / n_modes * h
        return p[:, up_axis] / L[up_axis] > h

    return seed(cfg, mask, jitter=jitter)


def surge_column(cfg, *, depth=0.3, top=0.1, u0=2.0, column_x=0.7, column_y=0.5, width=0.1,
                 jitter=0.5, up_axis=2):
    """Liquid slab of height ``depth`` whose upper ``top`` moves at +``u0``
    along x, hitting a square solid column of side ``width`` centred at
    (``column_x``, ``column_y``) that spans the height except the top two
    cells.  Positions in units of the domain size; z-up by default.

    Returns ``(pos, vel, ptype, solid_mask)``; particles seeded inside the
    column are dropped (as scenes.py:161-166 does for the tower).
    """
    L = cfg.domain_size
    res = cfg.res
    h_ax = [a for a in range(3) if a != up_axis]

    def mask(p):
        return p[:, up_axis] < depth * L[up_axis]

    def vel(p, liq):
        v = np.zeros_like(p)
        moving = liq & (p[:, up_axis] >= (depth - top) * L[up_axis])
        v[moving, h_ax[0]] = u0
        return v

    pos, pvel, ptype = seed(cfg, mask, vel, jitter=jitter)

    def in_column(u, w):
        return (np.abs(u - column_x) < 0.5 * width) & (np.abs(w - column_y) < 0.5 * width)

    c = [(np.arange(n) + 0.5) / n for n in res]
    U, W = np.meshgrid(c[h_ax[0]], c[h_ax[1]], indexing="ij")
    plate = in_column(U, W)
    shape = [1, 1, 1]
    solid = np.zeros(res, bool)
    idx = [slice(None)] * 3
    for a_i, a in enumerate(h_ax):
        shape[a] = res[a]
    solid |= plate.reshape(shape)
    idx[up_axis] = slice(res[up_axis] - 2, None)
    solid[tuple(idx)] = False
    keep = ~in_column(pos[:, h_ax[0]] / L[h_ax[0]], pos[:, h_ax[1]] / L[h_ax[1]])
    return pos[keep], pvel[keep], ptype[keep], solid


SCENES = {
    "dam_break": dam_break,
    "still_pool": still_pool,
    "asteroid": asteroid_impact,
    "rayleigh_taylor": rayleigh_taylor,
    "wave_vs_tower": wave_vs_tower,
    "rayleigh_taylor_modes": rayleigh_taylor_modes,
    "surge_column": surge_column,
}
