"""Pins for the oracle's solid-obstacle path and for the product's scene
seeding, CPU only.

* ``solid_face_mask`` / ``zero_solid_obstacle_faces`` (pfflip/kernels.py:601-627)
  against a brute-force walk over every face and its two cells.
* ``push_out_of_solids`` (pfflip/kernels.py:630-679) against hand-derived
  geometries: a particle whose nearest free cell lies t cells away along the
  shallowest axis moves by exactly (t + 0.5) dx that way, loses only the
  velocity component pointing back in, ties go to the first direction in the
  order -x, +x, -y, +y, -z, +z, and nothing moves when no free cell lies
  within 7 cells or when the particle is dead.
* ``scenes.dam_break`` / ``still_pool`` / ``surge_column`` of the product
  against the oracle's one-shot lattice (the reference formula,
  pfflip/scenes.py:17-45) bit for bit across several generation chunks.
* The reference's obstacle tests (tests/test_pfflip.py:193-242) restated on
  the oracle at a small size.
"""

import numpy as np
import pytest

from conftest import load_pkg
from oracle import ref_np as O


# --------------------------------------------------------------- obstacle faces
def _brute_blocked(solid, axis):
    """Walk every face of ``axis``: blocked if either of its (at most two)
    cells inside the grid is solid."""
    nx, ny, nz = solid.shape
    shape = [nx, ny, nz]
    shape[axis] += 1
    out = np.zeros(shape, bool)
    for i in range(shape[0]):
        for j in range(shape[1]):
            for k in range(shape[2]):
                cells = [(i, j, k)]
                lo = [i, j, k]
                lo[axis] -= 1
                cells.append(tuple(lo))
                for c in cells:
                    if all(0 <= c[a] < solid.shape[a] for a in range(3)) and solid[c] == 1:
                        out[i, j, k] = True
    return out


def test_solid_face_mask_brute_force():
    rng = np.random.default_rng(3)
    solid = (rng.random((7, 5, 6)) < 0.2).astype(np.int32)
    solid[0, 0, 0] = 1          # a solid cell on a domain corner
    solid[-1, -1, -1] = 1
    for axis in range(3):
        ref = _brute_blocked(solid, axis)
        assert ref.any() and not ref.all()
        assert np.array_equal(O.solid_face_mask(solid, axis), ref)


def test_zero_solid_obstacle_faces_zeroes_exactly_the_blocked_faces():
    rng = np.random.default_rng(4)
    solid = np.zeros((6, 7, 5), np.int32)
    solid[2:4, 1:5, 2] = 1
    for axis in range(3):
        shape = list(solid.shape)
        shape[axis] += 1
        beta = rng.random(shape) + 0.5
        vel = rng.standard_normal(shape)
        b0, v0 = beta.copy(), vel.copy()
        O.zero_solid_obstacle_faces(solid, beta, vel, axis)
        m = _brute_blocked(solid, axis)
        assert np.all(beta[m] == 0.0) and np.all(vel[m] == 0.0)
        assert np.array_equal(beta[~m], b0[~m]) and np.array_equal(vel[~m], v0[~m])
        # the slab's own faces along the axis are blocked, both sides
        if axis == 2:
            assert m[3, 2, 2] and m[3, 2, 3] and not m[3, 2, 1] and not m[3, 2, 4]


# --------------------------------------------------------------- push-out
DX = 0.125


def _push(solid, cells_frac, vel, state=None):
    """Particles at cell + fraction (in cells), through the oracle."""
    x = np.asarray(cells_frac, np.float64) * DX
    v = np.asarray(vel, np.float64)
    st = np.zeros(len(x), np.int32) if state is None else np.asarray(state, np.int32)
    return O.push_out_of_solids(x, v, st, solid, DX)


def test_push_out_moves_to_nearest_free_cell_along_x():
    solid = np.zeros((24, 8, 8), np.int32)
    solid[10:15] = 1                      # x slab, cells 10..14, full in y and z
    xs = [[11.3, 4.5, 4.5],               # -x: cell 9 is 2 away; +x: cell 15 is 4 away
          [13.6, 2.5, 6.5],               # -x: 4 away; +x: 2 away
          [12.2, 4.5, 4.5],               # 3 away either side: the tie goes to -x
          [11.3, 4.5, 4.5]]               # as the first, moving away from the slab
    vs = [[0.7, 0.2, -0.1], [-0.4, 0.3, 0.5], [0.6, -0.2, 0.1], [-0.7, 0.2, -0.1]]
    x, v = _push(solid, xs, vs)
    expect_x = np.array(xs, np.float64) * DX
    expect_x[0, 0] -= 2.5 * DX
    expect_x[1, 0] += 2.5 * DX
    expect_x[2, 0] -= 3.5 * DX
    expect_x[3, 0] -= 2.5 * DX
    assert np.allclose(x, expect_x, rtol=0, atol=1e-15)
    # only the component pointing back into the solid is removed
    expect_v = np.array(vs, np.float64)
    expect_v[0, 0] = 0.0                  # moved -x, was moving +x
    expect_v[1, 0] = 0.0                  # moved +x, was moving -x
    expect_v[2, 0] = 0.0
    assert np.array_equal(v, expect_v)    # particle 3 keeps its -x velocity


def test_push_out_along_y_and_z():
    solid = np.zeros((8, 10, 12), np.int32)
    solid[:, 3:5, :] = 1                  # thin in y
    x, v = _push(solid, [[4.5, 3.25, 6.5]], [[0.1, 0.9, -0.3]])
    assert np.allclose(x[0], np.array([4.5, 3.25 - 1.5, 6.5]) * DX, rtol=0, atol=1e-15)
    assert np.array_equal(v[0], [0.1, 0.0, -0.3])
    solid = np.zeros((8, 10, 12), np.int32)
    solid[:, :, 4:6] = 1                  # thin in z: +z is one cell away from k = 5
    x, v = _push(solid, [[4.5, 5.5, 5.75]], [[0.1, 0.9, -0.3]])
    assert np.allclose(x[0], np.array([4.5, 5.5, 5.75 + 1.5]) * DX, rtol=0, atol=1e-15)
    assert np.array_equal(v[0], [0.1, 0.9, 0.0])


def test_push_out_search_stops_at_walls_and_after_seven_cells():
    solid = np.zeros((24, 8, 8), np.int32)
    solid[0:3] = 1                        # against the low x wall
    x, v = _push(solid, [[0.5, 4.5, 4.5]], [[-0.2, 0.0, 0.0]])
    assert np.allclose(x[0], np.array([0.5 + 3.5, 4.5, 4.5]) * DX, rtol=0, atol=1e-15)
    assert v[0, 0] == 0.0
    solid = np.zeros((24, 8, 8), np.int32)
    solid[2:22] = 1                       # 10 solid cells either side of cell 12
    x, v = _push(solid, [[12.5, 4.5, 4.5]], [[0.3, 0.1, 0.2]])
    assert np.array_equal(x[0], np.array([12.5, 4.5, 4.5]) * DX)
    assert np.array_equal(v[0], [0.3, 0.1, 0.2])


def test_push_out_skips_dead_moves_droplets_and_ignores_free_cells():
    solid = np.zeros((24, 8, 8), np.int32)
    solid[10:15] = 1
    xs = [[11.3, 4.5, 4.5], [11.3, 4.5, 4.5], [9.9, 4.5, 4.5]]
    vs = [[0.7, 0.0, 0.0], [0.7, 0.0, 0.0], [0.7, 0.0, 0.0]]
    x, v = _push(solid, xs, vs, state=[2, 1, 0])
    assert np.array_equal(x[0], np.array(xs[0]) * DX) and v[0, 0] == 0.7    # dead
    assert np.isclose(x[1, 0], (11.3 - 2.5) * DX) and v[1, 0] == 0.0         # droplet
    assert np.array_equal(x[2], np.array(xs[2]) * DX) and v[2, 0] == 0.7    # free cell


def test_push_out_cell_index_clamps_outside_the_grid():
    solid = np.zeros((8, 8, 8), np.int32)
    solid[7] = 1                          # the high x layer is solid
    x, _ = _push(solid, [[8.4, 4.5, 4.5]], [[0.0, 0.0, 0.0]])   # beyond the wall: cell 7
    assert np.isclose(x[0, 0], (8.4 - 1.5) * DX)


# --------------------------------------------------------------- scene seeding
@pytest.fixture(scope="module")
def scenes():
    return load_pkg().scenes


@pytest.mark.parametrize("chunk", [None, 777])
def test_product_dam_break_equals_oracle_lattice(scenes, chunk, monkeypatch):
    """The product generates the lattice in row chunks; the result equals
    the one-shot lattice of the reference formula bit for bit.  At its own
    chunk size 64^3 x 27 ppc spans two chunks; 777 rows exercises many
    chunk boundaries inside a lattice row."""
    if chunk is not None:
        monkeypatch.setattr(scenes, "_CHUNK", chunk)
        n, ppc = 6, 27
    else:
        n, ppc = 64, 27
        assert n ** 3 * ppc > scenes._CHUNK
    cfg = O.OracleConfig(res=(n, n, n), dx=1.0 / n, ppc=ppc)
    pos, vel, pt = scenes.dam_break(cfg)
    opos, ovel, opt = O.dam_break(cfg)
    assert pos.dtype == opos.dtype == np.float32
    assert np.array_equal(pos, opos) and np.array_equal(pt, opt) and not vel.any()


def test_product_seed_with_jitter_equals_reference_formula(scenes, monkeypatch):
    """z-up dam break with the benchmark's 0.5 jitter, against the formula
    written out here: row r of the 'ij' lattice is site (r // YZ, r // Z % Y,
    r % Z), jittered by (U - 0.5) * step * jitter with U from PCG64(seed)."""
    monkeypatch.setattr(scenes, "_CHUNK", 1000)
    n, ppc = 5, 8
    cfg = O.OracleConfig(res=(n, n, n), dx=1.0 / n, ppc=ppc)
    pos, _, pt = scenes.dam_break(cfg, frac=(0.4, 1.0, 0.6), jitter=0.5)
    X = n * 2
    step = cfg.dx / 2
    u = np.random.default_rng(cfg.seed).random((X ** 3, 3))
    ref = np.empty((X ** 3, 3))
    for r in range(X ** 3):
        site = np.array([r // (X * X), (r // X) % X, r % X], np.float64)
        ref[r] = (site + 0.5) * step + (u[r] - 0.5) * step * 0.5
    assert np.array_equal(pos, ref.astype(np.float32))
    liq = (ref[:, 0] < 0.4) & (ref[:, 2] < 0.6)
    assert np.array_equal(pt, liq.astype(np.int32))


def test_surge_column_drops_particles_inside_the_column(scenes):
    n = 20
    cfg = O.OracleConfig(res=(n, n, n), dx=1.0 / n, ppc=8)
    pos, vel, pt, solid = scenes.surge_column(cfg)
    full, _, _ = scenes.dam_break(cfg, frac=(1.0, 1.0, 0.3), jitter=0.5)
    inside = (np.abs(full[:, 0] - 0.7) < 0.05) & (np.abs(full[:, 1] - 0.5) < 0.05)
    assert len(pos) == int((~inside).sum()) and np.array_equal(pos, full[~inside])
    # the column: cells whose centre lies in the square, all heights but the top two
    c = (np.arange(n) + 0.5) / n
    sq = (np.abs(c[:, None] - 0.7) < 0.05) & (np.abs(c[None, :] - 0.5) < 0.05)
    assert np.array_equal(solid[:, :, 0], sq) and not solid[:, :, -2:].any()
    moving = (pt == 1) & (pos[:, 2] >= 0.2)
    assert np.all(vel[moving, 0] == 2.0) and not vel[~moving].any()


# --------------------------------------------------------------- obstacle scene (oracle)
def test_oracle_obstacle_scene_keeps_blocked_faces_at_zero(scenes):
    """Reference tests/test_pfflip.py:193-222 and :225-242 on the oracle
    (16^3, 4 steps): every face with beta = 0 carries zero velocity, and
    the speed stays bounded by 12 v0."""
    n = 16
    cfg = O.OracleConfig(res=(n, n, n), dx=1.0 / n, ppc=8, droplets=False)
    pos, vel, pt, solid = scenes.wave_vs_tower(cfg)
    assert solid.sum() > 0
    sim = O.OracleSim(cfg)
    sim.set_particles(pos, vel, pt)
    sim.set_solids(solid)
    v0 = float(np.abs(vel).max())
    for _ in range(4):
        sim.step()
        speed = float(np.linalg.norm(sim.pv[sim.pstate == 0], axis=1).max())
        assert speed < 12 * v0
    for ax in range(3):
        blocked = sim.beta[ax] == 0.0
        assert blocked.any() and np.abs(sim.vel[ax][blocked]).max() == 0.0
        sm = O.solid_face_mask(sim.solid, ax)
        assert np.all(sim.beta[ax][sm] == 0.0)
    # no live particle ends inside a solid cell that has a free cell within reach
    i = np.clip((sim.px / cfg.dx).astype(np.int64), 0, n - 1)
    inside = solid[i[:, 0], i[:, 1], i[:, 2]]
    assert not inside[sim.pstate != 2].any()
