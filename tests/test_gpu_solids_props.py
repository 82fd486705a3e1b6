"""The solid-obstacle path of the HIP step against the oracle, and the
reference's whole-step property tests run on the HIP ``Simulator`` (MI355X).

Solids: the obstacle-face zeroing of the fused phase pass before and after
gravity (pfflip/sim.py:256-285, kernels.py:601-627), the solid push-out of
the fused G2P + RK3 pass (kernels.py:630-679) and whole steps of the two
obstacle scenes against the oracle; the reference's obstacle tests
(tests/test_pfflip.py:193-242).  Properties: tests/test_pfflip.py:44-120 of
the reference (phi exactly 0 in the air, a one-cell interface, exact mass
and liquid count, hydrostatic rest, projection removing the divergence).
"""

import sys

import numpy as np
import pytest

from conftest import gpu_available, load_pkg
from oracle import ref_np as O

pytestmark = pytest.mark.gpu

if gpu_available():
    import torch


def say(msg):
    print(f"[gpu-solids] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pf():
    if not gpu_available():
        pytest.skip("no GPU")
    return load_pkg()


def np_(t):
    return t.detach().cpu().numpy()


def f32(x):
    return float(np.float32(x))


def rel_err(got, ref):
    ref = np.asarray(ref, np.float64)
    scale = max(np.abs(ref).max(), 1e-30)
    return float(np.abs(np.asarray(got, np.float64) - ref).max() / scale)


def ocfg_of(cfg):
    keys = ["res", "dx", "ppc", "rho_l", "density_ratio", "gravity", "alpha_phi",
            "alpha_flip_liquid", "alpha_flip_air", "cfl", "dt_max", "dt_min", "k_div", "pcg_tol",
            "pcg_max_iter", "mg_pre_sweeps", "mg_post_sweeps", "mg_coarsest_sweeps", "mg_omega",
            "mg_min_dim", "extrap_cells", "seed", "droplets"]
    return O.OracleConfig(**{k: getattr(cfg, k) for k in keys})


def obstacle_sim(pf, scene, n, ppc=8, **kw):
    if scene == "wave_vs_tower":
        cfg = pf.Config(res=(n, n, n), dx=1.0 / n, ppc=ppc, droplets=False, **kw)
        pos, vel, pt, solid = pf.scenes.wave_vs_tower(cfg)
    else:
        cfg = pf.Config(res=(n, n, n), dx=1.0 / n, ppc=ppc, droplets=False,
                        gravity=(0.0, 0.0, -9.81), **kw)
        pos, vel, pt, solid = pf.scenes.surge_column(cfg)
    sim = pf.Simulator(cfg)
    sim.set_particles(pos, vel, pt)
    sim.set_solids(solid)
    assert sim.has_solids
    return cfg, sim, solid.astype(np.int32)


# ------------------------------------------------------------------ obstacle faces
@pytest.mark.parametrize("scene", ["wave_vs_tower", "surge_column"])
def test_obstacle_faces_before_and_after_gravity_match_oracle(pf, scene):
    say(f"obstacle faces {scene} 20^3")
    cfg, sim, solid = obstacle_sim(pf, scene, 20)
    rng = np.random.default_rng(11)
    sim.pv.copy_(torch.from_numpy(rng.standard_normal((sim.n_particles, 3)).astype(np.float32)))
    osim = O.OracleSim(ocfg_of(cfg))
    sim._p2g()
    sim._phase_field_and_coefficients()
    torch.cuda.synchronize()
    gtag = np_(sim.is_liquid)
    vel_ref = []
    for ax in range(3):
        m, mu = np_(sim.mass[ax]), np_(sim.mom[ax])
        v = O.face_velocity_from_momentum(m, mu, 1e-12 * osim.rho_0l)[0]
        # the kernel takes rho_min and rho_0l as FP32 arguments, as the
        # reference's Warp kernels do; with the same FP32 constants the
        # subtraction rho - rho_min is exact in both (Sterbenz), so phi near
        # rho_min is not an ill-conditioned comparison
        f, b = O.face_beta(m, gtag, ax, f32(osim.rho_min), cfg.alpha_phi, f32(osim.rho_0l),
                           cfg.rho_l, cfg.rho_g)
        O.zero_solid_faces(b, v, ax)
        O.zero_solid_obstacle_faces(solid, b, v, ax)
        blocked = O.solid_face_mask(solid, ax) | O.wall_mask(b.shape, ax)
        gb, gv = np_(sim.beta[ax]), np_(sim.vel[ax])
        assert O.solid_face_mask(solid, ax).any()
        assert np.all(gb[blocked] == 0.0) and np.all(gv[blocked] == 0.0), ax
        assert np.all(gb[~blocked] > 0.0)
        assert rel_err(gb, b) < 1e-6 and rel_err(gv, v) < 1e-6
        assert rel_err(np_(sim.phi_face[ax]), f) < 1e-6
        vel_ref.append(gv.astype(np.float64))
    dt = 0.01
    sim._add_forces(dt)
    torch.cuda.synchronize()
    for ax in range(3):
        v = vel_ref[ax] + cfg.gravity[ax] * dt
        b = np_(sim.beta[ax]).astype(np.float64)
        O.zero_solid_faces(b, v, ax)
        O.zero_solid_obstacle_faces(solid, b, v, ax)
        blocked = O.solid_face_mask(solid, ax)
        gv = np_(sim.vel[ax])
        # gravity reached every face and was taken off the blocked ones again
        assert np.all(gv[blocked] == 0.0)
        assert rel_err(gv, v) < 1e-6


# ------------------------------------------------------------------ push-out
def _near_face(x, dx, tol=1e-4):
    """Positions within ``tol`` cells of a cell boundary on some axis, where
    FP32 and FP64 may pick different cells."""
    c = np.asarray(x, np.float64) / dx
    return (np.abs(c - np.round(c)) < tol).any(axis=1)


def test_g2p_advect_push_out_matches_oracle(pf):
    """G2P + RK3 + reflection + solid push-out, per particle and through the
    bins, against the oracle's g2p / advect_rk3 / reflect_at_walls /
    push_out_of_solids on the same fields: a block, a one-cell wall (the
    search finds free space one cell away) and a slab against the domain
    wall (the search stops at the wall)."""
    say("g2p + advect + push-out 16^3")
    lib = sys.modules[pf.__name__ + "._lib"]
    n = 16
    cfg = pf.Config(res=(n, n, n), dx=1.0 / n, ppc=8, droplets=False)
    sim = pf.Simulator(cfg)
    pos, vel, pt = pf.scenes.dam_break(cfg)
    sim.set_particles(pos, vel, pt)
    solid = np.zeros((n, n, n), np.int32)
    solid[6:10, 5:9, 4:12] = 1
    solid[:, 12, 2:14] = 1
    solid[0:2, :, 13:] = 1
    sim.set_solids(solid)
    rng = np.random.default_rng(12)
    for lst in (sim.vel, sim.vel_star, sim.vel_liq):
        for t in lst:
            t.copy_(torch.from_numpy((0.6 * rng.standard_normal(tuple(t.shape))).astype(np.float32)))
    st = np.zeros(sim.n_particles, np.int32)
    st[::97] = 1
    st[::101] = 2
    sim.pstate.copy_(torch.from_numpy(st))
    sim.pv.copy_(torch.from_numpy(rng.standard_normal((sim.n_particles, 3)).astype(np.float32)))
    px, pv, ptype = np_(sim.px).copy(), np_(sim.pv).copy(), np_(sim.ptype)
    vel = [np_(v) for v in sim.vel]
    vstar = [np_(v) for v in sim.vel_star]
    vliq = [np_(v) for v in sim.vel_liq]
    dt = 0.02
    dvel = [vel[a].astype(np.float64) - vstar[a] for a in range(3)]
    pv_ref = O.g2p(px, pv, ptype, st, vel, dvel, cfg.dx, cfg.alpha_flip_liquid, cfg.alpha_flip_air)
    lo, hi = np.full(3, 1e-4 * cfg.dx), np.array(cfg.domain_size) - 1e-4 * cfg.dx
    xa = O.advect_rk3(px, ptype, st, vel, vliq, dt, cfg.dx, lo, hi)
    xr, vr = O.reflect_at_walls(xa, pv_ref, st, lo, hi)
    xp, vp = O.push_out_of_solids(xr, vr, st, solid, cfg.dx)
    moved = np.abs(xp - xr).max(axis=1) > 0
    assert moved.sum() > 100
    keep = ~_near_face(xr, cfg.dx)
    assert keep.mean() > 0.99

    # per-particle stage methods (dvel from vel - vel_star, then advect)
    sim._g2p()
    sim._advect(dt)
    torch.cuda.synchronize()
    gx, gv = np_(sim.px).copy(), np_(sim.pv).copy()
    assert np.abs(gx - xp)[keep].max() < 2e-6
    assert np.abs(gv - vp)[keep].max() < 2e-6 * max(1.0, np.abs(vp).max())

    # the fused pass through the bins, from the same start state
    sim.px.copy_(torch.from_numpy(px))
    sim.pv.copy_(torch.from_numpy(pv))
    for a in range(3):
        sim.dvel[a].copy_(torch.from_numpy(dvel[a].astype(np.float32)))
    sim._bin()
    a = sim._g2p_args(lib.GA_G2P | lib.GA_ADVECT | lib.GA_REFLECT, dt)
    lib.call("pf_g2p_advect_binned", __import__("ctypes").byref(a), lib.ptr(sim.perm),
             lib.ptr(sim.cell_start), lib.ptr(sim._n_active), None, None, lib.stream())
    torch.cuda.synchronize()
    assert np.abs(np_(sim.px) - xp)[keep].max() < 2e-6
    assert np.abs(np_(sim.pv) - vp)[keep].max() < 2e-6 * max(1.0, np.abs(vp).max())


# ------------------------------------------------------------------ whole steps with solids
@pytest.mark.parametrize("scene", ["wave_vs_tower", "surge_column"])
def test_step_with_solids_matches_oracle(pf, scene):
    """Three full steps of an obstacle scene at 24^3; the oracle restarts
    each step from the GPU's particle state (as the solid-free step test)."""
    say(f"full step with solids {scene} 24^3")
    cfg, sim, solid = obstacle_sim(pf, scene, 24, pcg_tol=1e-7, pcg_max_iter=400)
    ocfg = ocfg_of(cfg)
    n = cfg.res[0]
    for s in range(3):
        osim = O.OracleSim(ocfg)
        osim.set_particles(np_(sim.px), np_(sim.pv), np_(sim.ptype))
        osim.set_solids(solid)
        dt = sim.dt_from_cfl()
        assert abs(dt - osim.dt_from_cfl()) <= 1e-6 * dt
        rec = sim.step(dt)
        # the oracle step, split before the push-out to see where it decides
        osim._p2g()
        osim._phase_field_and_coefficients()
        osim.vel_star = [v.copy() for v in osim.vel]
        osim._add_forces(dt)
        osim._pressure_solve(dt)
        osim._project(dt)
        osim._g2p()
        osim._extrapolate_liquid_velocity()
        eps = 1e-4 * cfg.dx
        lo, hi = np.full(3, eps), np.array(cfg.domain_size) - eps
        xa = O.advect_rk3(osim.px, osim.ptype, osim.pstate, osim.vel, osim.vel_liq, dt, cfg.dx, lo, hi)
        xr, vr = O.reflect_at_walls(xa, osim.pv, osim.pstate, lo, hi)
        xp, vp = O.push_out_of_solids(xr, vr, osim.pstate, solid, cfg.dx)
        torch.cuda.synchronize()
        band = np.abs(osim.phi_cell - 0.5) < 1e-4
        assert np.array_equal(np_(sim.is_liquid)[~band], osim.is_liquid[~band])
        assert rel_err(np_(sim.pressure), osim.pressure) < 1e-4, s
        u_scale = max(float(np.abs(v).max()) for v in osim.vel_star) + 9.81 * dt
        for ax in range(3):
            gv = np_(sim.vel[ax])
            assert np.abs(gv - osim.vel[ax]).max() < 3e-4 * u_scale, (s, ax)
            assert np.all(gv[O.solid_face_mask(solid, ax)] == 0.0)
        keep = ~_near_face(xr, cfg.dx)
        assert np.abs(np_(sim.px) - xp)[keep].max() < 3e-4 * dt * u_scale + 1e-6, s
        assert np.abs(np_(sim.pv) - vp)[keep].max() < 3e-4 * u_scale, s
        assert rec["pcg_res"] < cfg.pcg_tol
    say(f"  {scene}: pushed-out particles in last step "
        f"{int((np.abs(xp - xr).max(axis=1) > 0).sum())} of {n ** 3 * 8}")


# ------------------------------------------------------------------ reference obstacle tests
def test_obstacle_does_not_leak_gravity(pf):
    """Reference tests/test_pfflip.py:193-222 on the HIP simulator."""
    say("obstacle does not leak gravity 32^3")
    cfg, sim, solid = obstacle_sim(pf, "wave_vs_tower", 32)
    assert solid.sum() > 0
    for _ in range(5):
        sim.step()
    torch.cuda.synchronize()
    for ax in range(3):
        beta = np_(sim.beta[ax])
        v = np_(sim.vel[ax])
        blocked = beta == 0.0
        assert blocked.any()
        assert np.abs(v[blocked]).max() == 0.0, ax


def test_obstacle_scene_stays_bounded(pf):
    """Reference tests/test_pfflip.py:225-242 on the HIP simulator."""
    say("obstacle scene stays bounded 32^3 x 40 steps")
    cfg, sim, solid = obstacle_sim(pf, "wave_vs_tower", 32)
    v0 = float(np.abs(np_(sim.pv)).max())
    speeds = []
    for _ in range(40):
        sim.step()
        st = np_(sim.pstate)
        speeds.append(float(np.linalg.norm(np_(sim.pv)[st == 0], axis=1).max()))
    assert np.isfinite(speeds).all()
    assert max(speeds) < 12 * v0, (max(speeds), v0)


# ------------------------------------------------------------------ reference step properties
@pytest.fixture(scope="module")
def pool(pf):
    cfg = pf.Config(res=(32, 32, 32), dx=1.0 / 32, cfl=1.0)
    sim = pf.Simulator(cfg)
    sim.set_particles(*pf.scenes.still_pool(cfg, fill=0.5))
    return sim


def test_phase_field_saturates(pool):
    """Reference tests/test_pfflip.py:44-62: phi is exactly 0.0 in the air
    (eq. 7's rho_min sits log(1000) above the air's density), near 1 in the
    liquid bulk and clear of the 0.5 tag threshold."""
    say("phi saturates (pool 32^3)")
    pool.step()
    phi = np_(pool.phi_cell)
    ny = phi.shape[1]
    air = phi[:, 3 * ny // 4:, :]
    liq = phi[:, : ny // 4, :]
    assert air.max() == 0.0
    assert liq.mean() > 0.95 and liq.min() > 0.6
    # the whole air side, not only its top quarter: every cell above the
    # interface band holds exactly zero
    assert phi[:, ny // 2 + 2:, :].max() == 0.0


def test_interface_thickness_one_cell(pool):
    """Reference tests/test_pfflip.py:65-74."""
    say("interface thickness")
    pool.step()
    phi = np_(pool.phi_cell)
    col = phi[16, :, 16]
    inside = np.where(col > 0.95)[0]
    outside = np.where(col < 0.05)[0]
    gap = outside[outside > inside.max()].min() - inside.max()
    assert gap <= 3


def test_hydrostatic_equilibrium_is_stable(pool):
    """Reference tests/test_pfflip.py:90-103 (30 steps)."""
    say("hydrostatic 30 steps")
    liq = np_(pool.ptype) == 1
    y0 = np_(pool.px)[liq, 1].mean()
    speeds = []
    for _ in range(30):
        pool.step()
        speeds.append(np.linalg.norm(np_(pool.pv)[liq], axis=1).max())
    y1 = np_(pool.px)[liq, 1].mean()
    g_dt = 9.81 * pool.stats[-1]["dt"]
    assert max(speeds) < 0.5 * g_dt, (max(speeds), g_dt)
    assert abs(y1 - y0) < 0.01 * pool.cfg.domain_size[1]
    assert speeds[-1] <= speeds[0] * 1.5


def test_mass_conserved_exactly(pf):
    """Reference tests/test_pfflip.py:77-87: 10 dam-break steps at 24^3."""
    say("mass conserved 24^3 x 10 steps")
    cfg = pf.Config(res=(24, 24, 24), dx=1.0 / 24, cfl=1.0)
    sim = pf.Simulator(cfg)
    sim.set_particles(*pf.scenes.dam_break(cfg))
    m0 = np_(sim.pmass).astype(np.float64).sum()
    n0 = (np_(sim.ptype) == 1).sum()
    for _ in range(10):
        sim.step()
    assert np_(sim.pmass).astype(np.float64).sum() == m0
    assert (np_(sim.ptype) == 1).sum() == n0
    assert np.isfinite(np_(sim.px)).all()


def test_projection_reduces_divergence(pf):
    """Reference tests/test_pfflip.py:106-120 on the HIP stage methods."""
    say("projection reduces divergence 32^3")
    cfg = pf.Config(res=(32, 32, 32), dx=1.0 / 32, cfl=1.0, pcg_tol=1e-5)
    sim = pf.Simulator(cfg)
    sim.set_particles(*pf.scenes.dam_break(cfg))
    sim._p2g()
    sim._phase_field_and_coefficients()
    for ax in range(3):
        sim.vel_star[ax].copy_(sim.vel[ax])
    dt = 1e-3
    sim._add_forces(dt)
    div_before = sim._max_divergence()
    sim._pressure_solve(dt)
    sim._project(dt)
    div_after = sim._max_divergence()
    assert div_after < 0.05 * div_before, (div_before, div_after)
