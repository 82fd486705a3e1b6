"""Escaped particles / droplets / spray on the GPU (MI355X).

Stage parity with the oracle (``oracle/ref_np.py``, which restates
pfflip/droplets.py and sim.py:343-401): particle states must match exactly
except for particles whose interpolated phi (or shear, or sub-step count)
sits within a rounding distance of a threshold, which FP32 and FP64 may
decide differently; those are counted and bounded.  Radii, positions and
velocities are compared with the tolerances written next to each check.
The reference's own droplet tests (tests/test_pfflip.py:123-186, 245-268)
are restated at the end.
"""

import math
import sys
from importlib import import_module

import numpy as np
import pytest

from conftest import gpu_available, load_pkg
from oracle import ref_np as O

pytestmark = pytest.mark.gpu

if gpu_available():
    import torch


def say(msg):
    print(f"[gpu-droplets] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pf():
    if not gpu_available():
        pytest.skip("no GPU")
    return load_pkg()


def np_(t):
    return t.detach().cpu().numpy()


def put(t, a):
    t.copy_(torch.from_numpy(np.ascontiguousarray(a)))


def ocfg_of(cfg):
    keys = ["res", "dx", "ppc", "rho_l", "density_ratio", "gravity", "alpha_phi",
            "alpha_flip_liquid", "alpha_flip_air", "cfl", "dt_max", "dt_min", "k_div", "pcg_tol",
            "pcg_max_iter", "mg_pre_sweeps", "mg_post_sweeps", "mg_coarsest_sweeps", "mg_omega",
            "mg_min_dim", "extrap_cells", "seed", "droplets", "delta_phi", "droplet_radius_mean",
            "droplet_radius_spread", "droplet_drag_cd", "evaporation_k", "spray_shear_threshold",
            "spray_emission_rate", "spray_resolution_factor", "droplet_max_substeps"]
    return O.OracleConfig(**{k: getattr(cfg, k) for k in keys})


def make(pf, n=16, ppc=8, scene="dam_break", **kw):
    kw.setdefault("gravity", (0.0, 0.0, -9.81))
    cfg = pf.Config(res=(n, n, n), dx=1.0 / n, ppc=ppc, droplets=True, **kw)
    sim = pf.Simulator(cfg)
    sim.set_particles(*getattr(pf.scenes, scene)(cfg))
    return cfg, sim


# ------------------------------------------------------------------ stage parity
@pytest.mark.parametrize("spray,evap", [(False, 0.0), (True, 0.0), (True, 2e-3)])
def test_droplet_stage_matches_oracle(pf, spray, evap):
    say(f"droplet stage 16^3 spray={spray} evap={evap}")
    rng = np.random.default_rng(21)
    cfg, sim = make(pf, 16, 8, spray_shear_threshold=1.0 if spray else 0.0,
                    spray_emission_rate=0.5, evaporation_k=evap)
    n, dx, dt = sim.n_particles, cfg.dx, 0.01
    phi = rng.random(sim.res).astype(np.float32)
    put(sim.phi_cell, phi)
    vel = [(rng.standard_normal(s) * 0.5).astype(np.float32) for s in sim.face_shapes]
    for t, a in zip(sim.vel, vel):
        put(t, a)
    pv = (rng.standard_normal((n, 3)) * 2.0).astype(np.float32)
    put(sim.pv, pv)
    st = rng.choice(3, size=n, p=[0.6, 0.3, 0.1]).astype(np.int32)
    put(sim.pstate, st)
    r0 = (rng.random(n) * 0.02 * dx + 0.02 * dx).astype(np.float32)
    put(sim.pdrop_r, r0)
    px, pt = np_(sim.px).copy(), np_(sim.ptype).copy()
    seed = cfg.seed + sim.step_index

    sim._droplets(dt)
    torch.cuda.synchronize()
    ocfg = ocfg_of(cfg)
    x_ref, v_ref, st_ref, r_ref, c_ref = O.droplet_stage(px, pv, pt, st, r0, phi, vel, ocfg, dt, seed)

    # particles whose decisions FP32 may take differently from FP64
    f = O.sample_cell(phi.astype(np.float64), px.astype(np.float64), dx)
    near = np.zeros(n, bool)
    for thr in (0.5 - cfg.delta_phi, 0.5 + cfg.delta_phi, 0.1, 0.9):
        near |= np.abs(f - thr) < 1e-5
    if spray:
        shear = O.velocity_gradient_magnitude(vel, dx)
        near |= np.abs(O.sample_cell(shear, px.astype(np.float64), dx) - 1.0) < 1e-4
    speed = np.linalg.norm(pv.astype(np.float64), axis=1) + 1e-8
    q = speed * dt / (0.5 * dx)
    near |= np.abs(q - np.round(q)) < 1e-4
    assert near.sum() < 0.001 * n + 5
    ok = ~near
    st_gpu = np_(sim.pstate)
    assert np.array_equal(st_gpu[ok], st_ref[ok]), int((st_gpu[ok] != st_ref[ok]).sum())
    assert (st_ref != st).any() and (st_ref == O.STATE_DROPLET).any()

    r_gpu = np_(sim.pdrop_r).astype(np.float64)
    # radii: drawn from the same stream; exp/log/cos in FP32 vs FP64
    drop = ok & (st_ref == O.STATE_DROPLET)
    assert np.abs(r_gpu[drop] - r_ref[drop]).max() <= 1e-5 * r_ref[drop].max()
    x_gpu, v_gpu = np_(sim.px).astype(np.float64), np_(sim.pv).astype(np.float64)
    # advanced droplets: up to 8 FP32 sub-steps of O(1) velocities
    assert np.abs(x_gpu[ok] - x_ref[ok]).max() < 2e-6
    assert np.abs(v_gpu[ok] - v_ref[ok]).max() < 2e-5 * max(np.abs(v_ref).max(), 1.0)
    # particles killed in advance_droplets keep their pre-advance position
    died = ok & (st == O.STATE_DROPLET) & (st_ref == O.STATE_DEAD)
    assert np.array_equal(x_gpu[died], px[died].astype(np.float64))
    counts = sim.droplet_stats
    got = [counts["emitted"], counts["bubbles"], counts["merged"], counts["spray"]]
    assert all(abs(a - b) <= int(near.sum()) for a, b in zip(got, c_ref)), (got, c_ref)
    if spray:
        assert c_ref[3] > 0
    if evap > 0.0:
        assert died.any()


def test_droplet_radii_stream_and_distribution(pf):
    """Every liquid FLIP particle in phi = 0 becomes a droplet.  Its radius
    is the oracle's draw from the same stream (to FP32 ulps of exp/log/cos),
    and the radii are log-normal with the configured mean."""
    say("radii 64^3 x 8")
    cfg, sim = make(pf, 64, 8)
    n = sim.n_particles
    put(sim.ptype, np.ones(n, np.int32))
    sim.phi_cell.zero_()
    for t in sim.vel:
        t.zero_()
    sim.pv.zero_()
    sim._droplets(1e-6)
    torch.cuda.synchronize()
    assert sim.droplet_stats["emitted"] == n
    r = np_(sim.pdrop_r).astype(np.float64)
    mu, sigma = O.lognormal_params(cfg.droplet_radius_mean * cfg.dx, cfg.droplet_radius_spread)
    lr = np.log(r)
    assert abs(lr.mean() - mu) < 4 * sigma / math.sqrt(n)
    assert abs(lr.std() - sigma) < 2e-3
    assert abs(r.mean() / (cfg.droplet_radius_mean * cfg.dx) - 1.0) < 2e-3
    assert r.min() >= 0.02 * cfg.dx * (1 - 1e-6) and r.max() <= 2.0 * cfg.dx * (1 + 1e-6)
    idx = np.arange(0, n, 997)
    _, r_ref = O._draw_radius(O.rng_init(cfg.seed, idx), mu, sigma, 0.02 * cfg.dx, 2.0 * cfg.dx)
    assert np.abs(r[idx] / r_ref - 1.0).max() < 2e-6


def test_velocity_gradient_matches_oracle(pf):
    say("velocity gradient 24x20x16")
    lib = import_module("implementation-of-adaptive-phase-field-flip_amd._lib")
    rng = np.random.default_rng(3)
    res, dx = (24, 20, 16), 0.05
    shapes = O.face_shapes(res)
    vel = [rng.standard_normal(s).astype(np.float32) for s in shapes]
    dv = [torch.from_numpy(v).cuda() for v in vel]
    out = torch.zeros(res, dtype=torch.float32, device="cuda")
    lib.call("pf_velocity_gradient", lib.ptr3(dv), lib.ptr(out), *res, dx, lib.stream())
    torch.cuda.synchronize()
    ref = O.velocity_gradient_magnitude(vel, dx)
    assert np.abs(np_(out) - ref).max() < 2e-6 * ref.max()


def test_rasterize_spray_matches_oracle(pf):
    say("rasterize spray")
    cfg, sim = make(pf, 16, 8)
    n = sim.n_particles
    rng = np.random.default_rng(4)
    st = np.where(rng.random(n) < 0.05, 1, 0).astype(np.int32)
    put(sim.pstate, st)
    r = (rng.random(n) * 0.5 * cfg.dx).astype(np.float32)
    put(sim.pdrop_r, r)
    dens = np_(sim.rasterize_spray()).astype(np.float64)
    ref = O.rasterize_spray(np_(sim.px), st, r, sim.spray_res, sim.dx_spray)
    assert sim.spray_density.dtype == torch.float16
    # each contribution is rounded to FP16 and so is each atomic sum
    assert np.abs(dens - ref).max() < 4e-3 * ref.max()
    assert abs(dens.sum() / ref.sum() - 1.0) < 1e-3


# ------------------------------------------------------------------ whole step
def test_step_with_droplets_matches_oracle(pf):
    """One step of a 16^3 x 27 ppc asteroid impact with the droplet stage on,
    the oracle started from the GPU's state; the stage sits between G2P and
    advection."""
    say("step with droplets 16^3 x 27")
    cfg, sim = make(pf, 16, 27, scene="asteroid_impact", pcg_tol=1e-7, pcg_max_iter=400,
                    gravity=(0.0, -9.81, 0.0))
    for _ in range(6):
        sim.step()
    ocfg = ocfg_of(cfg)
    osim = O.OracleSim(ocfg)
    osim.set_particles(np_(sim.px), np_(sim.pv), np_(sim.ptype))
    osim.pstate = np_(sim.pstate).copy()
    osim.pdrop_r = np_(sim.pdrop_r).astype(np.float64)
    osim.step_index = sim.step_index
    x0 = np_(sim.px).astype(np.float64)
    dt = sim.dt_from_cfl()
    assert abs(dt - osim.dt_from_cfl()) <= 1e-6 * dt
    rec = sim.step(dt)
    osim.step(dt)
    torch.cuda.synchronize()
    f = O.sample_cell(osim.phi_cell, x0, cfg.dx)
    near = (np.abs(f - 0.3) < 1e-4) | (np.abs(f - 0.7) < 1e-4)
    st_gpu = np_(sim.pstate)
    assert np.array_equal(st_gpu[~near], osim.pstate[~near])
    assert rec["droplets"] == int((st_gpu == 1).sum())
    assert sum(sim.droplet_stats.values()) > 0
    u_scale = max(float(np.abs(v).max()) for v in osim.vel_star) + 9.81 * dt
    flip = ~near & (st_gpu == 0)
    assert np.abs(np_(sim.px)[flip] - osim.px[flip]).max() < 3e-4 * dt * u_scale + 1e-6
    assert np.abs(np_(sim.pv)[flip] - osim.pv[flip]).max() < 3e-4 * u_scale


# ------------------------------------------------------------------ reference tests
def test_bubbles_are_deleted_and_droplets_merge_back(pf):
    """Reference tests/test_pfflip.py:124-137."""
    say("asteroid 32^3 x 27, 25 steps")
    cfg, sim = make(pf, 32, 27, scene="asteroid_impact", gravity=(0.0, -9.81, 0.0))
    for _ in range(25):
        sim.step()
    s = sim.droplet_stats
    assert s["bubbles"] > 0, "air particles trapped in liquid are never deleted"
    counts = sim.state_counts()
    assert counts["flip"] + counts["droplet"] + counts["dead"] == sim.n_particles
    assert counts["dead"] == s["bubbles"] or counts["dead"] >= s["bubbles"] - s["merged"]
    assert sim.stats[-1]["droplets"] == counts["droplet"]


def test_droplets_do_not_contribute_to_the_grid(pf):
    """Reference tests/test_pfflip.py:170-186."""
    say("droplets off the grid")
    cfg = pf.Config(res=(16,) * 3, dx=1 / 16, cfl=1.0, ppc=8, droplets=True)
    sim = pf.Simulator(cfg)
    sim.set_particles(*pf.scenes.still_pool(cfg, fill=0.5))
    sim.step()
    m_all = float(sim.cell_mass.sum())
    st = np_(sim.pstate).copy()
    liq = np.flatnonzero(np_(sim.ptype) == 1)
    st[liq[::10]] = 1
    put(sim.pstate, st)
    sim._p2g()
    m_flip = float(sim.cell_mass.sum())
    assert m_flip < m_all
    removed = 1.0 - m_flip / m_all
    assert abs(removed - 0.10) < 0.02, removed


def test_droplets_do_not_throttle_the_time_step(pf):
    """Reference tests/test_pfflip.py:245-268."""
    say("CFL over FLIP particles only")
    cfg = pf.Config(res=(32,) * 3, dx=1 / 32, cfl=1.0, ppc=8, droplets=True)
    sim = pf.Simulator(cfg)
    sim.set_particles(*pf.scenes.asteroid_impact(cfg))
    sim.step()
    v = np_(sim.pv).copy()
    st = np_(sim.pstate).copy()
    v[0] = (500.0, 0.0, 0.0)
    st[0] = 1
    v[1] = (500.0, 0.0, 0.0)
    st[1] = 2
    put(sim.pv, v)
    put(sim.pstate, st)
    assert sim.dt_from_cfl() > cfg.dx / 50.0
