"""The C++/OpenMP restatement of the reference step (``oracle/cpu_step.cpp``,
the CPU baseline) against the FP64 numpy oracle, CPU only.

Both restate the same reference kernels, so one checks the other: three
steps of the dam break (8 and 27 ppc) and of the obstacle scene, the numpy
oracle restarted each step from the C++ particle state, with the tolerances
of the GPU parity tests (FP32 against FP64).  The thread count changes only
the order of the float atomics.
"""

import os
import subprocess

import numpy as np
import pytest

from conftest import load_pkg
from oracle import ref_np as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def cs():
    from oracle import cpu_step
    if not os.path.exists(cpu_step.LIB):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    cpu_step.load()
    return cpu_step


def rel_err(got, ref):
    ref = np.asarray(ref, np.float64)
    return float(np.abs(np.asarray(got, np.float64) - ref).max() / max(np.abs(ref).max(), 1e-30))


def make(cs, n, ppc, scene="dam_break", **kw):
    cfg = O.OracleConfig(res=(n, n, n), dx=1.0 / n, ppc=ppc, pcg_tol=1e-7, pcg_max_iter=400,
                         droplets=False, **kw)
    scenes = load_pkg().scenes
    solid = None
    if scene == "dam_break":
        pos, vel, pt = scenes.dam_break(cfg)
    else:
        pos, vel, pt, solid = scenes.wave_vs_tower(cfg)
    sim = cs.CpuStepSim(cfg, O.derived_densities(cfg))
    sim.set_particles(pos, vel, pt)
    if solid is not None:
        sim.set_solids(solid)
    return cfg, sim, solid, pt


@pytest.mark.parametrize("n,ppc,scene", [(16, 8, "dam_break"), (12, 27, "dam_break"),
                                         (16, 8, "wave_vs_tower")])
def test_cpu_step_matches_oracle(cs, n, ppc, scene):
    cs.set_threads(4)
    cfg, sim, solid, pt = make(cs, n, ppc, scene)
    for s in range(3):
        osim = O.OracleSim(cfg)
        osim.set_particles(sim.get("px"), sim.get("pv"), pt)
        if solid is not None:
            osim.set_solids(solid)
        dt = sim.dt_from_cfl()
        assert abs(dt - osim.dt_from_cfl()) <= 1e-6 * dt
        rec = sim.step(dt)
        osim.step(dt)
        band = np.abs(osim.phi_cell - 0.5) < 1e-4
        assert np.array_equal(sim.get("is_liquid")[~band], osim.is_liquid[~band])
        assert rel_err(sim.get("pressure"), osim.pressure) < 1e-4, s
        u_scale = max(float(np.abs(v).max()) for v in osim.vel_star) + 9.81 * dt
        for ax in range(3):
            assert np.abs(sim.get(f"vel{ax}") - osim.vel[ax]).max() < 3e-4 * u_scale, (s, ax)
        assert np.abs(sim.get("px") - osim.px).max() < 3e-4 * dt * u_scale + 1e-6
        assert np.abs(sim.get("pv") - osim.pv).max() < 3e-4 * u_scale
        assert rec["pcg_res"] < cfg.pcg_tol
        assert abs(rec["pcg_iter"] - osim.stats[-1]["pcg_iter"]) <= 1


def test_cpu_step_thread_count_changes_only_atomic_order(cs):
    out = []
    for threads in (1, 4):
        cs.set_threads(threads)
        cfg, sim, _, _ = make(cs, 16, 8)
        for _ in range(2):
            sim.step()
        out.append((sim.get("px"), sim.get("pressure")))
    assert np.abs(out[0][0] - out[1][0]).max() < 1e-4 * cfg.dx
    assert rel_err(out[1][1], out[0][1]) < 1e-4
