"""The hand-written adjoint of the PF-FLIP step (MI355X).

Linear stages are checked by the transpose identity <J u, w> = <u, J^T w>
against the forward kernels themselves; the whole step and a short rollout
by central finite differences of a scalar loss along random directions
(dt fixed, FP32 forward, MG-PCG at a tight tolerance so that the solve's
stopping point does not move with the perturbation).
"""

import ctypes
import sys

import numpy as np
import pytest

from conftest import gpu_available, load_pkg

pytestmark = pytest.mark.gpu

if gpu_available():
    import torch


def say(msg):
    print(f"[gpu-adjoint] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pf():
    if not gpu_available():
        pytest.skip("no GPU")
    return load_pkg()


def _lib(pf):
    return sys.modules[pf.__name__ + "._lib"]


def make(pf, n=16, ppc=8, tol=1e-8, seed=0):
    cfg = pf.Config(res=(n, n, n), dx=1.0 / n, ppc=ppc, droplets=False, gravity=(0.0, 0.0, -9.81),
                    pcg_tol=tol, pcg_max_iter=400)
    pos, vel, pt = pf.scenes.dam_break(cfg, frac=(0.4, 1.0, 0.6), jitter=0.5)
    rng = np.random.default_rng(seed)
    vel = (0.3 * rng.standard_normal(pos.shape)).astype(np.float32) * (pt[:, None] == 1)
    sim = pf.Simulator(cfg)
    sim.set_particles(pos, vel, pt)
    return cfg, sim


def dot(a, b):
    return float((a.double() * b.double()).sum())


def test_p2g_momentum_transpose(pf):
    """mom(v) is linear in v: <mom(dv), w> == <dv, P2G^T w>."""
    say("adjoint P2G transpose")
    L = _lib(pf)
    cfg, sim = make(pf)
    adj = pf.StepAdjoint(sim)
    rng = np.random.default_rng(1)
    dv = torch.from_numpy(rng.standard_normal((sim.n_particles, 3)).astype(np.float32)).cuda()
    sim.pv.copy_(dv)
    sim._particle_bounds()
    sim._bin()
    sim._splat()
    w = [torch.from_numpy(rng.standard_normal(tuple(s)).astype(np.float32)).cuda() for s in sim.face_shapes]
    lhs = sum(dot(sim.mom[a], w[a]) for a in range(3))
    zero = [torch.zeros_like(t) for t in w]
    gx = torch.zeros_like(sim.px)
    gv = torch.zeros_like(sim.pv)
    nx, ny, nz = sim.res
    L.call("pf_adj_p2g", L.ptr(sim.px), L.ptr(sim.pv), L.ptr(sim.pmass), L.ptr(sim.pstate), sim.n_particles,
           L.ptr3(zero), L.ptr3(w), nx, ny, nz, cfg.dx, cfg.r_p, L.ptr(gx), L.ptr(gv), L.stream())
    rhs = dot(dv, gv)
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs), (lhs, rhs)
    del adj


def test_extrapolation_transpose(pf):
    say("adjoint extrapolation transpose")
    L = _lib(pf)
    cfg, sim = make(pf)
    adj = pf.StepAdjoint(sim)
    adj.rebuild(sim.px.clone(), sim.pv.clone(), sim.dt_from_cfl())
    rng = np.random.default_rng(2)
    u = [torch.from_numpy(rng.standard_normal(tuple(s)).astype(np.float32)).cuda() for s in sim.face_shapes]
    w = [torch.from_numpy(rng.standard_normal(tuple(s)).astype(np.float32)).cuda() for s in sim.face_shapes]
    for a in range(3):
        sim.vel[a].copy_(u[a])
    sim._extrapolate_liquid_velocity()
    lhs = sum(dot(sim.vel_liq[a], w[a]) for a in range(3))
    g = [torch.zeros_like(t) for t in w]
    wl = [t.clone() for t in w]
    nx, ny, nz = sim.res
    L.call("pf_adj_extrapolate", L.ptr(sim.is_liquid), L.ptr3(wl), L.ptr3(g), L.ptr3(adj.level),
           nx, ny, nz, int(cfg.extrap_cells), L.stream())
    rhs = sum(dot(u[a], g[a]) for a in range(3))
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), 1.0), (lhs, rhs)


def test_projection_transpose_in_p(pf):
    """vel3 = -dt beta grad p / dx (vel2 = 0) is linear in p."""
    say("adjoint projection transpose")
    L = _lib(pf)
    cfg, sim = make(pf)
    adj = pf.StepAdjoint(sim)
    dt = sim.dt_from_cfl()
    adj.rebuild(sim.px.clone(), sim.pv.clone(), dt)
    rng = np.random.default_rng(3)
    p = torch.from_numpy(rng.standard_normal(sim.res).astype(np.float32)).cuda()
    w = [torch.from_numpy(rng.standard_normal(tuple(s)).astype(np.float32)).cuda() for s in sim.face_shapes]
    nx, ny, nz = sim.res
    for a in range(3):
        sim.vel[a].zero_()
    L.call("pf_project", L.ptr(p), L.ptr3(sim.beta), L.ptr3(sim.vel), None, None, nx, ny, nz, cfg.dx,
           float(dt), None, L.stream())
    lhs = sum(dot(sim.vel[a], w[a]) for a in range(3))
    gv = [t.clone() for t in w]
    gd = [torch.zeros_like(t) for t in w]
    L.call("pf_adj_project", L.ptr(p), L.ptr3(sim.beta), L.ptr3(gv), L.ptr3(gd), L.ptr3(adj.g_beta),
           L.ptr(adj.g_p), nx, ny, nz, cfg.dx, float(dt), L.stream())
    rhs = dot(p, adj.g_p)
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs), (lhs, rhs)


def test_g2p_transpose_in_grid(pf):
    """v' = a (v + I(dvel)) + (1 - a) I(vel) is affine in the grid fields."""
    say("adjoint G2P transpose")
    L = _lib(pf)
    cfg, sim = make(pf)
    adj = pf.StepAdjoint(sim)
    rng = np.random.default_rng(4)
    fields = [[torch.from_numpy(rng.standard_normal(tuple(s)).astype(np.float32)).cuda()
               for s in sim.face_shapes] for _ in range(2)]
    v0 = sim.pv.clone()
    x0 = sim.px.clone()
    wv = torch.from_numpy(rng.standard_normal((sim.n_particles, 3)).astype(np.float32)).cuda()

    def run(vel, dvel):
        sim.pv.copy_(v0)
        for a in range(3):
            sim.vel[a].copy_(vel[a])
            sim.dvel[a].copy_(dvel[a])
        args = sim._g2p_args(L.GA_G2P)
        L.call("pf_g2p_advect", ctypes.byref(args), L.stream())
        return sim.pv.clone()

    zero = [torch.zeros_like(t) for t in fields[0]]
    lhs = dot(run(fields[0], fields[1]) - run(zero, zero), wv)
    sim.pv.copy_(v0)
    args = sim._g2p_args(L.GA_G2P)
    gx_in, gv_in = torch.empty_like(x0), torch.empty_like(v0)
    for t in adj.g_vel + adj.g_dvel + adj.g_vel_liq:
        t.zero_()
    L.call("pf_adj_g2p_advect", ctypes.byref(args), L.ptr(torch.zeros_like(x0)), L.ptr(wv), L.ptr(gx_in),
           L.ptr(gv_in), L.ptr3(adj.g_vel), L.ptr3(adj.g_dvel), L.ptr3(adj.g_vel_liq), L.stream())
    rhs = sum(dot(fields[0][a], adj.g_vel[a]) + dot(fields[1][a], adj.g_dvel[a]) for a in range(3))
    assert abs(lhs - rhs) <= 1e-4 * abs(lhs), (lhs, rhs)


def _smooth_dir(sim, rng, liquid_only=True):
    """A smooth random per-particle direction: a few low Fourier modes of
    position, max |d| = 1."""
    x = sim.px.double()
    d = torch.zeros_like(x)
    for _ in range(4):
        k = torch.from_numpy(rng.integers(-2, 3, size=3).astype(np.float64)).cuda()
        ph = float(rng.uniform(0, 2 * np.pi))
        amp = torch.from_numpy(rng.standard_normal(3)).cuda()
        d += torch.cos(2 * np.pi * (x @ k) + ph)[:, None] * amp[None, :]
    if liquid_only:
        d *= (sim.ptype == 1).double()[:, None]
    return (d / d.abs().max()).float()


def _loss(sim, x, target):
    liq = sim.ptype == 1
    c = x[liq].double().mean(0)
    return ((c - target) ** 2).sum()


@pytest.mark.parametrize("n_steps", [1, 4])
def test_rollout_matches_finite_differences(pf, n_steps):
    say(f"adjoint vs finite differences, {n_steps} steps")
    cfg, sim = make(pf, n=16, ppc=8, tol=1e-8, seed=5)
    x0, v0 = sim.px.clone(), sim.pv.clone()
    target = torch.tensor([0.6, 0.5, 0.2], dtype=torch.float64, device="cuda")
    # time steps from an unperturbed run, then held fixed
    sim.stats.clear()
    with torch.no_grad():
        pf.rollout(sim, x0, v0, n_steps)
    dts = [r["dt"] for r in sim.stats[-n_steps:]]
    v = v0.clone().requires_grad_(True)
    xT, _ = pf.rollout(sim, x0, v, n_steps, dts=dts)
    loss = _loss(sim, xT, target)
    loss.backward()
    g = v.grad.detach()
    rng = np.random.default_rng(6)
    eps = 1e-3
    errs = []
    for _ in range(3):
        sim.px.copy_(x0)
        d = _smooth_dir(sim, rng)
        with torch.no_grad():
            lp = float(_loss(sim, pf.rollout(sim, x0, v0 + eps * d, n_steps, dts=dts)[0], target))
            lm = float(_loss(sim, pf.rollout(sim, x0, v0 - eps * d, n_steps, dts=dts)[0], target))
        fd = (lp - lm) / (2 * eps)
        ad = dot(g, d)
        errs.append(abs(fd - ad) / max(abs(fd), 1e-12))
        say(f"  fd={fd:.6e} ad={ad:.6e} rel={errs[-1]:.2e}")
    # measured: at most 1.4e-4 over these directions
    assert max(errs) < 1e-3, errs


def test_rollout_with_solids_matches_finite_differences(pf):
    """The surge-on-a-column scene: solid faces, the push-out of particles
    entering the column, and the adjoint's treatment of both.  A push-out is
    a jump in position, so a central difference straddling one carries an
    O(1) error for that particle; measured errors are 3e-3 - 8e-3, against
    the 1e-2 bound."""
    say("adjoint with a solid column, 2 steps")
    n = 24
    cfg = pf.Config(res=(n, n, n), dx=1.0 / n, ppc=8, droplets=False, gravity=(0.0, 0.0, -9.81),
                    pcg_tol=1e-8, pcg_max_iter=400)
    pos, vel, pt, solid = pf.scenes.surge_column(cfg, width=0.2)
    sim = pf.Simulator(cfg)
    sim.set_solids(solid)
    sim.set_particles(pos, vel, pt)
    assert sim.has_solids
    x0, v0 = sim.px.clone(), sim.pv.clone()
    target = torch.tensor([0.6, 0.5, 0.2], dtype=torch.float64, device="cuda")
    sim.stats.clear()
    with torch.no_grad():
        pf.rollout(sim, x0, v0, 2)
    dts = [r["dt"] for r in sim.stats[-2:]]
    v = v0.clone().requires_grad_(True)
    xT, _ = pf.rollout(sim, x0, v, 2, dts=dts)
    _loss(sim, xT, target).backward()
    g = v.grad.detach()
    rng = np.random.default_rng(9)
    eps = 1e-3
    errs = []
    for _ in range(3):
        sim.px.copy_(x0)
        d = _smooth_dir(sim, rng)
        with torch.no_grad():
            lp = float(_loss(sim, pf.rollout(sim, x0, v0 + eps * d, 2, dts=dts)[0], target))
            lm = float(_loss(sim, pf.rollout(sim, x0, v0 - eps * d, 2, dts=dts)[0], target))
        fd = (lp - lm) / (2 * eps)
        ad = dot(g, d)
        errs.append(abs(fd - ad) / max(abs(fd), 1e-12))
        say(f"  fd={fd:.6e} ad={ad:.6e} rel={errs[-1]:.2e}")
    assert max(errs) < 1e-2, errs
