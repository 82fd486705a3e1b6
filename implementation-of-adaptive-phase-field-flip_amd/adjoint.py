"""Reverse-mode differentiation of PF-FLIP steps (hand-written adjoint).

The reference has no adjoint (its Warp kernels are never recorded on a
tape).  ``rollout`` runs ``Simulator.step`` ``n`` times inside a
``torch.autograd.Function``: the forward pass keeps one checkpoint of the
particle state (x, v) and the time step per step; the backward pass walks
the steps in reverse, rebuilds each step's grid fields from its checkpoint
(binning, P2G, eq. (7) / beta / forces / RHS, MG-PCG, projection,
extrapolation: the same kernels as the forward step, so the same bits) and
applies the transposed stages in libpfflip_hip.so:

    G2P + RK3 + reflect   pf_adj_g2p_advect   (per particle; scatter to the grid)
    extrapolation         pf_adj_extrapolate  (transposed Jacobi sweeps)
    projection            pf_adj_project
    pressure solve        MG-PCG on p-bar (the operator is symmetric), then
                          pf_adj_solve_beta
    RHS / forces / eq. 7  pf_adj_phase        (one-sided subgradients at clamps)
    P2G                   pf_adj_p2g          (per particle; gather from the grid)

dt is held at its forward value (the CFL step size is not differentiated),
and so are the liquid tags, wall / solid masks and extrapolation masks; the
solid push-out passes x-bar and zeroes the v-bar component it zeroes.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from .sim import Simulator

_ALL_PHASE = (_lib.PH_VEL_FROM_MOM | _lib.PH_PHASE | _lib.PH_ZERO | _lib.PH_STAR
              | _lib.PH_GRAVITY | _lib.PH_RHS)


class StepAdjoint:
    """Adjoint buffers of one simulator and the backward pass of one step."""

    def __init__(self, sim: Simulator, tol: float = 1e-7, max_iter: int = 400):
        if sim.cfg.k_div > 0.0:
            raise NotImplementedError("the adjoint does not cover k_div > 0")
        self.sim = sim
        self.tol, self.max_iter = tol, max_iter
        dev = sim.dev
        z = lambda shape: torch.zeros(shape, dtype=torch.float32, device=dev)  # noqa: E731
        fs = sim.face_shapes
        self.g_vel = [z(s) for s in fs]
        self.g_dvel = [z(s) for s in fs]
        self.g_vel_liq = [z(s) for s in fs]
        self.g_beta = [z(s) for s in fs]
        self.g_mass = [z(s) for s in fs]
        self.g_mom = [z(s) for s in fs]
        self.level = [torch.zeros(s, dtype=torch.uint8, device=dev) for s in fs]
        self.g_p = z(sim.res)
        self.lam = z(sim.res)
        self.last_solve = (0, 0.0)

    # ------------------------------------------------------------------ forward rebuild
    def rebuild(self, x: torch.Tensor, v: torch.Tensor, dt: float) -> None:
        """Grid fields of the step that starts at (x, v); particles stay put."""
        sim = self.sim
        cfg = sim.cfg
        sim.px.copy_(x)
        sim.pv.copy_(v)
        sim._particle_bounds()
        sim._bin()
        sim._splat()
        sim._phase(_ALL_PHASE, dt)
        sim.solver.build_hierarchy(*sim.beta)
        sim.solver.solve(sim.rhs, sim.pressure, tol=cfg.pcg_tol, max_iter=cfg.pcg_max_iter)
        nx, ny, nz = sim.res
        _lib.call("pf_project", _lib.ptr(sim.pressure), _lib.ptr3(sim.beta), _lib.ptr3(sim.vel),
                  _lib.ptr3(sim.vel_star), _lib.ptr3(sim.dvel), nx, ny, nz, cfg.dx, float(dt),
                  None, _lib.stream())
        sim._extrapolate_liquid_velocity()

    # ------------------------------------------------------------------ backward
    def backward(self, x, v, dt, gx_out, gv_out):
        """(x, v)-bar of the step from (x, v) with time step dt, given the
        adjoints of its output (x', v')."""
        sim = self.sim
        cfg = sim.cfg
        nx, ny, nz = sim.res
        self.rebuild(x, v, dt)
        for t in self.g_vel + self.g_dvel + self.g_vel_liq:
            t.zero_()
        gx_in = torch.empty_like(x)
        gv_in = torch.empty_like(v)
        a = sim._g2p_args(_lib.GA_G2P | _lib.GA_ADVECT | _lib.GA_REFLECT, dt)
        _lib.call("pf_adj_g2p_advect", ctypes.byref(a), _lib.ptr(gx_out.contiguous()),
                  _lib.ptr(gv_out.contiguous()), _lib.ptr(gx_in), _lib.ptr(gv_in),
                  _lib.ptr3(self.g_vel), _lib.ptr3(self.g_dvel), _lib.ptr3(self.g_vel_liq), _lib.stream())
        _lib.call("pf_adj_extrapolate", _lib.ptr(sim.is_liquid), _lib.ptr3(self.g_vel_liq),
                  _lib.ptr3(self.g_vel), _lib.ptr3(self.level), nx, ny, nz, int(cfg.extrap_cells),
                  _lib.stream())
        _lib.call("pf_adj_project", _lib.ptr(sim.pressure), _lib.ptr3(sim.beta), _lib.ptr3(self.g_vel),
                  _lib.ptr3(self.g_dvel), _lib.ptr3(self.g_beta), _lib.ptr(self.g_p), nx, ny, nz,
                  cfg.dx, float(dt), _lib.stream())
        # lam = A^+ p-bar: the same operator (symmetric), mean removed
        self.last_solve = sim.solver.solve(self.g_p, self.lam, tol=self.tol, max_iter=self.max_iter)
        _lib.call("pf_adj_solve_beta", _lib.ptr(sim.pressure), _lib.ptr(self.lam), _lib.ptr3(self.g_beta),
                  nx, ny, nz, _lib.stream())
        pa = sim._phase_args(_ALL_PHASE, dt)
        _lib.call("pf_adj_phase", ctypes.byref(pa), _lib.ptr3(self.g_vel), _lib.ptr3(self.g_dvel),
                  _lib.ptr3(self.g_beta), _lib.ptr(self.lam), _lib.ptr3(self.g_mass),
                  _lib.ptr3(self.g_mom), _lib.stream())
        _lib.call("pf_adj_p2g", _lib.ptr(x), _lib.ptr(v), _lib.ptr(sim.pmass), _lib.ptr(sim.pstate),
                  sim.n_particles, _lib.ptr3(self.g_mass), _lib.ptr3(self.g_mom), nx, ny, nz, cfg.dx,
                  cfg.r_p, _lib.ptr(gx_in), _lib.ptr(gv_in), _lib.stream())
        return gx_in, gv_in


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, sim, n_steps, dts, adj):
        sim.px.copy_(x0.detach())
        sim.pv.copy_(v0.detach())
        ckpt, used = [], []
        for k in range(n_steps):
            ckpt.append((sim.px.clone(), sim.pv.clone()))
            rec = sim.step(None if dts is None else float(dts[k]))
            used.append(rec["dt"])
        ctx.sim, ctx.ckpt, ctx.dts, ctx.adj = sim, ckpt, used, adj
        return sim.px.clone(), sim.pv.clone()

    @staticmethod
    def backward(ctx, gx, gv):
        sim = ctx.sim
        gx = torch.zeros_like(sim.px) if gx is None else gx.contiguous().float()
        gv = torch.zeros_like(sim.pv) if gv is None else gv.contiguous().float()
        adj = ctx.adj if ctx.adj is not None else StepAdjoint(sim)
        for k in reversed(range(len(ctx.ckpt))):
            x, v = ctx.ckpt[k]
            gx, gv = adj.backward(x, v, ctx.dts[k], gx, gv)
        return gx, gv, None, None, None, None


def rollout(sim: Simulator, x0: torch.Tensor, v0: torch.Tensor, n_steps: int, dts=None,
            adjoint: StepAdjoint | None = None):
    """Run ``n_steps`` steps from particle state (x0, v0) and return
    (x_T, v_T), differentiable with respect to x0 and v0.  ``dts`` fixes the
    time steps (default: CFL per step, then held fixed in the backward
    pass); the ones used are in ``sim.stats``."""
    if x0.shape != sim.px.shape or v0.shape != sim.pv.shape:
        raise ValueError("x0 / v0 must match the simulator's particle arrays")
    return _Rollout.apply(x0, v0, sim, int(n_steps), dts, adjoint)
