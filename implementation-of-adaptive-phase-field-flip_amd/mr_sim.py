"""Adaptive PF-FLIP simulator on an MSBG (drop-in for pfflip/mr_sim.py).

The dual particle-grid adaptivity loop of paper 5.4 (mr_sim.py:1-17):

 1. block refinement map ``M[i] = min(l_p, M[i])`` over the blocks each
    particle's CFL-enlarged box touches
 2. 2:1 grading (host, ``MSBG.build``)
 3. (re)allocation of the cell channels and the solver
 4. P2G on the multiresolution grid
 5. phase field, face coefficients, forces, pressure projection
 6. per-particle refinement / probabilistic coarsening from the
    block-resolution distance to the interface
 7. G2P (before 6, as in the reference) and RK3 advection

Every stage runs in the HIP kernels of ``csrc/pf_mr.hip`` and
``csrc/pf_msbg.hip``; the host grades the block map, builds the cell
allocation and drives the flexible PCG (``MRSolver``).  There is no CPU
fallback.
"""

from __future__ import annotations

import ctypes
import math
import time
import warnings

import numpy as np
import torch

from . import _lib
from .config import Config
from .mr_solver import MRSolver
from .msbg import MSBG
from .particles import Particles

_MAX_DIST = 64.0
_ACC_CHANNELS = 8


def _pow2_at_most(x: float) -> float:
    """Largest power of two not above ``x`` (x > 0)."""
    return 2.0 ** math.floor(math.log2(x))


def noise_margin_sigma(cfg: Config):
    """Margin, in standard deviations of particle noise, above the phi = 0.5
    tagging threshold under probabilistic coarsening (mr_sim.py:73-101):
    phi < 0.5 needs the weight sum below alpha_phi / 4 of nominal, with
    relative noise 1/sqrt(ppc).  None when no deficit can tag air."""
    deficit = 1.0 - cfg.alpha_phi / 4.0
    if deficit <= 0.0:
        return None
    return deficit * math.sqrt(cfg.ppc)


class AdaptiveSimulator:
    """mr_sim.py:40-310.  Warns (RuntimeWarning) when the noise margin is
    below 2.5 sigma."""

    def __init__(self, cfg: Config, nblocks, n0=16, max_level=3, beta_bw=1.0, frozen_levels=None, resample=True):
        self.cfg = cfg
        self.dev = torch.device(cfg.device)
        self.beta_bw = float(beta_bw)
        self.frozen_levels = None if frozen_levels is None else np.asarray(frozen_levels, np.int32).ravel()
        self.do_resample = bool(resample)
        self.defer_threshold = 3
        # finest spacing h0 fixes the domain: nblocks * n0 cells per axis
        self.grid = MSBG(nblocks, n0=n0, h0=cfg.dx, device=cfg.device, max_level=max_level)
        self.max_level = self.grid.max_level
        self.block_extent = n0 * cfg.dx
        self.domain = tuple(n * self.block_extent for n in self.grid.nb)
        self.particles = Particles(cfg.device, capacity=1024)
        # eq. (7) normalisation, calibrated once at level 0
        self.rho_tilde_0 = self._calibrate()
        self.rho_0l = self.rho_tilde_0 * cfg.rho_l
        self.rho_min = cfg.eta_phi * cfg.rho_g * self.rho_tilde_0
        self._check_noise_margin()
        self._cap = 0
        self._level_sig = None
        self._acc_dirty = False
        nblk = self.grid.nblk
        i32 = lambda m: torch.zeros(m, dtype=torch.int32, device=self.dev)  # noqa: E731
        self._level_map = i32(nblk)
        self._mixed = i32(nblk)
        self._edt_tmp = i32(2 * nblk)
        self._dist = torch.zeros(nblk, dtype=torch.float32, device=self.dev)
        self._scalars = torch.zeros(2, dtype=torch.float32, device=self.dev)
        # P2G over a cell-sorted order with per-workgroup LDS caches (same
        # sums as the per-particle atomics, bit for bit)
        self.p2g_sorted = True
        self._p2g_scratch = torch.zeros(0, dtype=torch.int32, device=self.dev)
        self.time = 0.0
        self.step_index = 0
        self.stats: list[dict] = []

    # ------------------------------------------------------------------ setup
    def _check_noise_margin(self):
        sigma = noise_margin_sigma(self.cfg)
        if sigma is not None and sigma < 2.5:
            cfg = self.cfg
            warnings.warn(
                f"ppc={cfg.ppc} with alpha_phi={cfg.alpha_phi} leaves only {sigma:.1f} sigma of margin above the "
                "phi=0.5 phase-tagging threshold under probabilistic coarsening: bulk liquid cells will be "
                "mis-tagged as air, flipping their density by the density ratio. Use ppc >= 27, or "
                "alpha_phi <= 0.25.", RuntimeWarning, stacklevel=3)

    def _calibrate(self) -> float:
        """rho~_0: the eq. (6) weight sum of one cell of a level-0 lattice of
        ``ppc`` particles (mr_sim.py:104-118)."""
        cfg = self.cfg
        per_axis = round(cfg.ppc ** (1.0 / 3.0))
        step = cfg.dx / per_axis
        r = cfg.r_p
        R = int(math.ceil(r / step)) + 1
        a = np.arange(-R, R + 1) + 0.5
        d2 = (a[:, None, None] ** 2 + a[None, :, None] ** 2 + a[None, None, :] ** 2) * step * step
        t = np.maximum(1.0 - d2 / (r * r), 0.0)
        return float((t * t * t).sum())

    def set_particles(self, pos, vel, ptype, level=None):
        if level is not None and len(level) and int(np.max(level)) > self.max_level:
            raise ValueError(f"particle levels above max_level={self.max_level}")
        self.particles.set(pos, vel, ptype, level)

    # ------------------------------------------------------------------ channels
    def _ensure_channels(self):
        """(Re)allocate channels when the cell count outgrows them and rebuild
        the solver when the refinement map changes (mr_sim.py:121-148)."""
        n = self.grid.ncells
        sig = hash(self.grid.level_h.tobytes())
        if n > self._cap:
            cap = max(n, int(self._cap * 1.25))
            z = lambda m: torch.zeros(m, dtype=torch.float32, device=self.dev)  # noqa: E731
            self._acc = torch.zeros(_ACC_CHANNELS * cap, dtype=torch.int64, device=self.dev)
            self._fmass, self._fvel, self._fvel_star, self._fdvel, self._fbeta = (z(3 * cap) for _ in range(5))
            self._cwsum, self._phi, self._pressure, self._rhs = (z(cap) for _ in range(4))
            self._is_liquid = torch.zeros(cap, dtype=torch.int32, device=self.dev)
            self._cap = cap
            self._acc_dirty = False
            self._level_sig = None
        if sig != self._level_sig:
            self._level_sig = sig
            # release the old hierarchy first: its memory then serves the new
            # one from the allocator's cache instead of fresh device mappings
            self.solver = None
            self.solver = MRSolver(self.grid, preconditioner="mg")
        f = _lib.MrFields()
        f.acc = self._acc.data_ptr()
        for k in ("cwsum", "phi", "is_liquid", "fmass", "fvel", "fvel_star", "fdvel", "fbeta", "rhs", "pressure"):
            setattr(f, k, getattr(self, "_" + k).data_ptr())
        self._fields = f

    # views of the live cells (the reference's channel names)
    def _view(name, width):  # noqa: N805
        return property(lambda self: getattr(self, "_" + name)[: width * self.grid.ncells])

    fmass = _view("fmass", 3)
    fvel = _view("fvel", 3)
    fvel_star = _view("fvel_star", 3)
    fdvel = _view("fdvel", 3)
    fbeta = _view("fbeta", 3)
    cwsum = _view("cwsum", 1)
    phi = _view("phi", 1)
    pressure = _view("pressure", 1)
    rhs = _view("rhs", 1)
    is_liquid = _view("is_liquid", 1)
    del _view

    # ------------------------------------------------------------------ steps 1-3
    def _rebuild_grid(self, dt_hint: float = 0.0):
        """Refinement map, grading, allocation (mr_sim.py:151-165)."""
        if self.frozen_levels is not None:
            self.grid.build(self.frozen_levels)
        else:
            cfl_eps = max(self.cfg.cfl, 1.0) * self.cfg.dx
            ps = self.particles
            src = ps.struct()
            nbx, nby, nbz = self.grid.nb
            _lib.call("pf_mr_refine_map", ctypes.byref(src), _lib.ptr(self._level_map), nbx, nby, nbz,
                      float(self.block_extent), float(cfl_eps), int(self.max_level), _lib.stream())
            self.grid.build(self._level_map.cpu().numpy())  # build() applies the 2:1 grading
        self._ensure_channels()

    # ------------------------------------------------------------------ params
    def _params(self, dt: float, vmax: float) -> _lib.MrParams:
        cfg = self.cfg
        p = _lib.MrParams()
        p.r0, p.rho_l, p.rho_g = cfg.r_p, cfg.rho_l, cfg.rho_g
        p.rho_min, p.alpha_phi, p.rho_0l = self.rho_min, cfg.alpha_phi, self.rho_0l
        p.rho_tilde_0, p.k_div = self.rho_tilde_0, cfg.k_div
        p.gravity[:] = [float(g) for g in cfg.gravity]
        p.dt = dt
        p.nu_l = (1.0 - cfg.alpha_flip_liquid) * cfg.dx ** 2 / (6.0 * dt)
        p.nu_g = (1.0 - cfg.alpha_flip_air) * cfg.dx ** 2 / (6.0 * dt)
        p.alpha_cap = 0.999
        eps = 1e-4 * cfg.dx
        p.lo[:] = [eps] * 3
        p.hi[:] = [s - eps for s in self.domain]
        p.max_level = self.max_level
        # fixed-point scales: the largest single contribution maps below 2^48,
        # leaving 2^15 of them per accumulator before the int64 range ends
        m_max = max(cfg.rho_l, cfg.rho_g) * 8.0 ** self.max_level
        p.mass_scale = 2.0 ** 48 / (2.0 * _pow2_at_most(m_max))
        p.mom_scale = 2.0 ** 48 / (2.0 * _pow2_at_most(m_max * vmax)) if vmax > 0 else p.mass_scale
        p.w_scale = 2.0 ** 48
        return p

    def _max_speed(self) -> float:
        ps = self.particles
        _lib.call("pf_mr_max_speed", _lib.ptr(ps.v), ps.n, _lib.ptr(self._scalars), _lib.stream())
        return float(self._scalars[0].item())

    def dt_from_cfl(self, vmax: float | None = None) -> float:
        """mr_sim.py:193-201."""
        cfg = self.cfg
        vmax = self._max_speed() if vmax is None else vmax
        g = float(np.linalg.norm(cfg.gravity))
        vmax = max(vmax, math.sqrt(2.0 * g * cfg.dx) if g > 0 else 0.0, 1e-8)
        return float(np.clip(cfg.cfl * cfg.dx / vmax, cfg.dt_min, cfg.dt_max))

    # ------------------------------------------------------------------ stages
    def _distance_field(self):
        """Block-resolution distance to the interface blocks (mr_sim.py:163-182)."""
        _lib.call("pf_mr_block_distance", self.grid.ref(), _lib.ptr(self._is_liquid), _lib.ptr(self._mixed),
                  _lib.ptr(self._edt_tmp), _lib.ptr(self._dist), _MAX_DIST, _lib.stream())
        return self._dist

    def stage_p2g(self, prm):
        """Steps 4-5a (mr_sim.py:216-231): P2G, then phase field, face mass, u*."""
        g, f, s = self.grid.ref(), self._fields, _lib.stream()
        if self._acc_dirty:  # an earlier step stopped between P2G and finish
            self._acc.zero_()
        src = self.particles.struct()
        self._acc_dirty = True
        if self.p2g_sorted:
            need = int(_lib.load().pf_mr_p2g_scratch_ints(self.particles.n, self.grid.ncells))
            if self._p2g_scratch.numel() < need:
                self._p2g_scratch = torch.empty(int(need * 1.25), dtype=torch.int32, device=self.dev)
            _lib.call("pf_mr_p2g_sorted", g, ctypes.byref(src), ctypes.byref(prm), ctypes.byref(f),
                      _lib.ptr(self._p2g_scratch), s)
        else:
            _lib.call("pf_mr_p2g", g, ctypes.byref(src), ctypes.byref(prm), ctypes.byref(f), s)
        _lib.call("pf_mr_finish", g, ctypes.byref(prm), ctypes.byref(f), s)
        self._acc_dirty = False

    def stage_forces(self, prm):
        """mr_sim.py:232-250: beta, walls, u*, gravity, right-hand side."""
        g, f, s = self.grid.ref(), self._fields, _lib.stream()
        _lib.call("pf_mr_face_prepare", g, ctypes.byref(prm), ctypes.byref(f), s)
        _lib.call("pf_mr_rhs", g, ctypes.byref(prm), ctypes.byref(f), s)

    def stage_solve(self):
        """mr_sim.py:251-252 -> (iterations, relative residual)."""
        cfg = self.cfg
        return self.solver.solve(self.fbeta, self.rhs, self.pressure, tol=cfg.pcg_tol, max_iter=cfg.pcg_max_iter)

    def stage_project(self, prm):
        """mr_sim.py:253-261 and the du of 264-266 -> max |div u|."""
        g, f, s = self.grid.ref(), self._fields, _lib.stream()
        _lib.call("pf_mr_project", g, ctypes.byref(prm), ctypes.byref(f), s)
        _lib.call("pf_mr_max_divergence", g, _lib.ptr(self._fvel), _lib.ptr(self._scalars[1:]), s)

    def stage_g2p(self, prm):
        """mr_sim.py:262-273: per-particle alpha_FLIP and G2P."""
        src = self.particles.struct()
        _lib.call("pf_mr_g2p", self.grid.ref(), ctypes.byref(src), ctypes.byref(prm), ctypes.byref(self._fields),
                  _lib.stream())

    def stage_resample(self):
        """Step 6 (mr_sim.py:280-289) -> (n_before, n_after)."""
        ps = self.particles
        if not self.do_resample:
            return ps.n, ps.n
        dist = self._distance_field()
        return ps.resample(dist, self.grid.nb, self.block_extent, beta_bw=self.beta_bw, max_level=self.max_level,
                           r0=self.cfg.r_p, defer_threshold=self.defer_threshold, seed=self.step_index)

    def stage_advect(self, prm):
        """Step 7 (mr_sim.py:291-298)."""
        src = self.particles.struct()
        _lib.call("pf_mr_advect", self.grid.ref(), ctypes.byref(src), ctypes.byref(prm), _lib.ptr(self._fvel),
                  _lib.stream())

    # ------------------------------------------------------------------ main
    def step(self, dt: float | None = None, verbose=False) -> dict:
        """mr_sim.py:203-310."""
        ps = self.particles
        vmax = self._max_speed()
        if dt is None:
            dt = self.dt_from_cfl(vmax)
        t0 = time.perf_counter()
        self._rebuild_grid(dt)  # (1-3) particles drive the grid
        prm = self._params(dt, vmax)
        ncell = self.grid.ncells
        self.stage_p2g(prm)  # (4-5)
        self.stage_forces(prm)
        it, rel = self.stage_solve()
        self.stage_project(prm)
        self.stage_g2p(prm)
        n_before, n_after = self.stage_resample()  # (6)
        self.stage_advect(prm)  # (7)
        div = float(self._scalars[1].item())
        torch.cuda.synchronize(self.dev)
        self.time += dt
        self.step_index += 1
        rec = {
            "step": self.step_index, "t": self.time, "dt": dt, "cells": ncell,
            "compression": self.grid.compression, "particles": ps.n, "dn": n_after - n_before,
            "cg_iter": it, "cg_res": rel, "max_div": div, "levels": self.grid.level_histogram(),
            "wall_s": time.perf_counter() - t0,
        }
        if verbose:
            print(f"step {self.step_index:4d}  dt {dt:.2e}  cells {ncell:,}  particles {ps.n:,}  "
                  f"cg {it} ({rel:.1e})  div {div:.2e}")
        self.stats.append(rec)
        return rec
