"""Adaptive particles and the dual particle-grid resample (drop-in for
pfflip/particles.py:279-384).

Each particle carries a level ``l_p``: radius ``r0 2**l_p``, mass
``rho 8**l_p`` (level 0 is finest, as in the MSBG).  ``resample`` runs the
classify / prefix-sum / emit kernels of ``csrc/pf_mr.hip``: refine (split
into 8 children one level finer), keep, or coarsen with deferred
probabilistic deletion (paper 5.1, 5.4 step 6).  Positions and velocities
are ``(capacity, 3)`` float32 device tensors; ``n`` of them are live.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .poisson import _require_cuda


class Particles:
    """Growable particle set with the resample step (particles.py:279-384)."""

    def __init__(self, device="cuda:0", capacity: int = 1024):
        self.dev = _require_cuda(device)
        self.n = 0
        self.capacity = 0
        self._alloc(max(int(capacity), 1))

    # ------------------------------------------------------------------ storage
    def _alloc(self, cap, keep=0):
        d = self.dev
        old = {k: getattr(self, k) for k in ("x", "v", "type", "level", "defer")} if keep else {}
        self.x = torch.zeros((cap, 3), dtype=torch.float32, device=d)
        self.v = torch.zeros((cap, 3), dtype=torch.float32, device=d)
        self.type, self.level, self.defer = (torch.zeros(cap, dtype=torch.int32, device=d) for _ in range(3))
        for k, a in old.items():
            getattr(self, k)[:keep].copy_(a[:keep])
        self._out = None  # resample's scatter side, sized on demand
        self.capacity = cap

    def ensure_capacity(self, need):
        if need > self.capacity:
            self._alloc(max(int(need * 1.3), self.capacity * 2), keep=self.n)

    def set(self, pos, vel, ptype, level=None):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        n = len(pos)
        if n >= 2**31 // 8:
            raise ValueError("at most 2^28 particles (a split multiplies the count by up to 8)")
        lv = np.zeros(n, np.int32) if level is None else np.ascontiguousarray(level, np.int32)
        if lv.shape != (n,) or (n and lv.min() < 0):
            raise ValueError("levels must be n non-negative integers")
        self.n = 0
        self.ensure_capacity(n)
        self.n = n
        t = lambda a: torch.from_numpy(a).to(self.dev)  # noqa: E731
        self.x[:n].copy_(t(pos))
        self.v[:n].copy_(t(np.ascontiguousarray(vel, np.float32).reshape(n, 3)))
        self.type[:n].copy_(t(np.ascontiguousarray(ptype, np.int32).reshape(n)))
        self.level[:n].copy_(t(lv))
        self.defer[:n].zero_()
        self._max_level = int(lv.max()) if n else 0

    def struct(self, n=None) -> _lib.MrParticles:
        s = _lib.MrParticles()
        s.x, s.v, s.type = self.x.data_ptr(), self.v.data_ptr(), self.type.data_ptr()
        s.level, s.defer = self.level.data_ptr(), self.defer.data_ptr()
        s.n = self.n if n is None else int(n)
        return s

    # ------------------------------------------------------------------ derived
    def mass_radius(self, r0, rho_l, rho_g):
        """(mass, radius) of the live particles (particles.py:234-248); the
        P2G kernel derives the same values itself."""
        lv = self.level[: self.n]
        liquid = self.type[: self.n] == 1
        rho = torch.where(liquid, torch.tensor(rho_l, device=self.dev), torch.tensor(rho_g, device=self.dev))
        return rho * torch.pow(8.0, lv.float()), r0 * torch.pow(2.0, lv.float())

    # ------------------------------------------------------------------ resample
    def resample(self, dist, nblocks, block_extent, *, beta_bw, max_level, r0, defer_threshold=3, seed=0):
        """Classify, prefix-sum and emit (particles.py:331-372).  ``dist`` is
        the block-resolution distance field (a device float tensor of
        ``nblocks`` entries).  Returns ``(n_before, n_after)``."""
        n = self.n
        if n == 0:
            return 0, 0
        d = self.dev
        i32 = lambda m: torch.empty(m, dtype=torch.int32, device=d)  # noqa: E731
        lnew, nout, offs = i32(n), i32(n), i32(n + 1)
        tmp = i32(int(_lib.load().pf_mr_scan_tmp_ints(n)))
        nbx, nby, nbz = (int(v) for v in nblocks)
        s = _lib.stream()
        src = self.struct()
        _lib.call("pf_mr_classify", ctypes.byref(src), _lib.ptr(dist), nbx, nby, nbz, float(block_extent),
                  float(beta_bw), int(max_level), int(defer_threshold), int(seed), _lib.ptr(lnew), _lib.ptr(nout),
                  _lib.ptr(offs), _lib.ptr(tmp), s)
        total = int(offs[n].item())
        cap = max(int(total * 1.3), self.capacity) if total > self.capacity else self.capacity
        out = self._out
        if out is None or out["x"].shape[0] < cap:
            out = {"x": torch.zeros((cap, 3), dtype=torch.float32, device=d),
                   "v": torch.zeros((cap, 3), dtype=torch.float32, device=d),
                   **{k: torch.zeros(cap, dtype=torch.int32, device=d) for k in ("type", "level", "defer")}}
        dst = _lib.MrParticles()
        dst.x, dst.v, dst.type = out["x"].data_ptr(), out["v"].data_ptr(), out["type"].data_ptr()
        dst.level, dst.defer = out["level"].data_ptr(), out["defer"].data_ptr()
        dst.n = out["x"].shape[0]
        _lib.call("pf_mr_emit", ctypes.byref(src), _lib.ptr(lnew), _lib.ptr(nout), _lib.ptr(offs),
                  int(defer_threshold), float(r0), int(seed) + 7919, ctypes.byref(dst), s)
        # swap the live and the scatter side
        self._out = {"x": self.x, "v": self.v, "type": self.type, "level": self.level, "defer": self.defer}
        self.x, self.v, self.type, self.level, self.defer = out["x"], out["v"], out["type"], out["level"], out["defer"]
        self.capacity = self.x.shape[0]
        self.n = total
        return n, total

    def numpy(self):
        """Live particles on the host: dict of x, v, type, level, defer."""
        n = self.n
        return {k: getattr(self, k)[:n].cpu().numpy() for k in ("x", "v", "type", "level", "defer")}
