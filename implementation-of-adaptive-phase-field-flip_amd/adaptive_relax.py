"""Adaptive relaxation for the MSBG multigrid smoother (drop-in for
pfflip/adaptive_relax.py; paper 6.4, Algorithm 2).

Red-black sweeps run over an active block set; after each sweep a block's
effective residual is ``max |delta| * diag`` over its cells (the update
already carries ``1/diag``, so the product puts air and liquid on one scale),
and the blocks above ``theta`` stay active together with their face
neighbours.  ``theta = exp(mean + theta_sigma * std)`` of the logarithms of
the first iteration's positive block residuals; the reference reads them
back to the host, here the threshold is formed on the device (FP64, fixed
order) so a smoothing pass needs no synchronisation.  With no positive
residual the threshold is infinite and the remaining sweeps change nothing,
where the reference leaves the loop.  All blocks start active
(adaptive_relax.py:52-81).  Kernels: ``csrc/pf_msbg.hip``.
"""

from __future__ import annotations

import torch

from . import _lib


class AdaptiveSmoother:
    """Per-multigrid-level state of Algorithm 2 (adaptive_relax.py:163-241)."""

    def __init__(self, grid, device=None, *, omega=6.0 / 7.0, n_iter_max=12, inner_sweeps=1, theta_sigma=2.0,
                 warmup_sweeps=2):
        self.g = grid
        self.omega = float(omega)
        self.n_iter_max = int(n_iter_max)
        self.inner_sweeps = int(inner_sweeps)
        self.theta_sigma = float(theta_sigma)
        self.warmup_sweeps = int(warmup_sweeps)
        ns, nc, dev = grid.nslots, grid.ncells, grid.dev
        self.active = torch.ones(ns, dtype=torch.int32, device=dev)
        self.active_next = torch.zeros(ns, dtype=torch.int32, device=dev)
        self.block_res = torch.zeros(ns, dtype=torch.float32, device=dev)
        self.theta = torch.zeros(1, dtype=torch.float32, device=dev)
        self.delta = torch.zeros(nc, dtype=torch.float32, device=dev)
        self.active_fraction: list[float] = []

    def _sweep(self, lv, reverse):
        """One pass of ``inner_sweeps`` sweeps; lv.p / lv.p2 alternate."""
        self.delta.zero_()
        g, s = self.g.ref(), _lib.stream()
        for _ in range(self.inner_sweeps):
            for c in ((1, 0) if reverse else (0, 1)):
                _lib.call("pf_msbg_rbgs_active", g, _lib.ptr(lv.p), _lib.ptr(lv.p2), _lib.ptr(lv.b),
                          _lib.ptr(lv.bface), _lib.ptr(lv.diag), _lib.ptr(self.active), c, self.omega,
                          _lib.ptr(self.delta), s)
                lv.p, lv.p2 = lv.p2, lv.p

    def smooth(self, lv, record=False):
        """The adaptive relaxation loop at one multigrid level
        (adaptive_relax.py:201-241); updates lv.p."""
        self.active.fill_(1)
        for w in range(self.warmup_sweeps):
            self._sweep(lv, reverse=(w % 2 == 1))
        g, s = self.g.ref(), _lib.stream()
        for it in range(self.n_iter_max):
            if record:  # synchronises: profiling only
                self.active_fraction.append(float(self.active.sum().item()) / max(self.g.nslots, 1))
            self._sweep(lv, reverse=(it % 2 == 1))
            _lib.call("pf_msbg_relax_update", g, _lib.ptr(self.delta), _lib.ptr(lv.diag), _lib.ptr(self.active),
                      _lib.ptr(self.block_res), _lib.ptr(self.theta), int(it == 0), self.theta_sigma,
                      _lib.ptr(self.active_next), s)
            self.active, self.active_next = self.active_next, self.active
