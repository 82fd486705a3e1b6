"""Multiresolution multigrid on an MSBG (drop-in for pfflip/mr_multigrid.py).

Multigrid level L puts a block of MSBG level l at resolution
``min(l + L, log2(n0))`` (2:1 grading is preserved); coarse face
coefficients are the mean of the 2x2 fine faces they cover; transfers are
unsmoothed aggregation (prolongation = injection, restriction = its
transpose); the coarsest level gets ``coarsest_sweeps`` red-black sweeps
and mean removal (mr_multigrid.py:1-43, 122-246).  Kernels: ``pf_msbg.hip``.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from . import mr_operator as mro
from .msbg import MSBG


class _MGLevel:
    def __init__(self, grid: MSBG):
        self.grid = grid
        n = grid.ncells
        z = lambda: torch.zeros(n, dtype=torch.float32, device=grid.dev)  # noqa: E731
        self.bface = torch.zeros(n * 3, dtype=torch.float32, device=grid.dev)
        self.diag, self.p, self.p2, self.b, self.r = z(), z(), z(), z(), z()
        self.n = n


def reduce(op, x, y, work):
    """FP64 sum (op 0), dot (1) or max |x| (2) of float32 device vectors."""
    _lib.call("pf_vec_reduce", int(op), _lib.ptr(x), _lib.ptr(y), x.numel(), _lib.ptr(work), _lib.stream())
    out = ctypes.c_double()
    _lib.call("pf_read_scalars", _lib.ptr(work), ctypes.cast(ctypes.byref(out), ctypes.c_void_p), 8,
              _lib.stream())
    return out.value


def axpby(a, x, b, y):
    """y = a x + b y (x None: y = b y + a)."""
    _lib.call("pf_vec_axpby", float(a), _lib.ptr(x), float(b), _lib.ptr(y), y.numel(), _lib.stream())


def remove_mean(a, work):
    """a -= mean(a) (mr_multigrid.py:181-184), the FP64 sum kept on the device."""
    _lib.call("pf_vec_remove_mean", _lib.ptr(a), a.numel(), _lib.ptr(work), _lib.stream())


class MRMultigrid:
    """V-cycle preconditioner over the MSBG multigrid hierarchy
    (mr_multigrid.py:139-246)."""

    def __init__(self, grid: MSBG, device=None, *, omega=6.0 / 7.0, pre=3, post=3,
                 coarsest_sweeps=80, min_cells=512, smoother="rbgs", n_iter_max=12, theta_sigma=2.0):
        if smoother not in ("rbgs", "adaptive"):
            raise ValueError(f"smoother {smoother!r}: 'rbgs' or 'adaptive'")
        self.omega, self.pre, self.post = float(omega), int(pre), int(post)
        self.coarsest_sweeps = int(coarsest_sweeps)
        self._work = torch.zeros(int(_lib.load().pf_vec_reduce_bytes()) // 8, dtype=torch.float64,
                                 device=grid.dev)
        base = np.asarray(grid.level_h, np.int32)
        cap = grid.sh0  # a block bottoms out at one cell
        self.levels = [_MGLevel(grid)]
        L = 1
        while True:
            lv = np.where(base >= 0, np.minimum(base + L, cap), -1).astype(np.int32)
            g = MSBG(grid.nb, n0=grid.n0, h0=grid.h0, device=grid.device, max_level=cap)
            g.build(lv, grade=False)  # min(l + L, cap) keeps the 2:1 grading
            if g.ncells == self.levels[-1].n:
                break
            self.levels.append(_MGLevel(g))
            if g.ncells <= min_cells:
                break
            L += 1
        self.nlev = len(self.levels)
        # Algorithm 2 on every level but the coarsest (mr_multigrid.py:170-178)
        self.adaptive = None
        if smoother == "adaptive":
            from .adaptive_relax import AdaptiveSmoother
            self.adaptive = [AdaptiveSmoother(lv.grid, omega=omega, n_iter_max=n_iter_max, theta_sigma=theta_sigma)
                             for lv in self.levels[:-1]]
        self._record = False
        # The coarsest level: its sweeps (from zero) and mean removal as an
        # FP64 matrix built once per set of coefficients ("matrix"), or the
        # sweeps in one workgroup from LDS ("lds"), or launched half-sweeps
        # ("launched"), whichever the level's size allows first.
        lib = _lib.load()
        nc = self.levels[-1].n
        if nc <= int(lib.pf_msbg_coarse_op_max_cells()):
            self.coarsest_mode = "matrix"
        elif nc <= int(lib.pf_msbg_coarsest_max_cells()):
            self.coarsest_mode = "lds"
        else:
            self.coarsest_mode = "launched"
        self._coarse_op = (torch.empty(nc * nc, dtype=torch.float64, device=grid.dev)
                           if self.coarsest_mode == "matrix" else None)

    def build_hierarchy(self, bface0):
        """Level-0 coefficients, then coarsen down the hierarchy (mr_multigrid.py:187-198)."""
        L0 = self.levels[0]
        L0.bface.copy_(bface0)
        for i, lv in enumerate(self.levels):
            if i > 0:
                f = self.levels[i - 1]
                _lib.call("pf_msbg_coarsen_beta", lv.grid.ref(), f.grid.ref(), _lib.ptr(f.bface),
                          _lib.ptr(lv.bface), _lib.stream())
            mro._diag(lv.grid, lv.bface, lv.diag)
        if self._coarse_op is not None:
            c = self.levels[-1]
            _lib.call("pf_msbg_coarse_op", c.grid.ref(), _lib.ptr(c.bface), _lib.ptr(c.diag), self.coarsest_sweeps,
                      self.omega, _lib.ptr(self._coarse_op), _lib.stream())

    def _smooth(self, lv: _MGLevel, n_sweeps, reverse=False, index=None):
        """n_sweeps red-black sweeps, or the adaptive relaxation loop on a
        non-coarsest level when the smoother is adaptive (mr_multigrid.py:200-211)."""
        if self.adaptive is not None and index is not None and index < self.nlev - 1:
            self.adaptive[index].smooth(lv, record=self._record)
            return
        colours = (1, 0) if reverse else (0, 1)
        for _ in range(n_sweeps):
            for c in colours:
                mro._rbgs(lv.grid, lv.p, lv.p2, lv.b, lv.bface, lv.diag, c, self.omega)
                lv.p, lv.p2 = lv.p2, lv.p

    def _vcycle(self, i=0):
        """mr_multigrid.py:213-235."""
        lv = self.levels[i]
        if i == self.nlev - 1:
            if self.coarsest_mode == "matrix" and self._coarse_op is not None:
                _lib.call("pf_msbg_coarse_apply", _lib.ptr(self._coarse_op), _lib.ptr(lv.b), _lib.ptr(lv.p), lv.n,
                          _lib.stream())
            elif self.coarsest_mode in ("matrix", "lds"):
                _lib.call("pf_msbg_coarsest_smooth", lv.grid.ref(), _lib.ptr(lv.p), _lib.ptr(lv.b),
                          _lib.ptr(lv.bface), _lib.ptr(lv.diag), self.coarsest_sweeps, self.omega, _lib.stream())
            else:
                self._smooth(lv, self.coarsest_sweeps)
            remove_mean(lv.p, self._work)
            return
        self._smooth(lv, self.pre, index=i)
        mro._residual(lv.grid, lv.p, lv.b, lv.bface, lv.r)
        c = self.levels[i + 1]
        c.p.zero_()
        _lib.call("pf_msbg_restrict", lv.grid.ref(), c.grid.ref(), _lib.ptr(lv.r), _lib.ptr(c.b), _lib.stream())
        self._vcycle(i + 1)
        _lib.call("pf_msbg_prolong_add", lv.grid.ref(), c.grid.ref(), _lib.ptr(c.p), _lib.ptr(lv.p),
                  _lib.stream())
        self._smooth(lv, self.post, reverse=True, index=i)

    def precondition(self, r_in, z_out):
        """z = V-cycle(r), mean removed (mr_multigrid.py:237-243)."""
        L0 = self.levels[0]
        L0.b.copy_(r_in)
        L0.p.zero_()
        self._vcycle(0)
        z_out.copy_(L0.p)
        remove_mean(z_out, self._work)

    def describe(self):
        return " -> ".join(f"{lv.n:,}" for lv in self.levels)
