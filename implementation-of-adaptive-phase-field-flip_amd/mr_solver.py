"""Flexible Polak-Ribiere PCG on an MSBG (drop-in for pfflip/mr_solver.py).

Host-driven like the reference, but the CG scalars stay on the device: the
host reads one state record per iteration (the reference reads every
reduction).  The vector kernels, reductions (FP64, fixed combination order)
and scalar steps run in ``pf_msbg.hip``.  Preconditioners: the multigrid V-cycle (``"mg"``), the
same with the adaptive relaxation smoother (``"mg-adaptive"``) or Jacobi
(``"jacobi"``).
"""

from __future__ import annotations

import torch

from . import _lib
from . import mr_operator as mro
from .mr_multigrid import MRMultigrid, remove_mean


class MRSolver:
    def __init__(self, grid, device=None, preconditioner="mg", **mg_kwargs):
        if preconditioner not in ("mg", "mg-adaptive", "jacobi"):
            raise ValueError(f"preconditioner {preconditioner!r}: 'mg', 'mg-adaptive' or 'jacobi'")
        self.grid = grid
        self.kind = preconditioner
        if preconditioner == "mg-adaptive":  # mr_solver.py:77-82
            mg_kwargs.setdefault("smoother", "adaptive")
        self.mg = MRMultigrid(grid, **mg_kwargs) if preconditioner != "jacobi" else None
        n = grid.ncells
        self.n = n
        z = lambda: torch.zeros(n, dtype=torch.float32, device=grid.dev)  # noqa: E731
        self.diag, self.r, self.zv, self.q, self.Ap, self.b, self.r_prev = (z() for _ in range(7))
        nw = int(_lib.load().pf_vec_reduce_bytes()) // 8
        self._work = torch.zeros(nw, dtype=torch.float64, device=grid.dev)
        # one reduction buffer per scalar of an iteration; the state and flag
        self._works = [torch.zeros(nw, dtype=torch.float64, device=grid.dev) for _ in range(6)]
        self._st = torch.zeros(8, dtype=torch.float64, device=grid.dev)
        self._flag = torch.zeros(2, dtype=torch.int32, device=grid.dev)
        self._host = torch.empty(8, dtype=torch.float64, pin_memory=True)  # the per-iteration state read

    def prepare(self, bface):
        mro._diag(self.grid, bface, self.diag)
        if self.mg is not None:
            self.mg.build_hierarchy(bface)

    def _apply_preconditioner(self, r, z):
        if self.mg is not None:
            self.mg.precondition(r, z)
        else:
            _lib.call("pf_vec_jacobi", _lib.ptr(r), _lib.ptr(self.diag), _lib.ptr(z), self.n, _lib.stream())
            remove_mean(z, self._work)

    def solve(self, bface, rhs, p_out, *, tol=1e-4, max_iter=2000, verbose=False):
        """mr_solver.py:118-165; returns (iterations, relative residual).

        The CG scalars (rz, alpha, beta, rel) stay on the device; the host
        reads the state once per iteration to stop (the reference reads
        every reduction).  The arithmetic is the reference's: FP64 scalars,
        each rounded to float where it scales a vector."""
        s = _lib.stream()
        W = self._works
        st, flag = self._st, self._flag
        self.prepare(bface)
        self.b.copy_(rhs)
        remove_mean(self.b, self._work)
        _lib.call("pf_vec_reduce", 2, _lib.ptr(self.b), None, self.n, _lib.ptr(W[0]), s)
        _lib.call("pf_mrcg_scalar", 0, _lib.ptr(st), _lib.ptr(flag), _lib.ptr(W[0]), None, float(tol), s)
        p_out.zero_()
        if float(W[0][0].item()) == 0.0:
            return 0, 0.0
        self.r.copy_(self.b)
        self._apply_preconditioner(self.r, self.zv)
        self.q.copy_(self.zv)
        _lib.call("pf_vec_reduce", 1, _lib.ptr(self.r), _lib.ptr(self.zv), self.n, _lib.ptr(W[1]), s)
        _lib.call("pf_mrcg_scalar", 1, _lib.ptr(st), _lib.ptr(flag), _lib.ptr(W[1]), None, 0.0, s)
        it, rel = 0, 1.0
        host = self._host
        for it in range(1, max_iter + 1):
            mro._apply(self.grid, self.q, bface, self.Ap)
            _lib.call("pf_vec_reduce", 1, _lib.ptr(self.q), _lib.ptr(self.Ap), self.n, _lib.ptr(W[2]), s)
            _lib.call("pf_mrcg_scalar", 2, _lib.ptr(st), _lib.ptr(flag), _lib.ptr(W[2]), None, 0.0, s)
            _lib.call("pf_vec_axpby_dev", _lib.ptr(st), 1, 1.0, _lib.ptr(self.q), -1, _lib.ptr(p_out), self.n,
                      _lib.ptr(flag), s)
            self.r_prev.copy_(self.r)
            _lib.call("pf_vec_axpby_dev", _lib.ptr(st), 1, -1.0, _lib.ptr(self.Ap), -1, _lib.ptr(self.r), self.n,
                      _lib.ptr(flag), s)
            _lib.call("pf_vec_reduce", 2, _lib.ptr(self.r), None, self.n, _lib.ptr(W[3]), s)
            _lib.call("pf_mrcg_scalar", 3, _lib.ptr(st), _lib.ptr(flag), _lib.ptr(W[3]), None, 0.0, s)
            host.copy_(st)  # one read per iteration: the scalars and the stop flag (st[7])
            rel = float(host[3])
            stop = int(host[7])
            if stop == 3:  # stopped at the previous iteration's beta
                it -= 1
            if stop:
                break
            if verbose and it % 50 == 0:
                print(f"    cg {it:5d}  {rel:.3e}")
            self._apply_preconditioner(self.r, self.zv)
            _lib.call("pf_vec_reduce", 1, _lib.ptr(self.r), _lib.ptr(self.zv), self.n, _lib.ptr(W[4]), s)
            _lib.call("pf_vec_reduce", 1, _lib.ptr(self.r_prev), _lib.ptr(self.zv), self.n, _lib.ptr(W[5]), s)
            _lib.call("pf_mrcg_scalar", 4, _lib.ptr(st), _lib.ptr(flag), _lib.ptr(W[4]), _lib.ptr(W[5]), 0.0, s)
            _lib.call("pf_vec_axpby_dev", _lib.ptr(st), -1, 1.0, _lib.ptr(self.zv), 2, _lib.ptr(self.q), self.n,
                      _lib.ptr(flag), s)
        remove_mean(p_out, self._work)
        return it, rel
