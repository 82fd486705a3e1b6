"""Flexible Polak-Ribiere PCG on an MSBG (drop-in for pfflip/mr_solver.py).

Host-driven like the reference (one device-to-host read per reduction);
the vector kernels and reductions (FP64, fixed combination order) run in
``pf_msbg.hip``.  Preconditioners: the multigrid V-cycle (``"mg"``), the
same with the adaptive relaxation smoother (``"mg-adaptive"``) or Jacobi
(``"jacobi"``).
"""

from __future__ import annotations

import torch

from . import _lib
from . import mr_operator as mro
from .mr_multigrid import MRMultigrid, axpby, reduce, remove_mean


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
        self._work = torch.zeros(int(_lib.load().pf_vec_reduce_bytes()) // 8, dtype=torch.float64,
                                 device=grid.dev)

    def prepare(self, bface):
        mro.compute_diag(self.grid, bface, self.diag)
        if self.mg is not None:
            self.mg.build_hierarchy(bface)

    def _apply_preconditioner(self, r, z):
        if self.mg is not None:
            self.mg.precondition(r, z)
        else:
            _lib.call("pf_vec_jacobi", _lib.ptr(r), _lib.ptr(self.diag), _lib.ptr(z), self.n, _lib.stream())
            remove_mean(z, self._work)

    def solve(self, bface, rhs, p_out, *, tol=1e-4, max_iter=2000, verbose=False):
        """mr_solver.py:118-165; returns (iterations, relative residual)."""
        w = self._work
        self.prepare(bface)
        self.b.copy_(rhs)
        remove_mean(self.b, w)
        bnorm = reduce(2, self.b, None, w)
        if bnorm == 0.0:
            p_out.zero_()
            return 0, 0.0
        p_out.zero_()
        self.r.copy_(self.b)
        self._apply_preconditioner(self.r, self.zv)
        self.q.copy_(self.zv)
        rz = reduce(1, self.r, self.zv, w)
        it, rel = 0, 1.0
        for it in range(1, max_iter + 1):
            mro.apply_A(self.grid, self.q, bface, self.Ap)
            qAq = reduce(1, self.q, self.Ap, w)
            if qAq <= 0.0:
                break
            alpha = rz / qAq
            axpby(alpha, self.q, 1.0, p_out)
            self.r_prev.copy_(self.r)
            axpby(-alpha, self.Ap, 1.0, self.r)
            rel = reduce(2, self.r, None, w) / bnorm
            if verbose and it % 50 == 0:
                print(f"    cg {it:5d}  {rel:.3e}")
            if rel < tol:
                break
            self._apply_preconditioner(self.r, self.zv)
            rz_new = reduce(1, self.r, self.zv, w)
            if rz == 0.0 or rz_new <= 0.0:
                break
            beta = max((rz_new - reduce(1, self.r_prev, self.zv, w)) / rz, 0.0)
            axpby(1.0, self.zv, beta, self.q)
            rz = rz_new
        remove_mean(p_out, w)
        return it, rel
