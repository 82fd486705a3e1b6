"""Variable-coefficient Poisson operator on an MSBG (drop-in for
pfflip/mr_operator.py).

Finite volume: the coupling of adjacent cells A and B is
``T = beta_face * min(hA, hB)**2 / ((hA + hB) / 2)`` and A's row is
``sum_f T_f (p_A - p_B)``; across a 2:1 jump a coarse cell sees the 2x2
finer cells behind a face (mr_operator.py:1-36).  Every function launches
the HIP kernels of ``pf_msbg.hip`` on the current stream.
"""

from __future__ import annotations

from . import _lib


def apply_A(grid, p, bface, out):
    """out = A p (mr_operator.py:169-177)."""
    _lib.call("pf_msbg_apply", grid.ref(), _lib.ptr(p), _lib.ptr(bface), _lib.ptr(out), 0, _lib.stream())


def compute_diag(grid, bface, diag):
    """diag = sum of the row's couplings (mr_operator.py:163-166)."""
    _lib.call("pf_msbg_apply", grid.ref(), None, _lib.ptr(bface), _lib.ptr(diag), 1, _lib.stream())


def residual(grid, p, b, bface, r):
    """r = b - A p (mr_operator.py:180-190)."""
    _lib.call("pf_msbg_residual", grid.ref(), _lib.ptr(p), _lib.ptr(b), _lib.ptr(bface), _lib.ptr(r),
              _lib.stream())


def rbgs_half_sweep(grid, p_in, p_out, b, bface, diag, colour, omega):
    """One colour of weighted Gauss-Seidel (mr_operator.py:193-212) from p_in
    into p_out (the cells of that colour read the values before the sweep)."""
    _lib.call("pf_msbg_rbgs", grid.ref(), _lib.ptr(p_in), _lib.ptr(p_out), _lib.ptr(b), _lib.ptr(bface),
              _lib.ptr(diag), int(colour), float(omega), _lib.stream())


def cell_volume(grid, vol):
    """vol = h^3 per cell (mr_operator.py:215-220)."""
    _lib.call("pf_msbg_geometry", grid.ref(), None, _lib.ptr(vol), None, _lib.stream())
