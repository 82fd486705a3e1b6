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


def _field(grid, t, name, per_cell=1):
    """Refuse what the kernels would read or write out of bounds: a cell
    field is ``grid.ncells`` contiguous float32 values on the grid's device,
    a face field three per cell (each cell owns its three low faces)."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch.Tensor, got {type(t).__name__}")
    want = per_cell * grid.ncells
    if t.dtype != torch.float32 or not t.is_contiguous() or t.numel() != want:
        raise ValueError(f"{name}: need {want} contiguous float32 values, got {t.numel()} "
                         f"{t.dtype} (contiguous={t.is_contiguous()})")
    dev = torch.device(grid.dev) if grid.dev is not None else None
    same = dev is not None and t.device.type == "cuda" and (
        dev.index is None and t.device.index == torch.cuda.current_device() or t.device.index == dev.index)
    if not same:
        raise ValueError(f"{name}: on {t.device}, the grid lives on {grid.dev}")
    return _lib.ptr(t)


# The multigrid and the solver call the unchecked forms below in their inner
# loops, on buffers they allocated for the grid themselves; the public
# functions check their arguments first.
def _apply(grid, p, bface, out):
    _lib.call("pf_msbg_apply", grid.ref(), _lib.ptr(p), _lib.ptr(bface), _lib.ptr(out), 0, _lib.stream())


def _diag(grid, bface, diag):
    _lib.call("pf_msbg_apply", grid.ref(), None, _lib.ptr(bface), _lib.ptr(diag), 1, _lib.stream())


def _residual(grid, p, b, bface, r):
    _lib.call("pf_msbg_residual", grid.ref(), _lib.ptr(p), _lib.ptr(b), _lib.ptr(bface), _lib.ptr(r),
              _lib.stream())


def _rbgs(grid, p_in, p_out, b, bface, diag, colour, omega):
    _lib.call("pf_msbg_rbgs", grid.ref(), _lib.ptr(p_in), _lib.ptr(p_out), _lib.ptr(b), _lib.ptr(bface),
              _lib.ptr(diag), int(colour), float(omega), _lib.stream())


def apply_A(grid, p, bface, out):
    """out = A p (mr_operator.py:169-177)."""
    _field(grid, p, "p"), _field(grid, bface, "bface", 3), _field(grid, out, "out")
    _apply(grid, p, bface, out)


def compute_diag(grid, bface, diag):
    """diag = sum of the row's couplings (mr_operator.py:163-166)."""
    _field(grid, bface, "bface", 3), _field(grid, diag, "diag")
    _diag(grid, bface, diag)


def residual(grid, p, b, bface, r):
    """r = b - A p (mr_operator.py:180-190)."""
    _field(grid, p, "p"), _field(grid, b, "b"), _field(grid, bface, "bface", 3), _field(grid, r, "r")
    _residual(grid, p, b, bface, r)


def rbgs_half_sweep(grid, p_in, p_out, b, bface, diag, colour, omega):
    """One colour of weighted Gauss-Seidel (mr_operator.py:193-212) from p_in
    into p_out (the cells of that colour read the values before the sweep).
    ``colour`` is the parity of the cell's level-0 index sum (0 or 1)."""
    if colour not in (0, 1):
        raise ValueError(f"colour must be 0 or 1, got {colour!r}")
    for t, name in ((p_in, "p_in"), (p_out, "p_out"), (b, "b"), (diag, "diag")):
        _field(grid, t, name)
    _field(grid, bface, "bface", 3)
    _rbgs(grid, p_in, p_out, b, bface, diag, colour, omega)


def cell_volume(grid, vol):
    """vol = h^3 per cell (mr_operator.py:215-220)."""
    _lib.call("pf_msbg_geometry", grid.ref(), None, _field(grid, vol, "vol"), None, _lib.stream())
