"""Variable-coefficient pressure Poisson solver (drop-in for pfflip/poisson.py).

``MGPCG`` keeps the reference's constructor, level table (``levels[i].bx``,
``.by``, ``.bz``, ``.diag``, ``.p``, ``.b``, ``.r``), ``nlev``, ``ncells``,
``build_hierarchy``, ``precondition`` and ``solve`` (pfflip/poisson.py:226-409);
the work is done by ``pf_mg_*`` / ``pf_mgpcg_solve`` in libpfflip_hip.so.

Differences from the reference that do not change the algorithm:

* the FPCG scalars stay on the GPU; ``solve`` synchronises with the host
  about once per two iterations (while the GPU has queued work) instead of
  six times per iteration; the
  solver owns its host state (pinned flags, captured HIP graphs) through
  ``pf_mgpcg_state_create`` and frees it with the solver;
* dot products and means are deterministic FP64 reductions instead of FP32
  atomics, and the Polak-Ribiere numerator uses r_prev = r + alpha Ap, so no
  r_prev copy is kept (``r_prev`` exists as an attribute but is unused);
* ``precision="bf16"`` / ``"fp16"`` stores the V-cycle's face coefficients and
  iterates in 16 bits on the levels that run out of HBM (FP32 arithmetic);
  the flexible CG absorbs the inexact preconditioner;
* a red-black sweep is one fused kernel that reads the iterate from one
  buffer and writes it to a second one (``levels[l].r``, and ``levels[0].p``
  in FP32); ``fused_smoother=False`` runs the two half-sweeps in place, with
  the same result bit for bit;
* on the small levels below level 0 (up to 32^3 cells) the pre-sweeps with
  the restriction, and the prolongation with the post-sweeps, are one kernel
  each over LDS-resident tiles with redundant halos; ``split_levels=True``
  runs them one kernel per sweep (the same result bit for bit);
* the CG's inner products after the preconditioner (sum z, sum r, sum Ap,
  <r, z>, <Ap, z>) are formed by the V-cycle's last level-0 sweep from the
  values it holds, one FP64 partial per tile; ``fold_dots=False`` forms
  them in a kernel of their own (the same sums in another order).
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib

PRECISIONS = {"fp32": 0, "bf16": 1, "fp16": 2}
_HALF = {1: torch.bfloat16, 2: torch.float16}


def level_dims(res, min_dim):
    """pfflip/poisson.py:256-263."""
    out = []
    r = tuple(int(n) for n in res)
    while True:
        out.append(r)
        if min(r) // 2 < min_dim or any(n % 2 for n in r):
            break
        r = tuple(n // 2 for n in r)
    return out


class _Level:
    """One level of the hierarchy (pfflip/poisson.py:226-237)."""

    def __init__(self, res, device):
        nx, ny, nz = res
        self.res = (nx, ny, nz)
        z = lambda shape: torch.zeros(shape, dtype=torch.float32, device=device)  # noqa: E731
        self.bx = z((nx + 1, ny, nz))
        self.by = z((nx, ny + 1, nz))
        self.bz = z((nx, ny, nz + 1))
        self.diag = z((nx, ny, nz))
        self.p = z((nx, ny, nz))
        self.b = z((nx, ny, nz))
        self.r = z((nx, ny, nz))
        self.bx_h = self.by_h = self.bz_h = self.p_h = None


def _require_cuda(device):
    dev = torch.device(device)
    if dev.type != "cuda" or not torch.cuda.is_available():
        raise RuntimeError(f"pfflip_amd runs on the GPU only (device {device!r} is not available)")
    return dev


class MGPCG:
    """Multigrid-preconditioned flexible CG (pfflip/poisson.py:240-409)."""

    def __init__(self, res, device, *, omega=6.0 / 7.0, pre=2, post=2, coarsest=60, min_dim=4,
                 precision="fp32", fused_smoother=True, fold_dots=True,
                 split_levels=False):
        self.device = _require_cuda(device)
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        self.omega, self.pre, self.post, self.coarsest = omega, pre, post, coarsest
        self.precision = precision
        dims = level_dims(res, min_dim)
        if len(dims) > _lib.MAX_LEVELS:
            raise ValueError("too many multigrid levels")
        self.levels = [_Level(r, self.device) for r in dims]
        self.nlev = len(self.levels)
        self.ncells = int(res[0] * res[1] * res[2])
        z = lambda: torch.zeros(tuple(res), dtype=torch.float32, device=self.device)  # noqa: E731
        self.rhs, self.res_v, self.z, self.q, self.Ap, self.r_prev = z(), z(), z(), z(), z(), z()
        lib = _lib.load()
        self._scalars = torch.zeros(int(lib.pf_mg_scalar_bytes()), dtype=torch.uint8,
                                    device=self.device)
        mg = _lib.MG()
        mg.nlev = self.nlev
        mg.pre, mg.post, mg.coarsest = int(pre), int(post), int(coarsest)
        mg.omega = float(omega)
        mg.precision = PRECISIONS[precision]
        # one fused kernel per red-black sweep (default) or one per half-sweep;
        # both give the same bits
        mg.options = (0 if fused_smoother else _lib.MG_UNFUSED_SMOOTHER) | \
            (0 if fold_dots else _lib.MG_SEPARATE_DOTS) | \
            (_lib.MG_SPLIT_LEVELS if split_levels else 0)
        for i, L in enumerate(self.levels):
            mg.lev[i].nx, mg.lev[i].ny, mg.lev[i].nz = L.res
        _lib.check(lib.pf_mg_setup(ctypes.byref(mg)), "pf_mg_setup")
        self.small_from = int(mg.small_from)
        if mg.precision:
            # 16-bit faces and iterates on the levels run out of HBM; level 0's
            # 16-bit iterate is the preconditioned vector z in every case
            hd = _HALF[mg.precision]
            for i, L in enumerate(self.levels):
                nx, ny, nz = L.res
                if i < self.small_from:
                    L.bx_h = torch.zeros((nx + 1, ny, nz), dtype=hd, device=self.device)
                    L.by_h = torch.zeros((nx, ny + 1, nz), dtype=hd, device=self.device)
                    L.bz_h = torch.zeros((nx, ny, nz + 1), dtype=hd, device=self.device)
                if i < self.small_from or i == 0:
                    L.p_h = torch.zeros((nx, ny, nz), dtype=hd, device=self.device)
        for i, L in enumerate(self.levels):
            lv = mg.lev[i]
            for name in ("bx", "by", "bz", "diag", "p", "b", "r", "bx_h", "by_h", "bz_h", "p_h"):
                t = getattr(L, name)
                setattr(lv, name, None if t is None else t.data_ptr())
        mg.rhs, mg.res, mg.z = self.rhs.data_ptr(), self.res_v.data_ptr(), self.z.data_ptr()
        mg.q, mg.Ap, mg.scalars = self.q.data_ptr(), self.Ap.data_ptr(), self._scalars.data_ptr()
        # host state of this solver (pinned done-flags, the captured graphs),
        # freed with the solver
        st = ctypes.c_void_p(None)
        _lib.call("pf_mgpcg_state_create", ctypes.byref(st))
        mg.host_state = st.value
        self._state = st.value
        self._lib = lib
        self._mg = mg

    def __del__(self):
        st, self._state = getattr(self, "_state", None), None
        if st:
            self._lib.pf_mgpcg_state_destroy(st)

    # ---------------------------------------------------------------- hierarchy
    def build_hierarchy(self, bx, by, bz):
        """Level 0 = beta (copied unless it already is level 0's storage),
        Galerkin coarsening all the way down (pfflip/poisson.py:295-309)."""
        for t, L in zip((bx, by, bz), (self.levels[0].bx, self.levels[0].by, self.levels[0].bz)):
            if tuple(t.shape) != tuple(L.shape) or t.dtype != torch.float32 or not t.is_contiguous():
                raise ValueError("build_hierarchy: coefficients must be contiguous float32 face arrays")
        _lib.call("pf_mg_build", ctypes.byref(self._mg), _lib.ptr(bx), _lib.ptr(by), _lib.ptr(bz),
                  _lib.stream())

    def precondition(self, r_in, z_out):
        """z = V-cycle(r) with the mean removed (pfflip/poisson.py:345-352)."""
        _lib.call("pf_mg_precondition", ctypes.byref(self._mg), _lib.ptr(r_in), _lib.ptr(z_out),
                  _lib.stream())

    def vcycle(self, r_in, z_out):
        """z = one V-cycle from a zero guess on the right-hand side r, the
        coarsest level's mean removal included but not the final one
        (``_vcycle`` on ``levels[0].b = r``, pfflip/poisson.py:323-343).
        FP32 hierarchies only; ``z_out`` must not be ``levels[0].p``."""
        _lib.call("pf_mg_vcycle", ctypes.byref(self._mg), _lib.ptr(r_in), _lib.ptr(z_out), _lib.stream())

    # ---------------------------------------------------------------- solve
    def solve(self, b, p_out, *, tol=1e-4, max_iter=200, verbose=False):
        """Flexible PCG with Polak-Ribiere beta (pfflip/poisson.py:355-409).

        Writes ``p_out`` (mean removed) and returns ``(iterations, rel)`` with
        ``rel = max|r| / max|b - mean(b)|`` of the recursive residual.
        """
        it = ctypes.c_int32(0)
        rel = ctypes.c_double(0.0)
        _lib.call("pf_mgpcg_solve", ctypes.byref(self._mg), _lib.ptr(b), _lib.ptr(p_out),
                  float(tol), int(max_iter), ctypes.byref(it), ctypes.byref(rel), _lib.stream())
        if verbose:
            print(f"    pcg it={it.value:3d}  |r|inf/|b|inf = {rel.value:.3e}")
        return int(it.value), float(rel.value)


def time_smoother(solver, level=0, reps=10):
    """Average milliseconds of one smoother sweep kernel of an HBM level (HIP
    events on the current stream): a fused full sweep, or one half-sweep if
    the solver was built with ``fused_smoother=False``.  Overwrites that
    level's iterate buffers, so call it after a solve."""
    ms = ctypes.c_float(0.0)
    _lib.call("pf_mg_time_smoother", ctypes.byref(solver._mg), int(level), int(reps),
              ctypes.byref(ms), _lib.stream())
    return float(ms.value)


def apply_A(p, bx, by, bz, diag, out):
    """``out = A p`` (the ``apply_A`` kernel, pfflip/poisson.py:54-88)."""
    nx, ny, nz = p.shape
    _lib.call("pf_apply_A", nx, ny, nz, _lib.ptr(p), _lib.ptr(bx), _lib.ptr(by), _lib.ptr(bz),
              _lib.ptr(diag), _lib.ptr(out), _lib.stream())
    return out


def build_rhs(vel, rhs, dx, dt, cell_wsum=None, k_div=0.0, rho_tilde_0=1.0):
    """``rhs = -(dx/dt) sum_a (u_a[i+1] - u_a[i])``, minus
    ``dx^2 k_div max((cell_wsum - rho_tilde_0)/rho_tilde_0, 0)/dt`` when
    ``k_div > 0`` (the ``build_rhs`` kernel, pfflip/kernels.py:338-363).
    ``vel`` is the three face arrays, ``rhs`` the cell array it writes."""
    nx, ny, nz = rhs.shape
    if k_div > 0.0 and cell_wsum is None:
        raise ValueError("build_rhs: k_div > 0 needs cell_wsum")
    _lib.call("pf_build_rhs", nx, ny, nz, _lib.ptr3(vel), _lib.ptr(cell_wsum), _lib.ptr(rhs), float(dx),
              float(dt), float(k_div), float(rho_tilde_0), _lib.stream())
    return rhs
