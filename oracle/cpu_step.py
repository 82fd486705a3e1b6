"""ctypes wrapper of the C++/OpenMP restatement of the reference step
(``oracle/cpu_step.cpp``, built by ``oracle/Makefile`` into
``oracle/_build/libcpu_step.so``).

TEST INFRASTRUCTURE ONLY: the CPU baseline timed by bench.py and a second
restatement checked against the FP64 numpy oracle in tests/.  FP32 like the
reference, per-particle scatter with float atomics, the same MG-PCG; the
droplet stage is not restated.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libcpu_step.so")

FIELDS = {"px": 0, "pv": 1, "pressure": 2, "phi_cell": 3, "is_liquid": 4, "vel0": 5, "vel1": 6,
          "vel2": 7, "beta0": 8, "beta1": 9, "beta2": 10, "rhs": 11, "cell_mass": 12}


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("dx", "r_p", "rho_l", "rho_g", "alpha_phi", "alpha_flip_liquid", "alpha_flip_air")] + \
               [("gravity", ctypes.c_double * 3)] + \
               [(n, ctypes.c_double) for n in
                ("cfl", "dt_min", "dt_max", "pcg_tol", "omega", "rho_tilde_0", "rho_0l", "rho_min", "k_div")] + \
               [(n, ctypes.c_int32) for n in
                ("pcg_max_iter", "pre", "post", "coarsest", "min_dim", "extrap_cells")]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built (make -C oracle)")
        L = ctypes.CDLL(LIB)
        L.cs_create.restype = ctypes.c_void_p
        L.cs_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Params)]
        L.cs_destroy.argtypes = [ctypes.c_void_p]
        L.cs_set_threads.restype = ctypes.c_int
        L.cs_set_threads.argtypes = [ctypes.c_int]
        L.cs_set_particles.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p]
        L.cs_set_solids.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.cs_dt_from_cfl.restype = ctypes.c_double
        L.cs_dt_from_cfl.argtypes = [ctypes.c_void_p]
        L.cs_step.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
        L.cs_get.restype = ctypes.c_int
        L.cs_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        _lib = L
    return _lib


def set_threads(n: int = 0) -> int:
    """Set the OpenMP thread count (0 keeps it); returns the count in use."""
    return int(load().cs_set_threads(int(n)))


class CpuStepSim:
    """One uniform-grid simulator of the C++ restatement.  ``cfg`` carries
    the reference Config fields; ``rho`` = (rho_tilde_0, rho_0l, rho_min)."""

    def __init__(self, cfg, rho):
        L = load()
        P = Params()
        P.dx, P.r_p, P.rho_l, P.rho_g = cfg.dx, cfg.r_p, cfg.rho_l, cfg.rho_g
        P.alpha_phi, P.alpha_flip_liquid, P.alpha_flip_air = cfg.alpha_phi, cfg.alpha_flip_liquid, cfg.alpha_flip_air
        P.gravity = (ctypes.c_double * 3)(*[float(g) for g in cfg.gravity])
        P.cfl, P.dt_min, P.dt_max, P.pcg_tol = cfg.cfl, cfg.dt_min, cfg.dt_max, cfg.pcg_tol
        P.omega = cfg.mg_omega
        P.rho_tilde_0, P.rho_0l, P.rho_min = rho
        P.k_div = cfg.k_div
        P.pcg_max_iter, P.pre, P.post = cfg.pcg_max_iter, cfg.mg_pre_sweeps, cfg.mg_post_sweeps
        P.coarsest, P.min_dim, P.extrap_cells = cfg.mg_coarsest_sweeps, cfg.mg_min_dim, cfg.extrap_cells
        self.res = tuple(int(n) for n in cfg.res)
        self._h = ctypes.c_void_p(L.cs_create(*self.res, ctypes.byref(P)))
        self.n = 0

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.cs_destroy(self._h)
            self._h = None

    def set_particles(self, pos, vel, ptype):
        pos = np.ascontiguousarray(pos, np.float32)
        vel = np.ascontiguousarray(vel, np.float32)
        pt = np.ascontiguousarray(ptype, np.int32)
        self.n = len(pos)
        load().cs_set_particles(self._h, self.n, pos.ctypes.data, vel.ctypes.data, pt.ctypes.data)

    def set_solids(self, mask):
        m = np.ascontiguousarray(mask, np.int32)
        assert m.shape == self.res
        load().cs_set_solids(self._h, m.ctypes.data)

    def dt_from_cfl(self):
        return float(load().cs_dt_from_cfl(self._h))

    def step(self, dt=None):
        out = (ctypes.c_double * 4)()
        load().cs_step(self._h, -1.0 if dt is None else float(dt), out)
        return {"dt": out[0], "pcg_iter": int(out[1]), "pcg_res": out[2], "max_div": out[3]}

    def get(self, name):
        nx, ny, nz = self.res
        shapes = {"px": (self.n, 3), "pv": (self.n, 3), "vel0": (nx + 1, ny, nz), "vel1": (nx, ny + 1, nz),
                  "vel2": (nx, ny, nz + 1), "beta0": (nx + 1, ny, nz), "beta1": (nx, ny + 1, nz),
                  "beta2": (nx, ny, nz + 1)}
        shape = shapes.get(name, self.res)
        out = np.empty(shape, np.int32 if name == "is_liquid" else np.float32)
        if load().cs_get(self._h, FIELDS[name], out.ctypes.data) != 0:
            raise KeyError(name)
        return out
