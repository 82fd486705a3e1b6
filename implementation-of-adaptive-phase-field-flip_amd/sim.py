"""PF-FLIP simulator on the uniform MAC grid (drop-in for pfflip/sim.py).

Same construction, attributes (``vel``, ``vel_star``, ``vel_liq``, ``dvel``,
``mass``, ``mom``, ``beta``, ``phi_face``, ``valid``, ``cell_mass``,
``cell_wsum``, ``phi_cell``, ``is_liquid``, ``solid``, ``pressure``, ``rhs``,
``px``, ``pv``, ``ptype``, ``pmass``, ``prad``, ``pstate``, ``pdrop_r``, ...),
stage methods and ``step`` record as the reference ``Simulator``
(pfflip/sim.py:37-424).  Arrays are torch tensors on the GPU with the
reference's shapes and dtypes; every stage runs in libpfflip_hip.so.

``step`` fuses stages that the reference launches separately (it produces
the same fields): binning + tiled P2G; eq. (7) / beta / walls / vel_star /
gravity / RHS in one grid pass; projection + FLIP delta; G2P + RK3 +
reflection in one particle pass (split around the droplet stage when
``cfg.droplets`` is on: the stage runs between G2P and advection).  The stage methods (``_p2g``,
``_phase_field_and_coefficients``, ``_add_forces``, ``_pressure_solve``,
``_project``, ``_max_divergence``, ``_g2p``, ``_extrapolate_liquid_velocity``,
``_advect``) keep their individual reference semantics.

The escaped-particle / droplet stage (``_droplets``, pfflip/sim.py:343-390)
and ``rasterize_spray`` (sim.py:392-401) run in ``pf_droplets`` /
``pf_velocity_gradient`` / ``pf_rasterize_spray``.  Droplet radii come from
a counter-based PCG-hash stream keyed by (cfg.seed + step_index, particle)
like the reference's Warp RNG; the streams are not the same numbers.
Kernel radius: ``prad`` must stay ``cfg.r_p`` (= dx), which is what the
reference's uniform path always sets (sim.py:153).
"""

from __future__ import annotations

import ctypes
import math
import time

import numpy as np
import torch

from . import _lib
from .config import Config
from .poisson import MGPCG, _require_cuda


def _rho_tilde_0(cfg: Config) -> float:
    """Kernel weight sum of a perfect lattice (pfflip/sim.py:119-140)."""
    dx, r = cfg.dx, cfg.r_p
    per_axis = round(cfg.ppc ** (1.0 / 3.0))
    step = dx / per_axis
    R = int(math.ceil(r / step)) + 1
    acc = 0.0
    for a in range(-R, R + 1):
        for b in range(-R, R + 1):
            for c in range(-R, R + 1):
                d2 = ((a + 0.5) ** 2 + (b + 0.5) ** 2 + (c + 0.5) ** 2) * step * step
                t = max(1.0 - d2 / (r * r), 0.0)
                acc += t * t * t
    return acc


class Simulator:
    def __init__(self, cfg: Config):
        self.cfg = cfg
        self.dev = _require_cuda(cfg.device)
        nx, ny, nz = (int(n) for n in cfg.res)
        self.res = (nx, ny, nz)
        dev = self.dev
        z = lambda shape: torch.zeros(shape, dtype=torch.float32, device=dev)  # noqa: E731
        zi = lambda shape: torch.zeros(shape, dtype=torch.int32, device=dev)  # noqa: E731
        zb = lambda shape: torch.zeros(shape, dtype=torch.uint8, device=dev)  # noqa: E731
        self.face_shapes = ((nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1))

        self.solver = MGPCG(self.res, cfg.device, omega=cfg.mg_omega, pre=cfg.mg_pre_sweeps,
                            post=cfg.mg_post_sweeps, coarsest=cfg.mg_coarsest_sweeps,
                            min_dim=cfg.mg_min_dim, precision=cfg.solver_precision)

        self.vel = [z(s) for s in self.face_shapes]
        self.vel_star = [z(s) for s in self.face_shapes]
        self.vel_liq = [z(s) for s in self.face_shapes]
        self.dvel = [z(s) for s in self.face_shapes]
        self.mass = [z(s) for s in self.face_shapes]
        self.mom = [z(s) for s in self.face_shapes]
        # beta is stored directly in the solver's level-0 coefficients, so
        # build_hierarchy does not copy it (the reference copies, poisson.py:298)
        L0 = self.solver.levels[0]
        self.beta = [L0.bx, L0.by, L0.bz]
        self.phi_face = [z(s) for s in self.face_shapes]
        self.valid = [zi(s) for s in self.face_shapes]
        self._known = [zb(s) for s in self.face_shapes]
        self._known2 = [zb(s) for s in self.face_shapes]
        self._tmp_face = [z(s) for s in self.face_shapes]

        self.cell_mass = z(self.res)
        self.cell_wsum = z(self.res)
        self.phi_cell = z(self.res)
        self.is_liquid = zi(self.res)
        self.solid = zi(self.res)
        self.pressure = z(self.res)
        self.rhs = z(self.res)
        # [max |div u| (float bits), 3 spare, droplet counters x5, spare]:
        # read back with one copy at the end of a step
        self._scalar = zi((10,))

        # spray density at spray_resolution_factor x the grid, FP16
        # (sim.py:55-60), and the cell |grad u| used for spray emission
        f = max(1, int(cfg.spray_resolution_factor))
        self.spray_res = tuple(n * f for n in self.res)
        self.dx_spray = cfg.dx / f
        ns = math.prod(self.spray_res)
        # FP16 atomics update aligned 4-byte words: an odd cell count gets
        # one FP16 of padding so the last word stays inside the allocation
        self._spray_buf = torch.zeros(ns + (ns & 1), dtype=torch.float16, device=dev)
        self.spray_density = self._spray_buf[:ns].view(self.spray_res)
        self.shear = z(self.res)

        self.px = self.pv = self.ptype = self.pmass = self.prad = self.pstate = self.pdrop_r = None
        self.n_particles = 0
        self.droplet_stats = {"emitted": 0, "bubbles": 0, "merged": 0, "spray": 0}

        self.rho_tilde_0 = _rho_tilde_0(cfg)
        self.rho_0l = self.rho_tilde_0 * cfg.rho_l
        self.rho_min = cfg.eta_phi * cfg.rho_g * self.rho_tilde_0

        # optional per-stage GPU timing: set to {} to accumulate milliseconds
        # per stage of step() (HIP events on the current stream)
        self.stage_ms = None
        self._marks = []

        self.has_solids = False
        self.time = 0.0
        self.step_index = 0
        self.stats: list[dict] = []

        lib = _lib.load()
        self._nkeys = int(lib.pf_bin_nkeys(nx, ny, nz))
        self.cell_start = torch.zeros(self._nkeys + 2, dtype=torch.int32, device=dev)
        nb = int(lib.pf_bin_scan_tmp_bytes(nx, ny, nz))
        self._scan_tmp = torch.zeros((nb + 3) // 4, dtype=torch.int32, device=dev)
        self._n_active = zi((1,))
        sb = int(lib.pf_p2g_stage_bytes(nx, ny, nz, 1 if cfg.k_div > 0.0 else 0))
        self._p2g_stage = torch.empty((sb + 7) // 8, dtype=torch.int64, device=dev)
        # [max|v|, max m, max m|v|_inf] of the FLIP particles (pf_max_speed);
        # the fused G2P pass of step() leaves the next step's values here, valid
        # while _bounds_token matches the particle tensors (_bounds_key)
        self._bounds = z((4,))
        self._bounds_token = None
        self._particles_epoch = 0

    # ------------------------------------------------------------------ setup
    def set_solids(self, mask):
        """Obstacle mask on the cell grid, True = solid (sim.py:113-116)."""
        m = np.ascontiguousarray(mask, np.int32)
        if m.shape != self.res:
            raise ValueError(f"solid mask shape {m.shape} != grid {self.res}")
        self.solid.copy_(torch.from_numpy(m))
        self.has_solids = bool(m.any())

    def set_particles(self, pos, vel, ptype):
        """Upload particles (sim.py:142-155); pmass = rho_phase, prad = r_p."""
        cfg = self.cfg
        pos = np.asarray(pos)
        n = len(pos)
        if n >= 2**31:
            raise ValueError("at most 2^31 - 1 particles")
        self.n_particles = n
        dev = self.dev
        self.px = torch.from_numpy(np.ascontiguousarray(pos, np.float32).reshape(n, 3)).to(dev)
        self.pv = torch.from_numpy(np.ascontiguousarray(vel, np.float32).reshape(n, 3)).to(dev)
        pt = np.ascontiguousarray(ptype, np.int32)
        self.ptype = torch.from_numpy(pt).to(dev)
        self.pmass = torch.from_numpy(np.where(pt == 1, cfg.rho_l, cfg.rho_g).astype(np.float32)).to(dev)
        self.prad = torch.full((n,), cfg.r_p, dtype=torch.float32, device=dev)
        self.pstate = torch.zeros(n, dtype=torch.int32, device=dev)
        self.pdrop_r = torch.zeros(n, dtype=torch.float32, device=dev)
        # binning buffers: keys and slot ranks per particle, cell-sorted order
        self._keys = torch.empty(n, dtype=torch.int32, device=dev)
        self._rank = torch.empty(n, dtype=torch.int32, device=dev)
        self.perm = torch.empty(n, dtype=torch.int32, device=dev)
        self._particles_epoch += 1
        self._bounds_token = None

    # ------------------------------------------------------------------ helpers
    def _read(self, t, count=1):
        out = (ctypes.c_float * count)()
        _lib.call("pf_read_scalars", _lib.ptr(t), ctypes.cast(out, ctypes.c_void_p),
                  4 * count, _lib.stream())
        return list(out)

    def _read_scalar_block(self):
        """(max |div u|, [emitted, bubbles, merged, spray, alive droplets])."""
        out = (ctypes.c_int32 * 10)()
        _lib.call("pf_read_scalars", _lib.ptr(self._scalar), ctypes.cast(out, ctypes.c_void_p),
                  40, _lib.stream())
        div = ctypes.c_float.from_buffer_copy(ctypes.c_int32(out[0])).value
        return div, list(out[4:9])

    def _lohi(self):
        eps = 1e-4 * self.cfg.dx
        return (eps, eps, eps), tuple(s - eps for s in self.cfg.domain_size)

    def _phase_args(self, flags, dt=0.0):
        cfg = self.cfg
        a = _lib.PhaseArgs()
        a.nx, a.ny, a.nz = self.res
        a.flags = flags
        a.mass = _lib.ptr3(self.mass)
        a.mom = _lib.ptr3(self.mom)
        a.vel = _lib.ptr3(self.vel)
        a.vel_star = _lib.ptr3(self.vel_star)
        a.valid = _lib.ptr3(self.valid)
        a.cell_mass = self.cell_mass.data_ptr()
        a.cell_wsum = self.cell_wsum.data_ptr()
        a.phi_cell = self.phi_cell.data_ptr()
        a.is_liquid = self.is_liquid.data_ptr()
        a.phi_face = _lib.ptr3(self.phi_face)
        a.beta = _lib.ptr3(self.beta)
        a.solid = self.solid.data_ptr() if self.has_solids else None
        a.rhs = self.rhs.data_ptr()
        a.mass_eps = 1e-12 * self.rho_0l
        a.rho_min = self.rho_min
        a.alpha_phi = cfg.alpha_phi
        a.rho_0l = self.rho_0l
        a.rho_l = cfg.rho_l
        a.rho_g = cfg.rho_g
        a.gravity = (ctypes.c_float * 3)(*[float(g) for g in cfg.gravity])
        a.dt = float(dt)
        a.dx = cfg.dx
        a.rho_tilde_0 = self.rho_tilde_0
        a.k_div = cfg.k_div
        return a

    def _phase(self, flags, dt=0.0):
        a = self._phase_args(flags, dt)
        _lib.call("pf_phase_beta_forces", ctypes.byref(a), _lib.stream())

    def _g2p_args(self, flags, dt=0.0):
        cfg = self.cfg
        a = _lib.G2PArgs()
        a.px, a.pv = self.px.data_ptr(), self.pv.data_ptr()
        a.ptype, a.pstate = self.ptype.data_ptr(), self.pstate.data_ptr()
        a.pflag = None
        a.n = self.n_particles
        a.nx, a.ny, a.nz = self.res
        a.flags = flags | (_lib.GA_SOLIDS if (self.has_solids and flags & _lib.GA_REFLECT) else 0)
        a.vel = _lib.ptr3(self.vel)
        a.dvel = _lib.ptr3(self.dvel)
        a.vel_liq = _lib.ptr3(self.vel_liq)
        a.solid = self.solid.data_ptr() if self.has_solids else None
        a.dx, a.dt = cfg.dx, float(dt)
        a.alpha_liquid, a.alpha_air = cfg.alpha_flip_liquid, cfg.alpha_flip_air
        lo, hi = self._lohi()
        a.lo = (ctypes.c_float * 3)(*lo)
        a.hi = (ctypes.c_float * 3)(*hi)
        return a

    def _bin(self, stable: bool = False):
        """Cell-sort the FLIP particles (pf_bin_sort) in particle-index order.
        The P2G sums do not depend on the order within a cell, so ties are
        left in atomic order unless ``stable`` asks for index order."""
        nx, ny, nz = self.res
        _lib.call("pf_bin_sort", _lib.ptr(self.px), _lib.ptr(self.pstate), self.n_particles,
                  nx, ny, nz, self.cfg.dx, None, _lib.ptr(self._keys),
                  _lib.ptr(self._rank), _lib.ptr(self.perm), _lib.ptr(self.cell_start),
                  _lib.ptr(self._scan_tmp), _lib.ptr(self._n_active), 1 if stable else 0,
                  _lib.stream())

    def _splat(self):
        nx, ny, nz = self.res
        # cell_wsum is only read when k_div > 0 (kernels.py:360-362); otherwise
        # it keeps its zeros
        wsum = self.cell_wsum if self.cfg.k_div > 0.0 else None
        _lib.call("pf_p2g", _lib.ptr(self.px), _lib.ptr(self.pv), _lib.ptr(self.pmass),
                  _lib.ptr(self.perm), _lib.ptr(self.cell_start), nx, ny, nz, self.cfg.dx,
                  self.cfg.r_p, _lib.ptr(self._bounds), *[_lib.ptr(t) for t in self.mass],
                  *[_lib.ptr(t) for t in self.mom], _lib.ptr(self.cell_mass), _lib.ptr(wsum),
                  _lib.ptr(self._p2g_stage), 8 * self._p2g_stage.numel(), _lib.stream())

    def _bounds_key(self):
        # torch bumps _version on every in-place write through torch; the
        # kernels here write through raw pointers and reset the token instead
        return (self._particles_epoch, self.pv._version, self.pstate._version, self.pmass._version)

    def _particle_bounds(self):
        """max |v| (CFL) and the P2G fixed-point bounds, on the device (kept
        from the previous step's G2P pass when the particles are unchanged)."""
        if self._bounds_token is not None and self._bounds_token == self._bounds_key():
            return
        _lib.call("pf_max_speed", _lib.ptr(self.pv), _lib.ptr(self.pmass), _lib.ptr(self.pstate),
                  self.n_particles, _lib.ptr(self._bounds), _lib.stream())

    def _mark(self, name):
        if self.stage_ms is not None:
            e = torch.cuda.Event(enable_timing=True)
            e.record()
            self._marks.append((name, e))

    def _collect_marks(self):
        if self.stage_ms is None or not self._marks:
            return
        for (_, e0), (name, e1) in zip(self._marks[:-1], self._marks[1:]):
            self.stage_ms[name] = self.stage_ms.get(name, 0.0) + e0.elapsed_time(e1)
        self._marks = []

    # ------------------------------------------------------------------ step
    def dt_from_cfl(self) -> float:
        """pfflip/sim.py:158-166 (speed over pstate == 0 particles only)."""
        cfg = self.cfg
        self._particle_bounds()
        vmax = float(self._read(self._bounds)[0])
        g = float(np.linalg.norm(cfg.gravity))
        vmax = max(vmax, math.sqrt(2.0 * g * cfg.dx) if g > 0 else 0.0, 1e-8)
        return float(np.clip(cfg.cfl * cfg.dx / vmax, cfg.dt_min, cfg.dt_max))

    def step(self, dt: float | None = None, verbose: bool = False, tags=None) -> dict:
        """One time step (pfflip/sim.py:168-202).

        ``tags`` (optional, an int32 cell array) replays the liquid tags of
        another run instead of deriving them from this step's cell mass,
        e.g. to difference a rollout along the branch its adjoint follows
        (the adjoint holds the tags fixed)."""
        cfg = self.cfg
        if dt is None:
            dt = self.dt_from_cfl()     # also leaves the P2G bounds on the device
        else:
            self._particle_bounds()
        t0 = time.perf_counter()
        self._mark("start")
        self._bin()
        self._mark("bin_sort")
        self._splat()
        self._mark("p2g")
        flags = (_lib.PH_VEL_FROM_MOM | _lib.PH_PHASE | _lib.PH_ZERO | _lib.PH_STAR
                 | _lib.PH_GRAVITY | _lib.PH_RHS)
        if tags is not None:
            if tuple(tags.shape) != self.res:
                raise ValueError("tags must be a cell array of the grid's shape")
            self.is_liquid.copy_(tags)
            flags |= _lib.PH_TAGS_IN
        self._phase(flags, dt)
        self._mark("phase_beta_rhs")
        self.solver.build_hierarchy(*self.beta)
        self._mark("mg_build")
        pcg_it, pcg_res = self.solver.solve(self.rhs, self.pressure, tol=cfg.pcg_tol,
                                            max_iter=cfg.pcg_max_iter, verbose=verbose)
        self._mark("mgpcg_solve")
        nx, ny, nz = self.res
        _lib.call("pf_project", _lib.ptr(self.pressure), _lib.ptr3(self.beta), _lib.ptr3(self.vel),
                  _lib.ptr3(self.vel_star), _lib.ptr3(self.dvel), nx, ny, nz, cfg.dx, float(dt),
                  _lib.ptr(self._scalar), _lib.stream())
        self._mark("project")
        self._extrapolate_liquid_velocity()
        self._mark("extrapolate")
        # the particles have not moved since _bin: G2P samples LDS tiles
        # through the same bins
        bins = (_lib.ptr(self.perm), _lib.ptr(self.cell_start), _lib.ptr(self._n_active))
        self._bounds_token = None
        if cfg.droplets:
            # G2P, then the droplet stage, then advection (sim.py:180-185)
            a = self._g2p_args(_lib.GA_G2P, dt)
            _lib.call("pf_g2p_advect_binned", ctypes.byref(a), *bins, None, None, _lib.stream())
            self._launch_droplets(dt)
            a = self._g2p_args(_lib.GA_ADVECT | _lib.GA_REFLECT, dt)
            _lib.call("pf_g2p_advect", ctypes.byref(a), _lib.stream())
        else:
            # the same pass leaves the next step's CFL / P2G bounds
            a = self._g2p_args(_lib.GA_G2P | _lib.GA_ADVECT | _lib.GA_REFLECT, dt)
            _lib.call("pf_g2p_advect_binned", ctypes.byref(a), *bins, _lib.ptr(self.pmass),
                      _lib.ptr(self._bounds), _lib.stream())
            self._bounds_token = self._bounds_key()
        self._mark("g2p_advect")
        div_after, counters = self._read_scalar_block()   # synchronises
        n_drop = 0
        if cfg.droplets:
            self._add_droplet_stats(counters)
            n_drop = counters[4]
        wall = time.perf_counter() - t0
        self._collect_marks()
        self.time += dt
        self.step_index += 1
        rec = {"step": self.step_index, "t": self.time, "dt": dt, "pcg_iter": pcg_it,
               "pcg_res": pcg_res, "max_div": div_after, "droplets": n_drop, "wall_s": wall}
        self.stats.append(rec)
        return rec

    # ------------------------------------------------------------------ stages
    def _p2g(self):
        """P2G + face velocities (pfflip/sim.py:205-236)."""
        self._particle_bounds()
        self._bin()
        self._splat()
        self._phase(_lib.PH_VEL_FROM_MOM)

    def _phase_field_and_coefficients(self):
        """Eq. (7), beta, walls and solids (pfflip/sim.py:238-268)."""
        self._phase(_lib.PH_PHASE | _lib.PH_ZERO)

    def _add_forces(self, dt: float):
        """Gravity, then walls and solids again (pfflip/sim.py:270-285)."""
        self._phase(_lib.PH_GRAVITY | _lib.PH_ZERO, dt)

    def _pressure_solve(self, dt: float, verbose: bool = False):
        """RHS, hierarchy, MG-PCG (pfflip/sim.py:287-299)."""
        cfg = self.cfg
        self._phase(_lib.PH_RHS, dt)
        self.solver.build_hierarchy(*self.beta)
        return self.solver.solve(self.rhs, self.pressure, tol=cfg.pcg_tol,
                                 max_iter=cfg.pcg_max_iter, verbose=verbose)

    def _project(self, dt: float):
        """u -= dt beta grad p / dx (pfflip/sim.py:301-308)."""
        nx, ny, nz = self.res
        _lib.call("pf_project", _lib.ptr(self.pressure), _lib.ptr3(self.beta), _lib.ptr3(self.vel),
                  None, None, nx, ny, nz, self.cfg.dx, float(dt), None, _lib.stream())

    def _max_divergence(self) -> float:
        """max |div u / dx| (pfflip/sim.py:310-314)."""
        nx, ny, nz = self.res
        _lib.call("pf_project", None, None, _lib.ptr3(self.vel), None, None, nx, ny, nz,
                  self.cfg.dx, 0.0, _lib.ptr(self._scalar), _lib.stream())
        return float(self._read(self._scalar)[0])

    def _g2p(self):
        """dvel = vel - vel_star, then eq. (12) (pfflip/sim.py:316-327)."""
        nx, ny, nz = self.res
        _lib.call("pf_project", None, None, _lib.ptr3(self.vel), _lib.ptr3(self.vel_star),
                  _lib.ptr3(self.dvel), nx, ny, nz, self.cfg.dx, 0.0, None, _lib.stream())
        a = self._g2p_args(_lib.GA_G2P)
        self._bounds_token = None
        _lib.call("pf_g2p_advect", ctypes.byref(a), _lib.stream())

    def _extrapolate_liquid_velocity(self):
        """Seed + Jacobi sweeps into vel_liq (pfflip/sim.py:329-341)."""
        nx, ny, nz = self.res
        _lib.call("pf_extrapolate", _lib.ptr3(self.vel), _lib.ptr(self.is_liquid),
                  _lib.ptr3(self.vel_liq), _lib.ptr3(self._tmp_face), _lib.ptr3(self._known),
                  _lib.ptr3(self._known2), nx, ny, nz, int(self.cfg.extrap_cells), _lib.stream())

    def _droplet_args(self, dt: float):
        cfg = self.cfg
        sigma = float(cfg.droplet_radius_spread)
        mu = math.log(cfg.droplet_radius_mean * cfg.dx) - 0.5 * sigma * sigma   # droplets.py:355-360
        a = _lib.DropletArgs()
        a.px, a.pv = self.px.data_ptr(), self.pv.data_ptr()
        a.ptype, a.pstate, a.pdrop_r = self.ptype.data_ptr(), self.pstate.data_ptr(), self.pdrop_r.data_ptr()
        a.n = self.n_particles
        a.nx, a.ny, a.nz = self.res
        a.seed = int(cfg.seed + self.step_index)
        a.phi = self.phi_cell.data_ptr()
        a.shear = self.shear.data_ptr() if cfg.spray_shear_threshold > 0.0 else None
        a.vel = _lib.ptr3(self.vel)
        a.dx, a.dt, a.delta_phi = cfg.dx, float(dt), cfg.delta_phi
        a.r_mu, a.r_sigma = mu, sigma
        a.r_min, a.r_max = 0.02 * cfg.dx, 2.0 * cfg.dx
        a.spray_threshold, a.spray_rate = cfg.spray_shear_threshold, cfg.spray_emission_rate
        a.gravity = (ctypes.c_float * 3)(*[float(g) for g in cfg.gravity])
        a.rho_g, a.rho_l, a.cd = cfg.rho_g, cfg.rho_l, cfg.droplet_drag_cd
        a.evap_k = cfg.evaporation_k * cfg.dx * cfg.dx
        lo, hi = self._lohi()
        a.lo = (ctypes.c_float * 3)(*lo)
        a.hi = (ctypes.c_float * 3)(*hi)
        a.max_substeps = int(cfg.droplet_max_substeps)
        a.counters = self._scalar.data_ptr() + 16
        return a

    def _launch_droplets(self, dt: float):
        """Shear field (if spray emission is on), then one pf_droplets pass;
        the counters land in the scalar block."""
        cfg = self.cfg
        self._scalar[4:9].zero_()
        if cfg.spray_shear_threshold > 0.0:
            nx, ny, nz = self.res
            _lib.call("pf_velocity_gradient", _lib.ptr3(self.vel), _lib.ptr(self.shear), nx, ny, nz,
                      cfg.dx, _lib.stream())
        a = self._droplet_args(dt)
        self._bounds_token = None
        _lib.call("pf_droplets", ctypes.byref(a), _lib.stream())

    def _add_droplet_stats(self, counters):
        for k, i in (("emitted", 0), ("bubbles", 1), ("merged", 2), ("spray", 3)):
            self.droplet_stats[k] += int(counters[i])

    def _droplets(self, dt: float):
        """Escaped particles -> droplets, droplet dynamics (pfflip/sim.py:343-390)."""
        if not self.cfg.droplets:
            return
        self._launch_droplets(dt)
        self._add_droplet_stats(self._read_scalar_block()[1])

    def rasterize_spray(self):
        """Droplet volume into the FP16 spray field (pfflip/sim.py:392-401)."""
        self.spray_density.zero_()
        sx, sy, sz = self.spray_res
        _lib.call("pf_rasterize_spray", _lib.ptr(self.px), _lib.ptr(self.pstate), _lib.ptr(self.pdrop_r),
                  self.n_particles, _lib.ptr(self.spray_density), sx, sy, sz, float(self.dx_spray),
                  _lib.stream())
        return self.spray_density

    def _advect(self, dt: float):
        """RK3, wall reflection, solid push-out (pfflip/sim.py:408-424)."""
        a = self._g2p_args(_lib.GA_ADVECT | _lib.GA_REFLECT, dt)
        self._bounds_token = None
        _lib.call("pf_g2p_advect", ctypes.byref(a), _lib.stream())

    def state_counts(self):
        st = self.pstate.cpu().numpy()
        return {"flip": int((st == 0).sum()), "droplet": int((st == 1).sum()),
                "dead": int((st == 2).sum())}
