"""Simulation parameters (drop-in for pfflip/config.py:14-135).

Every field and default of the reference ``Config`` is kept.  Fields added
here default to the reference's behaviour.
"""

from __future__ import annotations

import math
from dataclasses import dataclass


@dataclass
class Config:
    # ---- domain / discretisation (config.py:17-24)
    res: tuple = (64, 64, 64)
    dx: float = 1.0
    ppc: int = 8
    # ---- physics (config.py:27-33)
    rho_l: float = 1000.0
    density_ratio: float = 1000.0
    gravity: tuple = (0.0, -9.81, 0.0)
    # ---- phase field, eq. (7) (config.py:36)
    alpha_phi: float = 1.0
    # ---- FLIP blending, eq. (12) (config.py:41-42)
    alpha_flip_liquid: float = 0.985
    alpha_flip_air: float = 0.96
    # ---- time stepping (config.py:47-52)
    cfl: float = 1.0
    dt_max: float = 1.0 / 24.0
    dt_min: float = 1e-6
    # ---- divergence correction (config.py:55)
    k_div: float = 0.0
    # ---- pressure solve (config.py:60-73)
    pcg_tol: float = 1e-4
    pcg_max_iter: int = 200
    mg_pre_sweeps: int = 3
    mg_post_sweeps: int = 3
    mg_coarsest_sweeps: int = 60
    mg_omega: float = 6.0 / 7.0
    mg_min_dim: int = 4
    # ---- velocity extrapolation (config.py:77)
    extrap_cells: int = 3
    # ---- escaped particles / spray (config.py:82-110): the droplet stage of
    # Simulator.step (csrc/pf_droplets.hip)
    droplets: bool = True
    delta_phi: float = 0.2
    droplet_radius_mean: float = 0.25
    droplet_radius_spread: float = 0.5
    droplet_drag_cd: float = 0.5
    evaporation_k: float = 0.0
    spray_shear_threshold: float = 0.0
    spray_emission_rate: float = 0.02
    spray_resolution_factor: int = 2
    droplet_max_substeps: int = 8
    # ---- runtime (config.py:113-114)
    device: str = "cuda:0"
    seed: int = 42
    # ---- added: storage precision of the multigrid V-cycle ("fp32" is the
    # reference's arithmetic; "bf16" / "fp16" store stencil coefficients and
    # smoother iterates in 16 bits and accumulate in FP32)
    solver_precision: str = "fp32"

    # ---- derived (config.py:117-135)
    @property
    def rho_g(self) -> float:
        return self.rho_l / self.density_ratio

    @property
    def r_p(self) -> float:
        return self.dx

    @property
    def eta_phi(self) -> float:
        return math.log(self.density_ratio)

    @property
    def domain_size(self) -> tuple:
        return (self.res[0] * self.dx, self.res[1] * self.dx, self.res[2] * self.dx)
