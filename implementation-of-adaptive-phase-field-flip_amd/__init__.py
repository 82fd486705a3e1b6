"""MI355X-native PF-FLIP time step (drop-in for the reference ``pfflip`` package).

Exports ``Config``, ``Simulator``, ``MGPCG``, ``apply_A`` and ``scenes`` with
the reference's names and semantics (pfflip/__init__.py:7-10,
pfflip/poisson.py:78,240), plus ``rollout`` / ``StepAdjoint``: steps
differentiable through a hand-written adjoint (``adjoint.py``), and the
multiresolution sparse block grid's Poisson stack ``MSBG`` /
``MRMultigrid`` / ``MRSolver`` and the adaptive simulator
``AdaptiveSimulator`` (the reference keeps them in their own modules,
pfflip/msbg.py, mr_multigrid.py, mr_solver.py, mr_sim.py, and so does this
package; they are re-exported here too), frame output (``io``) and the
command line (``run``).  Compute
runs in hand-written gfx950 HIP kernels
(``csrc/``, built into ``libpfflip_hip.so`` and bound with ctypes); there is
no CPU fallback.
"""

from . import scenes
from .adjoint import StepAdjoint, rollout
from .config import Config
from . import io
from .mr_multigrid import MRMultigrid
from .mr_sim import AdaptiveSimulator
from .mr_solver import MRSolver
from .msbg import MSBG
from .poisson import MGPCG, apply_A
from .sim import Simulator


def build(verbose: bool = False) -> str:
    """Compile the HIP library for gfx950 (hipcc; no GPU needed)."""
    from .build import build as _build
    return _build(verbose=verbose)


__all__ = ["Config", "Simulator", "MGPCG", "apply_A", "scenes", "io", "build", "rollout", "StepAdjoint", "MSBG",
           "MRMultigrid", "MRSolver", "AdaptiveSimulator"]
