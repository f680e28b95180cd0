"""MI355X (gfx950) embedded Runge-Kutta integration of N-body gravity.

Host-side Python mirror of the reference's interfaces
(noether17/ode_prototyping): an execution policy (:class:`HipExecutor`, the
counterpart of ``src/par_exec/CudaExecutor.cuh``), device state containers
(:class:`HipState`, ``src/state/CudaState.cuh``), ODE functors
(:class:`NBodyODE`, :class:`ExponentialODE`, :class:`VanDerPolODE`,
``src/ode``), output policies (:class:`RawOutput`, ``src/output/RawOutput.hpp``;
DOP853 / DOPRI5 :class:`DenseOutput`, ``NumericalRecipes/Output.hpp``)
and the integrator (:func:`integrate`, ``RKEmbeddedParallel<BT>::integrate``,
``src/integrator/RKEmbeddedParallel/RKEmbeddedParallel.hpp:13-81``), plus
the leapfrog integrator (:func:`leapfrog`, ``LeapfrogParallel.hpp``), batched
ensembles (:mod:`.ensemble`), the learned step-size controller
(:mod:`.controller`, :mod:`.train_controller`) and trajectory files
(:mod:`.nbody_io`).  All numerical work runs in libodehip.so through its C ABI
(include/odehip.h).
"""
from ._lib import (CONTROLLERS, MIXED_NEAR_CLOSE, MIXED_NEAR_DEFAULT, NBODY_VARIANTS,
                   REDUCE_GROUP, STAGES, STATUS, TABLEAUS, lib)
from .core import (ButcherTableau, DenseOutput, ExponentialODE, HipExecutor, HipState,
                   MostRecentOutput, NBodyODE, OdeHipError, RawOutput, SolveResult,
                   VanDerPolODE, integrate, integrate_host_driven, leapfrog, tableau_id)
from . import nbody_io, scenarios

__all__ = [
    "HipExecutor", "HipState", "NBodyODE", "ExponentialODE", "VanDerPolODE", "ButcherTableau",
    "tableau_id",
    "RawOutput", "MostRecentOutput", "DenseOutput", "SolveResult", "integrate",
    "integrate_host_driven", "leapfrog", "scenarios", "nbody_io", "TABLEAUS", "STAGES", "STATUS",
    "NBODY_VARIANTS", "REDUCE_GROUP", "lib",
]
