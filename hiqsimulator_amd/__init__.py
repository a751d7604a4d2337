"""MI355X-native state-vector engine with HiQsimulator's SimulatorMPI surface.

``hiqsimulator_amd.SimulatorMPI`` is the raw engine class, method for method
the reference's ``_cppsim_mpi.SimulatorMPI`` (_cppsim_mpi.cpp:61-83).  It runs
hand-written gfx950 HIP kernels behind the C ABI in include/hiqsim_amd.h.
There is no CPU fallback: importing fails if the native library is missing.
"""
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhiqsim_amd.so")

try:
    from . import _cppsim_amd  # noqa: F401  (links libhiqsim_amd.so)
except ImportError as exc:  # pragma: no cover - build problem, reported loudly
    raise ImportError(
        "hiqsimulator_amd: the native engine is not built (%s). Run "
        "`python hiqsimulator_amd/_build.py` or __graft_entry__.build()." % exc) from exc

SimulatorMPI = _cppsim_amd.SimulatorMPI
device_count = _cppsim_amd.device_count
# The reference's `_sched_cpp` classes (_sched_cpp.cpp:28-40).
SwapScheduler = _cppsim_amd.SwapScheduler
ClusterScheduler = _cppsim_amd.ClusterScheduler
# Many independent circuits in one graph launch, and the request loop with
# continuous batching (no reference counterparts).
Batch = _cppsim_amd.Batch
Server = _cppsim_amd.Server

# Kernel classes reported by SimulatorMPI.kernel_time(cls).
KERNEL_CLASSES = ("dense_k1", "dense_k2", "dense_k3", "dense_k4", "dense_k5", "diag", "scale",
                  "permute", "reduce", "normalize", "remove_bit", "fill_copy", "pauli", "tile")

__all__ = ["SimulatorMPI", "SwapScheduler", "ClusterScheduler", "Batch", "Server", "device_count",
           "KERNEL_CLASSES", "LIB_PATH"]
