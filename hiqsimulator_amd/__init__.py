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



def _hip_runtimes():
    """Paths of the HIP runtime libraries mapped into this process."""
    try:
        with open("/proc/self/maps") as f:
            return {line.split()[-1] for line in f if "libamdhip64" in line}
    except OSError:
        return set()


def state_view(sim):
    """``sim``'s state as a torch tensor on its GPU, without a copy
    (DLPack): 2^n complex128 or complex64 amplitudes, index bit p = the
    qubit ``sim.get_qubits_ids()[p]``.  Pending gates are run first.  The
    tensor aliases the simulator's buffer until the qubit count changes;
    writes through it change the state (synchronize torch's stream before
    the next call on ``sim``).

    torch wheels carry their own HIP runtime.  The two libraries share one
    runtime (and so can share device memory) only when torch is imported
    before ``hiqsimulator_amd``; otherwise this raises RuntimeError."""
    import torch
    if len(_hip_runtimes()) > 1:
        raise RuntimeError("state_view(): torch and hiqsimulator_amd loaded different HIP runtimes; "
                           "import torch before hiqsimulator_amd to share one")
    return torch.from_dlpack(sim.state_dlpack())


__all__ = ["SimulatorMPI", "SwapScheduler", "ClusterScheduler", "Batch", "Server", "device_count",
           "KERNEL_CLASSES", "LIB_PATH", "state_view"]
