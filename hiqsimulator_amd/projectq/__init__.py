"""ProjectQ front end for the MI355X engine (drop-in for ``hiq.projectq``).

Mirrors the reference's ``hiq/projectq`` package:
  * ``SimulatorMPI``     backend engine   (backends/_sim/_simulator_mpi.py)
  * ``GreedyScheduler``  compiler engine  (cengines/_greedyscheduler.py)
  * ``HiQMainEngine``    main engine      (cengines/_hiq_main_engine.py)
  * ``MetaSwap``, ``AllocateQuregGate``   (ops/_gates.py)
ProjectQ itself is an external dependency (requirements.txt of the
reference: projectq>=0.4.0); importing this subpackage fails with an
ImportError when it is not installed.  The core engine
(``hiqsimulator_amd.SimulatorMPI``) does not need it.
"""
try:
    import projectq  # noqa: F401
except ImportError as exc:  # pragma: no cover - depends on the environment
    raise ImportError("hiqsimulator_amd.projectq needs ProjectQ (pip install projectq)") from exc

from .ops import AllocateQuregGate, MetaSwap, MetaSwapGate  # noqa: E402
from .backends import SimulatorMPI  # noqa: E402
from .cengines import GreedyScheduler, HiQMainEngine  # noqa: E402

__all__ = ["SimulatorMPI", "GreedyScheduler", "HiQMainEngine", "MetaSwap", "MetaSwapGate",
           "AllocateQuregGate"]
