"""State I/O on the GPU: the zero-copy DLPack view of the device state and
the streamed save/restore to a file.  The view must equal the copied-out
state and alias it (writes through torch change the simulator's state); a
saved state must load back bit for bit, in another qubit order and across
precisions (complex128 -> complex64 within FP32 rounding)."""
import numpy as np
import pytest

import hiqsimulator_amd as hq
from hiqsimulator_amd.circuits import apply_circuit_inorder, random_layers_circuit

pytestmark = pytest.mark.gpu


def prepared(n, precision, seed):
    sim = hq.SimulatorMPI(seed, n, 4, precision)
    sim.allocate_qureg(list(range(n)), 0)
    apply_circuit_inorder(sim, random_layers_circuit(n, 6, seed), 4)
    sim.run()
    return sim


VIEW_CHILD = r"""
import sys
import numpy as np
import torch  # first: hiqsimulator_amd then binds to torch's HIP runtime
import hiqsimulator_amd as hq
from hiqsimulator_amd.circuits import apply_circuit_inorder, random_layers_circuit
prec = sys.argv[1]
sim = hq.SimulatorMPI(5, 12, 4, prec)
sim.allocate_qureg(list(range(11)), 0)
apply_circuit_inorder(sim, random_layers_circuit(11, 6, 5), 4)
ref = sim.get_state()
view = hq.state_view(sim)
assert view.is_cuda and tuple(view.shape) == (2 ** 11,)
assert view.dtype == (torch.complex128 if prec == "c128" else torch.complex64)
got = view.cpu().numpy()
assert np.array_equal(got.astype(np.complex128), ref)
view.mul_(-1j)  # writes through the view change the simulator's state
torch.cuda.synchronize()
assert np.array_equal(sim.get_state(), (-1j * got).astype(np.complex128))
del view
sim.apply_controlled_gate([[0, 1], [1, 0]], [0], [])
sim.run()
assert abs(sim.get_probability([], []) - 1.0) < 1e-6
# While a view lives the buffer is pinned: a permutation and a Pauli sum
# (both computed out of place) write their results back into it, so the
# view keeps following the state.
view = hq.state_view(sim)
ptr = view.data_ptr()
order = list(reversed(sim.get_qubits_ids()))
sim.permute_qubits(order)
sim.synchronize()
assert sim.get_qubits_ids() == order
assert np.array_equal(view.cpu().numpy().astype(np.complex128), sim.get_state())
sim.apply_qubit_operator([([(0, "X"), (3, "Z")], 0.5), ([(5, "Y")], 0.25j)], list(range(11)))
sim.synchronize()
assert np.array_equal(view.cpu().numpy().astype(np.complex128), sim.get_state())
assert hq.state_view(sim).data_ptr() == ptr  # a second view shares the pinned buffer
# a new qubit moves the state to a larger buffer; the old one stays allocated
# until the view is gone (the Pauli sum above is not unitary: the norm is
# whatever it left)
before = sim.get_state()
norm = sim.get_probability([], [])
sim.allocate_qubit(11)
sim.synchronize()
_ = float(view.abs().sum().item())
after = sim.get_state()
assert np.array_equal(after[:2 ** 11], before) and not after[2 ** 11:].any()
del view
sim.permute_qubits(list(range(12)))
assert abs(sim.get_probability([], []) - norm) < 1e-5 * norm
print("VIEW OK")
"""


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_dlpack_view_aliases_the_state(engine_mod, precision):
    """In a child process that imports torch first (so both libraries share
    torch's HIP runtime): the view equals the copied-out state and aliases
    it, and it keeps following the state across a permutation and a Pauli
    sum (its buffer is pinned while it lives)."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    print("dlpack view", precision, flush=True)
    env = dict(os.environ, PYTHONPATH=root)
    r = subprocess.run([sys.executable, "-c", VIEW_CHILD, precision], cwd=root, env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "VIEW OK" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]


REFUSE_CHILD = r"""
import hiqsimulator_amd as hq  # first: torch then brings a second HIP runtime
sim = hq.SimulatorMPI(1, 6, 4)
sim.allocate_qureg(list(range(6)), 0)
try:
    hq.state_view(sim)
except RuntimeError as e:
    print("REFUSED", e)
"""


def test_view_refuses_two_runtimes(engine_mod):
    """In a child process that loads hiqsimulator_amd before torch, the view
    refuses instead of handing torch memory of another HIP runtime."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, PYTHONPATH=root)
    r = subprocess.run([sys.executable, "-c", REFUSE_CHILD], cwd=root, env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "REFUSED" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]


def test_save_load_round_trip(engine_mod, tmp_path):
    print("save/load round trip", flush=True)
    path = str(tmp_path / "state.hqsv")
    a = prepared(12, "c128", 9)
    a.permute_qubits([3, 1, 0, 2] + list(range(4, 12)))
    a.save_state(path)
    b = hq.SimulatorMPI(1, 12, 4, "c128")
    b.allocate_qureg(list(range(12)), 0)
    b.load_state(path)
    assert b.get_qubits_ids() == a.get_qubits_ids()
    assert np.array_equal(b.get_state(), a.get_state())
    # complex128 file into a complex64 simulator
    c = hq.SimulatorMPI(1, 12, 4, "c64")
    c.allocate_qureg(list(range(12)), 0)
    c.load_state(path)
    assert np.abs(c.get_state() - a.get_state()).max() < 1e-7
    # and back: complex64 file into complex128 equals the complex64 state
    path64 = str(tmp_path / "state64.hqsv")
    c.save_state(path64)
    d = hq.SimulatorMPI(1, 12, 4, "c128")
    d.allocate_qureg(list(range(12)), 0)
    d.load_state(path64)
    assert np.array_equal(d.get_state(), c.get_state())


def test_load_errors(engine_mod, tmp_path):
    path = str(tmp_path / "s.hqsv")
    a = prepared(6, "c128", 2)
    a.save_state(path)
    b = hq.SimulatorMPI(1, 7, 4, "c128")
    b.allocate_qureg(list(range(7)), 0)
    with pytest.raises(RuntimeError):
        b.load_state(path)  # other qubit count
    c = hq.SimulatorMPI(1, 6, 4, "c128")
    c.allocate_qureg(list(range(10, 16)), 0)
    with pytest.raises(RuntimeError):
        c.load_state(path)  # other qubit ids
    truncated = tmp_path / "short.hqsv"
    truncated.write_bytes(open(path, "rb").read()[:-16])
    before = a.get_state()
    with pytest.raises(RuntimeError):
        a.load_state(str(truncated))
    assert np.array_equal(a.get_state(), before)  # untouched
    bad = tmp_path / "bad.hqsv"
    bad.write_bytes(b"not a state")
    with pytest.raises(RuntimeError):
        a.load_state(str(bad))
    with pytest.raises(RuntimeError):
        a.load_state(str(tmp_path / "missing.hqsv"))
