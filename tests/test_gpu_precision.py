"""Precision of the 32-qubit QFT workload: QFT-like circuit on a random normalised state, in
complex128 and complex64, against FP64 references.

* 26 qubits: the final state against numpy's FFT (the QFT-like circuit is
  the orthonormal inverse DFT with the output index bit-reversed, pinned on
  the oracle in test_oracle_sim.test_qft_closed_form).
* 32 qubits (64 GiB complex128 / 32 GiB complex64): QFT followed by its
  inverse must return the initial state (size-independent property),
  compared on the device in FP64.
Bars: complex128 max-abs <= 1e-12; complex64 fidelity >= 1 - 1e-6, where
fidelity = |<ref|psi>|^2 / (<ref|ref><psi|psi>).
"""
import numpy as np
import pytest

from hiqsimulator_amd.circuits import qft_like_circuit
from hiqsimulator_amd.scheduling import apply_scheduled

pytestmark = pytest.mark.gpu


def inverse(gates):
    return [(m.conj().T, t, c) for (m, t, c) in reversed(gates)]


def test_random_state_properties(engine_mod):
    n = 20
    a = engine_mod.SimulatorMPI(1, n, 4)
    a.allocate_qureg(list(range(n)), 0)
    a.init_random_state(7)
    s1 = a.get_state()
    assert abs(np.linalg.norm(s1) - 1.0) < 1e-12
    a.init_random_state(7)
    assert np.array_equal(a.get_state(), s1)
    a.init_random_state(8)
    assert np.abs(a.get_state() - s1).max() > 1e-6
    # complex Gaussian: E|x|^2 = 1/N, Re/Im uncorrelated with equal variance
    x = s1 * np.sqrt(s1.size)
    assert abs(np.mean(np.abs(x) ** 2) - 1.0) < 1e-2
    assert abs(np.mean(x.real ** 2) - np.mean(x.imag ** 2)) < 2e-2
    assert abs(np.mean(x.real * x.imag)) < 1e-2
    c = engine_mod.SimulatorMPI(1, n, 4, "c64")
    c.allocate_qureg(list(range(n)), 0)
    c.init_random_state(7)
    assert np.abs(c.get_state() - s1).max() < 1e-6


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_qft_26q_against_fft(engine_mod, precision):
    n = 26
    g = engine_mod.SimulatorMPI(1, n, 4, precision)
    g.allocate_qureg(list(range(n)), 0)
    g.init_random_state(2024)
    st = g.get_state()
    apply_scheduled(g, qft_like_circuit(n), 4)
    got = g.get_state()
    ref = np.fft.ifft(st, norm="ortho")
    # output index bit-reversed: reverse the axis order of the 2x...x2 tensor
    got = got.reshape([2] * n).transpose(list(reversed(range(n)))).reshape(-1)
    err = np.abs(got - ref).max()
    fid = abs(np.vdot(ref, got)) ** 2
    print("qft26 %s: max-abs %.3e, 1 - fidelity %.3e" % (precision, err, 1 - fid))
    if precision == "c128":
        assert err <= 1e-12
    else:
        assert fid >= 1 - 1e-6


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_qft_round_trip_32q(engine_mod, precision):
    n = 32
    g = engine_mod.SimulatorMPI(1, n, 4, precision)
    g.allocate_qureg(list(range(n)), 0)
    g.init_random_state(32)
    g.snapshot()
    gates = qft_like_circuit(n)
    apply_scheduled(g, gates, 4)
    apply_scheduled(g, inverse(gates), 4)
    r = g.compare_snapshot()
    fid = r["fidelity"]
    norm_drift = abs(r["state_norm2"] / r["ref_norm2"] - 1.0)
    print("qft32 round trip %s: max-abs %.3e, 1 - fidelity %.3e, norm drift %.3e"
          % (precision, r["max_abs_diff"], 1 - fid, norm_drift))
    if precision == "c128":
        assert r["max_abs_diff"] <= 1e-12
        assert 1 - fid <= 1e-12
    else:
        # fidelity of the rays; the norm itself drifts by the complex64
        # rounding of the fused matrices (about 1e-8 per pass)
        assert fid >= 1 - 1e-6
        assert norm_drift <= 1e-5
