"""Pins the oracle's entropy (SimulatorMPI::Entropy, SimulatorMPI.cpp:681-694:
-sum_i p_i log2 p_i with p_i = |psi_i|^2, zero probabilities skipped) to
closed forms, CPU only:

* a basis state has entropy 0;
* H on n qubits gives the uniform distribution over 2^n outcomes: n bits;
* a GHZ state has two outcomes of probability 1/2: 1 bit;
* a product of Ry(theta_q)|0> has independent bits with
  P(1) = sin^2(theta_q / 2), so the entropy is the sum of the binary
  entropies H2(sin^2(theta_q / 2)) (Shannon entropy is additive over
  independent variables);
* a state set to probabilities (1/2, 1/4, 1/8, 1/8) has 1.75 bits.

Each closed form is independent of the summation itself.  The planted
mistakes below (natural log instead of log2, |a| instead of |a|^2, a lost
sign) are computed on the oracle's own state and must each fail one of the
closed-form checks, which shows the checks can tell them apart.
"""
import math

import numpy as np
import pytest

from hiqsimulator_amd.circuits import H, X


def ry(theta):
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return np.array([[c, -s], [s, c]], dtype=complex)


def h2(p):
    return sum(-x * math.log2(x) for x in (p, 1.0 - p) if x > 0)


def fresh(oracle_mod, n):
    sim = oracle_mod.OracleSim(1, n, 4)
    sim.allocate_qureg(list(range(n)), 0)
    return sim


def closed_form_cases(oracle_mod):
    """(name, simulator, exact entropy) for every closed form above."""
    cases = []
    sim = fresh(oracle_mod, 5)
    sim.apply_controlled_gate(X, [2], [])
    sim.run()
    cases.append(("basis", sim, 0.0))
    for n in (1, 3, 8):
        sim = fresh(oracle_mod, n)
        for q in range(n):
            sim.apply_controlled_gate(H, [q], [])
            sim.run()
        cases.append(("uniform%d" % n, sim, float(n)))
    sim = fresh(oracle_mod, 6)
    sim.apply_controlled_gate(H, [0], [])
    sim.run()
    for q in range(1, 6):
        sim.apply_controlled_gate(X, [q], [0])
        sim.run()
    cases.append(("ghz6", sim, 1.0))
    rng = np.random.default_rng(17)
    thetas = rng.uniform(0.1, math.pi - 0.1, size=7)
    sim = fresh(oracle_mod, 7)
    for q, t in enumerate(thetas):
        sim.apply_controlled_gate(ry(t), [q], [])
        sim.run()
    cases.append(("ry_product", sim, sum(h2(math.sin(t / 2) ** 2) for t in thetas)))
    sim = fresh(oracle_mod, 2)
    sim.set_state(np.sqrt(np.array([0.5, 0.25, 0.125, 0.125])).astype(complex))
    cases.append(("dyadic", sim, 1.75))
    return cases


def failures(oracle_mod, entropy_of, tol=1e-12):
    return [name for name, sim, want in closed_form_cases(oracle_mod)
            if abs(entropy_of(sim) - want) > tol]


def test_oracle_entropy_closed_forms(oracle_mod):
    assert failures(oracle_mod, lambda s: s.entropy()) == []


def test_ry_product_state_probabilities(oracle_mod):
    """The product state's amplitudes are the outer product of the single
    qubits' (cos, sin) pairs, so its entropy case does test the product."""
    thetas = [0.3, 1.1, 2.0]
    sim = fresh(oracle_mod, 3)
    for q, t in enumerate(thetas):
        sim.apply_controlled_gate(ry(t), [q], [])
        sim.run()
    want = np.ones(1)
    for t in thetas:  # qubit q is index bit q: later qubits vary slowest
        want = np.kron(np.array([math.cos(t / 2), math.sin(t / 2)]), want)
    assert np.abs(sim.get_state() - want).max() < 1e-15


def planted(kind):
    def entropy_of(sim):
        a = np.abs(sim.get_state())
        p = a * a
        nz = p > 0
        if kind == "ln":
            return float(-(p[nz] * np.log(p[nz])).sum())
        if kind == "abs":
            return float(-(a[nz] * np.log2(a[nz])).sum())
        if kind == "sign":
            return float((p[nz] * np.log2(p[nz])).sum())
        raise ValueError(kind)
    return entropy_of


@pytest.mark.parametrize("kind", ["ln", "abs", "sign"])
def test_planted_entropy_mistakes_fail(oracle_mod, kind):
    assert failures(oracle_mod, planted(kind)) != []
