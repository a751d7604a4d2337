"""Pins the oracle's dense/diagonal passes to the reference's own kernels.

* golden vectors: tests/golden/ref_kernels.npz, written by
  tests/golden/make_golden.py from oracle/_ref (the reference's kernel
  headers, src/simulator-mpi/kernels/intrin, compiled unmodified);
* live: oracle/_ref (intrin and nointrin builds) when it is built;
* mathematics: unitarity preserves the norm, identity is a no-op, a
  permutation matrix moves amplitudes exactly, the controlled pass equals a
  dense pass of the block matrix diag(I, U).
"""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_kernels.npz")


def _cases():
    z = np.load(GOLDEN)
    for i in range(int(z["ncases"])):
        p = "c%02d_" % i
        yield (list(z[p + "pos"]), int(z[p + "ctrlmask"]), bool(z[p + "diag"]), z[p + "matrix"],
               z[p + "state"], z[p + "expected"])


def test_golden_vectors(oracle_mod):
    n = 0
    for pos, cm, diag, m, state, expected in _cases():
        got = state.copy()
        (oracle_mod.diag_pass if diag else oracle_mod.dense_pass)(got, m, pos, cm)
        scale = max(1.0, np.abs(expected).max())
        assert np.abs(got - expected).max() <= 1e-14 * scale * m.shape[0], (pos, cm, diag)
        n += 1
    assert n == 40


@pytest.mark.parametrize("flavor", ["intrin", "nointrin"])
def test_against_reference_kernels(oracle_mod, flavor):
    if not oracle_mod.RefKernels.available(flavor):
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    ref = oracle_mod.RefKernels(flavor)
    rng = np.random.default_rng(7)
    for trial in range(60):
        n = int(rng.integers(5, 11))
        k = int(rng.integers(1, 6))
        pos = [int(p) for p in rng.permutation(n)[:k]]
        free = [p for p in range(n) if p not in pos]
        cm = 0
        for p in rng.permutation(free)[: int(rng.integers(0, min(3, len(free)) + 1))]:
            cm |= 1 << int(p)
        diag = bool(rng.random() < 0.3)
        dim = 1 << k
        m = rng.normal(size=(dim, dim)) + 1j * rng.normal(size=(dim, dim))
        if diag:
            m = np.diag(np.diag(m))
        st = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        a = st.copy()
        (oracle_mod.diag_pass if diag else oracle_mod.dense_pass)(a, m, pos, cm)
        b = ref.apply_to(st, m, pos, cm, diag)
        assert np.abs(a - b).max() <= 1e-13 * max(1.0, np.abs(b).max()), (n, k, pos, cm, diag)


def test_identity_and_permutation_matrices(oracle_mod):
    rng = np.random.default_rng(3)
    n = 7
    st = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    for k in range(1, 6):
        pos = [int(p) for p in rng.permutation(n)[:k]]
        a = st.copy()
        oracle_mod.dense_pass(a, np.eye(1 << k), pos)
        assert np.array_equal(a, st)
        # X on matrix bit 0 only: swaps amplitude pairs differing in pos[0].
        perm = np.zeros((1 << k, 1 << k))
        for j in range(1 << k):
            perm[j ^ 1, j] = 1
        a = st.copy()
        oracle_mod.dense_pass(a, perm, pos)
        idx = np.arange(1 << n) ^ (1 << pos[0])
        assert np.array_equal(a, st[idx])


def test_ctrl_mask_equals_block_matrix(oracle_mod):
    """A control bit c on a k-qubit pass == dense (k+1)-qubit pass of diag(I, U)
    with c as the highest matrix bit (fusion_mpi.hpp:167-187)."""
    rng = np.random.default_rng(11)
    n = 8
    st = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    for k in range(1, 5):
        pos = [int(p) for p in rng.permutation(n)[: k + 1]]
        tgt, c = pos[:k], pos[k]
        u = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
        a = st.copy()
        oracle_mod.dense_pass(a, u, tgt, 1 << c)
        big = np.eye(2 << k, dtype=np.complex128)
        big[1 << k:, 1 << k:] = u
        b = st.copy()
        oracle_mod.dense_pass(b, big, tgt + [c])
        assert np.abs(a - b).max() < 1e-12


def test_unitary_preserves_norm(oracle_mod):
    from hiqsimulator_amd.circuits import haar_unitary
    rng = np.random.default_rng(5)
    n = 9
    st = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    st /= np.linalg.norm(st)
    for k in range(1, 6):
        pos = [int(p) for p in rng.permutation(n)[:k]]
        oracle_mod.dense_pass(st, haar_unitary(1 << k, rng), pos)
        assert abs(np.linalg.norm(st) - 1.0) < 1e-13


def test_planted_mistake_is_caught(oracle_mod):
    """A transposed matrix (a plausible bug) must fail the golden check."""
    for pos, cm, diag, m, state, expected in _cases():
        if diag or m.shape[0] < 4:
            continue
        got = state.copy()
        oracle_mod.dense_pass(got, m.T, pos, cm)
        assert np.abs(got - expected).max() > 1e-6
        break
