"""Writes tests/golden/ref_kernels.npz from the reference's own kernels.

Run in the container that holds the reference checkout, after
`make -C oracle` has built oracle/_ref (the reference's kernel headers
src/simulator-mpi/kernels/{intrin,nointrin} compiled unmodified):

    python tests/golden/make_golden.py

Each case: a random 8-qubit complex128 state and matrix (seeded), the bit
positions (matrix bit l -> pos[l]), a control mask and the diag flag; the
expected output is what the reference's AVX2 build (kernels/intrin, the
default -DINTRIN configuration) produces when dispatched as
SimulatorMPI::Run does (SimulatorMPI.cpp:470-514).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def main():
    ref = oracle.RefKernels("intrin")
    rng = np.random.default_rng(20240611)
    n = 8
    cases = []
    for k in range(1, 6):
        for variant in range(8):
            pos = [int(p) for p in rng.permutation(n)[:k]]
            free = [p for p in range(n) if p not in pos]
            nctrl = [0, 0, 1, 2, 0, 1, 3, 0][variant]
            cm = 0
            for p in rng.permutation(free)[:nctrl]:
                cm |= 1 << int(p)
            diag = variant in (4, 5)
            dim = 1 << k
            m = rng.normal(size=(dim, dim)) + 1j * rng.normal(size=(dim, dim))
            if diag:
                m = np.diag(np.diag(m))
            state = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
            out = ref.apply_to(state, m, pos, cm, diag)
            cases.append((k, pos, cm, diag, m, state, out))
    arrs = {}
    for i, (k, pos, cm, diag, m, state, out) in enumerate(cases):
        arrs["c%02d_pos" % i] = np.array(pos, dtype=np.uint32)
        arrs["c%02d_ctrlmask" % i] = np.array(cm, dtype=np.uint64)
        arrs["c%02d_diag" % i] = np.array(int(diag), dtype=np.int8)
        arrs["c%02d_matrix" % i] = m.astype(np.complex128)
        arrs["c%02d_state" % i] = state.astype(np.complex128)
        arrs["c%02d_expected" % i] = out.astype(np.complex128)
    arrs["ncases"] = np.array(len(cases))
    np.savez_compressed(os.path.join(HERE, "ref_kernels.npz"), **arrs)
    print("wrote", len(cases), "cases")


if __name__ == "__main__":
    main()
