"""Kernel-level parity: each HIP pass vs the oracle on the same seeded input.

complex128: max-abs <= 1e-12 (scaled by the state's magnitude).
complex64:  max-abs relative to the FP64 oracle <= 2e-6 per pass.
Permutation: bit-exact.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rand_state(rng, n):
    st = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return st / np.linalg.norm(st)


def rand_case(rng, n, k, nctrl, diag):
    pos = [int(p) for p in rng.permutation(n)[:k]]
    free = [p for p in range(n) if p not in pos]
    cm = 0
    for p in rng.permutation(free)[:nctrl]:
        cm |= 1 << int(p)
    dim = 1 << k
    m = rng.normal(size=(dim, dim)) + 1j * rng.normal(size=(dim, dim))
    if diag:
        m = np.diag(np.diag(m))
    return pos, cm, m


def engine_with_state(eng, st, precision="c128"):
    n = int(np.log2(st.size))
    sim = eng.SimulatorMPI(1, 40, 5, precision)
    sim.allocate_qureg(list(range(n)), 0)
    sim.set_state(st)
    return sim


@pytest.mark.parametrize("n", [1, 2, 3, 5, 6, 7, 9, 12, 14, 17])
def test_dense_and_diag_c128(engine_mod, oracle_mod, n):
    rng = np.random.default_rng(100 + n)
    for k in range(1, min(5, n) + 1):
        for nctrl in sorted({0, min(1, n - k), min(2, n - k)}):
            for diag in (False, True):
                print("case n=%d k=%d nctrl=%d diag=%d" % (n, k, nctrl, diag), flush=True)
                pos, cm, m = rand_case(rng, n, k, nctrl, diag)
                st = rand_state(rng, n)
                sim = engine_with_state(engine_mod, st)
                sim.apply_matrix(m, pos, cm, diag)
                got = sim.get_state()
                ref = st.copy()
                (oracle_mod.diag_pass if diag else oracle_mod.dense_pass)(ref, m, pos, cm)
                err = np.abs(got - ref).max()
                assert err <= 1e-12 * max(1.0, np.abs(ref).max()), (n, k, pos, cm, diag, err)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_dense_positions_sweep_20q(engine_mod, oracle_mod, k):
    """All target placements that matter for coalescing: low, mixed, high."""
    n = 20
    rng = np.random.default_rng(7 * k)
    placements = [list(range(k)), list(range(n - k, n)), [int(p) for p in rng.permutation(n)[:k]],
                  [0] + list(range(n - k + 1, n))]
    st = rand_state(rng, n)
    for pos in placements:
        print("case k=%d pos=%s" % (k, pos), flush=True)
        m = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
        sim = engine_with_state(engine_mod, st)
        sim.apply_matrix(m, pos, 0, False)
        got = sim.get_state()
        ref = st.copy()
        oracle_mod.dense_pass(ref, m, pos, 0)
        assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), pos


@pytest.mark.parametrize("n", [4, 11, 16])
def test_dense_and_diag_c64(engine_mod, oracle_mod, n):
    from hiqsimulator_amd.circuits import haar_unitary
    rng = np.random.default_rng(200 + n)
    for k in range(1, min(5, n) + 1):
        for diag in (False, True):
            pos, cm, _ = rand_case(rng, n, k, 1 if n > k else 0, diag)
            m = haar_unitary(1 << k, rng)
            if diag:
                m = np.diag(np.exp(1j * rng.uniform(0, 6.28, size=1 << k)))
            st = rand_state(rng, n)
            sim = engine_with_state(engine_mod, st, "c64")
            sim.apply_matrix(m, pos, cm, diag)
            got = sim.get_state()
            ref = st.copy()
            (oracle_mod.diag_pass if diag else oracle_mod.dense_pass)(ref, m, pos, cm)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            assert err <= 2e-6, (n, k, diag, err)


@pytest.mark.parametrize("knobs", [(8, 256), (0, 256), (6, 512)])
def test_permutation_knobs_bit_exact(engine_mod, knobs):
    """The tiled permutation with and without its block swizzle and in both
    workgroup sizes: the result is the closed form out[sigma(i)] = in[i]."""
    import hiqsimulator_amd
    mod = hiqsimulator_amd._cppsim_amd
    old = (mod.get_tuning("permute_swizzle"), mod.get_tuning("permute_block"))
    try:
        mod.set_tuning("permute_swizzle", knobs[0])
        mod.set_tuning("permute_block", knobs[1])
        for precision in ("c128", "c64"):
            test_permutation_bit_exact(engine_mod, 20, precision)
    finally:
        mod.set_tuning("permute_swizzle", old[0])
        mod.set_tuning("permute_block", old[1])


@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("n", [1, 3, 8, 15, 21])
def test_permutation_bit_exact(engine_mod, n, precision):
    rng = np.random.default_rng(300 + n)
    st = rand_state(rng, n)
    if precision == "c64":
        st = st.astype(np.complex64).astype(np.complex128)
    sim = engine_with_state(engine_mod, st, precision)
    new = [int(x) for x in rng.permutation(n)]
    sim.permute_qubits(new)
    got = sim.get_state()
    idx = np.zeros(1 << n, dtype=np.int64)
    i = np.arange(1 << n, dtype=np.int64)
    for p, q in enumerate(new):
        idx |= ((i >> q) & 1) << p
    expect = np.empty_like(st)
    expect[idx] = st
    assert np.array_equal(got, expect)
    assert sim.get_qubits_ids() == new


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_dense_forms_22q(engine_mod, oracle_mod, precision):
    """Exercises both forms of the dense pass on a 22-qubit state: the LDS
    tile form (targets low or mixed), the direct form (all targets >= 20)
    and the direct form forced by a control bit below 6."""
    n = 22
    rng = np.random.default_rng(22)
    st = rand_state(rng, n)
    cases = [([20], 0), ([21, 20], 0), ([0, 1, 2, 3, 4], 0), ([5, 21, 0], 0), ([20, 21], 1 << 3),
             ([7, 13], 1 << 2), ([1, 0, 9, 4], (1 << 20) | (1 << 15)), ([21], (1 << 0) | (1 << 19))]
    tol = 1e-12 if precision == "c128" else 3e-6
    for pos, cm in cases:
        print("case pos=%s ctrl=%#x" % (pos, cm), flush=True)
        k = len(pos)
        m = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
        m /= np.linalg.norm(m, 2)
        sim = engine_with_state(engine_mod, st, precision)
        sim.apply_matrix(m, pos, cm, False)
        got = sim.get_state()
        ref = st.astype(np.complex64).astype(np.complex128) if precision == "c64" else st.copy()
        oracle_mod.dense_pass(ref, m, pos, cm)
        err = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300)
        assert err <= tol, (pos, cm, err)


@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("route", [1, 0])
def test_dense5_tile_route(engine_mod, oracle_mod, route, precision):
    """5-qubit passes without controls run as a one-op tile pass
    (dense5_tile = 1, the default) or in the dense forms (0): targets in any
    order and placement on 12 and 16 qubits; a pass with a control keeps
    the dense forms either way.  Relative error <= 1e-12 (complex128) and
    2e-6 (complex64)."""
    import hiqsimulator_amd
    mod = hiqsimulator_amd._cppsim_amd
    old = mod.get_tuning("dense5_tile")
    tol = 1e-12 if precision == "c128" else 2e-6
    rng = np.random.default_rng(55 + route)
    cases = [(12, [0, 1, 2, 3, 4], 0), (12, [11, 3, 7, 0, 9], 0), (16, [15, 14, 13, 12, 11], 0),
             (16, [8, 2, 15, 6, 10], 0), (16, [1, 3, 5, 7, 9], 1 << 12), (9, [8, 0, 4, 2, 6], 0)]
    try:
        mod.set_tuning("dense5_tile", route)
        for n, pos, cm in cases:
            print("dense5 %s route=%d n=%d pos=%s ctrl=%#x" % (precision, route, n, pos, cm), flush=True)
            st = rand_state(rng, n)
            m = rng.normal(size=(32, 32)) + 1j * rng.normal(size=(32, 32))
            m /= np.linalg.norm(m, 2)
            sim = engine_with_state(engine_mod, st, precision)
            sim.apply_matrix(m, pos, cm, False)
            ref = st.copy()
            oracle_mod.dense_pass(ref, m, pos, cm)
            err = np.abs(sim.get_state() - ref).max() / max(np.abs(ref).max(), 1e-300)
            assert err <= tol, (n, pos, cm, err)
    finally:
        mod.set_tuning("dense5_tile", old)


@pytest.mark.parametrize("block", [256, 512, 1024])
def test_direct_block_sizes(engine_mod, oracle_mod, block):
    """4-qubit passes in the direct form (tile_high_target_wide = 6 sends
    every 4-qubit pass with all targets at bit 6 or above there) with 256-,
    512- and 1024-thread blocks: whole and partial grids, controls, both
    precisions."""
    import hiqsimulator_amd
    mod = hiqsimulator_amd._cppsim_amd
    old = (mod.get_tuning("direct_block"), mod.get_tuning("tile_high_target_wide"))
    rng = np.random.default_rng(91 + block)
    try:
        mod.set_tuning("direct_block", block)
        mod.set_tuning("tile_high_target_wide", 6)
        for n in (10, 13, 18):
            for trial in range(3):
                pos = [int(p) for p in 6 + rng.permutation(n - 6)[:4]]
                free = [q for q in range(n) if q not in pos]
                cm = 0
                for q in rng.permutation(free)[:trial]:
                    cm |= 1 << int(q)
                for precision in ("c128", "c64"):
                    print("direct_block=%d n=%d pos=%s ctrl=%#x %s" % (block, n, pos, cm, precision), flush=True)
                    st = rand_state(rng, n)
                    m = rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16))
                    m /= np.linalg.norm(m, 2)
                    sim = engine_with_state(engine_mod, st, precision)
                    sim.apply_matrix(m, pos, cm, False)
                    ref = st.astype(np.complex64).astype(np.complex128) if precision == "c64" else st.copy()
                    oracle_mod.dense_pass(ref, m, pos, cm)
                    err = np.abs(sim.get_state() - ref).max() / max(np.abs(ref).max(), 1e-300)
                    assert err <= (3e-6 if precision == "c64" else 1e-12), (n, pos, cm, precision, err)
        with pytest.raises(RuntimeError):
            mod.set_tuning("direct_block", 300)
    finally:
        mod.set_tuning("direct_block", old[0])
        mod.set_tuning("tile_high_target_wide", old[1])


@pytest.mark.parametrize("swizzle,high", [(0, 20), (4, 20), (8, 64), (10, 64), (16, 0), (0, 0)])
def test_block_order_and_form_knobs(engine_mod, oracle_mod, swizzle, high):
    """Block-order swizzle and the tile/direct threshold change only the
    order and form of the work: results equal the oracle at every setting.
    The threshold is set for 4-qubit passes too (tile_high_target_wide), so
    with high = 0 the 4-qubit passes at low placements take the direct form,
    swizzled and not."""
    import hiqsimulator_amd
    mod = hiqsimulator_amd._cppsim_amd
    old = (mod.get_tuning("block_swizzle"), mod.get_tuning("tile_high_target"),
           mod.get_tuning("block_swizzle_wide"), mod.get_tuning("tile_high_target_wide"))
    n = 22
    rng = np.random.default_rng(7 + swizzle)
    st = rand_state(rng, n)
    try:
        mod.set_tuning("block_swizzle", swizzle)
        mod.set_tuning("block_swizzle_wide", swizzle)
        mod.set_tuning("tile_high_target", high)
        mod.set_tuning("tile_high_target_wide", high)
        print("swizzle", swizzle, "high", high, flush=True)
        for pos in ([0, 1, 2, 3], [18, 19, 20, 21], [2, 9, 15, 21], [20], [0, 5, 10, 15, 20]):
            k = len(pos)
            m = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
            m /= np.linalg.norm(m, 2)
            sim = engine_with_state(engine_mod, st, "c128")
            sim.apply_matrix(m, pos, 0, False)
            ref = st.copy()
            oracle_mod.dense_pass(ref, m, pos, 0)
            assert np.abs(sim.get_state() - ref).max() <= 1e-12 * np.abs(ref).max(), pos
    finally:
        mod.set_tuning("block_swizzle", old[0])
        mod.set_tuning("tile_high_target", old[1])
        mod.set_tuning("block_swizzle_wide", old[2])
        mod.set_tuning("tile_high_target_wide", old[3])
