"""hiqsim_sample_noalloc (the sampler entry point that allocates nothing)
against hiqsim_sample on the same states and seeds, through the C ABI.

Both run the same two kernels (block probabilities, then the Philox shot
loop), so the outcomes must be identical index for index.  Two calls in a
row: the second on a smaller state with fewer shots, so it reads the
static scratch array that the first call left behind."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C128, C64 = 0, 1


def _lib(engine_mod):
    lib = ctypes.CDLL(engine_mod.LIB_PATH)
    lib.hiqsim_last_error.restype = ctypes.c_char_p
    lib.hiqsim_sample_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                          ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]
    lib.hiqsim_create.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                  ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    lib.hiqsim_sample.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64,
                                  ctypes.POINTER(ctypes.c_uint64)]
    lib.hiqsim_pin_device_state.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                            ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int),
                                            ctypes.POINTER(ctypes.c_int)]
    lib.hiqsim_unpin_device_state.argtypes = [ctypes.c_void_p]
    lib.hiqsim_destroy.argtypes = [ctypes.c_void_p]
    lib.hiqsim_allocate_qureg.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.c_size_t,
                                          ctypes.c_double, ctypes.c_double]
    lib.hiqsim_init_random_state.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    lib.hiqsim_synchronize.argtypes = [ctypes.c_void_p]
    return lib


def _ok(lib, rc):
    assert rc == 0, lib.hiqsim_last_error()


def _engine(lib, n, prec, state_seed):
    h = ctypes.c_void_p()
    _ok(lib, lib.hiqsim_create(1, n, 4, prec, 0, ctypes.byref(h)))
    ids = (ctypes.c_int64 * n)(*range(n))
    _ok(lib, lib.hiqsim_allocate_qureg(h, ids, n, 0.0, 0.0))
    _ok(lib, lib.hiqsim_init_random_state(h, state_seed))
    _ok(lib, lib.hiqsim_synchronize(h))
    return h


def _hip():
    """The HIP runtime the engine library is linked against (the same
    loaded instance), for the caller-side device buffers of the test."""
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


HOST_TO_DEVICE, DEVICE_TO_HOST = 1, 2


class DeviceArray:
    def __init__(self, hip, nbytes):
        self.hip, self.nbytes, self.ptr = hip, nbytes, ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), nbytes) == 0

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.hip.hipFree(self.ptr)


def _both(lib, hip, n, prec, shots, seed, state_seed):
    """(outcomes of hiqsim_sample, outcomes of hiqsim_sample_noalloc)."""
    h = _engine(lib, n, prec, state_seed)
    try:
        ref = (ctypes.c_uint64 * shots)()
        _ok(lib, lib.hiqsim_sample(h, shots, seed, ref))
        ptr, amps, p, dev = ctypes.c_void_p(), ctypes.c_uint64(), ctypes.c_int(), ctypes.c_int()
        _ok(lib, lib.hiqsim_pin_device_state(h, ctypes.byref(ptr), ctypes.byref(amps), ctypes.byref(p),
                                             ctypes.byref(dev)))
        assert amps.value == 1 << n and p.value == prec and dev.value == 0
        got = np.zeros(shots, dtype=np.uint64)
        with DeviceArray(hip, shots * 8) as out:
            assert hip.hipMemset(out.ptr, 0xFF, shots * 8) == 0
            assert hip.hipDeviceSynchronize() == 0
            print("sample_noalloc: n=%d prec=%d shots=%d seed=%d" % (n, prec, shots, seed), flush=True)
            _ok(lib, lib.hiqsim_sample_noalloc(ptr, n, prec, shots, seed, out.ptr, None))
            assert hip.hipDeviceSynchronize() == 0
            assert hip.hipMemcpy(got.ctypes.data, out.ptr, shots * 8, DEVICE_TO_HOST) == 0
        _ok(lib, lib.hiqsim_unpin_device_state(h))
        return np.ctypeslib.as_array(ref).copy(), got
    finally:
        lib.hiqsim_destroy(h)


@pytest.mark.parametrize("prec", [C128, C64])
def test_noalloc_sampler_matches_engine_sampler(engine_mod, prec):
    lib, hip = _lib(engine_mod), _hip()
    # first call: 22 qubits (2^14 probability blocks), then a smaller state
    # with fewer shots: 13 qubits (2^5 blocks) reads a scratch area the first
    # call filled further
    for n, shots, seed, state_seed in ((22, 4096, 7, 101), (13, 1000, 8, 202)):
        ref, got = _both(lib, hip, n, prec, shots, seed, state_seed)
        assert got.shape == ref.shape
        assert np.array_equal(got, ref), (n, int(np.sum(got != ref)))
        assert int(got.max()) < (1 << n)


def _hip_graph_api(hip):
    vp, pvp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)
    hip.hipStreamCreate.argtypes = [pvp]
    hip.hipStreamDestroy.argtypes = [vp]
    hip.hipStreamSynchronize.argtypes = [vp]
    hip.hipStreamBeginCapture.argtypes = [vp, ctypes.c_int]
    hip.hipStreamEndCapture.argtypes = [vp, pvp]
    hip.hipGraphInstantiate.argtypes = [pvp, vp, vp, ctypes.c_char_p, ctypes.c_size_t]
    hip.hipGraphLaunch.argtypes = [vp, vp]
    hip.hipGraphExecDestroy.argtypes = [vp]
    hip.hipGraphDestroy.argtypes = [vp]


def test_noalloc_sampler_under_graph_capture(engine_mod):
    """The entry point captured into a HIP graph on the caller's stream (the
    use it is for: no allocation may happen while a stream is captured).
    Three replays, with an eager call for another state size and seed
    between them (it leaves other contents in the shared scratch array),
    each give the outcomes of hiqsim_sample."""
    lib, hip = _lib(engine_mod), _hip()
    _hip_graph_api(hip)
    n, shots, seed = 16, 2048, 5
    h = _engine(lib, n, C128, 303)
    h2 = _engine(lib, 12, C128, 404)
    stream, graph, gexec = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    try:
        ref = (ctypes.c_uint64 * shots)()
        _ok(lib, lib.hiqsim_sample(h, shots, seed, ref))
        ref = np.ctypeslib.as_array(ref).copy()
        ref2 = (ctypes.c_uint64 * 300)()
        _ok(lib, lib.hiqsim_sample(h2, 300, 9, ref2))
        ref2 = np.ctypeslib.as_array(ref2).copy()
        ptr, ptr2 = ctypes.c_void_p(), ctypes.c_void_p()
        amps, p, dev = ctypes.c_uint64(), ctypes.c_int(), ctypes.c_int()
        _ok(lib, lib.hiqsim_pin_device_state(h, ctypes.byref(ptr), ctypes.byref(amps), ctypes.byref(p),
                                             ctypes.byref(dev)))
        _ok(lib, lib.hiqsim_pin_device_state(h2, ctypes.byref(ptr2), ctypes.byref(amps), ctypes.byref(p),
                                             ctypes.byref(dev)))
        assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
        with DeviceArray(hip, shots * 8) as out, DeviceArray(hip, 300 * 8) as out2:
            # an eager call first: the entry point looks up its scratch
            # array's address on its first call on a device
            _ok(lib, lib.hiqsim_sample_noalloc(ptr, n, C128, shots, seed, out.ptr, stream))
            assert hip.hipStreamSynchronize(stream) == 0
            print("sample_noalloc: capture n=%d shots=%d" % (n, shots), flush=True)
            assert hip.hipStreamBeginCapture(stream, 0) == 0  # hipStreamCaptureModeGlobal
            rc = lib.hiqsim_sample_noalloc(ptr, n, C128, shots, seed, out.ptr, stream)
            assert hip.hipStreamEndCapture(stream, ctypes.byref(graph)) == 0
            _ok(lib, rc)
            assert hip.hipGraphInstantiate(ctypes.byref(gexec), graph, None, None, 0) == 0
            got, got2 = np.zeros(shots, dtype=np.uint64), np.zeros(300, dtype=np.uint64)
            for replay in range(3):
                print("sample_noalloc: replay %d" % replay, flush=True)
                assert hip.hipMemset(out.ptr, 0xFF, shots * 8) == 0
                assert hip.hipDeviceSynchronize() == 0
                assert hip.hipGraphLaunch(gexec, stream) == 0
                assert hip.hipStreamSynchronize(stream) == 0
                assert hip.hipMemcpy(got.ctypes.data, out.ptr, shots * 8, DEVICE_TO_HOST) == 0
                assert np.array_equal(got, ref), (replay, int(np.sum(got != ref)))
                # eager work between the replays
                _ok(lib, lib.hiqsim_sample_noalloc(ptr2, 12, C128, 300, 9, out2.ptr, None))
                assert hip.hipDeviceSynchronize() == 0
                assert hip.hipMemcpy(got2.ctypes.data, out2.ptr, 300 * 8, DEVICE_TO_HOST) == 0
                assert np.array_equal(got2, ref2), replay
        _ok(lib, lib.hiqsim_unpin_device_state(h))
        _ok(lib, lib.hiqsim_unpin_device_state(h2))
    finally:
        if gexec.value:
            hip.hipGraphExecDestroy(gexec)
        if graph.value:
            hip.hipGraphDestroy(graph)
        if stream.value:
            hip.hipStreamDestroy(stream)
        lib.hiqsim_destroy(h)
        lib.hiqsim_destroy(h2)


def test_noalloc_sampler_rejects_out_of_range(engine_mod):
    lib, hip = _lib(engine_mod), _hip()
    with DeviceArray(hip, 32) as state, DeviceArray(hip, 4 * 8) as out:
        rc = lib.hiqsim_sample_noalloc(state.ptr, 36, C128, 4, 1, out.ptr, None)
        assert rc != 0 and b"n_qubits" in lib.hiqsim_last_error()
        rc = lib.hiqsim_sample_noalloc(state.ptr, 1, C128, (1 << 20) + 1, 1, out.ptr, None)
        assert rc != 0 and b"2^20" in lib.hiqsim_last_error()
        rc = lib.hiqsim_sample_noalloc(None, 1, C128, 4, 1, out.ptr, None)
        assert rc != 0 and b"d_state" in lib.hiqsim_last_error()
        rc = lib.hiqsim_sample_noalloc(state.ptr, 1, 5, 4, 1, out.ptr, None)
        assert rc != 0 and b"precision" in lib.hiqsim_last_error()
        # a valid call after the rejected ones: |1> in two complex128 amplitudes
        one = np.array([0, 0, 1, 0], dtype=np.float64)
        assert hip.hipMemcpy(state.ptr, one.ctypes.data, 32, HOST_TO_DEVICE) == 0
        _ok(lib, lib.hiqsim_sample_noalloc(state.ptr, 1, C128, 4, 3, out.ptr, None))
        assert hip.hipDeviceSynchronize() == 0
        got = np.zeros(4, dtype=np.uint64)
        assert hip.hipMemcpy(got.ctypes.data, out.ptr, 32, DEVICE_TO_HOST) == 0
        assert got.tolist() == [1, 1, 1, 1]
