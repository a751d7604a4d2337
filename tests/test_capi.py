"""CPU checks of the drop-in boundary: the C-ABI library loads and exports
every entry point include/hiqsim_amd.h declares, and the pybind module
exposes every method of the reference's `_cppsim_mpi.SimulatorMPI`
(_cppsim_mpi.cpp:63-82).  No compute calls (no GPU here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "hiqsim_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hiqsim_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import hiqsimulator_amd
    lib = ctypes.CDLL(hiqsimulator_amd.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 29
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_error_path_without_device():
    import hiqsimulator_amd
    lib = ctypes.CDLL(hiqsimulator_amd.LIB_PATH)
    lib.hiqsim_last_error.restype = ctypes.c_char_p
    h = ctypes.c_void_p()
    rc = lib.hiqsim_create(ctypes.c_uint64(1), ctypes.c_int64(10), ctypes.c_int64(4), 7, 0,
                           ctypes.byref(h))
    assert rc != 0 and b"precision" in lib.hiqsim_last_error()
    rc = lib.hiqsim_run(None)
    assert rc != 0 and b"null" in lib.hiqsim_last_error()


def test_pybind_surface_matches_reference():
    import hiqsimulator_amd
    ref_methods = ["get_qubits_ids", "get_local_qubits_ids", "get_global_qubits_ids",
                   "set_qubits_perm", "swap_qubits", "allocate_qureg", "allocate_qubit",
                   "deallocate_qubit", "measure_qubits", "apply_controlled_gate", "emulate_math",
                   "get_amplitude", "get_probability", "run", "entropy", "cheat_local",
                   "collapse_wavefunction"]
    extra = ["get_expectation_value", "apply_qubit_operator", "set_wavefunction",
             "permute_qubits", "get_state", "set_state", "state_dlpack", "save_state", "load_state",
             "emulate_time_evolution"]
    for name in ref_methods + extra:
        assert hasattr(hiqsimulator_amd.SimulatorMPI, name), name


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "hiqsimulator_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "from oracle" not in text, f
                assert "liboracle" not in text and "_ref/" not in text, f


def test_tuning_parameters_round_trip():
    """Every tuning parameter reads back what was set; out-of-range values
    and unknown names raise and leave the setting as it was."""
    import pytest
    from hiqsimulator_amd import _cppsim_amd as m
    ranges = {"tile_high_target": (0, 64), "block_swizzle": (0, 16),
              "block_swizzle_wide": (0, 16), "tile_mfma": (0, 1)}
    for name, (lo, hi) in ranges.items():
        old = m.get_tuning(name)
        try:
            for v in (lo, hi):
                m.set_tuning(name, v)
                assert m.get_tuning(name) == v
            for bad in (lo - 1, hi + 1):
                with pytest.raises(RuntimeError):
                    m.set_tuning(name, bad)
                assert m.get_tuning(name) == hi
        finally:
            m.set_tuning(name, old)
    with pytest.raises(RuntimeError):
        m.get_tuning("no_such_knob")


def test_enumerated_tuning_parameters():
    """Knobs with a set of values: each value reads back, a value outside
    the set raises and leaves the setting as it was."""
    import pytest
    from hiqsimulator_amd import _cppsim_amd as m
    sets = {"tile_threads": (0, 256, 512), "direct_block": (256, 512, 1024), "permute_block": (256, 512)}
    for name, values in sets.items():
        old = m.get_tuning(name)
        try:
            for v in values:
                m.set_tuning(name, v)
                assert m.get_tuning(name) == v
            for bad in (-1, 300, 2048):
                with pytest.raises(RuntimeError):
                    m.set_tuning(name, bad)
                assert m.get_tuning(name) == values[-1]
        finally:
            m.set_tuning(name, old)
