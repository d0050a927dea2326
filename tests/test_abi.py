"""The C ABI library (CPU only: loads, exports what include/revsim.h declares, validates input).
No compute call runs here: there is no GPU in this environment."""
import ctypes

import pytest


def test_library_exports_every_declared_symbol(gpu_lib):
    import revsim

    declared = revsim._abi.declared_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(gpu_lib, name), name
    bound = {n for n, _, _ in revsim._abi.SIGNATURES}
    assert set(declared) == bound


def test_config_layout_matches(gpu_lib):
    import revsim

    assert gpu_lib.rs_config_size() == ctypes.sizeof(revsim.RsConfig)
    assert gpu_lib.rs_version() == revsim._abi.RS_ABI_VERSION


def test_default_config_round_trip(gpu_lib):
    import revsim

    c = revsim.RsConfig()
    gpu_lib.rs_default_config(ctypes.byref(c))
    assert c.abi_version == revsim._abi.RS_ABI_VERSION and c.n_molecules == 216 and c.rdf_bins == 200
    assert c.cutoff == 0.8 and c.eps_rf == 78.3
    # enthalpy-of-vaporization loss off by default, reference constants (water/water.jl:741,751)
    assert c.w_dhvap == 0.0 and c.dhvap_u_gas == 3.69 and abs(c.dhvap_target - 46.92) < 1e-12
    m = revsim.make_config(216, 1.862, 0.8)
    assert (m.w_dhvap, m.dhvap_u_gas, m.dhvap_target) == (c.w_dhvap, c.dhvap_u_gas, c.dhvap_target)


def test_create_rejects_bad_config(gpu_lib):
    import revsim

    h = ctypes.c_void_p()
    bad = revsim.make_config(216, 1.862, 1.0)  # cutoff above half the box
    assert gpu_lib.rs_create(ctypes.byref(bad), 0, ctypes.byref(h)) == -1
    assert b"cutoff" in gpu_lib.rs_last_error(None)
    bad = revsim.make_config(216, 1.862, 0.8, abi_version=99)
    assert gpu_lib.rs_create(ctypes.byref(bad), 0, ctypes.byref(h)) == -1
    # the enthalpy-of-vaporization loss is classical-only; FP32 pair arithmetic needs a cubic box
    bad = revsim.make_config(216, 1.862, 0.8, use_mlp=1, w_dhvap=1.0)
    assert gpu_lib.rs_create(ctypes.byref(bad), 0, ctypes.byref(h)) == -1
    assert b"use_mlp" in gpu_lib.rs_last_error(None)
    bad = revsim.make_config(216, [1.862, 1.862, 1.9], 0.8, precision=revsim.RS_PRECISION_FP32)
    assert gpu_lib.rs_create(ctypes.byref(bad), 0, ctypes.byref(h)) == -1
    assert b"cubic" in gpu_lib.rs_last_error(None)


def test_null_context_is_an_error_not_a_crash(gpu_lib):
    import numpy as np

    x = np.zeros(9)
    assert gpu_lib.rs_forces(None, x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                             x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == -1


def test_no_device_is_reported(gpu_lib):
    """In this CPU-only container rs_create reports RS_E_NODEVICE; the Python API raises."""
    import revsim

    try:
        sim = revsim.Simulator(revsim.make_config(216, 1.862, 0.8))
    except revsim.RevSimError as e:
        assert e.code == -7
        return
    sim.close()
    pytest.skip("a GPU is visible")
