"""rs_msd_noalloc (the MSD loss entry point that allocates nothing; its kernels' scratch is a
static device array) against rs_msd on the same frames: two calls in a row, the second smaller,
results equal bit for bit; a call beyond the scratch's limit is refused. Device arrays for the
frames and the gradient come from the HIP runtime through ctypes."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    for p in ("/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"):
        try:
            return ctypes.CDLL(p)
        except OSError:
            continue
    pytest.fail("libamdhip64.so not found")


class DevArray:
    def __init__(self, hip, n):
        self.hip, self.n = hip, n
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(8 * n)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(8 * self.n), 1) == 0

    def get(self):
        out = np.empty(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, ctypes.c_size_t(8 * self.n), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _frames(M, nf, box, seed):
    rng = np.random.default_rng(seed)
    base = rng.uniform(0, box, 3 * M)
    steps = 0.02 * rng.standard_normal((nf - 1, 3 * M))
    fr = np.concatenate([base[None, :], base[None, :] + np.cumsum(steps, axis=0)])
    return np.ascontiguousarray(np.mod(fr, box).reshape(-1))


def test_msd_noalloc_matches_msd(gpu_lib):
    import revsim

    hip = _hip()
    M, box = 3000, 4.48
    sim = revsim.Simulator(revsim.make_config(M, box, 1.0, msd_window_steps=40, msd_every=5, w_msd=1.0), device=0)
    for nf, seed in ((16, 1), (6, 2)):  # the second call is smaller
        fr = _frames(M, nf, box, seed)
        D, L, d = sim.msd(fr)
        fdev, gdev = DevArray(hip, fr.size), DevArray(hip, fr.size)
        fdev.put(fr)
        D2, L2 = ctypes.c_double(), ctypes.c_double()
        rc = gpu_lib.rs_msd_noalloc(sim.h, fdev.p, nf, ctypes.byref(D2), ctypes.byref(L2), gdev.p)
        assert rc == 0, gpu_lib.rs_last_error(sim.h)
        g2 = gdev.get()
        fdev.free()
        gdev.free()
        assert D2.value == D and L2.value == L, (nf, D, D2.value, L, L2.value)
        assert np.array_equal(g2, d), nf
    # beyond the static scratch: 64 origins of 32,768 molecules
    big = revsim.Simulator(revsim.make_config(32768, 9.95, 1.0, msd_window_steps=640, msd_every=5, w_msd=1.0),
                           device=0)
    fdev = DevArray(hip, 128 * 3 * 32768)
    D2, L2 = ctypes.c_double(), ctypes.c_double()
    rc = gpu_lib.rs_msd_noalloc(big.h, fdev.p, 128, ctypes.byref(D2), ctypes.byref(L2), fdev.p)
    fdev.free()
    assert rc == -1  # RS_E_CONFIG
