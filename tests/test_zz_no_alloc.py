"""The allocation-free labelling entry point (swmi_label_bonds_noalloc) against the engine's own labelling
(swmi_label_bonds + swmi_download_labels) and the oracle: two calls in a row through the C ABI, the second
on a smaller lattice, so that the second call runs on scratch memory that still holds the first call's
contents. The test's device buffers come from the HIP runtime that libswmi.so itself links (through
ctypes), on a stream of that runtime."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    """The few HIP runtime calls the test needs for its own buffers and stream."""

    def __init__(self):
        import computational_physics_amd as cpa
        cpa.load_library()   # libswmi.so loads libamdhip64 first: the same runtime is used below
        self.lib = ctypes.CDLL("libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        self.lib.hipFree.argtypes = [vp]
        self.lib.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        self.lib.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        self.lib.hipStreamSynchronize.argtypes = [vp]
        self.lib.hipStreamDestroy.argtypes = [vp]

    def check(self, status, what):
        assert status == 0, f"{what}: hip error {status}"

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        self.check(self.lib.hipMalloc(ctypes.byref(p), nbytes), "hipMalloc")
        return p

    def h2d(self, dst, arr):
        self.check(self.lib.hipMemcpy(dst, arr.ctypes.data, arr.nbytes, 1), "hipMemcpy H2D")

    def d2h(self, arr, src):
        self.check(self.lib.hipMemcpy(arr.ctypes.data, src, arr.nbytes, 2), "hipMemcpy D2H")


def _bonds(rng, n0, n1, density):
    return ((rng.random(n0 * n1) < density).astype(np.uint8) | ((rng.random(n0 * n1) < density).astype(np.uint8) << 1))


@pytest.mark.parametrize("shapes", [[(512, 256, 0.55), (128, 64, 0.5)], [(1024, 1024, 0.6), (32, 16, 0.45)]])
def test_noalloc_labels_equal_the_engine_labelling(swmi, oracle, shapes):
    hip = _Hip()
    stream = ctypes.c_void_p()
    hip.check(hip.lib.hipStreamCreate(ctypes.byref(stream)), "hipStreamCreate")
    rng = np.random.default_rng(17)
    n_max = max(n0 * n1 for n0, n1, _ in shapes)
    d_bonds, d_labels = hip.malloc(n_max), hip.malloc(4 * n_max)
    try:
        for n0, n1, density in shapes:
            print(f"noalloc labelling {n0}x{n1} density {density}", flush=True)
            bonds = _bonds(rng, n0, n1, density)
            hip.h2d(d_bonds, bonds)
            swmi.label_bonds_noalloc(d_bonds.value, n0, n1, d_labels.value, stream.value)
            hip.check(hip.lib.hipStreamSynchronize(stream), "hipStreamSynchronize")
            got = np.empty(n0 * n1, dtype=np.uint32)
            hip.d2h(got, d_labels)
            with swmi.Engine(swmi.LatticeConfig(n0, n1, seed=1)) as e:
                want = e.label_bonds(bonds)
                assert e.error_flags(0) == 0
            np.testing.assert_array_equal(got, want, err_msg=f"{n0}x{n1}")
            if n0 * n1 <= 1 << 17:
                np.testing.assert_array_equal(got, oracle.label(n0, n1, bonds))
    finally:
        hip.lib.hipFree(d_bonds)
        hip.lib.hipFree(d_labels)
        hip.lib.hipStreamDestroy(stream)
