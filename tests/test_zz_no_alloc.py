"""GPU: the allocation-free force call (cosmosim_ext_forces_noalloc) against
the entry point it varies, cosmosim_ext_compute_forces: the same tree build and
walk kernels on the same inputs, so the accelerations must be bit-identical.
Two calls in a row per precision, the second with a smaller input (the scratch
keeps the first call's contents); sizes beyond the maximum are refused.
Device arrays come from the HIP runtime the library itself links (the test
allocates them; the call under test does not); the call goes through the C-ABI.
"""
import ctypes

import numpy as np
import pytest

import cosmosim_amd as ca
from tests._util import same_bits

pytestmark = pytest.mark.gpu
MAX_BODIES = 16384
H2D, D2H = 1, 2


def _hip():
    ca.lib()                                     # loads the library and its HIP runtime
    hip = ctypes.CDLL("libamdhip64.so.7")        # the same runtime (by soname)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def _dev(hip, nbytes):
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), nbytes) == 0
    return p


def _noalloc(bodies, G, eps, theta, precision):
    hip = _hip()
    n = len(bodies)
    cols = []
    for f in ("x", "y", "z", "mass"):
        a = np.ascontiguousarray(bodies[f], dtype=np.float64)
        d = _dev(hip, a.nbytes)
        assert hip.hipMemcpy(d, a.ctypes.data, a.nbytes, H2D) == 0
        cols.append(d)
    nan = np.full(n, np.nan)
    out = []
    for _ in range(3):
        d = _dev(hip, nan.nbytes)
        assert hip.hipMemcpy(d, nan.ctypes.data, nan.nbytes, H2D) == 0
        out.append(d)
    st = np.array([12345], dtype=np.int32)
    dst = _dev(hip, 4)
    assert hip.hipMemcpy(dst, st.ctypes.data, 4, H2D) == 0
    rc = ca.lib().cosmosim_ext_forces_noalloc(*cols, n, G, eps, theta, ca.PRECISIONS[precision],
                                              *out, dst, None)
    assert rc == 0
    assert hip.hipDeviceSynchronize() == 0
    res = np.zeros((3, n))
    for k in range(3):
        assert hip.hipMemcpy(res[k].ctypes.data, out[k], res[k].nbytes, D2H) == 0
    assert hip.hipMemcpy(st.ctypes.data, dst, 4, D2H) == 0
    for d in cols + out + [dst]:
        hip.hipFree(d)
    assert int(st[0]) == 0
    return np.ascontiguousarray(res.T)


@pytest.mark.parametrize("precision", ["fp64_strict", "fp64", "fp32"])
def test_noalloc_forces_equal_compute_forces(precision):
    for n, seed in ((MAX_BODIES, 31), (5000, 32)):
        b = ca.generate(ca.IC_MERGER, n, seed)
        b["mass"][::97] = 0.0                 # dead bodies get a = 0
        got = _noalloc(b, 1.0, 0.01, 0.5, precision)
        sim = ca.Simulation(b, softening=0.01, theta=0.5, precision=precision)
        want = sim.accelerations()
        sim.close()
        assert np.all(got[b["mass"] <= 0] == 0.0)
        assert same_bits(got, want), (precision, n)


def test_noalloc_refuses_oversize_and_bad_arguments():
    L = ca.lib()
    p = ctypes.c_void_p(1)
    assert L.cosmosim_ext_forces_noalloc(p, p, p, p, MAX_BODIES + 1, 1.0, 0.01, 0.5, 2,
                                         p, p, p, p, None) == -5
    assert L.cosmosim_ext_forces_noalloc(p, p, p, p, 10, 1.0, 0.01, 0.5, 7,
                                         p, p, p, p, None) == -1
    assert L.cosmosim_ext_forces_noalloc(None, p, p, p, 10, 1.0, 0.01, 0.5, 2,
                                         p, p, p, p, None) == -1
