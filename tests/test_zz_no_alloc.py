"""ccgpu_energy_noalloc (the energy sum on caller-owned device buffers, with its
block partials in the library's static per-device array) against ccgpu_energy,
the entry point it is a variant of: two calls in a row on buffers the test
allocates with the HIP runtime the library itself uses, the second call on a
smaller input, plus the size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip(ccnm):
    ccnm.lib()   # the library loads the image's HIP runtime; the test uses the same one
    rt = C.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipDeviceSynchronize.argtypes = []
    return rt


H2D, D2H = 1, 2


def device_copy(hip, arr):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), max(arr.nbytes, 16)) == 0
    assert hip.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D) == 0
    return p


def test_energy_noalloc_matches_energy(ccnm, hip):
    for A, nmax in ((66, 2), (14, 1)):   # the second call has the smaller input
        s = ccnm.Solver(A=A, matter=ccnm.PNM, rho=0.16, nmax=nmax)
        s.set_chiral(ccnm.NNLO, 450.0)
        s.generate_interaction(ccnm.INT_CHIRAL)
        s.setup_t_amplitudes()
        s.ccd_iter(subspace=8, diis_step=8, tol=1e-8, maxit=100)
        want = s.ccd_energy_save()
        n = s.dims["t2_unique"]
        t = np.ascontiguousarray(s.debug_fetch("T", n))
        r0 = np.ascontiguousarray(s.debug_fetch("R0", n))
        dt, dr, de = device_copy(hip, t), device_copy(hip, r0), device_copy(hip, np.zeros(2))
        try:
            ccnm.energy_noalloc(n, dt.value, dr.value, de.value, None)
            assert hip.hipDeviceSynchronize() == 0
            out = np.zeros(2)
            assert hip.hipMemcpy(out.ctypes.data, de, 16, D2H) == 0
        finally:
            for p in (dt, dr, de):
                hip.hipFree(p)
        got = complex(out[0], out[1])
        assert abs(got - want) <= 1e-12 * abs(want), (got, want)
        s.close()


def test_energy_noalloc_rejects_sizes_above_its_scratch(ccnm):
    with pytest.raises(RuntimeError, match="above the maximum"):
        ccnm.energy_noalloc(1 << 40, 16, 16, 16, None)
