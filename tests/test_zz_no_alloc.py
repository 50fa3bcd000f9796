"""gfx_calc_functional_dev, the entry point that allocates nothing: F, dAdA
and max dAdA of caller-owned device fields (reference device layout) with
the reduction's partials in the library's static device array.  Checked
against calc_functional on the library's own fields (the existing entry
point, same inputs) on two calls in a row, the second on a smaller lattice,
to 1e-12 relative (F and dAdA are sums in a different order; max dAdA is a
maximum of per-site values computed in the same order, equal to rounding).
Device buffers of the test are allocated by the test itself through the
same HIP runtime the library uses."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HIP = None


def hip():
    global HIP
    if HIP is None:
        HIP = C.CDLL("libamdhip64.so.7")
        HIP.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        HIP.hipFree.argtypes = [C.c_void_p]
    return HIP


def dev_array(a):
    h = hip()
    p = C.c_void_p()
    assert h.hipMalloc(C.byref(p), a.nbytes) == 0
    assert h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
    return p


def reference_functional(gfx, dims, u, g):
    gfx.set_lattice(*dims)
    gfx.set_precision(gfx.PREC_FP64)
    gfx.init_gaugefixing(u, g)
    gfx.load().calc_functional()
    F, D, M = gfx.results()
    gfx.finalize_gaugefixing()
    return np.array([F, D, M])


def no_alloc_functional(gfx, orc, dims, u, g):
    lib = gfx.load()
    fn = lib.gfx_calc_functional_dev
    fn.argtypes = [C.c_void_p, C.c_void_p] + [C.c_int] * 4 + [C.c_void_p, C.c_void_p]
    fn.restype = C.c_int
    u2 = np.ascontiguousarray(orc.su3_to_2v(u))  # reference device layout
    g2 = np.ascontiguousarray(orc.su3_to_2v(g))
    du, dg = dev_array(u2), dev_array(g2)
    res = np.full(3, -1.0)
    dres = dev_array(res)
    h = hip()
    try:
        rc = fn(du, dg, *dims, dres, None)
        assert rc == 0, rc
        assert h.hipDeviceSynchronize() == 0
        assert h.hipMemcpy(res.ctypes.data, dres, res.nbytes, 2) == 0  # device to host
    finally:
        for p in (du, dg, dres):
            h.hipFree(p)
    return res


def test_no_alloc_functional_matches_calc_functional(gfx, orc):
    for dims, seed in (((8, 8, 8, 12), 7101), ((4, 6, 4, 8), 7102)):
        V = orc.volume(dims)
        u = orc.random_init_gauge(V, seed)
        g = orc.random_init_trafo(V, seed + 50)
        want = reference_functional(gfx, dims, u, g)
        got = no_alloc_functional(gfx, orc, dims, u, g)
        print(f"{dims}: calc_functional {want}, no-alloc {got}", flush=True)
        assert np.all(np.abs(got - want) <= 1e-12 * np.abs(want)), (dims, got, want)
        assert want[1] > 1.0


def test_no_alloc_functional_rejects_oversize(gfx):
    """Sizes beyond the static array are refused before any device work."""
    lib = gfx.load()
    fn = lib.gfx_calc_functional_dev
    fn.argtypes = [C.c_void_p, C.c_void_p] + [C.c_int] * 4 + [C.c_void_p, C.c_void_p]
    fn.restype = C.c_int
    assert fn(C.c_void_p(16), C.c_void_p(16), 64, 64, 64, 66, C.c_void_p(16), None) == 900
    assert fn(None, C.c_void_p(16), 4, 4, 4, 4, C.c_void_p(16), None) == 100
