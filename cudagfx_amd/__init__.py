"""cudagfx_amd -- Python mirror of the cudagfx gauge-fixing interface on MI355X.

Thin ctypes layer over ``libcudagfx.so`` (built from ``cudagfx_amd/csrc`` and
``cudagfx_amd/host`` by ``cudagfx_amd/Makefile``).  The functions keep the
reference's names and argument meaning (cudagaugefix.cu, overrelax.cu,
simulated_annealing.cu, heatbath_thermalization.cu, MersenneTwister.cu); the
lattice extents and the results FUNC/DADA/maxDADA/PLAQ are the library's
exported globals, as in the reference's global.h.

Host arrays are numpy float64 in the reference layout: an su3 matrix is 18
doubles (row-major complex), the gauge field holds 4*V matrices at 4*s+mu,
g holds V matrices.  There is no fallback: if the HIP library is missing or
no GPU is present, the compute calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcudagfx.so")
DRIVER_PATH = os.path.join(_HERE, "bin", "cudagaugefix")

PREC_FP64 = 0
PREC_MIXED = 1
# the sweep forms of a mixed run on packed links, for time_sweeps /
# sweep_variant / set_sweep_variant: FP64 polish on 24-bit high + 16-bit low
# parts, FP32 phase on the 21-bit links
SWEEP_PACKED64 = 2
SWEEP_PACKED32 = 3
# the FP64 polish of the mixed Landau run on 32-bit fixed-point links and g
SWEEP_FIXED64 = 4
GAUGE_LANDAU = 0
GAUGE_COULOMB = 1

# every C-ABI symbol of include/cudagfx.h and include/cudagfx_host.h
ABI_FUNCTIONS = [
    "init_gaugefixing", "finalize_gaugefixing", "init_thermalization", "finalize_thermalization",
    "overrelax_gauge", "calc_functional", "calc_plaquette", "simannealing_gauge",
    "thermalize_gauge", "init_MT", "update_MT", "finalize_MT", "get_trafo", "set_trafo",
    "get_gauge", "apply_trafo", "gfx_set_omega", "gfx_set_precision", "gfx_set_fp32_stall", "gfx_set_packed_links", "gfx_set_verbose",
    "gfx_set_mt_seed", "gfx_coulomb_gauge", "gfx_functional_slices", "gfx_halfsweep",
    "gfx_time_sweeps", "gfx_last_run_stats", "gfx_error_string",
    "read_input", "read_gf_ildg", "read_ildg_dims", "write_gf_ildg", "Random", "PlantSeeds", "PutSeed", "GetSeed",
    "SelectStream", "unit_init_trafo", "random_init_trafo", "unit_init_gauge",
    "random_init_gauge", "g_trafo", "gauge_functional", "dAdA", "mean_plaq",
]
ABI_GLOBALS = ["T", "L", "LX", "LY", "LZ", "VOLUME", "FUNC", "DADA", "maxDADA", "PLAQ", "BETA",
               "randseed", "orxmaxit", "orxcheckinterval", "orxflag", "saflag", "thermflag",
               "orxeps", "saparam", "thermparam"]


class NativeMissing(RuntimeError):
    pass


class SAParam(C.Structure):
    """SAPARAM_LIST (global.h:50-56)"""
    _fields_ = [("Tmin", C.c_double), ("Tmax", C.c_double), ("expo", C.c_double),
                ("N", C.c_int), ("checkint", C.c_int)]


class ThermParam(C.Structure):
    """THERMPARAM_LIST (global.h:60-66)"""
    _fields_ = [("Nsweep", C.c_int), ("Nsave", C.c_int), ("beta", C.c_double),
                ("checkint", C.c_int), ("startcond", C.c_int)]


_lib = None


def load():
    """Load libcudagfx.so (raises NativeMissing if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeMissing(f"{LIB_PATH} not built (run make -C cudagfx_amd)")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "init_gaugefixing": (C.c_int, [vp, vp]),
        "finalize_gaugefixing": (None, []),
        "init_thermalization": (C.c_int, [vp]),
        "finalize_thermalization": (None, []),
        "overrelax_gauge": (C.c_int, [C.c_int, C.c_double, C.c_int]),
        "calc_functional": (C.c_double, []),
        "calc_plaquette": (C.c_double, [C.c_int]),
        "simannealing_gauge": (None, []),
        "thermalize_gauge": (None, []),
        "init_MT": (C.c_int, [C.c_int, C.c_int]),
        "update_MT": (None, []),
        "finalize_MT": (None, []),
        "get_trafo": (C.c_int, [vp]),
        "set_trafo": (C.c_int, [vp]),
        "get_gauge": (C.c_int, [vp]),
        "apply_trafo": (C.c_int, [vp]),
        "gfx_set_omega": (C.c_int, [C.c_double]),
        "gfx_set_precision": (C.c_int, [C.c_int, C.c_double]),
        "gfx_set_fp32_stall": (C.c_int, [C.c_int, C.c_double]),
        "gfx_set_packed_links": (C.c_int, [C.c_int]),
        "gfx_set_verbose": (C.c_int, [C.c_int]),
        "gfx_set_mt_seed": (C.c_int, [C.c_long]),
        "gfx_set_mt_file": (C.c_int, [C.c_char_p]),
        "gfx_set_mt_params": (C.c_int, [vp]),
        "gfx_get_mt_fields": (C.c_int, [vp, vp, vp, vp]),
        "gfx_mt_sizes": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "gfx_set_sa": (C.c_int, [C.c_int, C.c_int, C.c_int]),
        "gfx_sa_temperature": (C.c_double, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int]),
        "gfx_heatbath_sweeps": (C.c_int, [C.c_int, C.c_double]),
        "gfx_micro_sweeps": (C.c_int, [C.c_int, C.c_int]),
        "gfx_set_autotune": (C.c_int, [C.c_int]),
        "gfx_stream_copy": (C.c_int, [C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
        "gfx_batch_create": (C.c_int, []),
        "gfx_batch_destroy": (None, []),
        "gfx_batch_size": (C.c_int, []),
        "gfx_batch_add": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, vp, vp]),
        "gfx_batch_add_copy": (C.c_int, [C.c_int, vp]),
        "gfx_batch_add_random_copy": (C.c_int, [C.c_int]),
        "gfx_batch_run": (C.c_int, [C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int]),
        "gfx_batch_set_persistent": (C.c_int, [C.c_int]),
        "gfx_batch_persistent_launches": (C.c_int, []),
        "gfx_batch_result": (C.c_int, [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.POINTER(C.c_int)]),
        "gfx_batch_get_trafo": (C.c_int, [C.c_int, vp]),
        "gfx_batch_set_trafo": (C.c_int, [C.c_int, vp]),
        "gfx_batch_stats": (C.c_int, [C.POINTER(C.c_double), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_double)]),
        "gfx_batch_variant": (C.c_int, [C.c_int, C.POINTER(C.c_int), vp]),
        "gfx_su2_sample": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_int, vp]),
        "gfx_su2_sample4": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_int, vp]),
        "gfx_batch_anneal": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                       C.c_int, C.c_int]),
        "gfx_batch_best_copy": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
        "gfx_get_sweep_variant": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int),
                                            C.POINTER(C.c_int), C.POINTER(C.c_int), vp]),
        "gfx_set_sweep_variant": (C.c_int, [C.c_int, C.c_int, C.c_int]),
        "gfx_sync": (C.c_int, []),
        "gfx_coulomb_gauge": (C.c_int, [C.c_int, C.c_double, C.c_int, vp]),
        "gfx_functional_slices": (C.c_int, [vp, vp, vp]),
        "gfx_halfsweep": (C.c_int, [C.c_int, C.c_int]),
        "gfx_time_sweeps": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_double)]),
        "gfx_last_run_stats": (C.c_int, [C.POINTER(C.c_double), C.POINTER(C.c_int),
                                         C.POINTER(C.c_int)]),
        "gfx_set_sweep_sampling": (C.c_int, [C.c_int]),
        "gfx_sweep_sample_stats": (C.c_int, [C.c_int, C.POINTER(C.c_double),
                                             C.POINTER(C.c_double), C.POINTER(C.c_int)]),
        "gfx_error_string": (C.c_char_p, [C.c_int]),
        "read_input": (C.c_int, [C.c_char_p]),
        "read_gf_ildg": (C.c_int, [vp, C.c_char_p]),
        "read_ildg_dims": (C.c_int, [C.c_char_p, C.POINTER(C.c_int)]),
        "write_gf_ildg": (C.c_int, [vp, C.c_char_p, C.c_int]),
        "Random": (C.c_double, []),
        "PlantSeeds": (None, [C.c_long]),
        "PutSeed": (None, [C.c_long]),
        "GetSeed": (None, [C.POINTER(C.c_long)]),
        "SelectStream": (None, [C.c_int]),
        "unit_init_trafo": (None, [vp]),
        "random_init_trafo": (None, [vp]),
        "unit_init_gauge": (None, [vp]),
        "random_init_gauge": (None, [vp]),
        "g_trafo": (None, [vp, vp]),
        "gauge_functional": (C.c_double, [vp]),
        "dAdA": (C.c_double, [vp]),
        "mean_plaq": (C.c_double, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    assert isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.c_void_p)


def _gi(name):
    return C.c_int.in_dll(load(), name)


def _gd(name):
    return C.c_double.in_dll(load(), name)


def set_lattice(LX, LY, LZ, LT):
    """Set the lattice globals (the reference sets them in read_input /
    read_gf_ildg, read_input.l:276-279)."""
    for n, v in zip(["LX", "LY", "LZ", "T"], [LX, LY, LZ, LT]):
        _gi(n).value = v
    _gi("L").value = LX
    _gi("VOLUME").value = LX * LY * LZ * LT


def lattice():
    return tuple(_gi(n).value for n in ("LX", "LY", "LZ", "T"))


def results():
    """(FUNC, DADA, maxDADA) as left by the last functional evaluation."""
    return _gd("FUNC").value, _gd("DADA").value, _gd("maxDADA").value


def saparam():
    return SAParam.in_dll(load(), "saparam")


def thermparam():
    return ThermParam.in_dll(load(), "thermparam")


def _check(rc, what):
    if rc != 0:
        msg = load().gfx_error_string(rc).decode()
        raise RuntimeError(f"{what} failed: {rc} ({msg})")


# ------------------------------------------------------------------ reference API
def init_gaugefixing(gf, trafo):
    _check(load().init_gaugefixing(_ptr(gf), _ptr(trafo)), "init_gaugefixing")


def finalize_gaugefixing():
    load().finalize_gaugefixing()


def init_thermalization(gf):
    _check(load().init_thermalization(_ptr(gf)), "init_thermalization")


def finalize_thermalization():
    load().finalize_thermalization()


def overrelax_gauge(maxit, eps, checkinterval):
    return load().overrelax_gauge(int(maxit), float(eps), int(checkinterval))


def calc_functional():
    return load().calc_functional()


def calc_plaquette(outputyn=0):
    return load().calc_plaquette(int(outputyn))


def simannealing_gauge():
    load().simannealing_gauge()


def thermalize_gauge():
    load().thermalize_gauge()


def init_MT(n_gauss, n_unif):
    _check(load().init_MT(int(n_gauss), int(n_unif)), "init_MT")


def update_MT():
    load().update_MT()


def finalize_MT():
    load().finalize_MT()


def get_trafo():
    V = _gi("VOLUME").value
    g = np.zeros(18 * V)
    _check(load().get_trafo(_ptr(g)), "get_trafo")
    return g


def set_trafo(g):
    _check(load().set_trafo(_ptr(g)), "set_trafo")


def get_gauge():
    V = _gi("VOLUME").value
    u = np.zeros(18 * 4 * V)
    _check(load().get_gauge(_ptr(u)), "get_gauge")
    return u


def apply_trafo(gf):
    """Device g_trafo of host links gf (observables.c:17-47); returns a new array."""
    out = np.array(gf, dtype=np.float64, copy=True, order="C")
    _check(load().apply_trafo(_ptr(out)), "apply_trafo")
    return out


# ------------------------------------------------------------------ extensions
def set_omega(omega):
    load().gfx_set_omega(float(omega))


def set_precision(mode, switch_theta=1e-6):
    load().gfx_set_precision(int(mode), float(switch_theta))


def set_fp32_stall(iters=300, gate=1e-8):
    """The mixed path also leaves FP32 once its best max dAdA is below `gate`
    and has not halved for `iters` iterations; iters=0 turns this off."""
    _check(load().gfx_set_fp32_stall(int(iters), float(gate)), "set_fp32_stall")


def set_packed_links(on=1):
    """Links of a mixed Landau, Coulomb or batch run: 1 packed (FP32 phase on
    21-bit links within 2^-21, FP64 polish on 32-bit fixed-point links and g,
    or below its reach on 24-bit high + 16-bit low parts within 2^-38 of the
    FP64 links), 0 the FP32 and FP64 links."""
    _check(load().gfx_set_packed_links(int(on)), "set_packed_links")


def set_verbose(v):
    load().gfx_set_verbose(int(v))


def set_mt_seed(seed):
    load().gfx_set_mt_seed(int(seed))


def set_mt_params(params):
    p = np.ascontiguousarray(params, dtype=np.uint32)
    assert p.size == 4 * 4096
    load().gfx_set_mt_params(p.ctypes.data_as(C.c_void_p))


def get_mt_fields():
    ng, nu = C.c_int(), C.c_int()
    load().gfx_mt_sizes(C.byref(ng), C.byref(nu))
    gauss = np.zeros(4096 * ng.value, dtype=np.float32)
    unif = np.zeros(4096 * nu.value, dtype=np.float32)
    state = np.zeros(4096 * 19, dtype=np.uint32)
    params = np.zeros(4096 * 4, dtype=np.uint32)
    _check(load().gfx_get_mt_fields(gauss.ctypes.data_as(C.c_void_p),
                                    unif.ctypes.data_as(C.c_void_p),
                                    state.ctypes.data_as(C.c_void_p),
                                    params.ctypes.data_as(C.c_void_p)), "get_mt_fields")
    return gauss, unif, state, params


def set_sa(sampler=0, nmicro=0, ndir=4):
    load().gfx_set_sa(int(sampler), int(nmicro), int(ndir))


def heatbath_sweeps(nsweep, beta):
    _check(load().gfx_heatbath_sweeps(int(nsweep), float(beta)), "heatbath")


class Batch:
    """The batching form (gfx_batch_*): Gribov copies sharing links and
    lattices of different sizes fixed together."""

    def __init__(self):
        _check(load().gfx_batch_create(), "batch_create")
        self.dims = []

    def add(self, dims, gf, trafo):
        q = load().gfx_batch_add(*[int(d) for d in dims], _ptr(gf), _ptr(trafo))
        if q < 0:
            _check(-q, "batch_add")
        self.dims.append(tuple(dims))
        return q

    def add_copy(self, base, trafo=None):
        q = load().gfx_batch_add_copy(int(base), None if trafo is None else _ptr(trafo))
        if q < 0:
            _check(-q, "batch_add_copy")
        self.dims.append(self.dims[base])
        return q

    def add_random_copy(self, base):
        q = load().gfx_batch_add_random_copy(int(base))
        if q < 0:
            _check(-q, "batch_add_random_copy")
        self.dims.append(self.dims[base])
        return q

    def run(self, maxit, eps, checkinterval, precision=PREC_FP64, switch_theta=1e-6,
            periods_per_graph=4):
        return load().gfx_batch_run(int(maxit), float(eps), int(checkinterval), int(precision),
                                    float(switch_theta), int(periods_per_graph))

    def result(self, q):
        F, D, M, it = C.c_double(), C.c_double(), C.c_double(), C.c_int()
        _check(load().gfx_batch_result(int(q), C.byref(F), C.byref(D), C.byref(M), C.byref(it)),
               "batch_result")
        return F.value, D.value, M.value, it.value

    def trafo(self, q):
        LX, LY, LZ, LT = self.dims[q]
        g = np.zeros(18 * LX * LY * LZ * LT)
        _check(load().gfx_batch_get_trafo(int(q), _ptr(g)), "batch_get_trafo")
        return g

    def set_trafo(self, q, g):
        """Lattice q's g for the next run or anneal (V su3, host layout)."""
        LX, LY, LZ, LT = self.dims[q]
        g = np.ascontiguousarray(g, dtype=np.float64)
        if g.size != 18 * LX * LY * LZ * LT:
            raise ValueError("set_trafo: g has the wrong size for lattice %d" % q)
        _check(load().gfx_batch_set_trafo(int(q), _ptr(g)), "batch_set_trafo")

    def stats(self):
        ms, gl, a, b, su = C.c_double(), C.c_int(), C.c_int(), C.c_int(), C.c_double()
        load().gfx_batch_stats(C.byref(ms), C.byref(gl), C.byref(a), C.byref(b), C.byref(su))
        return dict(ms=ms.value, graph_launches=gl.value, iters_fp32=a.value, iters_fp64=b.value,
                    site_updates=su.value,
                    persistent_launches=load().gfx_batch_persistent_launches())

    def set_persistent(self, on):
        """Check periods of single-precision stretches in persistent launches
        (see gfx_batch_set_persistent)."""
        _check(load().gfx_batch_set_persistent(int(on)), "batch_set_persistent")

    def variant(self, precision=PREC_FP64):
        v = C.c_int()
        ms = np.zeros(32)
        rc = load().gfx_batch_variant(int(precision), C.byref(v), ms.ctypes.data_as(C.c_void_p))
        return (v.value, ms) if rc == 0 else None

    def set_anneal_precision(self, mode):
        """PREC_FP64 (default) or PREC_MIXED: the anneal on FP32 links and g."""
        _check(load().gfx_batch_set_anneal_precision(int(mode)), "anneal precision")

    def anneal(self, nsteps, Tmin, Tmax, expo=-1.0, nmicro=0, sampler=0, steps_per_graph=8):
        """Simulated annealing of every lattice's g (see gfx_batch_anneal)."""
        _check(load().gfx_batch_anneal(int(nsteps), float(Tmin), float(Tmax), float(expo),
                                       int(nmicro), int(sampler), int(steps_per_graph)),
               "batch_anneal")

    def best_copy(self, base):
        """(index, F) of the lattice with the largest F among the copies of base."""
        q, F = C.c_int(), C.c_double()
        _check(load().gfx_batch_best_copy(int(base), C.byref(q), C.byref(F)), "best_copy")
        return q.value, F.value

    def close(self):
        load().gfx_batch_destroy()


def stream_copy(nbytes, nrep=1):
    ms = C.c_double()
    _check(load().gfx_stream_copy(int(nbytes), int(nrep), C.byref(ms)), "stream_copy")
    return ms.value


def set_autotune(on):
    load().gfx_set_autotune(int(on))


def sweep_variant(precision=PREC_FP64, ndir=4):
    """(variant, block, rolled, ms per variant) chosen by the autotuner, or None."""
    v, b, r = C.c_int(), C.c_int(), C.c_int()
    ms = np.zeros(32)
    rc = load().gfx_get_sweep_variant(int(precision), int(ndir), C.byref(v), C.byref(b),
                                      C.byref(r), ms.ctypes.data_as(C.c_void_p))
    if rc != 0:
        return None
    return v.value, b.value, bool(r.value), ms


def set_sweep_variant(precision, ndir, variant):
    _check(load().gfx_set_sweep_variant(int(precision), int(ndir), int(variant)), "variant")


def su2_sample(sampler, beta, wnorm, n):
    """n draws of the SA sweep's SU(2) sampler (alpha_0 for w = (wnorm,0,0,0))."""
    out = np.zeros(int(n))
    _check(load().gfx_su2_sample(int(sampler), float(beta), float(wnorm), int(n), _ptr(out)),
           "su2_sample")
    return out


def su2_sample4(sampler, beta, wnorm, n):
    """The same draws with all four quaternion components, shape (n, 4)."""
    out = np.zeros(4 * int(n))
    _check(load().gfx_su2_sample4(int(sampler), float(beta), float(wnorm), int(n), _ptr(out)),
           "su2_sample4")
    return out.reshape(int(n), 4)


def micro_sweeps(n, ndir=4):
    _check(load().gfx_micro_sweeps(int(n), int(ndir)), "micro_sweeps")


def sync():
    _check(load().gfx_sync(), "sync")


def coulomb_gauge(maxit, eps, checkinterval):
    LT = _gi("T").value
    iters = np.zeros(LT, dtype=np.int32)
    rc = load().gfx_coulomb_gauge(int(maxit), float(eps), int(checkinterval),
                                  iters.ctypes.data_as(C.c_void_p))
    return rc, iters


def functional_slices():
    LT = _gi("T").value
    F, D, M = np.zeros(LT), np.zeros(LT), np.zeros(LT)
    _check(load().gfx_functional_slices(_ptr(F), _ptr(D), _ptr(M)), "functional_slices")
    return F, D, M


def halfsweep(parity, ndir=4):
    _check(load().gfx_halfsweep(int(parity), int(ndir)), "halfsweep")


def time_sweeps(nsweeps, precision=PREC_FP64):
    ms = C.c_double()
    _check(load().gfx_time_sweeps(int(nsweeps), int(precision), C.byref(ms)), "time_sweeps")
    return ms.value


def set_sweep_sampling(on):
    """Bracket each run of consecutive check-free half-sweep launches of the
    following overrelax_gauge calls with a pair of HIP events (0: off)."""
    _check(load().gfx_set_sweep_sampling(int(on)), "set_sweep_sampling")


def sweep_sample_stats(precision):
    """(mean ms per launch, smallest segment mean, launches) of the bracketed
    launches of one sweep form."""
    mean, mn, n = C.c_double(), C.c_double(), C.c_int()
    _check(load().gfx_sweep_sample_stats(int(precision), C.byref(mean), C.byref(mn), C.byref(n)),
           "sweep_sample_stats")
    return mean.value, mn.value, n.value


def last_run_stats():
    ms, a, b = C.c_double(), C.c_int(), C.c_int()
    load().gfx_last_run_stats(C.byref(ms), C.byref(a), C.byref(b))
    return ms.value, a.value, b.value


def sa_temperature(i, Tmin, Tmax, expo, N):
    return load().gfx_sa_temperature(int(i), float(Tmin), float(Tmax), float(expo), int(N))


# ------------------------------------------------------------------ host services
def plant_seeds(seed):
    load().PlantSeeds(int(seed))


def random_init_trafo_host(V=None):
    if V is None:
        V = _gi("VOLUME").value
    g = np.zeros(18 * V)
    load().random_init_trafo(_ptr(g))
    return g


def random_init_gauge_host():
    V = _gi("VOLUME").value
    u = np.zeros(18 * 4 * V)
    load().random_init_gauge(_ptr(u))
    return u


def unit_gauge():
    V = _gi("VOLUME").value
    u = np.zeros(18 * 4 * V)
    load().unit_init_gauge(_ptr(u))
    return u


def unit_trafo():
    V = _gi("VOLUME").value
    g = np.zeros(18 * V)
    load().unit_init_trafo(_ptr(g))
    return g


def read_input(path):
    return load().read_input(str(path).encode())


def read_gf_ildg(path):
    V = _gi("VOLUME").value
    u = np.zeros(18 * 4 * V)
    rc = load().read_gf_ildg(_ptr(u), str(path).encode())
    return rc, u


def ildg_dims(path):
    """(lx, ly, lz, lt) from an ILDG file's ildg-format record, or None."""
    d = (C.c_int * 4)()
    rc = load().read_ildg_dims(str(path).encode(), d)
    return tuple(d) if rc == 0 else None


def write_gf_ildg(gf, path, prec=64):
    return load().write_gf_ildg(_ptr(gf), str(path).encode(), int(prec))


def host_gauge_functional(gf):
    return load().gauge_functional(_ptr(gf))


def host_dAdA(gf):
    d = load().dAdA(_ptr(gf))
    return d, _gd("maxDADA").value


def host_mean_plaq(gf):
    return load().mean_plaq(_ptr(gf))


def host_g_trafo(gf, g):
    out = np.array(gf, copy=True)
    load().g_trafo(_ptr(out), _ptr(np.ascontiguousarray(g)))
    return out
