"""CPU oracle for the cudagfx gauge-fixing path -- TEST INFRASTRUCTURE ONLY.

This package is imported only by ``tests/``, by ``__graft_entry__.smoke()``
(to check the GPU result) and by ``bench.py``'s ``cpu_baseline`` leg.  The shipped
library (``cudagfx_amd``) never imports it.

It wraps ``oracle/liboracle.so`` (built from ``oracle/oracle.c`` by
``oracle/Makefile``), a C restatement of the reference arithmetic with the
reference file:line cited per function, and -- where it was built --
``oracle/_ref/libcudagfx_ref.so``, the reference's own host C routines
(rngs.c, su3manip.c, observables.c, read_input.c) compiled from the reference
checkout.  Arrays are numpy float64 in the reference layouts: su3 = 18
doubles, 2-row = 12 doubles, link index 4*s+mu.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_REF = os.path.join(_HERE, "_ref", "libcudagfx_ref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


class MTParams(C.Structure):
    _fields_ = [("matrix_a", C.c_uint32), ("mask_b", C.c_uint32),
                ("mask_c", C.c_uint32), ("seed", C.c_uint32)]


def _load():
    if not os.path.exists(_LIB):
        raise RuntimeError(f"oracle library missing: {_LIB} (run make -C oracle)")
    lib = C.CDLL(_LIB)
    sig = {
        "orc_init_nn": (None, [C.c_int] * 4 + [_ip]),
        "orc_init_eo": (None, [C.c_int] * 4 + [_ip, _ip]),
        "orc_lehmer_random": (C.c_double, []),
        "orc_lehmer_plant_seeds": (None, [C.c_long]),
        "orc_lehmer_select_stream": (None, [C.c_int]),
        "orc_lehmer_get_seed": (C.c_long, []),
        "orc_lehmer_put_seed": (None, [C.c_long]),
        "orc_random_init_trafo": (None, [_dp, C.c_int]),
        "orc_random_init_gauge": (None, [_dp, C.c_int]),
        "orc_su3_to_2v": (None, [_dp, _dp, C.c_long]),
        "orc_2v_to_su3": (None, [_dp, _dp, C.c_long]),
        "orc_or_halfsweep": (None, [_dp, _dp, _ip, _ip, C.c_int, C.c_int, C.c_double,
                                    C.c_int, C.c_void_p, C.c_int]),
        "orc_functional": (C.c_double, [_dp, _dp, _ip, C.c_int, C.c_int,
                                        C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "orc_functional_slices": (None, [_dp, _dp, _ip, C.c_int, C.c_int, C.c_int,
                                         _dp, _dp, _dp]),
        "orc_overrelax_gauge": (C.c_int, [_dp, _dp] + [C.c_int] * 4 +
                                [C.c_int, C.c_double, C.c_int, C.c_double,
                                 C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
        "orc_coulomb_gauge": (C.c_int, [_dp, _dp] + [C.c_int] * 4 +
                              [C.c_int, C.c_double, C.c_int, C.c_double, _ip]),
        "orc_plaquette_2v": (C.c_double, [_dp, _ip] + [C.c_int] * 4),
        "orc_g_trafo": (None, [_dp, _dp, _ip, C.c_int]),
        "orc_gauge_functional": (C.c_double, [_dp, C.c_int]),
        "orc_dAdA": (C.c_double, [_dp, _ip, C.c_int, C.POINTER(C.c_double)]),
        "orc_mean_plaq": (C.c_double, [_dp, _ip, C.c_int]),
        "orc_su3_normalize": (None, [_dp]),
        "orc_cm_relax": (None, [_dp, _dp]),
        "orc_overrelax_step": (None, [_dp, _dp, C.c_double, C.c_int]),
        "orc_overrelax_step_short": (None, [_dp, _dp, C.c_double]),
        "orc_reconstruct": (None, [_dp, _dp]),
        "orc_reconstruct_dagger": (None, [_dp, _dp]),
        "orc_mt_generate": (None, [C.POINTER(MTParams), _u32p, C.c_int, C.c_int, C.c_int,
                                   C.c_void_p, C.c_void_p]),
        "orc_mt19937_nth": (C.c_uint32, [C.c_uint32, C.c_int]),
        "orc_box_muller_exact": (None, [C.c_double, C.c_double, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]),
        "orc_mt_nper": (C.c_int, [C.c_int, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def ref_lib():
    """The reference's own host C routines, or None where they were not built."""
    if not os.path.exists(_REF):
        return None
    return C.CDLL(_REF)


# ---------------------------------------------------------------- geometry
def volume(dims):
    LX, LY, LZ, LT = dims
    return LX * LY * LZ * LT


def neighbours(dims):
    V = volume(dims)
    nn = np.zeros(8 * V, dtype=np.int32)
    lib.orc_init_nn(*dims, nn)
    return nn


def eo_lists(dims):
    V = volume(dims)
    ev = np.zeros(V // 2, dtype=np.int32)
    od = np.zeros(V // 2, dtype=np.int32)
    lib.orc_init_eo(*dims, ev, od)
    return ev, od


# ---------------------------------------------------------------- RNG / init
def random_init_trafo(V, seed):
    """PlantSeeds(seed) then random_init_trafo (su3manip.c:539-641)."""
    lib.orc_lehmer_plant_seeds(seed)
    g = np.zeros(18 * V)
    lib.orc_random_init_trafo(g, V)
    return g


def random_init_gauge(V, seed):
    lib.orc_lehmer_plant_seeds(seed)
    u = np.zeros(18 * 4 * V)
    lib.orc_random_init_gauge(u, V)
    return u


def su3_to_2v(a):
    n = a.size // 18
    out = np.zeros(12 * n)
    lib.orc_su3_to_2v(np.ascontiguousarray(a), out, n)
    return out


def from_2v(a):
    n = a.size // 12
    out = np.zeros(18 * n)
    lib.orc_2v_to_su3(np.ascontiguousarray(a), out, n)
    return out


# ---------------------------------------------------------------- sweeps
def halfsweep(u2v, g2v, dims, parity, ndir=4, omega=1.68, nterms=3, active=None):
    """One half sweep (overrelax.cu:335-446) on g2v in place; nterms = -3:
    the 3-term series in the short form of the library's polish sweeps, on
    h as the relaxation leaves it (not normalised)."""
    nn = neighbours(dims)
    ev, od = eo_lists(dims)
    sites = ev if parity == 0 else od
    V = volume(dims)
    vs = V // dims[3]
    act = None
    if active is not None:
        act = np.ascontiguousarray(active, dtype=np.uint8)
    lib.orc_or_halfsweep(u2v, g2v, nn, sites, V // 2, ndir, omega, nterms,
                         act.ctypes.data if act is not None else None, vs)


def ref_halfsweep(u_su3, g_su3, dims, parity, ndir=4, omega=1.68):
    """One half sweep in place on the reference's host layout (u: 4V su3,
    g: V su3, 18 doubles each) by oracle/ref_sweep.c, built from the
    reference's own host SU(3) routines; None where oracle/_ref is absent."""
    ref = ref_lib()
    if ref is None:
        return None
    nn = neighbours(dims)
    ev, od = eo_lists(dims)
    sites = ev if parity == 0 else od
    vp = C.c_void_p
    f = ref.ref_or_halfsweep
    f.restype = None
    f.argtypes = [vp, vp, vp, vp, C.c_int, C.c_int, C.c_double]
    f(u_su3.ctypes.data, g_su3.ctypes.data, nn.ctypes.data, sites.ctypes.data,
      volume(dims) // 2, ndir, omega)
    return g_su3


def functional(u2v, g2v, dims, ndir=4):
    """Returns (F, dAdA, maxdAdA) as calc_functional (overrelax.cu:610-671)."""
    nn = neighbours(dims)
    F = C.c_double()
    D = C.c_double()
    m = lib.orc_functional(u2v, g2v, nn, volume(dims), ndir, C.byref(F), C.byref(D))
    return F.value, D.value, m


def functional_slices(u2v, g2v, dims, ndir=3):
    nn = neighbours(dims)
    LT = dims[3]
    F = np.zeros(LT)
    D = np.zeros(LT)
    M = np.zeros(LT)
    lib.orc_functional_slices(u2v, g2v, nn, volume(dims), LT, ndir, F, D, M)
    return F, D, M


def overrelax_gauge(u2v, g2v, dims, maxit, eps, ck, omega=1.68, hist_cap=100000):
    """overrelax_gauge (overrelax.cu:677-823) on g2v in place.
    Returns (stop iteration or -1, history array (n,3) of F, dAdA, max)."""
    hist = np.zeros(3 * hist_cap)
    nh = C.c_int()
    it = lib.orc_overrelax_gauge(u2v, g2v, *dims, maxit, eps, ck, omega,
                                 hist.ctypes.data, hist_cap, C.byref(nh))
    n = min(nh.value, hist_cap)
    return it, hist[: 3 * n].reshape(n, 3)


def coulomb_gauge(u2v, g2v, dims, maxit, eps, ck, omega=1.68):
    iters = np.zeros(dims[3], dtype=np.int32)
    rc = lib.orc_coulomb_gauge(u2v, g2v, *dims, maxit, eps, ck, omega, iters)
    return rc, iters


def plaquette_2v(u2v, dims):
    return lib.orc_plaquette_2v(u2v, neighbours(dims), *dims)


# ---------------------------------------------------------------- host observables
def g_trafo(u_su3, g_su3, dims):
    u = np.array(u_su3, copy=True)
    lib.orc_g_trafo(u, np.ascontiguousarray(g_su3), neighbours(dims), volume(dims))
    return u


def gauge_functional(u_su3, dims):
    return lib.orc_gauge_functional(np.ascontiguousarray(u_su3), volume(dims))


def dAdA(u_su3, dims):
    m = C.c_double()
    d = lib.orc_dAdA(np.ascontiguousarray(u_su3), neighbours(dims), volume(dims), C.byref(m))
    return d, m.value


def mean_plaq(u_su3, dims):
    return lib.orc_mean_plaq(np.ascontiguousarray(u_su3), neighbours(dims), volume(dims))


# ---------------------------------------------------------------- MT
def mt_generate(params, state, nper, initialized, want_words=False):
    """RandomGPU restated (MersenneTwister.cu:142-199).  params: array of
    MTParams (modified: seed <- state[0]); state: uint32 (count*19)."""
    count = len(params)
    out = np.zeros(count * nper, dtype=np.float32)
    words = np.zeros(count * nper, dtype=np.uint32) if want_words else None
    lib.orc_mt_generate(params, state, count, nper, initialized, out.ctypes.data,
                        words.ctypes.data if words is not None else None)
    return (out, words) if want_words else out


# ---------------------------------------------------------------- SA sampler
def sa_sample_ref(gauss, unif, beta, wnorm, n, components=1):
    """alpha of dev_heatbath_su2 for w = (wnorm, 0, 0, 0), draw i using
    gauss[i] and unif[4i..4i+3] (simulated_annealing.cu:1-73): dev_get_a0 in
    float arithmetic (y = (|g|^2 - log(1 - u0)) / (|w|^2 beta), accepted when
    2 u1^2 <= 2 - y, then a0 = 1 - y), else alpha = w.  With v = w/|w| =
    (1, 0, 0, 0) the product a v is a itself.  components=1: alpha_0 (n,);
    components=4: (n, 4), the direction part as simulated_annealing.cu:47-55
    forms it: norm = sqrt(1 - a0^2), cos_theta = 2 u2 - 1, phi = 2 pi u3,
    (a1, a2, a3) = norm (sin_theta cos phi, sin_theta sin phi, cos_theta)."""
    f32 = np.float32
    g = np.abs(np.asarray(gauss[:n], dtype=f32))
    u = np.asarray(unif[:4 * n], dtype=f32).reshape(n, 4)
    b = -np.log(f32(1.0) - u[:, 0])
    y = g * g + b
    y = (y.astype(np.float64) / (wnorm * wnorm * beta)).astype(f32)
    acc = f32(2.0) * (u[:, 1] * u[:, 1]) <= f32(2.0) - y
    a0 = np.where(acc, (f32(1.0) - y).astype(np.float64), wnorm)
    if components == 1:
        return a0
    norm = np.sqrt(np.maximum(0.0, 1.0 - a0 * a0))
    ct = 2.0 * u[:, 2].astype(np.float64) - 1.0
    st = np.sqrt(1.0 - ct * ct)
    phi = 6.2831853071795862 * u[:, 3].astype(np.float64)
    out = np.zeros((n, 4))
    out[:, 0] = a0
    out[:, 1] = np.where(acc, norm * st * np.cos(phi), 0.0)
    out[:, 2] = np.where(acc, norm * st * np.sin(phi), 0.0)
    out[:, 3] = np.where(acc, norm * ct, 0.0)
    return out
