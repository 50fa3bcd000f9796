"""Benchmark of the gauge-fixing hot path on one MI355X.

Workload (one GPU): Landau gauge fixing of a 32^4 SU(3) configuration at
beta = 6.0, heatbath-thermalized on the GPU from a cold start (500 sweeps,
twister seeds from srand(4357)), omega = 1.7, initial gauge transformation
from the Lehmer RNG with RandSeed = 351, stop at max_x dAdA(x) < 1e-12,
checked every 10 iterations.

A *step* is one complete gauge-fixing run from that fixed start to the stop
decision in the mixed path (FP32 12-real sweeps until max dAdA < 1e-6,
then FP64 reunitarisation and FP64 polish sweeps; both phases read packed
links, each real as a 24-bit fixed-point high part (the FP32 phase, within
2^-24) plus a 16-bit low part (the polish, within 2^-38 of the FP64 links);
F/theta reductions in FP64; the stop confirmed by a functional on the FP64
links).
value = site-updates/s of those steps (iterations x V / wall time).  The
plain FP64 run to the same bound is timed beside it (time-to-theta,
speedup), with the error of F and of the links of the mixed result against
it.  The roofline entry is the sweep kernel that dominates the mixed run,
timed with HIP events on the library's stream; algorithmic bytes per
site-update are 8 links + 3 g (FP32 g on packed high parts: 432 B, FP64 g
on packed links: 768 B, FP32 on FP32 links: 528 B, FP64: 1056 B).

cpu_baseline: an OpenMP CPU half sweep built from the reference's own host
SU(3) routines (su3manip.c, complex.c) on its 18-real AoS layout
(oracle/ref_sweep.c, compiled into oracle/_ref), timed on the host cores on
a bounded sample of the same 32^4 problem.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

DIMS = (32, 32, 32, 32)
BETA = 6.0
NTHERM = 500
OMEGA = 1.7
EPS = float(np.float32(1e-12))  # the input grammar reads eps through a C float
SWITCH = 1e-6
SWITCH_LOW = 1e-10
CK = 10
RANDSEED = 351
MTSEED = 4357
HBM_PEAK_GBS = 8000.0
BYTES_FP32 = 8 * 48 + 3 * 48
BYTES_FP64 = 8 * 96 + 3 * 96
BYTES_P32 = 8 * 36 + 3 * 48
BYTES_P64 = 8 * 60 + 3 * 96
# the roofline kernels by sweep form
KERNEL_P32 = "k_or_sweep_packed<float4,float>"
KERNEL_P64 = "k_or_sweep_packed<double2,double>"
# sweep forms timed: name -> precision argument of time_sweeps (the values of
# PREC_FP64, PREC_MIXED, SWEEP_PACKED64, SWEEP_PACKED32 in cudagfx_amd)
FORMS = {"fp64": 0, "fp32": 1, "packed64": 2, "packed32": 3}
FORM_BYTES = {"fp64": BYTES_FP64, "fp32": BYTES_FP32, "packed64": BYTES_P64, "packed32": BYTES_P32}


def synthetic_mt_params():
    """Twister parameter records (4096 x {matrix_a, mask_b, mask_c, seed}):
    the reference loads them from MersenneTwister.dat, which does not ship
    with this repository; these are generated from a fixed splitmix64 stream
    with the top bit of matrix_a set and the MT19937 tempering masks."""
    out = np.zeros((4096, 4), dtype=np.uint32)
    x = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        for i in range(4096):
            x = np.uint64(x + np.uint64(0x9E3779B97F4A7C15))
            z = x
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
            out[i] = (np.uint32(int(z) & 0xFFFFFFFF) | np.uint32(0x80000000), 0x9D2C5680,
                      0xEFC60000, 0)
    return out.ravel()


def thermalized_config(gfx):
    gfx.set_lattice(*DIMS)
    u = gfx.unit_gauge()
    gfx.init_thermalization(u)
    gfx.set_mt_params(synthetic_mt_params())
    gfx.set_mt_seed(MTSEED)
    gfx.heatbath_sweeps(NTHERM, BETA)
    plaq = gfx.calc_plaquette(0)
    u = gfx.get_gauge()
    gfx.finalize_MT()
    gfx.finalize_thermalization()
    return u, plaq


def read_traffic(kernel):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", "sweep_traffic_32x4.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get(kernel, {}).get("bytes_per_launch")
    except Exception:
        return None


def _omp_threads(n):
    """Set the OpenMP team size of later parallel regions in this thread."""
    import ctypes
    try:
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))
        return True
    except OSError:
        return False


def cpu_baseline(u_su3, g_su3, budget_s=10.0, budget_1core_s=5.0):
    """CPU half sweeps of the same 32^4 problem on the host cores: the sweep
    built from the reference's own host SU(3) routines on its 18-real AoS
    layout (oracle/ref_sweep.c in oracle/_ref, kind "reference"), or, where
    that was not built, the oracle's 2-row restatement (kind "port")."""
    import oracle

    V = int(np.prod(DIMS))
    ref = oracle.ref_lib() is not None
    if ref:
        u = np.ascontiguousarray(u_su3).copy()
        g = np.ascontiguousarray(g_su3).copy()
        sweep = lambda p: oracle.ref_halfsweep(u, g, DIMS, p, 4, OMEGA)
    else:
        u2 = oracle.su3_to_2v(u_su3)
        g2 = oracle.su3_to_2v(g_su3)
        sweep = lambda p: oracle.halfsweep(u2, g2, DIMS, p, 4, OMEGA)

    def timed(budget):
        n = 0
        t0 = time.perf_counter()
        while True:
            sweep(n % 2)
            n += 1
            if time.perf_counter() - t0 > budget and n >= 2:
                break
        return n, time.perf_counter() - t0

    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    _omp_threads(cores)
    n, dt = timed(budget_s)
    rate = n * (V // 2) / dt
    what = ("oracle/ref_sweep.c ref_or_halfsweep (the reference's host su3mult/su3dagger/su3add "
            "routines on its 18-real AoS layout)" if ref else
            "oracle/oracle.c orc_or_halfsweep (2-row restatement)")
    out = {"value": rate, "unit": "site-updates/s", "cores": cores,
           "kind": "reference" if ref else "port",
           "sample": f"{n} FP64 half-sweeps of the 32^4 problem ({n * (V // 2)} site-updates) "
                     f"by {what} with OpenMP, {dt:.1f} s",
           "GBs_at_1584B" if ref else "GBs_at_1056B": rate * (1584 if ref else 1056) / 1e9}
    if _omp_threads(1):
        n1, dt1 = timed(budget_1core_s)
        out["value_1core"] = n1 * (V // 2) / dt1
        out["sample_1core"] = f"{n1} half-sweeps on 1 thread, {dt1:.1f} s"
        _omp_threads(cores)
    return out


def align_global(g_a, g_b):
    """Links are fixed up to a constant gauge rotation h (F is invariant under
    g -> h g); return max |g_a - h g_b| with h the best unitary fit."""
    A = g_a.reshape(-1, 3, 3, 2)
    B = g_b.reshape(-1, 3, 3, 2)
    A = A[..., 0] + 1j * A[..., 1]
    B = B[..., 0] + 1j * B[..., 1]
    M = np.einsum("nij,nkj->ik", A, B.conj())  # sum_x A(x) B(x)^+
    U, _, Vh = np.linalg.svd(M)
    h = U @ Vh
    return float(np.abs(A - np.einsum("ij,njk->nik", h, B)).max())


def profile_run(gfx):
    """The kernels of the timed region in isolation, for counter collection:
    the autotuner's variant timing for the four sweep forms, then 20 FP64,
    20 FP32, 20 packed FP64 and 20 packed FP32 sweeps of the picked variants
    (the last 40 launches of each picked kernel), 5 standalone FP64 checks, a 31-iteration FP64 run with fused
    checks every 10 iterations, and a plain copy of 512 MiB whose bytes are
    known."""
    global NTHERM
    NTHERM = 20  # counters do not depend on the field values
    u, _ = thermalized_config(gfx)
    gfx.plant_seeds(RANDSEED)
    V = int(np.prod(DIMS))
    g0 = gfx.random_init_trafo_host(V)
    gfx.init_gaugefixing(u, g0)
    gfx.set_omega(OMEGA)
    for form in FORMS.values():
        gfx.time_sweeps(1, form)
    t = {name: gfx.time_sweeps(20, form) for name, form in FORMS.items()}
    for _ in range(5):
        gfx.calc_functional()
    gfx.set_trafo(g0)
    gfx.set_precision(gfx.PREC_FP64)
    gfx.overrelax_gauge(31, EPS, CK)
    tc = gfx.stream_copy(512 << 20, 5)
    out = {"profile": True, "sweep_ms": t, "copy_512MiB_ms": tc,
           "variants": {name: list(gfx.sweep_variant(form, 4)[:3]) for name, form in FORMS.items()}}
    gfx.finalize_gaugefixing()
    return out


def measure(args):
    import cudagfx_amd as gfx

    gfx.load()
    gfx.set_verbose(0)
    assert FORMS == {"fp64": gfx.PREC_FP64, "fp32": gfx.PREC_MIXED,
                     "packed64": gfx.SWEEP_PACKED64, "packed32": gfx.SWEEP_PACKED32}
    if args.profile:
        return profile_run(gfx)
    V = int(np.prod(DIMS))
    t_setup = time.perf_counter()
    u, plaq = thermalized_config(gfx)
    gfx.plant_seeds(RANDSEED)
    g0 = gfx.random_init_trafo_host(V)
    t_up = time.perf_counter()
    gfx.init_gaugefixing(u, g0)
    gfx.sync()
    t_up = time.perf_counter() - t_up
    gfx.set_omega(OMEGA)
    # pick the sweep variants (autotuner) before anything is timed
    for form in FORMS.values():
        gfx.time_sweeps(1, form)
    setup_s = time.perf_counter() - t_setup

    def run(prec, switch=SWITCH, packed=1):
        gfx.set_trafo(g0)
        gfx.set_precision(prec, switch)
        gfx.set_packed_links(packed)
        gfx.sync()
        t0 = time.perf_counter()
        it = gfx.overrelax_gauge(20000, EPS, CK)
        gfx.sync()
        wall = time.perf_counter() - t0
        ms, i32, i64 = gfx.last_run_stats()
        F, D, M = gfx.results()
        return dict(iter=it, wall_s=wall, lib_ms=ms, it32=i32, it64=i64, F=F, dAdA=D, max=M)

    # plain FP64 reference run
    r64 = run(gfx.PREC_FP64)
    g_fp64 = gfx.get_trafo()
    # mixed steps
    for _ in range(args.warmup):
        run(gfx.PREC_MIXED)
    walls, last = [], None
    for _ in range(args.steps):
        last = run(gfx.PREC_MIXED)
        walls.append(last["wall_s"])
    t_down = time.perf_counter()
    g_mixed = gfx.get_trafo()
    t_down = time.perf_counter() - t_down
    ms_step = 1000.0 * max(walls)
    # the same run with the switch at 1e-10 (FP32 rounding floor of this
    # lattice ~5e-11; the stall guard covers lattices whose floor is higher)
    tuned = run(gfx.PREC_MIXED, SWITCH_LOW)
    # the mixed step on the FP32 and FP64 links, for comparison
    plain_links = run(gfx.PREC_MIXED, SWITCH, packed=0)
    gfx.set_packed_links(1)
    iters = last["it32"] + last["it64"]
    value = iters * V / (ms_step / 1000.0)

    # kernel timing of the sweeps (HIP events on the library stream)
    gfx.set_trafo(g0)
    tk = {name: gfx.time_sweeps(50, form) for name, form in FORMS.items()}
    gbs = {name: (V // 2) * FORM_BYTES[name] / (tk[name] * 1e-3) / 1e9 for name in FORMS}
    # the sweep kernel with the larger share of the mixed step
    dom32 = last["it32"] * tk["packed32"] >= last["it64"] * tk["packed64"]
    name = "packed32" if dom32 else "packed64"
    kernel = KERNEL_P32 if dom32 else KERNEL_P64
    roof = {"bound": "hbm", "achieved": gbs[name], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": gbs[name] / HBM_PEAK_GBS, "traffic": read_traffic(kernel), "kernel": kernel,
            "ms_per_launch": tk[name],
            "algorithmic_bytes_per_launch": (V // 2) * FORM_BYTES[name]}

    link_err = align_global(g_mixed, g_fp64)
    variants = {}
    for name, prec in FORMS.items():
        sv = gfx.sweep_variant(prec, 4)
        if sv is not None:
            variants[name] = {"variant": sv[0], "block": sv[1], "rolled": sv[2],
                              "ms_per_variant": [round(x, 5) for x in sv[3]]}
    gfx.finalize_gaugefixing()

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(u, g0)
        except Exception as e:  # the oracle is optional on a box without it
            cpu = {"error": str(e)}

    out = {
        "metric": "site-updates/s + time-to-theta<1e-12, 32^4 Landau, mixed vs FP64; HBM GB/s",
        "value": value,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 sweeps + f64 reductions/polish",
        "data": "synthetic: cold start + 500 GPU heatbath sweeps at beta=6.0 (twister params "
                "generated, srand(4357) seeds), Lehmer RandSeed=351 initial g",
        "config": {"workload": "Landau gauge fixing 32^4 beta=6.0, omega=1.7, theta<1e-12, ck=10",
                   "lattice": list(DIMS), "site_updates_per_step": iters * V},
        "time_to_theta_s": {"mixed": ms_step / 1000.0, "fp64": r64["wall_s"],
                            "speedup": r64["wall_s"] / (ms_step / 1000.0)},
        "mixed_switch_1e-10": {"time_s": tuned["wall_s"], "speedup": r64["wall_s"] / tuned["wall_s"],
                               "iters_fp32": tuned["it32"], "iters_fp64": tuned["it64"],
                               "stop": tuned["iter"], "theta": tuned["max"],
                               "F_rel": abs(tuned["F"] - r64["F"]) / r64["F"]},
        "mixed_on_fp32_fp64_links": {"time_s": plain_links["wall_s"],
                                     "speedup": r64["wall_s"] / plain_links["wall_s"],
                                     "iters_fp32": plain_links["it32"], "iters_fp64": plain_links["it64"],
                                     "stop": plain_links["iter"],
                                     "F_rel": abs(plain_links["F"] - r64["F"]) / r64["F"]},
        "iterations": {"mixed_fp32": last["it32"], "mixed_fp64": last["it64"],
                       "fp64": r64["it64"], "stop_mixed": last["iter"], "stop_fp64": r64["iter"]},
        "sweep_kernels": {name: {"ms": tk[name], "GBs": gbs[name], "bytes_per_site_update": FORM_BYTES[name],
                                 "site_updates_per_s": (V // 2) / (tk[name] * 1e-3)} for name in FORMS},
        "accuracy_mixed_vs_fp64": {"F_rel": abs(last["F"] - r64["F"]) / r64["F"],
                                   "theta_mixed": last["max"], "theta_fp64": r64["max"],
                                   "dAdA_mixed": last["dAdA"], "dAdA_fp64": r64["dAdA"],
                                   "dAdA_over_3_fp64": r64["dAdA"] / 3.0,
                                   "F_mixed": last["F"], "F_fp64": r64["F"],
                                   "g_maxabs_after_global_rotation": link_err},
        "autotuned_sweep": variants,
        "host_boundary": {"upload_links_and_g_s": t_up, "download_g_s": t_down,
                          "pcie_inclusive_site_updates_per_s":
                              iters * V / (ms_step / 1000.0 + t_up + t_down)},
        "plaquette": plaq,
        "setup_s": setup_s,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    return out


class _CStdoutToStderr:
    """The library's own messages (C printf, as the reference prints them) go
    to stderr while this is active, so that stdout carries the JSON line only."""

    def __enter__(self):
        import ctypes
        self.libc = ctypes.CDLL(None)
        sys.stdout.flush()
        self.libc.fflush(None)
        self.saved = os.dup(1)
        os.dup2(2, 1)
        return self

    def __exit__(self, *exc):
        self.libc.fflush(None)
        sys.stdout.flush()
        os.dup2(self.saved, 1)
        os.close(self.saved)
        return False


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="short run for rocprofv3: autotuning, 20 sweeps of each form, "
                         "checks, and a 512 MiB streaming copy as the byte-counter reference")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; the path would shard by "
                          "time-slab domain decomposition across GPUs, which is not built"}))
        sys.exit(2)
    with _CStdoutToStderr():
        out = measure(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
