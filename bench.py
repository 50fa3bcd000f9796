"""Benchmark of the gauge-fixing hot path on one MI355X.

Workload (one GPU): Landau gauge fixing of a 32^4 SU(3) configuration at
beta = 6.0, heatbath-thermalized on the GPU from a cold start (500 sweeps,
twister seeds from srand(4357)), omega = 1.7, initial gauge transformation
from the Lehmer RNG with RandSeed = 351, stop at max_x dAdA(x) < 1e-12,
checked every 10 iterations.

A *step* is one complete gauge-fixing run from that fixed start to the stop
decision in the mixed path: FP32 sweeps until max dAdA < 1e-6 on packed
links (each real of a link as a 21-bit fixed-point integer, within 2^-21 of
the FP64 link, 32 bytes per link), then an FP64 reunitarisation of g and FP64 polish sweeps
with links and g in 32-bit fixed point (within 2^-32; FP64 arithmetic),
F/theta reductions in FP64, the stop confirmed by a functional on the FP64
links and FP64 g.  value = site-updates/s of those steps (iterations x V /
wall time).  The plain FP64 run to the same bound is timed beside it
(time-to-theta, speedup), with the error of F and of g against it.

roofline: the sweep kernel with the larger share of the mixed step
(launches x time per launch), its launch time the mean over HIP event pairs
on the library's stream around the runs of consecutive half sweeps between
two checks (about 17 launches each) of one more (untimed) mixed step run
right after the timed ones;
algorithmic bytes per site-update are 8 links + 3 g: FP32 g on 21-bit links
400 B, fixed-point links and g 528 B, FP32 on FP32 links 528 B, FP64 g on
24+16-bit links 768 B, FP64 1056 B.  The entry names the autotuned variant
and takes its HBM traffic and its rocprofv3 mean duration from
profiles/sweep_variants_32x4.json (made by --profile under rocprofv3 and
--summarize-profile).

gribov: the 64-copy Gribov workload at its stated schedule, run once: 64
Haar-random copies of one 24^4 beta = 6.0 configuration, simulated annealing
from T = 0.45 to 0.01 over 4000 steps of one SA sweep and 3 microcanonical
sweeps (in the mixed form: FP32 links and g), then mixed OR to theta < 1e-12
and the copy with the largest F; then the plain FP64 OR from the same
annealed copies, timed beside it.

cpu_baseline: an OpenMP CPU half sweep built from the reference's own host
SU(3) routines (su3manip.c, complex.c) on its 18-real AoS layout
(oracle/ref_sweep.c, compiled into oracle/_ref), timed on the host cores on
a bounded sample of the same 32^4 problem.

uneven: the uneven-size batch, run once: 16 configurations (4 x 16^4,
4 x 24^4, 4 x 32^3x64, 4 x 48^4) at beta 5.8, 6.0, 6.2 in turn, 200
heatbath sweeps each, fixed together in the mixed form (FP32 on 21-bit links,
FP64 polish on 32-bit fixed-point links and g) to theta < 1e-12.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-gribov] [--no-uneven]
       python bench.py --profile --plan FILE     (kernels for rocprofv3 counters)
       python bench.py --trace-step --plan FILE  (one mixed and one FP64 step, bracketed)
       python bench.py --only gribov|uneven [--persistent]  (one batch workload alone)
       python bench.py --only coulomb            (Coulomb 48^3x96: FP64, mixed at switch 1e-6, 1e-10)
       python bench.py --summarize-profile FILE STATS FETCH WRITE
       python bench.py --summarize-step FILE STATS
"""
import argparse
import csv
import glob
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
# sweep forms timed: name -> precision argument of time_sweeps (PREC_FP64,
# PREC_MIXED, SWEEP_PACKED64, SWEEP_PACKED32, SWEEP_FIXED64 in cudagfx_amd)
FORMS = {"fp64": 0, "fp32": 1, "packed64": 2, "packed32": 3, "fixed64": 4}
FORM_BYTES = {"fp64": 8 * 96 + 3 * 96, "fp32": 8 * 48 + 3 * 48, "packed64": 8 * 60 + 3 * 96,
              "packed32": 8 * 32 + 3 * 48, "fixed64": 8 * 48 + 3 * 48}
FORM_KERNEL = {"fp64": "k_or_sweep<double2, double>", "fp32": "k_or_sweep<float4, float>",
               "packed64": "k_or_sweep_packed<double2, double>",
               "packed32": "k_or_sweep_packed<float4, float>", "fixed64": "k_or_sweep<int4, double>"}
# the two sweep forms of the default mixed Landau run
MIXED_FORMS = ("packed32", "fixed64")
PROFILE_JSON = os.path.join(ROOT, "profiles", "sweep_variants_32x4.json")
GRIBOV = dict(dims=(24, 24, 24, 24), beta=6.0, ntherm=500, copies=64, steps=4000, Tmin=0.01,
              Tmax=0.45, micro=3, switch=1e-6, mtseed=62)
# the uneven batch: 4 lattices of each shape, beta cycling through BETAS,
# twister seed 800 + k and Lehmer RandSeed 900 + k for configuration k
UNEVEN = dict(shapes=[(16, 16, 16, 16)] * 4 + [(24, 24, 24, 24)] * 4 + [(32, 32, 32, 64)] * 4
              + [(48, 48, 48, 48)] * 4, betas=(5.8, 6.0, 6.2), ntherm=200, mtseed=800,
              randseed=900, switch=1e-6, periods=8)


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


def thermalized_config(gfx, dims=DIMS, beta=BETA, nsweep=None, mtseed=MTSEED):
    gfx.set_lattice(*dims)
    u = gfx.unit_gauge()
    gfx.init_thermalization(u)
    gfx.set_mt_params(synthetic_mt_params())
    gfx.set_mt_seed(mtseed)
    gfx.heatbath_sweeps(NTHERM if nsweep is None else nsweep, beta)
    plaq = gfx.calc_plaquette(0)
    u = gfx.get_gauge()
    gfx.finalize_MT()
    gfx.finalize_thermalization()
    return u, plaq


def profiled_variant(name, variant):
    """Traffic and rocprofv3 duration of one sweep variant from the committed
    profile, or None."""
    if not os.path.exists(PROFILE_JSON):
        return None
    try:
        return json.load(open(PROFILE_JSON))["forms"][name]["variants"].get(str(variant))
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


def _write_plan(path, plan):
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    with open(path, "w") as f:
        json.dump(plan, f, indent=1)


def _marker(gfx):
    """A small streaming copy (k_stream_copy of 4 KiB): brackets a section
    in a kernel trace."""
    gfx.stream_copy(4096, 1)
    gfx.sync()


def _setup_32x4(gfx, nsweep=None, dims=DIMS):
    V = int(np.prod(dims))
    u, plaq = thermalized_config(gfx, dims=dims, nsweep=nsweep)
    gfx.plant_seeds(RANDSEED)
    g0 = gfx.random_init_trafo_host(V)
    gfx.init_gaugefixing(u, g0)
    gfx.set_omega(OMEGA)
    return u, g0, plaq


def profile_run(gfx, plan_path, dims=DIMS, stochastic=True):
    """The kernels for counter collection, in a fixed order recorded in
    plan_path: the autotuner's timing of every variant of the five sweep
    forms, a marker copy, then 10 sweeps (20 launches) of every variant of
    every form in turn (tiling is data, so the order tells the variants of
    one kernel apart), a marker copy, 5 standalone FP64 checks, a
    31-iteration FP64 run with fused checks, the stochastic kernels (5 link
    heatbath sweeps with their twister refills, 5 SA steps with 3
    microcanonical sweeps each), and 5 copies of 512 MiB whose bytes are
    known (the counter calibration)."""
    u, g0, _ = _setup_32x4(gfx, nsweep=20, dims=dims)  # counters do not depend on the field values
    plan = {"lattice": list(dims), "forms": {}, "sequence": []}
    for name, form in FORMS.items():
        gfx.time_sweeps(1, form)
        v, blk, rolled, ms = gfx.sweep_variant(form, 4)
        plan["forms"][name] = {"picked": v, "tune_ms": [float(x) for x in ms],
                               "valid": [k for k in range(32) if ms[k] > 0]}
    _marker(gfx)
    for name, form in FORMS.items():
        for v in plan["forms"][name]["valid"]:
            gfx.set_sweep_variant(form, 4, v)
            ms = gfx.time_sweeps(10, form)
            plan["sequence"].append({"form": name, "variant": v, "launches": 20, "hip_event_ms": ms})
        gfx.set_sweep_variant(form, 4, plan["forms"][name]["picked"])
    _marker(gfx)
    for _ in range(5):
        gfx.calc_functional()
    gfx.set_trafo(g0)
    gfx.set_precision(gfx.PREC_FP64)
    gfx.overrelax_gauge(31, EPS, CK)
    gfx.sync()
    if not stochastic:
        tc = gfx.stream_copy(512 << 20, 5)
        gfx.finalize_gaugefixing()
        plan["copy_512MiB_ms"] = tc
        _write_plan(plan_path, plan)
        return {"profile": True, "plan": plan_path}
    # stochastic kernels: link heatbath (8 launches per sweep, a twister
    # refill after each), then SA (an even and an odd heatbath half sweep of
    # g per step, a refill after each, then the microcanonical sweeps)
    gfx.finalize_gaugefixing()
    gfx.set_lattice(*dims)
    gfx.init_thermalization(u)
    gfx.set_mt_params(synthetic_mt_params())
    gfx.set_mt_seed(MTSEED)
    gfx.heatbath_sweeps(5, BETA)
    gfx.finalize_MT()
    gfx.finalize_thermalization()
    gfx.set_lattice(*dims)
    gfx.init_gaugefixing(u, g0)
    gfx.set_mt_params(synthetic_mt_params())
    gfx.set_mt_seed(MTSEED)
    sa = gfx.saparam()
    sa.Tmin, sa.Tmax, sa.expo, sa.N, sa.checkint = 0.01, 0.45, -1.0, 5, 1000
    gfx.set_sa(0, 3, 4)
    gfx.simannealing_gauge()
    gfx.set_sa(0, 0, 4)
    gfx.finalize_MT()
    gfx.sync()
    tc = gfx.stream_copy(512 << 20, 5)
    gfx.finalize_gaugefixing()
    plan["copy_512MiB_ms"] = tc
    plan["stochastic"] = {"heatbath_sweeps": 5, "sa_steps": 5, "sa_micro": 3}
    _write_plan(plan_path, plan)
    return {"profile": True, "plan": plan_path}


def trace_step(gfx, plan_path):
    """One mixed step and one FP64 step of the 32^4 workload, each between
    two marker copies, after the autotuner has run (for a kernel trace: the
    kernels between the markers are exactly one step)."""
    u, g0, _ = _setup_32x4(gfx)
    for form in FORMS.values():
        gfx.time_sweeps(1, form)
    walls = {}
    for name, prec in (("mixed", gfx.PREC_MIXED), ("fp64", gfx.PREC_FP64)):
        gfx.set_trafo(g0)
        gfx.set_precision(prec, SWITCH)
        gfx.set_packed_links(1)
        gfx.sync()
        _marker(gfx)
        t0 = time.perf_counter()
        it = gfx.overrelax_gauge(20000, EPS, CK)
        gfx.sync()
        w = time.perf_counter() - t0
        _marker(gfx)
        ms, i32, i64 = gfx.last_run_stats()
        walls[name] = {"wall_s": w, "lib_ms": ms, "stop": it, "iters_fp32": i32, "iters_fp64": i64}
    variants = {}
    for n, p in FORMS.items():
        sv = gfx.sweep_variant(p, 4)
        variants[n] = sv[0] if sv else None
    gfx.finalize_gaugefixing()
    _write_plan(plan_path, {"steps": walls, "variants": variants})
    return {"trace_step": True, "plan": plan_path, "steps": walls}


# ---------------------------------------------------------------- rocprofv3 summaries
def _rows(d, suffix):
    files = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not files:
        raise SystemExit(f"no *{suffix} under {d}")
    rows = []
    for fn in files:
        with open(fn) as f:
            rows += list(csv.DictReader(f))
    return rows


def _dispatches(d):
    """Kernel dispatches of a --kernel-trace run in launch order: (name, us)."""
    rows = _rows(d, "kernel_trace.csv")
    rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    return [(r["Kernel_Name"], (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3, r)
            for r in rows]


def _counter_dispatches(d, counter):
    """Per dispatch (in launch order) the value of one counter of a --pmc run."""
    rows = [r for r in _rows(d, "counter_collection.csv") if r["Counter_Name"].startswith(counter)]
    acc = {}
    for r in rows:
        k = int(r["Dispatch_Id"])
        a = acc.setdefault(k, [r["Kernel_Name"], 0.0])
        a[1] += float(r["Counter_Value"])
    return [(acc[k][0], acc[k][1]) for k in sorted(acc)]


def _short(name):
    return name.split("(")[0].replace("void ", "")


def _marker_groups(seq):
    """(first, end) index ranges of the runs of consecutive k_stream_copy
    dispatches (a marker or the calibration copies) in a dispatch list."""
    groups = []
    for i, x in enumerate(seq):
        if "k_stream_copy" in x[0]:
            if groups and groups[-1][1] == i:
                groups[-1][1] = i + 1
            else:
                groups.append([i, i + 1])
    return groups


def summarize_profile(plan_path, stats_dir, fetch_dir, write_dir):
    plan = json.load(open(plan_path))
    V = int(np.prod(plan["lattice"]))
    trace = _dispatches(stats_dir)
    fetch = _counter_dispatches(fetch_dir, "FETCH_SIZE")
    write = _counter_dispatches(write_dir, "WRITE_SIZE")
    out = {"lattice": plan["lattice"], "forms": {}, "stochastic": {}}
    # calibration: the last 5 copies (512 MiB read and written each)
    fc = [v for n, v in fetch if "k_stream_copy" in n][-5:]
    wc = [v for n, v in write if "k_stream_copy" in n][-5:]
    nbytes = 512 << 20
    fscale = nbytes / (1024.0 * np.mean(fc))
    wscale = nbytes / (1024.0 * np.mean(wc))
    copy_us = [t for n, t, _ in trace if "k_stream_copy" in n][-5:]
    out["calibration"] = {"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                          "FETCH_SIZE_KiB": float(np.mean(fc)), "WRITE_SIZE_KiB": float(np.mean(wc)),
                          "fetch_scale": fscale, "write_scale": wscale,
                          "copy_us": float(np.mean(copy_us)),
                          "copy_GBs": 2 * nbytes / (np.mean(copy_us) * 1e-6) / 1e9}

    def between_markers(seq, key):
        g = _marker_groups(seq)
        return [x for x in seq[g[0][1]:g[1][0]] if key(x[0])]

    is_sweep = lambda n: "k_or_sweep" in n
    tr = between_markers(trace, is_sweep)
    fe = between_markers(fetch, is_sweep)
    wr = between_markers(write, is_sweep)
    k = 0
    for ent in plan["sequence"]:
        n = ent["launches"]
        name, v = ent["form"], ent["variant"]
        durs = [t for _, t, _ in tr[k:k + n]]
        f = np.mean([x[1] for x in fe[k:k + n]]) * 1024 * fscale
        w = np.mean([x[1] for x in wr[k:k + n]]) * 1024 * wscale
        kern = _short(tr[k][0])
        alg = (V // 2) * FORM_BYTES[name]
        us = float(np.mean(durs))
        out["forms"].setdefault(name, {"picked_on_profile_box": plan["forms"][name]["picked"],
                                       "algorithmic_bytes_per_launch": alg, "variants": {}})
        out["forms"][name]["variants"][str(v)] = {
            "kernel": kern, "tiling": v >> 3, "launches": n, "us_rocprof": us,
            "us_hip_events": ent["hip_event_ms"] * 1e3,
            "fetch_bytes": float(f), "write_bytes": float(w), "traffic_bytes": float(f + w),
            "traffic_over_algorithmic": float((f + w) / alg),
            "GBs_algorithmic": alg / (us * 1e-6) / 1e9, "frac_of_8TBs": alg / (us * 1e-6) / 8e12}
        k += n
    assert k == len(tr) == len(fe) == len(wr), (k, len(tr), len(fe), len(wr))
    # stochastic kernels and checks (after the second marker, before the calibration copies)
    gt, gf, gw = _marker_groups(trace), _marker_groups(fetch), _marker_groups(write)
    if len(gt) < 3:
        return out
    tail_t = trace[gt[1][1]:gt[2][0]]
    tail_f = fetch[gf[1][1]:gf[2][0]]
    tail_w = write[gw[1][1]:gw[2][0]]
    # grouped by kernel name in each run (the sweeps of the fused-check run
    # may use other variants from run to run: they are left out here)
    def group(seq):
        gr = {}
        for x in seq:
            n = _short(x[0])
            if "k_or_sweep" in n:
                continue
            gr.setdefault(n, []).append(x[1])
        return gr
    gt_, gf_, gw_ = group(tail_t), group(tail_f), group(tail_w)
    for n, us in gt_.items():
        if n not in gf_ or n not in gw_ or not (len(us) == len(gf_[n]) == len(gw_[n])):
            continue
        t = float(np.mean(us))
        fb = float(np.mean(gf_[n])) * 1024 * fscale
        wb = float(np.mean(gw_[n])) * 1024 * wscale
        out["stochastic"][n] = {"launches": len(us), "us_rocprof": t, "fetch_bytes": fb,
                                "write_bytes": wb, "traffic_GBs": (fb + wb) / (t * 1e-6) / 1e9}
    return out


def summarize_step(plan_path, stats_dir):
    plan = json.load(open(plan_path))
    trace = _dispatches(stats_dir)
    g = _marker_groups(trace)
    out = {"variants": plan["variants"], "steps": {}}
    for k, name in enumerate(("mixed", "fp64")):
        sec = trace[g[2 * k][1]:g[2 * k + 1][0]]
        first, last = sec[0][2], sec[-1][2]
        span_us = (int(last["End_Timestamp"]) - int(first["Start_Timestamp"])) / 1e3
        per = {}
        for n, t, _ in sec:
            p = per.setdefault(_short(n), [0, 0.0])
            p[0] += 1
            p[1] += t
        busy = sum(p[1] for p in per.values())
        out["steps"][name] = {"wall_s_host": plan["steps"][name]["wall_s"],
                              "gpu_span_s": span_us * 1e-6, "kernel_time_s": busy * 1e-6,
                              "stop": plan["steps"][name]["stop"],
                              "iters_fp32": plan["steps"][name]["iters_fp32"],
                              "iters_fp64": plan["steps"][name]["iters_fp64"],
                              "kernels": {n: {"launches": p[0], "total_ms": p[1] / 1e3,
                                              "mean_us": p[1] / p[0], "share": p[1] / busy}
                                          for n, p in sorted(per.items(), key=lambda kv: -kv[1][1])}}
    return out


# ---------------------------------------------------------------- the bench line
def gribov_run(gfx, persistent=False):
    """The Gribov-copy workload at its stated schedule, once (persistent:
    the check periods of single-precision stretches in persistent launches
    instead of graph replays)."""
    G = GRIBOV
    t0 = time.perf_counter()
    u, plaq = thermalized_config(gfx, G["dims"], G["beta"], G["ntherm"], G["mtseed"])
    gfx.set_omega(OMEGA)
    gfx.set_mt_params(synthetic_mt_params())
    gfx.set_mt_seed(G["mtseed"])
    b = gfx.Batch()
    # the mixed form throughout: the anneal on FP32 links and g as well
    b.set_anneal_precision(gfx.PREC_MIXED)
    b.set_persistent(int(persistent))
    base = b.add(G["dims"], u, gfx.unit_trafo())
    for _ in range(G["copies"] - 1):
        b.add_random_copy(base)
    gfx.sync()
    setup = time.perf_counter() - t0
    t1 = time.perf_counter()
    b.anneal(G["steps"], G["Tmin"], G["Tmax"], -1.0, G["micro"], 0, steps_per_graph=20)
    gfx.sync()
    t_sa = time.perf_counter() - t1
    # the annealed copies, so that the FP64 OR starts where the mixed one did
    annealed = [b.trafo(q) for q in range(G["copies"])]
    t2 = time.perf_counter()
    rc = b.run(20000, EPS, CK, gfx.PREC_MIXED, G["switch"], 8)
    gfx.sync()
    t_or = time.perf_counter() - t2
    res = [b.result(q) for q in range(G["copies"])]
    st = b.stats()
    qb, Fb = b.best_copy(base)
    # the plain FP64 OR of the same annealed copies, for comparison
    for q, g in enumerate(annealed):
        b.set_trafo(q, g)
    del annealed
    gfx.sync()
    t3 = time.perf_counter()
    rc64 = b.run(20000, EPS, CK, gfx.PREC_FP64, G["switch"], 8)
    gfx.sync()
    t_or64 = time.perf_counter() - t3
    res64 = [b.result(q) for q in range(G["copies"])]
    qb64, Fb64 = b.best_copy(base)
    b.set_anneal_precision(gfx.PREC_FP64)
    b.set_persistent(0)
    b.close()
    gfx.finalize_MT()
    V = int(np.prod(G["dims"]))
    stops = [r[3] for r in res]
    return {"workload": "64 Haar-random copies of one 24^4 beta=6.0 configuration (500 heatbath "
                        "sweeps), SA T=0.45->0.01 over 4000 steps of 1 SA + 3 microcanonical sweeps "
                        "(FP32 links and g, FP64 sampling), mixed OR (switch 1e-6) to theta<1e-12, "
                        "best copy by F",
            "rc": rc, "setup_s": setup, "sa_s": t_sa, "sa_ms_per_step": 1e3 * t_sa / G["steps"],
            "sa_site_sweeps_per_s": G["copies"] * V * 4 * G["steps"] / t_sa,
            "or_s": t_or, "or_persistent_launches": st["persistent_launches"],
            "or_graph_launches": st["graph_launches"], "or_site_updates": st["site_updates"],
            "or_site_updates_per_s": st["site_updates"] / t_or,
            "iters_fp32": st["iters_fp32"], "iters_fp64": st["iters_fp64"],
            "stop_min": int(min(stops)), "stop_max": int(max(stops)), "stop_mean": float(np.mean(stops)),
            "converged": int(sum(1 for r in res if r[2] < EPS and r[3] >= 0)),
            "best_copy": qb, "best_F": Fb, "F_min": float(min(r[0] for r in res)),
            "or_fp64": {"rc": rc64, "or_s": t_or64, "speedup_mixed": t_or64 / t_or,
                        "stop_mean": float(np.mean([r[3] for r in res64])),
                        "converged": int(sum(1 for r in res64 if r[2] < EPS and r[3] >= 0)),
                        "best_copy": qb64, "best_F": Fb64,
                        "max_F_rel_diff_per_copy": float(max(abs(a[0] - c[0]) / c[0]
                                                             for a, c in zip(res, res64)))},
            "plaquette": plaq}


def coulomb_run(gfx):
    """Coulomb gauge of a 48^3 x 96 beta = 6.2 configuration (100 heatbath
    sweeps from a cold start), every time slice to max dAdA < 1e-12 with
    checks every 10 iterations, all 96 slices per launch: plain FP64, then
    mixed with the switch at 1e-6 and at 1e-10 (FP32 slices on the 21-bit
    links, FP64 slices polished on the fixed-point links and g), each from
    the same initial g."""
    dims = (48, 48, 48, 96)
    V = int(np.prod(dims))
    t0 = time.perf_counter()
    u, plaq = thermalized_config(gfx, dims, 6.2, 100, 1001)
    gfx.plant_seeds(1002)
    g = gfx.random_init_trafo_host(V)
    gfx.set_lattice(*dims)
    gfx.init_gaugefixing(u, g)
    gfx.set_omega(OMEGA)
    # the autotuner times every sweep variant of a form on the form's first
    # use on a lattice; two checks' worth of each form first (the mixed one
    # with a switch that every slice passes at its first check, so that the
    # FP64 polish form is tuned as well) keep that one-time cost out of the
    # timed runs, as the warmup steps do for the 32^4 workload
    tw = time.perf_counter()
    for prec, switch in ((gfx.PREC_FP64, 1e-6), (gfx.PREC_MIXED, 1e30)):
        gfx.set_trafo(g)
        gfx.set_precision(prec, switch)
        gfx.coulomb_gauge(2 * CK, EPS, CK)
    gfx.sync()
    tune = time.perf_counter() - tw
    setup = time.perf_counter() - t0
    out = {"workload": "Coulomb gauge 48^3x96 beta=6.2 (100 heatbath sweeps), omega=1.7, per-slice "
                       "theta<1e-12, ck=10, all 96 slices per launch", "setup_s": setup,
           "autotune_warmup_s": tune, "plaquette": plaq}
    F64 = None
    for name, prec, switch in (("fp64", gfx.PREC_FP64, 1e-6), ("mixed", gfx.PREC_MIXED, 1e-6),
                               ("mixed_switch_1e-10", gfx.PREC_MIXED, 1e-10)):
        gfx.set_trafo(g)
        gfx.set_precision(prec, switch)
        gfx.sync()
        t1 = time.perf_counter()
        rc, iters = gfx.coulomb_gauge(30000, EPS, CK)
        gfx.sync()
        wall = time.perf_counter() - t1
        F, D, M = gfx.functional_slices()
        if F64 is None:
            F64 = F
        out[name] = {"rc": rc, "wall_s": wall, "iters_min": int(np.min(iters)),
                     "iters_mean": float(np.mean(iters)), "iters_max": int(np.max(iters)),
                     "max_theta": float(np.max(M)),
                     "F_rel_vs_fp64": float(np.max(np.abs(F - F64) / F64))}
    gfx.set_precision(gfx.PREC_FP64)
    gfx.finalize_gaugefixing()
    for name in ("mixed", "mixed_switch_1e-10"):
        out[name]["speedup"] = out["fp64"]["wall_s"] / out[name]["wall_s"]
    return out


def uneven_run(gfx, persistent=False):
    """The uneven-size batch, once: 16 heatbath-thermalized configurations
    of four shapes fixed together in the mixed form (persistent: as in
    gribov_run)."""
    U = UNEVEN
    t0 = time.perf_counter()
    cases = []
    for k, dims in enumerate(U["shapes"]):
        u, _ = thermalized_config(gfx, dims, U["betas"][k % 3], U["ntherm"], U["mtseed"] + k)
        gfx.plant_seeds(U["randseed"] + k)
        cases.append((dims, u, gfx.random_init_trafo_host(int(np.prod(dims)))))
    gfx.set_omega(OMEGA)
    b = gfx.Batch()
    b.set_persistent(int(persistent))
    for dims, u, g in cases:
        b.add(dims, u, g)
    gfx.sync()
    setup = time.perf_counter() - t0
    t1 = time.perf_counter()
    rc = b.run(60000, EPS, CK, gfx.PREC_MIXED, U["switch"], U["periods"])
    gfx.sync()
    wall = time.perf_counter() - t1
    res = [b.result(q) for q in range(len(cases))]
    st = b.stats()
    b.set_persistent(0)
    b.close()
    return {"workload": "16 configurations 4x16^4, 4x24^4, 4x32^3x64, 4x48^4 (31.2 M sites), beta "
                        "5.8/6.0/6.2 in turn, 200 heatbath sweeps each from per-configuration "
                        "twister seeds; mixed Landau OR (switch 1e-6) to theta<1e-12 in one batch",
            "rc": rc, "setup_s": setup, "wall_s": wall, "site_updates": st["site_updates"],
            "site_updates_per_s": st["site_updates"] / wall, "graph_launches": st["graph_launches"],
            "persistent_launches": st["persistent_launches"],
            "iters_fp32": st["iters_fp32"], "iters_fp64": st["iters_fp64"],
            "converged": int(sum(1 for r in res if r[2] < EPS and r[3] >= 0)),
            "stops": [int(r[3]) for r in res], "F": [r[0] for r in res],
            "max_dAdA": max(r[2] for r in res)}


def measure(args):
    import cudagfx_amd as gfx

    gfx.load()
    gfx.set_verbose(0)
    assert FORMS == {"fp64": gfx.PREC_FP64, "fp32": gfx.PREC_MIXED, "packed64": gfx.SWEEP_PACKED64,
                     "packed32": gfx.SWEEP_PACKED32, "fixed64": gfx.SWEEP_FIXED64}
    if args.profile:
        L = args.lattice
        return profile_run(gfx, args.plan, dims=(L, L, L, L), stochastic=L == DIMS[0])
    if args.trace_step:
        return trace_step(gfx, args.plan)
    if args.only == "gribov":
        return {"gribov": gribov_run(gfx, args.persistent)}
    if args.only == "uneven":
        return {"uneven": uneven_run(gfx, args.persistent)}
    if args.only == "coulomb":
        return {"coulomb": coulomb_run(gfx)}
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
    # launch times inside a step: one more mixed step (untimed) with an event
    # pair around each run of consecutive half sweeps between two checks
    gfx.set_sweep_sampling(1)
    inst = run(gfx.PREC_MIXED)
    in_step = {}
    for name in ("packed32", "fixed64"):
        mean, mn, n = gfx.sweep_sample_stats(FORMS[name])
        if n == 0:
            raise SystemExit(f"no sampled {name} launches in the instrumented step")
        in_step[name] = {"ms_mean": mean, "ms_min": mn, "launches_sampled": n}
    gfx.set_sweep_sampling(0)
    # back-to-back launches of every form from the start g (HIP events on the
    # library stream), each after 20 untimed launches of the same form
    gfx.set_trafo(g0)
    tk = {}
    for name in ("packed32", "fixed64", "fp64", "fp32", "packed64"):
        gfx.time_sweeps(10, FORMS[name])
        tk[name] = gfx.time_sweeps(50, FORMS[name])
    # the same run with the switch at 1e-10 (FP32 rounding floor of this
    # lattice ~5e-11; the stall guard covers lattices whose floor is higher)
    tuned = run(gfx.PREC_MIXED, SWITCH_LOW)
    # the mixed step on the FP32 and FP64 links, for comparison
    plain_links = run(gfx.PREC_MIXED, SWITCH, packed=0)
    gfx.set_packed_links(1)
    iters = last["it32"] + last["it64"]
    value = iters * V / (ms_step / 1000.0)
    gbs = {name: (V // 2) * FORM_BYTES[name] / (tk[name] * 1e-3) / 1e9 for name in FORMS}
    variants = {}
    for name, prec in FORMS.items():
        sv = gfx.sweep_variant(prec, 4)
        if sv is not None:
            variants[name] = {"variant": sv[0], "block": sv[1], "rolled": sv[2],
                              "ms_per_variant": [round(x, 5) for x in sv[3]]}
    # the sweep kernel with the larger share of the mixed step, by its launch
    # time inside the instrumented step
    share = {"packed32": 2 * inst["it32"] * in_step["packed32"]["ms_mean"],
             "fixed64": 2 * inst["it64"] * in_step["fixed64"]["ms_mean"]}
    name = max(share, key=share.get)
    v = variants[name]["variant"]
    prof = profiled_variant(name, v)
    alg = (V // 2) * FORM_BYTES[name]
    ach = alg / (in_step[name]["ms_mean"] * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS,
            "traffic": prof["traffic_bytes"] if prof else None,
            "kernel": FORM_KERNEL[name], "form": name, "variant": v,
            "ms_per_launch_in_step": in_step[name]["ms_mean"],
            "ms_per_launch_in_step_min": in_step[name]["ms_min"],
            "launches_sampled": in_step[name]["launches_sampled"],
            "ms_per_launch_back_to_back": tk[name],
            "frac_back_to_back": gbs[name] / HBM_PEAK_GBS,
            "ms_per_launch_rocprof_profile": prof["us_rocprof"] / 1e3 if prof else None,
            "frac_rocprof_profile": prof["frac_of_8TBs"] if prof else None,
            "traffic_over_algorithmic": prof["traffic_over_algorithmic"] if prof else None,
            "algorithmic_bytes_per_launch": (V // 2) * FORM_BYTES[name],
            "bytes_per_site_update": FORM_BYTES[name],
            "step_share_ms": {k: x for k, x in share.items()}}
    link_err = align_global(g_mixed, g_fp64)
    gfx.finalize_gaugefixing()

    gribov = None
    if not args.no_gribov:
        try:
            gribov = gribov_run(gfx)
        except Exception as e:  # reported, the Landau line stands on its own
            gribov = {"error": str(e)}

    uneven = None
    if not args.no_uneven:
        try:
            uneven = uneven_run(gfx)
        except Exception as e:  # reported, the Landau line stands on its own
            uneven = {"error": str(e)}

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
        "dtype": "f32 sweeps on 21-bit fixed-point links, then f64 arithmetic on 32-bit "
                 "fixed-point links and g; f64 reductions and stop check on the f64 links",
        "data": "synthetic: cold start + 500 GPU heatbath sweeps at beta=6.0 (twister params "
                "generated, srand(4357) seeds), Lehmer RandSeed=351 initial g",
        "config": {"workload": "Landau gauge fixing 32^4 beta=6.0, omega=1.7, theta<1e-12, ck=10",
                   "lattice": list(DIMS), "site_updates_per_step": iters * V,
                   "link_formats": {"fp32_phase": "21-bit fixed point per real (32 B per link)",
                                    "polish": "32-bit fixed point per real, links and g (48 B per matrix)"}},
        "time_to_theta_s": {"mixed": ms_step / 1000.0, "fp64": r64["wall_s"],
                            "speedup": r64["wall_s"] / (ms_step / 1000.0),
                            "mixed_median": float(np.median(walls))},
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
        "sweep_kernels_in_step": {n: dict(in_step[n], kernel=FORM_KERNEL[n],
                                          GBs=(V // 2) * FORM_BYTES[n] / (in_step[n]["ms_mean"] * 1e-3) / 1e9,
                                          frac=(V // 2) * FORM_BYTES[n] / (in_step[n]["ms_mean"] * 1e-3)
                                          / 1e9 / HBM_PEAK_GBS)
                                  for n in in_step},
        "sweep_kernels": {n: {"kernel": FORM_KERNEL[n], "ms": tk[n], "GBs": gbs[n],
                              "frac": gbs[n] / HBM_PEAK_GBS,
                              "bytes_per_site_update": FORM_BYTES[n],
                              "site_updates_per_s": (V // 2) / (tk[n] * 1e-3),
                              "variant": variants.get(n, {}).get("variant")} for n in FORMS},
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
        "gribov": gribov,
        "uneven": uneven,
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
    ap.add_argument("--no-gribov", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--only", choices=("gribov", "uneven", "coulomb"),
                    help="run one of the batch workloads alone (for a kernel trace)")
    ap.add_argument("--persistent", action="store_true",
                    help="with --only: persistent launches instead of graph replays")
    ap.add_argument("--profile", action="store_true",
                    help="kernels for rocprofv3 counters: every sweep variant, checks, stochastic "
                         "kernels, and 512 MiB streaming copies as the byte-counter reference")
    ap.add_argument("--trace-step", action="store_true",
                    help="one mixed and one FP64 step between marker copies, for a kernel trace")
    ap.add_argument("--plan", default=os.path.join(ROOT, "gpurun_out", "profile_plan.json"))
    ap.add_argument("--lattice", type=int, default=DIMS[0],
                    help="--profile on an L^4 lattice (the stochastic kernels at 32^4 only)")
    ap.add_argument("--summarize-profile", nargs=4, metavar=("PLAN", "STATS", "FETCH", "WRITE"))
    ap.add_argument("--summarize-step", nargs=2, metavar=("PLAN", "STATS"))
    args = ap.parse_args()
    if args.summarize_profile:
        print(json.dumps(summarize_profile(*args.summarize_profile), indent=1))
        return
    if args.summarize_step:
        print(json.dumps(summarize_step(*args.summarize_step), indent=1))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; the path would shard by "
                          "time-slab domain decomposition across GPUs, which is not built"}))
        sys.exit(2)
    with _CStdoutToStderr():
        out = measure(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
