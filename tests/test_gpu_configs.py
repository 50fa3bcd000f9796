"""GPU tests of the workloads beside the 32^4 benchmark, at their real shapes
where the run fits a test and at small shapes against the oracle:

* uneven batch: 16 configurations of 4 shapes at beta in {5.8, 6.0, 6.2},
  heatbath-thermalized from per-configuration twister seeds, mixed-precision
  Landau OR in one batch;
* Coulomb gauge on 48^3 x 96 at beta = 6.2, all 96 time slices in one
  launch per half sweep, FP32 bulk + FP64 polish.

Results of the full shapes are checked by size-independent properties (every
lattice / slice reaches the bound, the fixed links reproduce the reported F
and theta on the host, mixed and FP64 give the same F); the small shapes are
checked against single FP64 oracle runs.
"""
import time

import numpy as np
import pytest

from test_gpu_or import EPS12
from test_gpu_stoch import synthetic_params

pytestmark = pytest.mark.gpu

BETAS = (5.8, 6.0, 6.2)


def _thermalized(gfx, dims, beta, nsweep, seed):
    gfx.set_lattice(*dims)
    gfx.init_thermalization(gfx.unit_gauge())
    gfx.set_mt_params(synthetic_params(seed).ravel())
    gfx.set_mt_seed(seed)
    gfx.heatbath_sweeps(nsweep, beta)
    u = gfx.get_gauge()
    gfx.finalize_MT()
    gfx.finalize_thermalization()
    return u


def _uneven_batch(gfx, orc, shapes, nsweep, seed0):
    cases = []
    for k, dims in enumerate(shapes):
        beta = BETAS[k % 3]
        u = _thermalized(gfx, dims, beta, nsweep, seed0 + k)
        g = orc.random_init_trafo(orc.volume(dims), seed0 + 100 + k)
        cases.append((dims, beta, u, g))
    return cases


def test_uneven_batch_small_shapes_vs_oracle(gfx, orc):
    """16 configurations, 4 shapes x 4, beta in {5.8, 6.0, 6.2}, per-config
    seeds; the mixed batch matches each configuration's FP64 oracle run in F
    (1e-10 relative) and reaches the bound."""
    shapes = [(8, 8, 8, 8)] * 4 + [(6, 6, 6, 6)] * 4 + [(4, 4, 4, 8)] * 4 + [(8, 8, 4, 6)] * 4
    gfx.set_omega(1.7)
    cases = _uneven_batch(gfx, orc, shapes, 100, 700)
    b = gfx.Batch()
    for dims, _, u, g in cases:
        b.add(dims, u, g)
    rc = b.run(20000, EPS12, 10, gfx.PREC_MIXED, 1e-6, 4)
    res = [b.result(q) for q in range(len(cases))]
    st = b.stats()
    b.close()
    assert rc == 0 and st["iters_fp32"] > 0 and st["iters_fp64"] > 0
    for q, (dims, beta, u, g) in enumerate(cases):
        u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
        it, hist = orc.overrelax_gauge(u2, g2, dims, 20000, EPS12, 10, 1.7)
        assert it > 0 and res[q][2] < EPS12, (q, dims, beta)
        assert abs(res[q][0] - hist[-1, 0]) / hist[-1, 0] < 1e-10, (q, dims, beta)


@pytest.mark.timeout(900)
def test_uneven_batch_full_shapes(gfx, orc):
    """The uneven batch at its real shapes: 4 x 16^4, 4 x 24^4, 4 x 32^3x64,
    4 x 48^4 (31.2 M sites), 200 heatbath sweeps per configuration from a
    cold start, mixed OR to theta < 1e-12 in one batch with the autotuned
    sweep; every fixed configuration is checked with the oracle's host
    functional (F to 1e-12, max dAdA below the bound)."""
    shapes = ([(16, 16, 16, 16)] * 4 + [(24, 24, 24, 24)] * 4 + [(32, 32, 32, 64)] * 4
              + [(48, 48, 48, 48)] * 4)
    gfx.set_omega(1.7)
    cases = []
    for k, dims in enumerate(shapes):
        u = _thermalized(gfx, dims, BETAS[k % 3], 200, 800 + k)
        gfx.plant_seeds(900 + k)
        cases.append((dims, u, gfx.random_init_trafo_host(orc.volume(dims))))
    b = gfx.Batch()
    for dims, u, g in cases:
        b.add(dims, u, g)
    t0 = time.perf_counter()
    rc = b.run(60000, EPS12, 10, gfx.PREC_MIXED, 1e-6, 8)
    wall = time.perf_counter() - t0
    res = [b.result(q) for q in range(len(cases))]
    st = b.stats()
    v = b.variant(gfx.PREC_MIXED)
    gs = [b.trafo(q) for q in range(len(cases))]
    b.close()
    print(f"uneven batch: {wall:.1f} s, {st}, FP32 variant {v[0] if v else None}")
    assert rc == 0 and st["graph_launches"] > 0
    assert all(r[2] < EPS12 and r[3] >= 0 for r in res)
    # every configuration: the oracle's host functional of its fixed links
    for q, (dims, u, _) in enumerate(cases):
        uf = orc.g_trafo(u, gs[q], dims)
        assert abs(orc.gauge_functional(uf, dims) - res[q][0]) < 1e-12, (q, dims)
        assert orc.dAdA(uf, dims)[1] < 2 * EPS12, (q, dims)


@pytest.mark.timeout(900)
def test_coulomb_48x96_mixed_vs_fp64(gfx, orc):
    """Coulomb gauge of a 48^3 x 96 beta = 6.2 configuration: all 96 slices
    reach max dAdA < 1e-12 in both precision forms, their F agree to 1e-10
    relative, and the host functional of the fixed slices agrees."""
    dims = (48, 48, 48, 96)
    V = orc.volume(dims)
    u = _thermalized(gfx, dims, 6.2, 100, 1001)
    gfx.plant_seeds(1002)
    g = gfx.random_init_trafo_host(V)
    gfx.set_lattice(*dims)
    gfx.set_verbose(0)
    gfx.init_gaugefixing(u, g)
    gfx.set_omega(1.7)
    out = {}
    for name, prec in (("fp64", gfx.PREC_FP64), ("mixed", gfx.PREC_MIXED)):
        gfx.set_trafo(g)
        gfx.set_precision(prec, 1e-6)
        t0 = time.perf_counter()
        rc, iters = gfx.coulomb_gauge(30000, EPS12, 10)
        wall = time.perf_counter() - t0
        F, D, M = gfx.functional_slices()
        out[name] = (rc, iters, F, M, wall, gfx.get_trafo() if name == "mixed" else None)
    gfx.set_precision(gfx.PREC_FP64)
    gfx.finalize_gaugefixing()
    print({k: (v[0], int(np.min(v[1])), int(np.mean(v[1])), int(np.max(v[1])), v[4])
           for k, v in out.items()})
    for name in out:
        rc, iters, F, M, _, _ = out[name]
        assert rc == 0 and np.all(iters >= 0) and np.all(M < EPS12), name
    assert np.max(np.abs(out["mixed"][2] - out["fp64"][2]) / out["fp64"][2]) < 1e-10
    # host check of the mixed result: per-slice F and max dAdA of the spatial links
    u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(out["mixed"][5])
    Fo, Do, Mo = orc.functional_slices(u2, g2, dims, 3)
    assert np.max(np.abs(Fo - out["mixed"][2])) < 1e-12
    assert np.all(Mo < 2 * EPS12)
