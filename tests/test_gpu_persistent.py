"""GPU tests of the batching form's persistent launches: the check periods of
single-precision stretches run inside one cooperative launch.  Every lattice
must end where the graph replays leave it (same stop iteration, g to
rounding) and where the single-lattice oracle run ends (+-2, 1e-10)."""
import numpy as np
import pytest

from test_gpu_or import EPS12, smooth_config

pytestmark = pytest.mark.gpu


def _oracle_run(orc, dims, u, g, omega, ck):
    u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
    it, hist = orc.overrelax_gauge(u2, g2, dims, 4000, EPS12, ck, omega)
    return it, orc.from_2v(g2), hist


def _run(gfx, cases, persistent, prec, ck, periods, eps=EPS12, copies=()):
    b = gfx.Batch()
    b.set_persistent(persistent)
    for dims, u, g in cases:
        b.add(dims, u, g)
    for base, g in copies:
        b.add_copy(base, g)
    n = len(cases) + len(copies)
    print(f"batch run: {n} lattices, persistent {persistent}, precision {prec}, ck {ck}", flush=True)
    rc = b.run(8000, eps, ck, prec, 1e-6, periods)
    res = [b.result(q) for q in range(n)]
    gs = [b.trafo(q) for q in range(n)]
    st = b.stats()
    b.set_persistent(0)
    b.close()
    return rc, res, gs, st


@pytest.mark.parametrize("ck", [1, 5])
def test_persistent_fp64_batch_matches_graphs_and_oracle(gfx, orc, ck):
    gfx.set_omega(1.7)
    shapes = [(8, 8, 8, 8), (4, 4, 4, 8), (8, 4, 6, 4)]
    cases = []
    for k, dims in enumerate(shapes):
        u = smooth_config(orc, dims, 100 + k, eps=0.3)
        g = orc.random_init_trafo(orc.volume(dims), 200 + k)
        cases.append((dims, u, g))
    rc_g, res_g, gs_g, st_g = _run(gfx, cases, 0, gfx.PREC_FP64, ck, 2)
    rc_p, res_p, gs_p, st_p = _run(gfx, cases, 1, gfx.PREC_FP64, ck, 2)
    assert rc_g == 0 and rc_p == 0
    assert st_p["persistent_launches"] > 0 and st_g["persistent_launches"] == 0
    for q, (dims, u, g) in enumerate(cases):
        assert res_p[q][3] == res_g[q][3], (q, res_p[q][3], res_g[q][3])
        assert np.abs(gs_p[q] - gs_g[q]).max() < 1e-13
        assert abs(res_p[q][0] - res_g[q][0]) < 1e-14
        it, g_cpu, hist = _oracle_run(orc, dims, u, g, 1.7, ck)
        assert abs(res_p[q][3] - it) <= 2, (q, res_p[q][3], it)
        assert np.abs(gs_p[q] - g_cpu).max() < 1e-10


def test_persistent_gribov_copies(gfx, orc):
    """Copies sharing their links: persistent launches end where the graph
    replays do."""
    gfx.set_omega(1.7)
    dims = (8, 8, 8, 8)
    V = orc.volume(dims)
    u = smooth_config(orc, dims, 300, eps=0.4)
    starts = [orc.random_init_trafo(V, 400 + k) for k in range(4)]
    cases = [(dims, u, starts[0])]
    copies = [(0, starts[k]) for k in range(1, 4)]
    rc_g, res_g, gs_g, _ = _run(gfx, cases, 0, gfx.PREC_FP64, 2, 2, copies=copies)
    rc_p, res_p, gs_p, st_p = _run(gfx, cases, 1, gfx.PREC_FP64, 2, 2, copies=copies)
    assert rc_g == 0 and rc_p == 0 and st_p["persistent_launches"] > 0
    for q in range(4):
        assert res_p[q][3] == res_g[q][3]
        assert np.abs(gs_p[q] - gs_g[q]).max() < 1e-13


def test_persistent_mixed_batch(gfx, orc):
    """Mixed batch (FP32 on the 21-bit links, the polish on fixed-point links
    and g): the stretches with every lattice in one precision run
    persistent; the stops and F equal the graph replays', every g is below
    eps on the exact links."""
    gfx.set_omega(1.7)
    shapes = [(8, 8, 8, 8), (8, 8, 4, 8)]
    cases = [(dims, smooth_config(orc, dims, 500 + k, eps=0.35),
              orc.random_init_trafo(orc.volume(dims), 600 + k)) for k, dims in enumerate(shapes)]
    rc_g, res_g, gs_g, st_g = _run(gfx, cases, 0, gfx.PREC_MIXED, 10, 2)
    rc_p, res_p, gs_p, st_p = _run(gfx, cases, 1, gfx.PREC_MIXED, 10, 2)
    assert rc_g == 0 and rc_p == 0
    assert st_p["persistent_launches"] > 0
    assert st_p["iters_fp32"] > 0 and st_p["iters_fp64"] > 0
    for q, (dims, u, _) in enumerate(cases):
        assert res_p[q][3] == res_g[q][3], (q, res_p[q][3], res_g[q][3])
        assert abs(res_p[q][0] - res_g[q][0]) < 1e-13
        assert res_p[q][2] < EPS12
        uf = orc.g_trafo(u, gs_p[q], dims)
        assert orc.dAdA(uf, dims)[1] < EPS12 * (1 + 1e-9)


def test_persistent_run_cut_at_maxit(gfx, orc):
    """A run cut at maxit (no lattice reaches eps; 203 is not a whole number
    of check periods, so the last iterations run eagerly after the
    persistent launches) leaves every lattice where the graph replays leave
    it."""
    gfx.set_omega(1.7)
    shapes = [(8, 8, 8, 8), (4, 4, 4, 8)]
    cases = [(dims, smooth_config(orc, dims, 800 + k, eps=0.3),
              orc.random_init_trafo(orc.volume(dims), 810 + k)) for k, dims in enumerate(shapes)]
    out = []
    for persistent in (0, 1):
        b = gfx.Batch()
        b.set_persistent(persistent)
        for dims, u, g in cases:
            b.add(dims, u, g)
        print(f"batch run cut at maxit 203, persistent {persistent}", flush=True)
        rc = b.run(203, 1e-30, 5, gfx.PREC_FP64, 1e-6, 2)
        out.append((rc, [b.result(q) for q in range(2)], [b.trafo(q) for q in range(2)], b.stats()))
        b.set_persistent(0)
        b.close()
    (rc_g, res_g, gs_g, st_g), (rc_p, res_p, gs_p, st_p) = out
    assert rc_g == rc_p == 300  # GFX_ERR_NOCONV
    assert st_p["persistent_launches"] > 0
    for q in range(2):
        assert res_p[q][3] == res_g[q][3] == -1
        assert abs(res_p[q][0] - res_g[q][0]) < 1e-14
        assert np.abs(gs_p[q] - gs_g[q]).max() < 1e-13
