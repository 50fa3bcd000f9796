"""GPU tests of the batching form: every lattice of a batch (uneven sizes,
Gribov copies sharing links) must follow the single-lattice oracle run from
the same start to 1e-10 and stop at the same iteration (+-2)."""
import numpy as np
import pytest

from test_gpu_or import EPS12, smooth_config
from test_gpu_stoch import synthetic_params

pytestmark = pytest.mark.gpu


def _oracle_run(orc, dims, u, g, omega, ck):
    u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
    it, hist = orc.overrelax_gauge(u2, g2, dims, 4000, EPS12, ck, omega)
    return it, orc.from_2v(g2), hist


@pytest.mark.parametrize("ck,periods", [(1, 1), (5, 3)])
def test_uneven_batch_matches_single_runs(gfx, orc, ck, periods):
    gfx.set_omega(1.7)
    shapes = [(8, 8, 8, 8), (4, 4, 4, 8), (8, 4, 6, 4)]
    cases = []
    b = gfx.Batch()
    for k, dims in enumerate(shapes):
        u = smooth_config(orc, dims, 100 + k, eps=0.3)
        g = orc.random_init_trafo(orc.volume(dims), 200 + k)
        b.add(dims, u, g)
        cases.append((dims, u, g))
    rc = b.run(4000, EPS12, ck, gfx.PREC_FP64, 1e-6, periods)
    res = [b.result(q) for q in range(len(shapes))]
    gs = [b.trafo(q) for q in range(len(shapes))]
    st = b.stats()
    b.close()
    assert rc == 0
    if ck > 1:
        assert st["graph_launches"] > 0
    for q, (dims, u, g) in enumerate(cases):
        it, g_cpu, hist = _oracle_run(orc, dims, u, g, 1.7, ck)
        assert abs(res[q][3] - it) <= 2, (q, res[q][3], it)
        assert np.abs(gs[q] - g_cpu).max() < 1e-10
        assert res[q][2] < EPS12 and abs(res[q][0] - hist[-1, 0]) < 1e-12


def test_gribov_copies_share_links(gfx, orc):
    """Copies are the same links with different initial g: each copy equals
    the single run from its own start."""
    gfx.set_omega(1.7)
    dims = (8, 8, 8, 8)
    V = orc.volume(dims)
    u = smooth_config(orc, dims, 300, eps=0.4)
    starts = [orc.random_init_trafo(V, 400 + k) for k in range(4)]
    b = gfx.Batch()
    base = b.add(dims, u, starts[0])
    for k in range(1, 4):
        b.add_copy(base, starts[k])
    rc = b.run(4000, EPS12, 2, gfx.PREC_FP64, 1e-6, 2)
    res = [b.result(q) for q in range(4)]
    gs = [b.trafo(q) for q in range(4)]
    b.close()
    assert rc == 0
    for k in range(4):
        it, g_cpu, _ = _oracle_run(orc, dims, u, starts[k], 1.7, 2)
        assert abs(res[k][3] - it) <= 2
        assert np.abs(gs[k] - g_cpu).max() < 1e-10


def test_mixed_batch_matches_fp64_batch(gfx, orc):
    gfx.set_omega(1.7)
    shapes = [(8, 8, 8, 8), (8, 8, 4, 8)]
    runs = []
    for prec in (gfx.PREC_FP64, gfx.PREC_MIXED):
        b = gfx.Batch()
        for k, dims in enumerate(shapes):
            b.add(dims, smooth_config(orc, dims, 500 + k, eps=0.35),
                  orc.random_init_trafo(orc.volume(dims), 600 + k))
        rc = b.run(5000, EPS12, 10, prec, 1e-6, 2)
        runs.append(([b.result(q) for q in range(len(shapes))], b.stats()))
        b.close()
        assert rc == 0
    (r64, s64), (rmx, smx) = runs
    assert smx["iters_fp32"] > 0 and smx["iters_fp64"] > 0
    for q in range(len(shapes)):
        assert rmx[q][2] < EPS12
        assert abs(rmx[q][0] - r64[q][0]) / r64[q][0] < 1e-10


def test_random_gribov_copies(gfx, orc):
    """Haar-random starts from the twisters: unitary, distinct, all converge;
    the copy with the largest F is the one to select."""
    gfx.set_omega(1.7)
    gfx.set_mt_params(synthetic_params(77).ravel())
    gfx.set_mt_seed(77)
    dims = (8, 8, 8, 8)
    V = orc.volume(dims)
    u = smooth_config(orc, dims, 700, eps=0.5)
    b = gfx.Batch()
    base = b.add(dims, u, orc.random_init_trafo(V, 701))
    copies = [b.add_random_copy(base) for _ in range(3)]
    g_start = [b.trafo(q) for q in copies]
    rc = b.run(5000, EPS12, 5, gfx.PREC_FP64, 1e-6, 2)
    res = [b.result(q) for q in range(4)]
    b.close()
    gfx.finalize_MT()
    assert rc == 0 and all(r[2] < EPS12 for r in res)
    for g in g_start:
        m = g.reshape(-1, 3, 3, 2)
        m = m[..., 0] + 1j * m[..., 1]
        assert np.abs(m @ np.conj(np.swapaxes(m, -1, -2)) - np.eye(3)).max() < 1e-13
        assert np.abs(np.linalg.det(m) - 1).max() < 1e-13
    assert np.abs(g_start[0] - g_start[1]).max() > 0.1
    # Haar: the mean of Re tr g over sites is ~0 (it is 3 for the identity)
    tr = g_start[0].reshape(-1, 18)[:, [0, 8, 16]].sum(axis=1)
    assert abs(tr.mean()) < 0.1


@pytest.mark.parametrize("maxit,ck,periods", [(41, 5, 2), (9, 1, 3)])
def test_batch_state_at_maxit(gfx, orc, maxit, ck, periods):
    """Runs cut at maxit (eager head, graph periods, eager tail) leave every
    lattice in the state after iteration maxit-1, as the single-run oracle."""
    gfx.set_omega(1.7)
    shapes = [(8, 8, 4, 4), (4, 6, 4, 8)]
    b = gfx.Batch()
    cases = []
    for k, dims in enumerate(shapes):
        u = orc.random_init_gauge(orc.volume(dims), 920 + k)
        g = orc.random_init_trafo(orc.volume(dims), 930 + k)
        b.add(dims, u, g)
        cases.append((dims, u, g))
    rc = b.run(maxit, EPS12, ck, gfx.PREC_FP64, 1e-6, periods)
    res = [b.result(q) for q in range(len(shapes))]
    gs = [b.trafo(q) for q in range(len(shapes))]
    b.close()
    assert rc != 0  # nothing converges in so few iterations
    for q, (dims, u, g) in enumerate(cases):
        u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
        it, hist = orc.overrelax_gauge(u2, g2, dims, maxit, EPS12, ck, 1.7)
        assert it == -1 and res[q][3] == -1
        assert np.abs(gs[q] - orc.from_2v(g2)).max() < 1e-12
        assert abs(res[q][0] - hist[-1, 0]) < 1e-13
        assert abs(res[q][2] - hist[-1, 2]) / hist[-1, 2] < 1e-11
