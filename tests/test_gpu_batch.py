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


def test_batch_set_trafo_restarts_a_lattice(gfx, orc):
    """set_trafo replaces a lattice's g exactly; a second run from it ends
    where a fresh batch from that g ends."""
    gfx.set_omega(1.7)
    dims = (8, 8, 8, 8)
    V = orc.volume(dims)
    u = smooth_config(orc, dims, 900, eps=0.3)
    g0 = orc.random_init_trafo(V, 901)
    g1 = orc.random_init_trafo(V, 902)
    b = gfx.Batch()
    b.add(dims, u, g0)
    assert b.run(4000, EPS12, 5, gfx.PREC_FP64, 1e-6, 2) == 0
    b.set_trafo(0, g1)
    assert np.array_equal(b.trafo(0), g1)
    assert b.run(4000, EPS12, 5, gfx.PREC_FP64, 1e-6, 2) == 0
    r_again, g_again = b.result(0), b.trafo(0)
    b.close()
    b = gfx.Batch()
    b.add(dims, u, g1)
    assert b.run(4000, EPS12, 5, gfx.PREC_FP64, 1e-6, 2) == 0
    r_fresh, g_fresh = b.result(0), b.trafo(0)
    b.close()
    assert r_again[3] == r_fresh[3]
    assert np.abs(g_again - g_fresh).max() < 1e-13


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


@pytest.mark.parametrize("eps", [EPS12, 1e-14])
def test_mixed_batch_matches_fp64_batch(gfx, orc, eps):
    """The mixed batch reaches the FP64 batch's F; its FP64 phase runs on
    fixed-point links and g at eps = 1e-12 and on 24 + 16-bit links with FP64
    g at 1e-14 (below 3.1e-13 the fixed-point stop bound would take more
    than a quarter of sqrt(eps)).  Every returned g is below eps on the
    exact links, with the reported F."""
    gfx.set_omega(1.7)
    shapes = [(8, 8, 8, 8), (8, 8, 4, 8)]
    cases = [(dims, smooth_config(orc, dims, 500 + k, eps=0.35),
              orc.random_init_trafo(orc.volume(dims), 600 + k)) for k, dims in enumerate(shapes)]
    runs = []
    for prec in (gfx.PREC_FP64, gfx.PREC_MIXED):
        b = gfx.Batch()
        for dims, u, g in cases:
            b.add(dims, u, g)
        rc = b.run(8000, eps, 10, prec, 1e-6, 2)
        runs.append(([b.result(q) for q in range(len(shapes))], b.stats(),
                     [b.trafo(q) for q in range(len(shapes))]))
        b.close()
        assert rc == 0
    (r64, s64, _), (rmx, smx, gmx) = runs
    assert smx["iters_fp32"] > 0 and smx["iters_fp64"] > 0
    for q, (dims, u, _) in enumerate(cases):
        assert rmx[q][2] < eps
        assert abs(rmx[q][0] - r64[q][0]) / r64[q][0] < 1e-10
        uf = orc.g_trafo(u, gmx[q], dims)
        assert orc.dAdA(uf, dims)[1] < eps * (1 + 1e-9)
        assert abs(orc.gauge_functional(uf, dims) - rmx[q][0]) < 1e-12


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
    # Haar SU(3): E tr g = 0 and E |tr g|^2 = 1 (E |tr g|^4 = 2, so the mean
    # over the 3 V samples of the three copies has standard deviation
    # 1/sqrt(3 V)); the identity would give 3 and 9
    a = np.concatenate([g.reshape(-1, 18) for g in g_start])
    tr = a[:, [0, 8, 16]].sum(axis=1) + 1j * a[:, [1, 9, 17]].sum(axis=1)
    n = tr.size
    assert abs(tr.mean()) < 4 / np.sqrt(n)
    assert abs((np.abs(tr) ** 2).mean() - 1.0) < 0.05 and 5 / np.sqrt(n) < 0.05


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


def _host_F(orc, u, g, dims):
    return orc.gauge_functional(orc.g_trafo(u, g, dims), dims)


@pytest.mark.parametrize("sampler,dims", [(0, (8, 8, 8, 8)), (1, (8, 8, 8, 8)), (0, (16, 16, 16, 16))])
def test_batch_anneal_of_one_lattice_equals_simannealing(gfx, orc, sampler, dims):
    """A batch holding one lattice draws the same twister numbers in the same
    order as simannealing_gauge, so both give the same g (at 16^4 the single
    lattice's first SA launches run under each workgroup order it times)."""
    V = orc.volume(dims)
    u = _thermalized(gfx, dims, 6.0, 30, 41)
    g0 = orc.random_init_trafo(V, 42)
    steps, Tmin, Tmax, nmicro = 40, 0.01, 0.45, 3
    gfx.set_mt_params(synthetic_params(43).ravel())
    gfx.set_mt_seed(43)
    b = gfx.Batch()
    b.add(dims, u, g0)
    b.anneal(steps, Tmin, Tmax, -1.0, nmicro, sampler, steps_per_graph=8)
    g_batch = b.trafo(0)
    b.close()
    gfx.finalize_MT()
    gfx.set_lattice(*dims)
    gfx.set_verbose(0)
    gfx.init_gaugefixing(u, g0)
    sa = gfx.saparam()
    sa.Tmin, sa.Tmax, sa.expo, sa.N, sa.checkint = Tmin, Tmax, -1.0, steps, 1000
    gfx.set_mt_seed(43)
    gfx.set_sa(sampler, nmicro, 4)
    gfx.simannealing_gauge()
    g_single = gfx.get_trafo()
    gfx.finalize_MT()
    gfx.finalize_gaugefixing()
    assert np.abs(g_batch - g_single).max() < 1e-12
    assert _host_F(orc, u, g_batch, dims) > _host_F(orc, u, g0, dims) + 0.2


def test_gribov_batch_anneal_then_fix(gfx, orc):
    """Haar-random copies of one configuration, annealed and fixed together;
    every copy reaches the bound (checked on the host from the fixed links)
    and the best copy is the one with the largest F."""
    dims = (8, 8, 8, 8)
    V = orc.volume(dims)
    u = _thermalized(gfx, dims, 6.0, 40, 51)
    gfx.set_omega(1.7)
    gfx.set_mt_params(synthetic_params(52).ravel())
    gfx.set_mt_seed(52)
    b = gfx.Batch()
    base = b.add(dims, u, orc.random_init_trafo(V, 53))
    for _ in range(7):
        b.add_random_copy(base)
    starts = [b.trafo(q) for q in range(8)]
    b.anneal(200, 0.01, 0.45, -1.0, 3, 0, steps_per_graph=16)
    annealed = [b.trafo(q) for q in range(8)]
    rc = b.run(8000, EPS12, 10, gfx.PREC_MIXED, 1e-6, 4)
    res = [b.result(q) for q in range(8)]
    fixed = [b.trafo(q) for q in range(8)]
    qb, Fb = b.best_copy(base)
    b.close()
    gfx.finalize_MT()
    assert rc == 0
    for q in range(8):
        assert _host_F(orc, u, annealed[q], dims) > _host_F(orc, u, starts[q], dims) + 0.3
        uf = orc.g_trafo(u, fixed[q], dims)
        _, mx = orc.dAdA(uf, dims)
        assert mx < 2 * EPS12 and res[q][2] < EPS12
        assert abs(orc.gauge_functional(uf, dims) - res[q][0]) < 1e-12
    assert Fb == max(r[0] for r in res) and res[qb][0] == Fb


def test_gribov_batch_fp32_anneal_then_fix(gfx, orc):
    """The anneal on FP32 links and g (gfx_batch_set_anneal_precision): every
    annealed g comes back unitary in FP64, F rises as much as in the FP64
    anneal from the same starts and twister numbers (mean over the copies
    within 0.02), and every copy then reaches the bound in the mixed OR."""
    dims = (8, 8, 8, 8)
    V = orc.volume(dims)
    u = _thermalized(gfx, dims, 6.0, 40, 51)
    gfx.set_omega(1.7)
    out = {}
    for mode in (gfx.PREC_FP64, gfx.PREC_MIXED):
        gfx.set_mt_params(synthetic_params(52).ravel())
        gfx.set_mt_seed(52)
        b = gfx.Batch()
        b.set_anneal_precision(mode)
        base = b.add(dims, u, orc.random_init_trafo(V, 53))
        for _ in range(7):
            b.add_random_copy(base)
        starts = [b.trafo(q) for q in range(8)]
        b.anneal(200, 0.01, 0.45, -1.0, 3, 0, steps_per_graph=16)
        annealed = [b.trafo(q) for q in range(8)]
        rc = b.run(8000, EPS12, 10, gfx.PREC_MIXED, 1e-6, 4)
        res = [b.result(q) for q in range(8)]
        b.set_anneal_precision(gfx.PREC_FP64)
        b.close()
        gfx.finalize_MT()
        out[mode] = (starts, annealed, rc, res)
    F_after = {}
    for mode, (starts, annealed, rc, res) in out.items():
        assert rc == 0 and all(r[2] < EPS12 for r in res)
        F_after[mode] = np.mean([_host_F(orc, u, a, dims) for a in annealed])
        for q in range(8):
            assert _host_F(orc, u, annealed[q], dims) > _host_F(orc, u, starts[q], dims) + 0.3
            m = annealed[q].reshape(-1, 3, 3, 2)
            m = m[..., 0] + 1j * m[..., 1]
            assert np.abs(m @ np.conj(np.swapaxes(m, -1, -2)) - np.eye(3)).max() < 1e-13
    assert np.array_equal(out[gfx.PREC_FP64][0][1], out[gfx.PREC_MIXED][0][1])  # same starts
    assert abs(F_after[gfx.PREC_MIXED] - F_after[gfx.PREC_FP64]) < 0.02, F_after


def test_gribov_batch_full_shape(gfx, orc):
    """The Gribov-copy workload at its real shape: 64 Haar-random copies of
    one 24^4 beta = 6.0 configuration, annealed from T = 0.45 to 0.01 with 3
    microcanonical steps per step (400 of the 4000 steps of the full schedule,
    to keep the test short), then mixed-precision OR to theta < 1e-12 in one
    batch; every copy converges and the best copy is selected."""
    dims = (24, 24, 24, 24)
    u = _thermalized(gfx, dims, 6.0, 100, 61)
    gfx.set_omega(1.7)
    gfx.set_mt_params(synthetic_params(62).ravel())
    gfx.set_mt_seed(62)
    b = gfx.Batch()
    base = b.add(dims, u, gfx.unit_trafo())
    for _ in range(63):
        b.add_random_copy(base)
    b.anneal(400, 0.01, 0.45, -1.0, 3, 0, steps_per_graph=20)
    rc = b.run(20000, EPS12, 10, gfx.PREC_MIXED, 1e-6, 8)
    res = [b.result(q) for q in range(64)]
    st = b.stats()
    qb, Fb = b.best_copy(base)
    g_best = b.trafo(qb)
    b.close()
    gfx.finalize_MT()
    assert rc == 0 and st["graph_launches"] > 0
    assert all(r[2] < EPS12 and r[3] >= 0 for r in res)
    assert Fb == max(r[0] for r in res)
    # the selected copy's fixed links have the reported F
    assert abs(_host_F(orc, u, g_best, dims) - Fb) < 1e-10


def test_mixed_batch_lattices_switch_on_their_own(gfx, orc):
    """In a mixed batch every lattice leaves FP32 at its own check and goes
    on in FP64 while the others are still in FP32; each reaches the bound
    with the F of its FP64 oracle run, and its reported stop iteration is
    the number of iterations it ran (as the single-lattice mixed run)."""
    gfx.set_omega(1.7)
    shapes = [(4, 4, 4, 4), (8, 8, 8, 8), (6, 6, 6, 8), (8, 8, 8, 16)]
    cases = []
    b = gfx.Batch()
    for k, dims in enumerate(shapes):
        u = smooth_config(orc, dims, 950 + k, eps=0.3 + 0.05 * k)
        g = orc.random_init_trafo(orc.volume(dims), 960 + k)
        b.add(dims, u, g)
        cases.append((dims, u, g))
    rc = b.run(20000, EPS12, 10, gfx.PREC_MIXED, 1e-6, 4)
    res = [b.result(q) for q in range(len(shapes))]
    gs = [b.trafo(q) for q in range(len(shapes))]
    st = b.stats()
    b.close()
    assert rc == 0 and st["iters_fp32"] > 0 and st["iters_fp64"] > 0
    for q, (dims, u, g) in enumerate(cases):
        u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
        it, hist = orc.overrelax_gauge(u2, g2, dims, 20000, EPS12, 10, 1.7)
        assert res[q][2] < EPS12
        assert abs(res[q][0] - hist[-1, 0]) / hist[-1, 0] < 1e-10
        # below the bound on the exact links with the returned FP64 g
        assert orc.dAdA(orc.g_trafo(u, gs[q], dims), dims)[1] < EPS12
        # the same lattice alone, mixed, single-lattice path
        gfx.set_lattice(*dims)
        gfx.set_verbose(0)
        gfx.init_gaugefixing(u, g)
        gfx.set_precision(gfx.PREC_MIXED, 1e-6)
        it1 = gfx.overrelax_gauge(20000, EPS12, 10)
        gfx.set_precision(gfx.PREC_FP64)
        gfx.finalize_gaugefixing()
        assert abs(res[q][3] - it1) <= 20, (q, res[q][3], it1)
