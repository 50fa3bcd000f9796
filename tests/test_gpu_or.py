"""GPU parity of the overrelaxation path against the oracle (through the C ABI).

Inputs are seeded (Lehmer RNG, rngs.c) and identical for both sides; the
FP64 path must agree with the oracle to 1e-10 max-abs per matrix element and
stop at the same iteration (+-2) with checkinterval 1.
"""
import numpy as np
import pytest

from test_oracle import fx_round, p21_round, packed_round

pytestmark = pytest.mark.gpu

EPS12 = float(np.float32(1e-12))  # (double)(float)1e-12, read_input.l:189-192


def rotation_distance(ga, gb):
    """max |ga - h gb| over the sites for the constant unitary h that fits
    best (F is invariant under g -> h g)."""
    A = ga.reshape(-1, 3, 3, 2)
    B = gb.reshape(-1, 3, 3, 2)
    A = A[..., 0] + 1j * A[..., 1]
    B = B[..., 0] + 1j * B[..., 1]
    W, _, Vh = np.linalg.svd(np.einsum("nij,nkj->ik", A, B.conj()))
    return np.abs(A - np.einsum("ij,njk->nik", W @ Vh, B)).max()


def _setup(gfx, dims, u, g, omega=1.68, prec=0):
    gfx.set_lattice(*dims)
    gfx.set_verbose(0)
    gfx.set_omega(omega)
    gfx.set_precision(prec)
    gfx.init_gaugefixing(u, g)


def smooth_config(orc, dims, seed, eps=0.25):
    """A gauge copy of a near-unit field: U = h(x) W h^+(x+mu) with W close
    to 1, built from the oracle's random SU(3) matrices."""
    V = orc.volume(dims)
    w = orc.random_init_gauge(V, seed)
    unit = np.zeros(18)
    unit[[0, 8, 16]] = 1.0
    # pull every link towards 1 and reunitarise
    m = (unit[None, :] + eps * w.reshape(-1, 18)).copy()
    r12 = orc.su3_to_2v(m.ravel())
    # Gram-Schmidt through the oracle normalisation
    out = np.zeros_like(m)
    for k in range(m.shape[0]):
        a = orc.from_2v(r12[12 * k:12 * k + 12]).copy()
        orc.lib.orc_su3_normalize(a)
        out[k] = a
    h = orc.random_init_trafo(V, seed + 1)
    return orc.g_trafo(out.ravel(), h, dims)


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (8, 8, 8, 8), (4, 6, 8, 2)])
@pytest.mark.parametrize("parity", [0, 1])
def test_halfsweep_matches_oracle(gfx, orc, dims, parity):
    V = orc.volume(dims)
    u = orc.random_init_gauge(V, 4357)
    g = orc.random_init_trafo(V, 351)
    _setup(gfx, dims, u, g)
    gfx.halfsweep(parity, 4)
    g_gpu = gfx.get_trafo()
    gfx.finalize_gaugefixing()
    u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
    orc.halfsweep(u2, g2, dims, parity, 4, 1.68)
    g_cpu = orc.from_2v(g2)
    assert np.abs(g_gpu - g_cpu).max() < 1e-13


def test_functional_matches_oracle(gfx, orc):
    dims = (8, 8, 8, 8)
    V = orc.volume(dims)
    u = orc.random_init_gauge(V, 11)
    g = orc.random_init_trafo(V, 12)
    _setup(gfx, dims, u, g)
    m = gfx.calc_functional()
    F, D, M = gfx.results()
    plaq = gfx.calc_plaquette(0)
    gfx.finalize_gaugefixing()
    Fo, Do, Mo = orc.functional(orc.su3_to_2v(u), orc.su3_to_2v(g), dims)
    assert abs(F - Fo) < 1e-13 and abs(D - Do) / Do < 1e-13 and m == M
    assert abs(M - Mo) / Mo < 1e-13
    assert abs(plaq - orc.plaquette_2v(orc.su3_to_2v(u), dims)) < 1e-13


@pytest.mark.parametrize("dims,omega", [((8, 8, 8, 8), 1.68), ((8, 8, 8, 8), 1.7),
                                        ((4, 4, 4, 8), 1.0)])
def test_overrelax_run_matches_oracle(gfx, orc, dims, omega):
    V = orc.volume(dims)
    u = smooth_config(orc, dims, 77)
    g = orc.random_init_trafo(V, 351)
    _setup(gfx, dims, u, g, omega)
    it_gpu = gfx.overrelax_gauge(3000, EPS12, 1)
    F_gpu = gfx.results()[0]
    g_gpu = gfx.get_trafo()
    fixed_gpu = gfx.apply_trafo(u)
    gfx.finalize_gaugefixing()
    u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
    it_cpu, hist = orc.overrelax_gauge(u2, g2, dims, 3000, EPS12, 1, omega)
    g_cpu = orc.from_2v(g2)
    fixed_cpu = orc.g_trafo(u, g_cpu, dims)
    assert it_cpu > 0 and it_gpu > 0
    assert abs(it_gpu - it_cpu) <= 2
    assert np.abs(g_gpu - g_cpu).max() < 1e-10
    assert np.abs(fixed_gpu - fixed_cpu).max() < 1e-10
    assert abs(F_gpu - hist[-1, 0]) < 1e-12


def test_cold_links_stop_at_zero(gfx, orc):
    """U = 1 and g = 1: F = 3, dAdA = 0, so the first check stops at i = 0."""
    dims = (4, 4, 4, 4)
    V = orc.volume(dims)
    u = np.zeros(18 * 4 * V)
    u.reshape(-1, 18)[:, [0, 8, 16]] = 1.0
    g = np.zeros(18 * V)
    g.reshape(-1, 18)[:, [0, 8, 16]] = 1.0
    _setup(gfx, dims, u, g)
    assert gfx.overrelax_gauge(100, EPS12, 1) == 0
    F, D, M = gfx.results()
    gfx.finalize_gaugefixing()
    assert abs(F - 3.0) < 1e-15 and D == 0.0 and M == 0.0


def test_pure_gauge_fixes_to_unit(gfx, orc):
    """U = h(x) h^+(x+mu) is a pure gauge: the fixed field has F = 3."""
    dims = (8, 8, 8, 8)
    V = orc.volume(dims)
    unit = np.zeros(18 * 4 * V)
    unit.reshape(-1, 18)[:, [0, 8, 16]] = 1.0
    u = orc.g_trafo(unit, orc.random_init_trafo(V, 5), dims)
    g = orc.random_init_trafo(V, 6)
    _setup(gfx, dims, u, g)
    it = gfx.overrelax_gauge(5000, EPS12, 1)
    F, D, M = gfx.results()
    fixed = gfx.apply_trafo(u)
    gfx.finalize_gaugefixing()
    assert it > 0 and M < EPS12
    assert abs(F - 3.0) < 1e-12
    assert abs(orc.gauge_functional(fixed, dims) - 3.0) < 1e-12


def test_get_set_trafo_roundtrip(gfx, orc):
    dims = (4, 4, 6, 8)
    V = orc.volume(dims)
    u = orc.random_init_gauge(V, 1)
    g = orc.random_init_trafo(V, 2)
    _setup(gfx, dims, u, g)
    back = gfx.get_trafo()
    links = gfx.get_gauge()
    g2 = orc.random_init_trafo(V, 3)
    gfx.set_trafo(g2)
    back2 = gfx.get_trafo()
    gfx.finalize_gaugefixing()
    # rows 0,1 exact; row 2 rebuilt from rows 0,1 (from2vtosu3_trafo)
    assert np.array_equal(back, orc.from_2v(orc.su3_to_2v(g)))
    assert np.array_equal(links, orc.from_2v(orc.su3_to_2v(u)))
    assert np.array_equal(back2, orc.from_2v(orc.su3_to_2v(g2)))


def test_mixed_precision_matches_fp64(gfx, orc):
    """FP32 12-real bulk sweeps to max dAdA < 1e-6, then FP64 polish: the same
    stop bound, F within 1e-10 relative of the FP64 run (stated error)."""
    dims = (8, 8, 8, 16)
    V = orc.volume(dims)
    u = smooth_config(orc, dims, 90, eps=0.35)
    g = orc.random_init_trafo(V, 351)
    _setup(gfx, dims, u, g, 1.7)
    it64 = gfx.overrelax_gauge(5000, EPS12, 10)
    F64, _, M64 = gfx.results()
    g64 = gfx.get_trafo()
    gfx.set_trafo(g)
    gfx.set_precision(gfx.PREC_MIXED, 1e-6)
    itm = gfx.overrelax_gauge(5000, EPS12, 10)
    Fm, _, Mm = gfx.results()
    ms, i32, i64 = gfx.last_run_stats()
    gm = gfx.get_trafo()
    gfx.set_precision(gfx.PREC_FP64)
    gfx.finalize_gaugefixing()
    assert it64 > 0 and itm > 0 and M64 < EPS12 and Mm < EPS12
    assert i32 > 0 and i64 > 0 and i32 + i64 == itm + 1
    assert abs(Fm - F64) / F64 < 1e-10
    # g agrees up to a constant rotation (F is invariant under g -> h g)
    assert rotation_distance(gm, g64) < 1e-5


def test_coulomb_matches_oracle(gfx, orc):
    dims = (8, 8, 8, 4)
    V = orc.volume(dims)
    u = smooth_config(orc, dims, 31, eps=0.3)
    g = orc.random_init_trafo(V, 32)
    _setup(gfx, dims, u, g, 1.68)
    rc, it_gpu = gfx.coulomb_gauge(3000, EPS12, 1)
    g_gpu = gfx.get_trafo()
    F, D, M = gfx.functional_slices()
    gfx.finalize_gaugefixing()
    u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
    rc_o, it_cpu = orc.coulomb_gauge(u2, g2, dims, 3000, EPS12, 1, 1.68)
    assert rc == 0 and rc_o == 0
    assert np.all(np.abs(it_gpu - it_cpu) <= 2), (it_gpu, it_cpu)
    assert np.all(M < EPS12)
    assert np.abs(g_gpu - orc.from_2v(g2)).max() < 1e-10
    Fo, Do, Mo = orc.functional_slices(u2, g2, dims, 3)
    assert np.abs(F - Fo).max() < 1e-12


@pytest.mark.parametrize("maxit,ck", [(7, 3), (12, 1), (30, 10)])
def test_fused_check_state_at_maxit(gfx, orc, maxit, ck):
    """The checks ride on the sweeps (odd sweep of iteration i, even sweep of
    i+1 into a spare field).  A run cut at maxit must still leave the state
    after iteration maxit-1, and its last check must report F, dAdA and max
    dAdA of the state it checked, as the oracle's per-link functional does."""
    dims = (8, 4, 6, 8)
    V = orc.volume(dims)
    u = orc.random_init_gauge(V, 901)
    g = orc.random_init_trafo(V, 902)
    _setup(gfx, dims, u, g, 1.7)
    it = gfx.overrelax_gauge(maxit, EPS12, ck)
    F, D, M = gfx.results()
    g_gpu = gfx.get_trafo()
    gfx.finalize_gaugefixing()
    u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
    it_o, hist = orc.overrelax_gauge(u2, g2, dims, maxit, EPS12, ck, 1.7)
    assert it == -1 and it_o == -1
    assert np.abs(g_gpu - orc.from_2v(g2)).max() < 1e-12
    last = hist[-1]
    assert abs(F - last[0]) < 1e-13
    assert abs(D - last[1]) / last[1] < 1e-11 and abs(M - last[2]) / last[2] < 1e-11


def test_fused_check_matches_functional_at_stop(gfx, orc):
    """At the stop (max dAdA ~ 1e-12) the fused check and the standalone
    per-link functional (calc_functional) of the returned g agree."""
    dims = (8, 8, 8, 8)
    V = orc.volume(dims)
    u = smooth_config(orc, dims, 907, eps=0.3)
    g = orc.random_init_trafo(V, 908)
    _setup(gfx, dims, u, g, 1.7)
    it = gfx.overrelax_gauge(5000, EPS12, 5)
    F, D, M = gfx.results()
    m2 = gfx.calc_functional()
    F2, D2, M2 = gfx.results()
    gfx.finalize_gaugefixing()
    assert it > 0 and M < EPS12 and m2 == M2
    assert abs(F - F2) < 1e-14
    assert abs(D - D2) / D2 < 1e-6 and abs(M - M2) / M2 < 1e-6


def test_mixed_stall_guard_leaves_fp32(gfx, orc):
    """A switch bound below the FP32 rounding floor: the stall guard ends the
    FP32 phase, the FP64 polish reaches the bound, F agrees with FP64."""
    dims = (8, 8, 8, 8)
    V = orc.volume(dims)
    u = smooth_config(orc, dims, 911, eps=0.35)
    g = orc.random_init_trafo(V, 912)
    _setup(gfx, dims, u, g, 1.7)
    it64 = gfx.overrelax_gauge(8000, EPS12, 10)
    F64 = gfx.results()[0]
    gfx.set_trafo(g)
    gfx.set_precision(gfx.PREC_MIXED, 1e-15)
    gfx.set_fp32_stall(200, 1e-8)
    itm = gfx.overrelax_gauge(8000, EPS12, 10)
    Fm, _, Mm = gfx.results()
    _, i32, i64 = gfx.last_run_stats()
    gfx.set_fp32_stall(300, 1e-8)
    gfx.set_precision(gfx.PREC_FP64)
    gfx.finalize_gaugefixing()
    assert it64 > 0 and itm > 0 and Mm < EPS12
    assert i32 > 0 and i64 > 0
    assert abs(Fm - F64) / F64 < 1e-10


def test_every_sweep_variant_matches_oracle(gfx, orc):
    """All 32 sweep variants (kernel form x workgroup order) give the same
    half sweeps; 16^4 admits every workgroup order."""
    dims = (16, 16, 16, 16)
    V = orc.volume(dims)
    u = orc.random_init_gauge(V, 4361)
    g = orc.random_init_trafo(V, 4362)
    u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
    for parity in (0, 1):
        orc.halfsweep(u2, g2, dims, parity, 4, 1.68)
    g_cpu = orc.from_2v(g2)
    _setup(gfx, dims, u, g)
    worst = 0.0
    for v in range(32):
        gfx.set_sweep_variant(gfx.PREC_FP64, 4, v)
        gfx.set_trafo(g)
        gfx.halfsweep(0, 4)
        gfx.halfsweep(1, 4)
        worst = max(worst, np.abs(gfx.get_trafo() - g_cpu).max())
        assert worst < 1e-13, v
    gfx.finalize_gaugefixing()


@pytest.mark.parametrize("ck", [3, 7])
def test_coulomb_fused_checks_match_oracle(gfx, orc, ck):
    """Slices stop at different checks; with the checks fused into the sweeps
    each stopped slice keeps the state of its check."""
    dims = (8, 8, 8, 6)
    V = orc.volume(dims)
    u = smooth_config(orc, dims, 41 + ck, eps=0.35)
    g = orc.random_init_trafo(V, 42 + ck)
    _setup(gfx, dims, u, g, 1.7)
    rc, it_gpu = gfx.coulomb_gauge(5000, EPS12, ck)
    g_gpu = gfx.get_trafo()
    F, D, M = gfx.functional_slices()
    gfx.finalize_gaugefixing()
    u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
    rc_o, it_cpu = orc.coulomb_gauge(u2, g2, dims, 5000, EPS12, ck, 1.7)
    assert rc == 0 and rc_o == 0
    assert len(set(it_cpu.tolist())) > 1  # the slices do stop at different checks
    assert np.all(np.abs(it_gpu - it_cpu) <= 2), (it_gpu, it_cpu)
    assert np.abs(g_gpu - orc.from_2v(g2)).max() < 1e-10
    assert np.all(M < EPS12)


@pytest.mark.parametrize("rolled", [True, False])
def test_fp32_halfsweep_close_to_oracle(gfx, orc, rolled):
    """The FP32 sweep uses a shorter but equal arithmetic (two rows of each
    SU(3) product, the series applied to g directly); a half sweep agrees
    with the oracle's FP64 one to FP32 rounding."""
    dims = (8, 8, 8, 8)
    V = orc.volume(dims)
    u = orc.random_init_gauge(V, 4371)
    g = orc.random_init_trafo(V, 4372)
    _setup(gfx, dims, u, g, 1.7, prec=gfx.PREC_MIXED)
    gfx.set_sweep_variant(gfx.PREC_MIXED, 4, 0 if rolled else 2)
    gfx.halfsweep(0, 4)
    gfx.halfsweep(1, 4)
    g_gpu = gfx.get_trafo()
    gfx.set_precision(gfx.PREC_FP64)
    gfx.finalize_gaugefixing()
    u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
    orc.halfsweep(u2, g2, dims, 0, 4, 1.7)
    orc.halfsweep(u2, g2, dims, 1, 4, 1.7)
    err = np.abs(g_gpu - orc.from_2v(g2)).max()
    assert 1e-9 < err < 2e-5, err


def test_coulomb_mixed_slices_switch_on_their_own(gfx, orc):
    """Mixed Coulomb: each slice leaves FP32 at its own check and is polished
    in FP64 while others are still in FP32; every slice ends below the bound
    with the F of the oracle's FP64 fixing of that slice."""
    dims = (8, 8, 8, 6)
    V = orc.volume(dims)
    u = smooth_config(orc, dims, 61, eps=0.4)
    g = orc.random_init_trafo(V, 62)
    _setup(gfx, dims, u, g, 1.7, prec=gfx.PREC_MIXED)
    rc, it_gpu = gfx.coulomb_gauge(8000, EPS12, 5)
    F, D, M = gfx.functional_slices()
    gfx.set_precision(gfx.PREC_FP64)
    gfx.finalize_gaugefixing()
    u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
    rc_o, it_cpu = orc.coulomb_gauge(u2, g2, dims, 8000, EPS12, 5, 1.7)
    Fo, Do, Mo = orc.functional_slices(u2, g2, dims, 3)
    assert rc == 0 and rc_o == 0 and np.all(M < EPS12)
    assert np.max(np.abs(F - Fo) / Fo) < 1e-10
    assert len(set(it_gpu.tolist())) > 1


def test_packed_sweeps_match_oracle(gfx, orc):
    """The sweep of the mixed run's FP64 polish (FP64 g and arithmetic, links
    as 24-bit high + 16-bit low parts), each of its 16 variants, against the
    oracle's FP64 sweep on the same packed-rounded links: equal to FP64
    rounding.  Against the oracle on the exact links the difference is the
    effect of the 2^-39 link rounding (below 1e-10 on these random links)."""
    dims = (16, 16, 16, 16)
    V = orc.volume(dims)
    u = orc.random_init_gauge(V, 4381)
    g = orc.random_init_trafo(V, 4382)
    u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
    g2s = g2.copy()
    u2s = packed_round(u2)
    for parity in (0, 1):
        orc.halfsweep(u2, g2, dims, parity, 4, 1.68)
        orc.halfsweep(u2s, g2s, dims, parity, 4, 1.68)
    g_cpu, g_cpu_packed = orc.from_2v(g2), orc.from_2v(g2s)
    _setup(gfx, dims, u, g)
    for v in [v for v in range(32) if not v & 2]:
        gfx.set_sweep_variant(gfx.SWEEP_PACKED64, 4, v)
        gfx.set_trafo(g)
        gfx.time_sweeps(1, gfx.SWEEP_PACKED64)
        g_gpu = gfx.get_trafo()
        assert np.abs(g_gpu - g_cpu_packed).max() < 1e-13, v
        assert np.abs(g_gpu - g_cpu).max() < 1e-10, v
    gfx.finalize_gaugefixing()
    assert np.abs(g_cpu_packed - g_cpu).max() > 1e-14  # the rounding is visible


def test_fixed_point_sweeps_match_oracle(gfx, orc):
    """The sweep of the mixed Landau run's FP64 polish on 32-bit fixed-point
    links and g (FP64 arithmetic, the series in the short form, g rounded to
    the fixed-point grid after every update), each of its 32 variants (the
    rolled direction loop and the star one term ahead: the same terms in the
    same order), against the oracle's FP64 sweeps with the short series on the same
    rounded links, g rounded the same way after each half sweep and
    reunitarised at the end: equal to FP64 rounding but for the rare element
    whose value sits within rounding of a grid midpoint (one grid step,
    2^-31, and its effect on the neighbours).  Against the oracle on the
    exact links and g the difference is the effect of the 2^-32 rounding."""
    dims = (16, 16, 16, 16)
    V = orc.volume(dims)
    u = orc.random_init_gauge(V, 4385)
    g = orc.random_init_trafo(V, 4386)
    u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
    g2x = fx_round(g2)
    u2x = fx_round(u2)
    for parity in (0, 1):
        orc.halfsweep(u2, g2, dims, parity, 4, 1.7)
        orc.halfsweep(u2x, g2x, dims, parity, 4, 1.7, nterms=-3)
        g2x = fx_round(g2x)
    g_cpu = orc.from_2v(g2)
    g_cpu_fx = orc.from_2v(g2x).reshape(-1, 18)
    for k in range(V):
        orc.lib.orc_su3_normalize(g_cpu_fx[k])
    g_cpu_fx = g_cpu_fx.ravel()
    _setup(gfx, dims, u, g, 1.7)
    for v in range(32):
        gfx.set_sweep_variant(gfx.SWEEP_FIXED64, 4, v)
        gfx.set_trafo(g)
        gfx.time_sweeps(1, gfx.SWEEP_FIXED64)
        g_gpu = gfx.get_trafo()
        d = np.abs(g_gpu - g_cpu_fx).reshape(-1, 18).max(axis=1)
        print(f"variant {v}: max {d.max():.3e}, sites above 1e-13: {np.count_nonzero(d > 1e-13)}",
              flush=True)
        assert d.max() < 4 * 2.0 ** -31, (v, d.max())
        assert np.count_nonzero(d > 1e-13) <= V // 200, (v, np.count_nonzero(d > 1e-13))
        assert np.abs(g_gpu - g_cpu).max() < 1e-7, v
    gfx.finalize_gaugefixing()
    assert np.abs(g_cpu_fx - g_cpu).max() > 1e-11  # the rounding is visible


def test_packed_fp32_sweeps_close_to_oracle(gfx, orc):
    """The FP32 sweep of the mixed run's first phase on the 21-bit links
    (q = rint(u 2^20) per real), each of its 32 variants, against the
    oracle's FP64 sweep on the same rounded links: equal to FP32 rounding."""
    dims = (8, 8, 8, 16)
    V = orc.volume(dims)
    u = orc.random_init_gauge(V, 4391)
    g = orc.random_init_trafo(V, 4392)
    u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
    links = p21_round(u2)
    g2e = g2.copy()
    for parity in (0, 1):
        orc.halfsweep(links, g2, dims, parity, 4, 1.7)
        orc.halfsweep(u2, g2e, dims, parity, 4, 1.7)
    g_cpu = orc.from_2v(g2)
    g_exact = orc.from_2v(g2e)
    _setup(gfx, dims, u, g, 1.7)
    for v in range(32):
        gfx.set_sweep_variant(gfx.SWEEP_PACKED32, 4, v)
        gfx.set_trafo(g)
        gfx.time_sweeps(1, gfx.SWEEP_PACKED32)
        g_gpu = gfx.get_trafo()
        err = np.abs(g_gpu - g_cpu).max()
        assert 1e-9 < err < 2e-5, (v, err)
    assert np.abs(g_cpu - g_exact).max() > 1e-8  # the link rounding is visible
    gfx.finalize_gaugefixing()


@pytest.mark.parametrize("dims,seed", [((8, 8, 8, 16), 90), ((8, 8, 8, 8), 911),
                                       ((6, 6, 6, 6), 912)])
def test_packed_run_matches_plain_mixed_run(gfx, orc, dims, seed):
    """The mixed run on packed links (FP32 phase on the 21-bit links, FP64
    polish on 32-bit fixed-point links and g) and on the FP32 and FP64
    links: both stop
    below the bound, at iterations within 10%, with F equal to 1e-12
    relative and g equal up to a constant rotation; the bound holds on the
    FP64 links (host functional of the packed run's result)."""
    V = orc.volume(dims)
    u = smooth_config(orc, dims, seed, eps=0.35)
    g = orc.random_init_trafo(V, seed + 1)
    _setup(gfx, dims, u, g, 1.7, prec=gfx.PREC_MIXED)
    runs = []
    for packed in (0, 1):
        gfx.set_packed_links(packed)
        gfx.set_trafo(g)
        it = gfx.overrelax_gauge(8000, EPS12, 1)
        runs.append((it, gfx.results(), gfx.last_run_stats(), gfx.get_trafo()))
    gfx.set_packed_links(1)
    gfx.set_precision(gfx.PREC_FP64)
    gfx.finalize_gaugefixing()
    (it0, r0, s0, g0), (it1, r1, s1, g1) = runs
    assert it0 > 0 and it1 > 0 and abs(it0 - it1) <= 0.1 * it0, (it0, it1)
    assert s1[1] > 0 and s1[2] > 0  # an FP32 phase and a polish ran
    assert r1[2] < EPS12 and abs(r1[0] - r0[0]) / r0[0] < 1e-12
    assert rotation_distance(g1, g0) < 1e-5
    _, mx = orc.dAdA(orc.g_trafo(u, g1, dims), dims)
    assert mx < EPS12 * (1 + 1e-9)


def test_packed_polish_falls_back_below_its_floor(gfx, orc, capfd):
    """With eps far below what the packed links resolve (their effect on
    theta is ~1e-22), the polish reaches eps on the packed links, the check
    on the FP64 links refuses the stop (its log line), and the run goes on
    in plain FP64 until theta < eps holds on the FP64 links."""
    import ctypes
    dims = (8, 8, 8, 8)
    V = orc.volume(dims)
    u = smooth_config(orc, dims, 95, eps=0.35)
    g = orc.random_init_trafo(V, 96)
    eps = 1e-26
    _setup(gfx, dims, u, g, 1.7, prec=gfx.PREC_MIXED)
    capfd.readouterr()
    gfx.set_verbose(1)
    it = gfx.overrelax_gauge(20000, eps, 5)
    gfx.set_verbose(0)
    ctypes.CDLL(None).fflush(None)
    log = capfd.readouterr().out
    F, D, M = gfx.results()
    _, i32, i64 = gfx.last_run_stats()
    gq = gfx.get_trafo()
    gfx.set_packed_links(0)
    gfx.set_trafo(g)
    it0 = gfx.overrelax_gauge(20000, eps, 5)
    F0 = gfx.results()[0]
    gfx.set_packed_links(1)
    gfx.set_precision(gfx.PREC_FP64)
    gfx.finalize_gaugefixing()
    assert it > 0 and it0 > 0 and i32 > 0 and i64 > 0 and M < eps
    assert abs(F - F0) / F0 < 1e-13
    assert orc.dAdA(orc.g_trafo(u, gq, dims), dims)[1] < 10 * eps
    confirm = [float(l.split("max(dAdA) =")[1].split()[0]) for l in log.splitlines()
               if "(FP64 links)" in l]
    assert len(confirm) >= 1 and confirm[0] >= eps, confirm


@pytest.mark.parametrize("packed", [1, 0])
def test_mixed_stop_inside_the_fp32_phase(gfx, orc, packed, capfd):
    """eps above the FP32 floor and the switch below it: the run stops in
    its FP32 phase at its first check below eps, and that stop is decided on
    the FP64 links (the checked state converted to FP64 g, the per-link
    functional): the state it returns satisfies the bound on the host, with
    F close to the FP64 run's."""
    import ctypes
    dims = (8, 8, 8, 8)
    V = orc.volume(dims)
    u = smooth_config(orc, dims, 93, eps=0.35)
    g = orc.random_init_trafo(V, 94)
    eps = 1e-8
    _setup(gfx, dims, u, g, 1.7)
    it64 = gfx.overrelax_gauge(20000, eps, 5)
    F64 = gfx.results()[0]
    gfx.set_trafo(g)
    gfx.set_precision(gfx.PREC_MIXED, 1e-20)
    gfx.set_packed_links(packed)
    capfd.readouterr()
    gfx.set_verbose(1)
    it = gfx.overrelax_gauge(20000, eps, 5)
    gfx.set_verbose(0)
    ctypes.CDLL(None).fflush(None)
    log = capfd.readouterr().out
    F, D, M = gfx.results()
    _, i32, i64 = gfx.last_run_stats()
    gq = gfx.get_trafo()
    gfx.set_packed_links(1)
    gfx.set_precision(gfx.PREC_FP64)
    gfx.finalize_gaugefixing()
    assert it64 > 0 and it > 0 and M < eps
    lines = [l for l in log.splitlines() if l.startswith("iter")]
    val = lambda l: float(l.split("max(dAdA) =")[1].split()[0])
    fp32_checks = [val(l) for l in lines[: i32 // 5 + 1] if "(FP64 links)" not in l]
    confirmed = [val(l) for l in lines if "(FP64 links)" in l]
    # the FP32 phase ran until its first check below eps, which was then
    # checked on the FP64 links; the run ends on a confirmed state
    assert fp32_checks[-1] < eps and all(c >= eps for c in fp32_checks[:-1])
    assert confirmed and confirmed[-1] < eps
    if i64 == 0:
        assert i32 == it + 1 and len(confirmed) == 1
    assert abs(F - F64) / F64 < 1e-6
    assert orc.dAdA(orc.g_trafo(u, gq, dims), dims)[1] < eps


def test_coulomb_tiny_eps_on_plain_links(gfx, orc):
    """A Coulomb run whose eps is below the packing bound's reach runs on
    the FP32 and FP64 link planes and stops every slice below eps."""
    dims = (8, 8, 8, 4)
    V = orc.volume(dims)
    u = smooth_config(orc, dims, 97, eps=0.4)
    g = orc.random_init_trafo(V, 98)
    eps = 1e-20
    _setup(gfx, dims, u, g, 1.7, prec=gfx.PREC_MIXED)
    rc, it = gfx.coulomb_gauge(20000, eps, 5)
    F, D, M = gfx.functional_slices()
    g1 = gfx.get_trafo()
    gfx.set_precision(gfx.PREC_FP64)
    gfx.finalize_gaugefixing()
    assert rc == 0 and np.all(it > 0) and np.all(M < eps)
    Fo, Do, Mo = orc.functional_slices(orc.su3_to_2v(u), orc.su3_to_2v(g1), dims, 3)
    assert np.all(Mo < 10 * eps)


@pytest.mark.parametrize("eps", [1e-14, 1e-13])
def test_coulomb_mixed_eps_below_fixed_point_reach(gfx, orc, eps):
    """For eps below 16 (600 2^-32)^2 = 3.1e-13 the FP64 slices of a mixed
    Coulomb run are not polished on fixed-point fields, whose stop bound
    (600 2^-32 added to sqrt(max dAdA)) would take most or all of sqrt(eps):
    at eps = 1e-14 no check could stop a slice.  Every slice stops below eps,
    also on the exact links."""
    dims = (8, 8, 8, 4)
    V = orc.volume(dims)
    u = smooth_config(orc, dims, 131, eps=0.4)
    g = orc.random_init_trafo(V, 132)
    _setup(gfx, dims, u, g, 1.7, prec=gfx.PREC_MIXED)
    rc, it = gfx.coulomb_gauge(20000, eps, 5)
    F, D, M = gfx.functional_slices()
    g1 = gfx.get_trafo()
    st = gfx.last_run_stats()
    gfx.set_precision(gfx.PREC_FP64)
    gfx.finalize_gaugefixing()
    assert rc == 0 and np.all(it > 0) and np.all(M < eps) and st[1] > 0 and st[2] > 0
    Fo, Do, Mo = orc.functional_slices(orc.su3_to_2v(u), orc.su3_to_2v(g1), dims, 3)
    assert np.all(Mo < eps * (1 + 1e-9)), Mo


@pytest.mark.parametrize("dims,ck", [((8, 8, 8, 6), 5), ((8, 8, 8, 4), 1)])
def test_coulomb_packed_matches_plain_links(gfx, orc, dims, ck):
    """Mixed Coulomb on packed links (FP32 slices on the 21-bit links, FP64
    slices on 32-bit fixed-point links and g, the slices holding whole
    workgroups) and on the FP32 and FP64 links: every slice stops below the
    bound on the exact links (host functional of the packed run's result),
    per-slice F equal to 1e-12 relative, stops within two checks of each
    other (within 5 iterations at checkinterval 1: a slice stops once its
    max dAdA is below eps with the rounding's worst effect added, about 0.74
    eps at eps = 1e-12)."""
    V = orc.volume(dims)
    u = smooth_config(orc, dims, 63, eps=0.4)
    g = orc.random_init_trafo(V, 64)
    _setup(gfx, dims, u, g, 1.7, prec=gfx.PREC_MIXED)
    runs = []
    for packed in (0, 1):
        gfx.set_packed_links(packed)
        gfx.set_trafo(g)
        rc, it = gfx.coulomb_gauge(8000, EPS12, ck)
        runs.append((rc, it, gfx.functional_slices(), gfx.get_trafo(), gfx.last_run_stats()))
    gfx.set_packed_links(1)
    gfx.set_precision(gfx.PREC_FP64)
    gfx.finalize_gaugefixing()
    (rc0, it0, (F0, _, M0), _, _), (rc1, it1, (F1, _, M1), g1, st1) = runs
    assert rc0 == 0 and rc1 == 0 and st1[1] > 0 and st1[2] > 0
    assert np.all(M0 < EPS12) and np.all(M1 < EPS12)
    assert np.max(np.abs(F1 - F0) / F0) < 1e-12
    assert np.all(np.abs(it1 - it0) <= max(2 * ck, 5)), (it0, it1)
    Fo, Do, Mo = orc.functional_slices(orc.su3_to_2v(u), orc.su3_to_2v(g1), dims, 3)
    assert np.all(Mo < EPS12 * (1 + 1e-9)) and np.max(np.abs(Fo - F1)) < 1e-12


@pytest.mark.parametrize("prec", [0, 1])
def test_coulomb_slices_without_whole_blocks(gfx, orc, prec):
    """6^3 x 4: a slice holds 108 sites of a parity, not a whole number of
    sweep blocks, so the checks run as standalone per-slice functionals (the
    fused form needs whole blocks); FP64 matches the oracle's stops and g, the
    mixed form reaches the bound with the oracle's per-slice F."""
    dims = (6, 6, 6, 4)
    V = orc.volume(dims)
    u = smooth_config(orc, dims, 71, eps=0.35)
    g = orc.random_init_trafo(V, 72)
    _setup(gfx, dims, u, g, 1.7, prec=prec)
    rc, it_gpu = gfx.coulomb_gauge(6000, EPS12, 3)
    g_gpu = gfx.get_trafo()
    F, D, M = gfx.functional_slices()
    gfx.set_precision(gfx.PREC_FP64)
    gfx.finalize_gaugefixing()
    u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
    rc_o, it_cpu = orc.coulomb_gauge(u2, g2, dims, 6000, EPS12, 3, 1.7)
    Fo, Do, Mo = orc.functional_slices(u2, g2, dims, 3)
    assert rc == 0 and rc_o == 0 and np.all(M < EPS12)
    assert np.max(np.abs(F - Fo) / Fo) < 1e-10
    if prec == 0:
        assert np.all(np.abs(it_gpu - it_cpu) <= 2), (it_gpu, it_cpu)
        assert np.abs(g_gpu - orc.from_2v(g2)).max() < 1e-10
