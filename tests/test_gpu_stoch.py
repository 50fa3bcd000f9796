"""GPU tests of the stochastic parts: the Mersenne Twister fields bit for bit
against the oracle, Box-Muller to its error bound, the link heatbath against
known plaquette values, simulated annealing invariants."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIBC = C.CDLL("libc.so.6")


def synthetic_params(seed=12345):
    """Twister records for the test: the reference reads them from
    MersenneTwister.dat, which is not shipped; any matrix_a with the top bit
    set and any tempering masks exercise the same recurrence."""
    rng = np.random.default_rng(seed)
    p = np.zeros((4096, 4), dtype=np.uint32)
    p[:, 0] = rng.integers(0, 2 ** 32, 4096, dtype=np.uint64).astype(np.uint32) | np.uint32(0x80000000)
    p[:, 1] = rng.integers(0, 2 ** 32, 4096, dtype=np.uint64).astype(np.uint32)
    p[:, 2] = rng.integers(0, 2 ** 32, 4096, dtype=np.uint64).astype(np.uint32)
    return p


def oracle_params(orc, p, srand_seed):
    """seedMTGPU (MersenneTwister.cu:98-114): srand(seed), one rand() per twister."""
    LIBC.srand(srand_seed)
    arr = (orc.MTParams * 4096)()
    for i in range(4096):
        arr[i].matrix_a, arr[i].mask_b, arr[i].mask_c = int(p[i, 0]), int(p[i, 1]), int(p[i, 2])
        arr[i].seed = LIBC.rand() & 0xFFFFFFFF
    return arr


def with_real_records(p):
    """p with the 16 records of the reference's MersenneTwister.dat held by
    the words fixture (tests/golden/mt_dat_words.json) in their slots."""
    import json
    import os
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "mt_dat_words.json")))
    p = p.copy()
    for i, rec in zip(fx["indices"], fx["records"]):
        p[i, :3] = rec
    return p


@pytest.mark.parametrize("n_gauss,n_unif", [(4096 * 6, 4096 * 20), (5000, 19 * 4096 + 7)])
def test_mt_fields_bitexact(gfx, orc, n_gauss, n_unif):
    """The GPU twister fields bit for bit against the oracle (whose words
    for the reference's own records are pinned by the fixture), with 16 real
    records among synthetic ones."""
    p = with_real_records(synthetic_params())
    gfx.set_mt_params(p.ravel())
    gfx.set_mt_seed(4357)
    gfx.init_MT(n_gauss, n_unif)
    params = oracle_params(orc, p, 4357)
    state = np.zeros(4096 * 19, dtype=np.uint32)
    ng = orc.lib.orc_mt_nper(n_gauss, 4096)
    nu = orc.lib.orc_mt_nper(n_unif, 4096)
    for step in range(3):
        init = 0 if step == 0 else 1
        ug = orc.mt_generate(params, state, ng, init)
        uu = orc.mt_generate(params, state, nu, init)
        if step > 0:
            gfx.update_MT()
        gauss, unif, st, pr = gfx.get_mt_fields()
        assert np.array_equal(unif, uu), step
        assert np.array_equal(st, state), step
        assert np.array_equal(pr.reshape(-1, 4)[:, 3], np.array([q.seed for q in params],
                                                               dtype=np.uint32))
        # Box-Muller against the exact transform of the same uniforms
        u1 = ug.reshape(ng, 4096)[0::2].astype(np.float64)
        u2 = ug.reshape(ng, 4096)[1::2].astype(np.float64)
        r = np.sqrt(-2.0 * np.log(u1))
        z1, z2 = r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)
        g = gauss.reshape(ng, 4096)
        assert np.abs(g[0::2] - z1).max() < 2e-5 and np.abs(g[1::2] - z2).max() < 2e-5
    gfx.finalize_MT()


def test_mt_uniform_range_and_moments(gfx):
    gfx.set_mt_params(synthetic_params(7).ravel())
    gfx.set_mt_seed(1)
    gfx.init_MT(4096 * 2, 4096 * 256)
    gauss, unif, _, _ = gfx.get_mt_fields()
    gfx.finalize_MT()
    assert unif.min() > 0.0 and unif.max() <= 1.0
    assert abs(unif.mean() - 0.5) < 2e-3 and abs(unif.var() - 1 / 12) < 2e-3


def _thermalize(gfx, dims, beta, nsweep, seed, start_hot=False):
    gfx.set_lattice(*dims)
    u = gfx.unit_gauge()
    if start_hot:
        gfx.plant_seeds(seed)
        u = gfx.random_init_gauge_host()
    gfx.init_thermalization(u)
    gfx.set_mt_params(synthetic_params(seed).ravel())
    gfx.set_mt_seed(seed)
    return u


@pytest.mark.parametrize("beta,want,tol", [(6.0, 0.5937, 0.006), (0.5, 0.5 / 18 + 0.25 / 216, 0.00087)])
def test_heatbath_plaquette(gfx, orc, beta, want, tol):
    """SU(3) Wilson plaquette: beta = 6.0 gives 0.5937 (large volume; 8^4 is
    within the tolerance); strong coupling <P> = beta/18 + beta^2/216 + O(beta^3),
    checked to 3% at beta = 0.5 (the statistical error of the mean of 60
    sweeps over the 6 x 8^4 plaquettes, each of standard deviation about
    1/sqrt(18), is about 2e-4)."""
    dims = (8, 8, 8, 8)
    _thermalize(gfx, dims, beta, 0, 4357)
    gfx.heatbath_sweeps(150, beta)
    vals = []
    for _ in range(60):
        gfx.heatbath_sweeps(1, beta)
        vals.append(gfx.calc_plaquette(0))
    u = gfx.get_gauge()
    gfx.finalize_MT()
    gfx.finalize_thermalization()
    assert abs(np.mean(vals) - want) < tol, np.mean(vals)
    # links stay in SU(3); the host plaquette of the downloaded field agrees
    m = u.reshape(-1, 3, 3, 2)
    m = m[..., 0] + 1j * m[..., 1]
    assert np.abs(m @ np.conj(np.swapaxes(m, -1, -2)) - np.eye(3)).max() < 1e-13
    assert abs(orc.mean_plaq(u, dims) - vals[-1]) < 1e-13


def test_heatbath_cold_and_hot_agree(gfx):
    """Equilibrium does not depend on the start."""
    dims = (8, 8, 8, 8)
    res = []
    for hot in (False, True):
        _thermalize(gfx, dims, 5.7, 0, 11, start_hot=hot)
        gfx.heatbath_sweeps(150, 5.7)
        v = []
        for _ in range(40):
            gfx.heatbath_sweeps(1, 5.7)
            v.append(gfx.calc_plaquette(0))
        gfx.finalize_MT()
        gfx.finalize_thermalization()
        res.append(np.mean(v))
    assert abs(res[0] - res[1]) < 0.006


@pytest.mark.parametrize("sampler", [0, 1])
def test_simulated_annealing_raises_functional(gfx, orc, sampler):
    dims = (8, 8, 8, 8)
    V = orc.volume(dims)
    _thermalize(gfx, dims, 6.0, 0, 99)
    gfx.heatbath_sweeps(60, 6.0)
    u = gfx.get_gauge()
    gfx.finalize_MT()
    gfx.finalize_thermalization()
    g = orc.random_init_trafo(V, 351)
    gfx.set_lattice(*dims)
    gfx.init_gaugefixing(u, g)
    gfx.calc_functional()
    F0 = gfx.results()[0]
    sa = gfx.saparam()
    sa.Tmin, sa.Tmax, sa.expo, sa.N, sa.checkint = 0.01, 0.45, -1.0, 300, 50
    gfx.set_mt_params(synthetic_params(3).ravel())
    gfx.set_mt_seed(3)
    gfx.set_sa(sampler, 3, 4)
    gfx.simannealing_gauge()
    gfx.calc_functional()
    F1 = gfx.results()[0]
    gsa = gfx.get_trafo()
    it = gfx.overrelax_gauge(5000, float(np.float32(1e-12)), 10)
    F2 = gfx.results()[0]
    gfx.finalize_MT()
    gfx.finalize_gaugefixing()
    m = gsa.reshape(-1, 3, 3, 2)
    m = m[..., 0] + 1j * m[..., 1]
    assert np.abs(m @ np.conj(np.swapaxes(m, -1, -2)) - np.eye(3)).max() < 1e-13
    assert F1 > F0 + 0.3 and it > 0 and F2 >= F1 - 1e-3


def test_micro_steps_preserve_functional(gfx, orc):
    """The microcanonical reflection changes g but keeps F (per subgroup
    Re tr(alpha X) is unchanged, heatbath_thermalization.cu:318-338)."""
    dims = (8, 8, 8, 8)
    V = orc.volume(dims)
    _thermalize(gfx, dims, 6.0, 0, 5)
    gfx.heatbath_sweeps(30, 6.0)
    u = gfx.get_gauge()
    gfx.finalize_MT()
    gfx.finalize_thermalization()
    gfx.set_lattice(*dims)
    gfx.init_gaugefixing(u, orc.random_init_trafo(V, 8))
    gfx.calc_functional()
    F0 = gfx.results()[0]
    g0 = gfx.get_trafo()
    gfx.micro_sweeps(3, 4)
    gfx.calc_functional()
    F1 = gfx.results()[0]
    g1 = gfx.get_trafo()
    gfx.finalize_gaugefixing()
    assert abs(F1 - F0) < 1e-12 * abs(F0)
    assert np.abs(g1 - g0).max() > 1e-2


def _a0_cdf(a, x):
    """CDF on [-1, 1] of the density sqrt(1 - a0^2) exp(a a0)."""
    from scipy import integrate
    f = lambda t: np.sqrt(max(0.0, 1.0 - t * t)) * np.exp(a * (t - 1.0))
    grid = np.linspace(-1.0, 1.0, 4001)
    pdf = np.array([f(t) for t in grid])
    cdf = integrate.cumulative_trapezoid(pdf, grid, initial=0.0)
    cdf /= cdf[-1]
    return np.interp(x, grid, cdf)


@pytest.mark.parametrize("a", [0.7, 3.0, 12.0])
def test_kennedy_pendleton_sampler_density(gfx, a):
    """alpha_0 of the Kennedy-Pendleton sampler follows
    sqrt(1 - a0^2) exp(a a0), a = beta |w| (Kolmogorov-Smirnov)."""
    from scipy import stats
    gfx.set_mt_params(synthetic_params(91).ravel())
    gfx.set_mt_seed(91)
    a0 = gfx.su2_sample(1, a / 2.0, 2.0, 200000)
    gfx.finalize_MT()
    acc = a0[~np.isnan(a0)]
    # four trials; at small a the proposals often fall outside [-1, 1] and all
    # four can fail (the subgroup then stays as it is)
    assert acc.size > (0.8 if a < 1 else 0.95) * a0.size
    assert np.all(np.abs(acc) <= 1.0)
    res = stats.kstest(acc, lambda x: _a0_cdf(a, x))
    assert res.pvalue > 1e-3, res


@pytest.mark.parametrize("a", [0.7, 3.0, 12.0])
def test_kennedy_pendleton_sampler_directions(gfx, a):
    """All four components of the Kennedy-Pendleton draws: unit quaternions
    whose direction (alpha_1, alpha_2, alpha_3)/|.| is uniform on the sphere
    and independent of alpha_0; with the alpha_0 law above this is the SU(2)
    conditional exp(a alpha_0) of the Haar measure, the distribution that
    gives the heatbath detailed balance."""
    from scipy import stats
    from test_oracle import _sphere_ks
    gfx.set_mt_params(synthetic_params(93).ravel())
    gfx.set_mt_seed(93)
    al = gfx.su2_sample4(1, a / 2.0, 2.0, 200000)
    gfx.finalize_MT()
    acc = al[~np.isnan(al[:, 0])]
    assert np.abs(np.sum(acc * acc, axis=1) - 1.0).max() < 1e-12
    vec = acc[:, 1:] / np.linalg.norm(acc[:, 1:], axis=1)[:, None]
    p3, pphi = _sphere_ks(vec[:, 2], np.arctan2(vec[:, 1], vec[:, 0]))
    assert p3 > 1e-3 and pphi > 1e-3, (p3, pphi)
    assert abs(stats.pearsonr(acc[:, 0], vec[:, 2])[0]) < 0.01


def test_reference_sampler_components_follow_its_rule(gfx, orc):
    """All four components of the reference rule's draws against the
    oracle's restatement on the same twister numbers."""
    gfx.set_mt_params(synthetic_params(94).ravel())
    gfx.set_mt_seed(94)
    n, beta, wnorm = 50000, 0.8, 1.7
    gfx.su2_sample(0, beta, wnorm, n)  # sizes the fields, then refills them
    gauss, unif = gfx.get_mt_fields()[:2]
    al = gfx.su2_sample4(0, beta, wnorm, n)
    gfx.finalize_MT()
    want = orc.sa_sample_ref(gauss, unif, beta, wnorm, n, components=4)
    same = np.all(np.abs(al - want) < 1e-5, axis=1)
    assert same.mean() > 0.9999, same.mean()


def test_reference_sampler_follows_its_rule(gfx, orc):
    """The reference rule (simulated_annealing.cu:1-73) on the same twister
    numbers as the oracle's restatement."""
    gfx.set_mt_params(synthetic_params(92).ravel())
    gfx.set_mt_seed(92)
    n, beta, wnorm = 50000, 0.8, 1.7
    gfx.su2_sample(0, beta, wnorm, n)  # sizes the fields, then refills them
    gauss, unif = gfx.get_mt_fields()[:2]
    a0 = gfx.su2_sample(0, beta, wnorm, n)
    gfx.finalize_MT()
    want = orc.sa_sample_ref(gauss, unif, beta, wnorm, n)
    same = np.abs(a0 - want) < 1e-5
    assert same.mean() > 0.9999, same.mean()
    assert 0.05 < np.mean(a0 != wnorm) < 1.0
