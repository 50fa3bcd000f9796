"""Pins of the CPU oracle (oracle/oracle.c): golden vectors of the reference's
own host C routines, the live reference build where present, known-answer
tests and the closed forms / invariants of the algorithm."""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _gold():
    return np.load(os.path.join(GOLD, "ref_fields.npz"))


# ------------------------------------------------------------------ golden vectors
def test_lehmer_kat(orc):
    """TestRandom (rngs.c:153-179): seed 1, 10000 draws -> 399268537;
    PlantSeeds(1) -> stream 1 = 22925."""
    lib = orc.lib
    lib.orc_lehmer_select_stream(0)
    lib.orc_lehmer_put_seed(1)
    for _ in range(10000):
        lib.orc_lehmer_random()
    assert lib.orc_lehmer_get_seed() == 399268537
    lib.orc_lehmer_select_stream(1)
    lib.orc_lehmer_plant_seeds(1)
    assert lib.orc_lehmer_get_seed() == 22925
    lib.orc_lehmer_select_stream(0)
    gold = json.load(open(os.path.join(GOLD, "lehmer.json")))
    assert gold["check_after_10000"] == 399268537 and gold["stream1_after_plant1"] == 22925
    lib.orc_lehmer_plant_seeds(gold["seed"])
    assert [lib.orc_lehmer_random() for _ in range(20)] == gold["draws"]


def test_random_init_bitexact(orc):
    gd = _gold()
    dims = tuple(gd["dims"])
    V = orc.volume(dims)
    assert np.array_equal(orc.random_init_trafo(V, 351), gd["trafo"])
    assert np.array_equal(orc.random_init_gauge(V, 4357), gd["gauge"])


def test_host_observables_bitexact(orc):
    gd = _gold()
    dims = tuple(gd["dims"])
    u, g = gd["gauge"], gd["trafo"]
    ut = orc.g_trafo(u, g, dims)
    assert np.array_equal(ut, gd["gauge_transformed"])
    for tag, a in (("u", u), ("ut", ut)):
        assert orc.gauge_functional(a, dims) == gd[tag + "_F"]
        d, m = orc.dAdA(a, dims)
        assert d == gd[tag + "_dAdA"] and m == gd[tag + "_maxdAdA"]
        assert orc.mean_plaq(a, dims) == gd[tag + "_plaq"]


def test_reference_routine_sweep_matches_oracle(orc):
    """The CPU baseline sweep built from the reference's own host SU(3)
    routines on its 18-real layout (oracle/ref_sweep.c) and the oracle's
    2-row sweep give the same g after an even and an odd half sweep (1e-13;
    they differ in summation order and in rebuilding row 2)."""
    if orc.ref_lib() is None:
        pytest.skip("oracle/_ref not built (no reference checkout)")
    dims = (4, 4, 6, 8)
    V = orc.volume(dims)
    u = orc.random_init_gauge(V, 4501)
    g = orc.random_init_trafo(V, 4502)
    g_ref = g.copy()
    for parity in (0, 1):
        orc.ref_halfsweep(u, g_ref, dims, parity, 4, 1.7)
    u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
    for parity in (0, 1):
        orc.halfsweep(u2, g2, dims, parity, 4, 1.7)
    err = np.abs(g_ref - orc.from_2v(g2)).max()
    assert err < 1e-13 and np.abs(g_ref - g).max() > 1e-3


def test_against_live_reference_build(orc):
    """Where oracle/_ref was built from the reference checkout: bigger lattice."""
    import ctypes as C

    ref = orc.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (no reference checkout)")
    dims = (6, 4, 4, 8)
    V = orc.volume(dims)
    for n, v in zip(["LX", "LY", "LZ", "T", "VOLUME"], list(dims) + [V]):
        C.c_int.in_dll(ref, n).value = v
    nn = orc.neighbours(dims)
    C.c_void_p.in_dll(ref, "nn").value = nn.ctypes.data
    vp = C.c_void_p
    ref.PlantSeeds.argtypes = [C.c_long]
    ref.random_init_gauge.argtypes = [vp]
    ref.random_init_trafo.argtypes = [vp]
    ref.g_trafo.argtypes = [vp, vp]
    for f in ("gauge_functional", "dAdA", "mean_plaq"):
        getattr(ref, f).restype = C.c_double
        getattr(ref, f).argtypes = [vp]
    ref.PlantSeeds(99)
    u = np.zeros(18 * 4 * V)
    ref.random_init_gauge(u.ctypes.data_as(vp))
    ref.PlantSeeds(98)
    g = np.zeros(18 * V)
    ref.random_init_trafo(g.ctypes.data_as(vp))
    assert np.array_equal(u, orc.random_init_gauge(V, 99))
    assert np.array_equal(g, orc.random_init_trafo(V, 98))
    ut = u.copy()
    ref.g_trafo(ut.ctypes.data_as(vp), g.ctypes.data_as(vp))
    assert np.array_equal(ut, orc.g_trafo(u, g, dims))
    assert ref.dAdA(ut.ctypes.data_as(vp)) == orc.dAdA(ut, dims)[0]
    assert ref.mean_plaq(ut.ctypes.data_as(vp)) == orc.mean_plaq(ut, dims)


# ------------------------------------------------------------------ helpers (numpy, independent)
def mats(a):
    r = a.reshape(-1, 3, 3, 2)
    return r[..., 0] + 1j * r[..., 1]


def flat(m):
    return np.stack([m.real, m.imag], axis=-1).reshape(-1)


def unitary_err(m):
    e = m @ np.conj(np.swapaxes(m, -1, -2)) - np.eye(3)
    return np.abs(e).max(), np.abs(np.linalg.det(m) - 1).max()


def rand_su3(orc, n, seed):
    return mats(orc.random_init_trafo(n, seed))


# ------------------------------------------------------------------ closed forms
def test_reconstruct_and_normalize_give_su3(orc):
    m = orc.random_init_trafo(64, 7)
    for k in range(64):
        r = orc.su3_to_2v(m[18 * k:18 * k + 18])
        full = np.zeros(18)
        orc.lib.orc_reconstruct(r, full)
        dag = np.zeros(18)
        orc.lib.orc_reconstruct_dagger(r, dag)
        M = mats(full)[0]
        assert np.allclose(mats(dag)[0], M.conj().T, atol=0, rtol=0)
        ue, de = unitary_err(M)
        assert ue < 1e-14 and de < 1e-14
    # normalise a perturbed matrix
    a = m[:18] + 1e-3 * np.sin(np.arange(18))
    orc.lib.orc_su3_normalize(a)
    ue, de = unitary_err(mats(a)[0])
    assert ue < 1e-14 and de < 1e-14


def test_su2_relax_is_the_maximiser(orc):
    """One CM step on a star that lives in the (0,1) subgroup: g K becomes the
    identity in that block (Re tr(alpha w) is maximal at alpha = w/|w|)."""
    rng = np.random.default_rng(1)
    for _ in range(20):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        s = rng.uniform(0.5, 2.0)
        su2 = np.array([[q[0] + 1j * q[3], q[2] + 1j * q[1]], [-q[2] + 1j * q[1], q[0] - 1j * q[3]]])
        K = np.eye(3, dtype=complex)
        K[:2, :2] = s * su2
        g = np.eye(3, dtype=complex)
        ga = flat(g)
        orc.lib.orc_cm_relax(ga, flat(K))
        gk = mats(ga)[0] @ K
        assert np.allclose(gk[:2, :2], s * np.eye(2), atol=1e-13)
        # and no unit SU(2) rotation does better
        F = np.real(np.trace(gk))
        for _ in range(50):
            r = rng.normal(size=4)
            r /= np.linalg.norm(r)
            R = np.eye(3, dtype=complex)
            R[:2, :2] = [[r[0] + 1j * r[3], r[2] + 1j * r[1]], [-r[2] + 1j * r[1], r[0] - 1j * r[3]]]
            assert np.real(np.trace(R @ mats(ga)[0] @ K)) <= F + 1e-12


def test_cm_relax_does_not_decrease_functional(orc):
    g = rand_su3(orc, 30, 3)
    K = rand_su3(orc, 30, 4) + rand_su3(orc, 30, 5) + 0.3 * rand_su3(orc, 30, 6)
    for k in range(30):
        ga = flat(g[k])
        before = np.real(np.trace(g[k] @ K[k]))
        orc.lib.orc_cm_relax(ga, flat(K[k]))
        assert np.real(np.trace(mats(ga)[0] @ K[k])) >= before - 1e-12


@pytest.mark.parametrize("omega", [1.68, 1.7, 1.0])
def test_short_series_equals_reference_series(orc, omega):
    """The short form of the 3-term series (rows 0, 1 of g + c1 (h - g) +
    c1 c2 (h g^+ h - 2 h + g), then normalised) equals the reference's
    normalize(1 + c1 R + c1 c2 R^2) g with R = h g^+ - 1 for unitary g (to
    rounding), and differs from it by the non-unitarity when g is not
    unitary (the two forms then have to be told apart)."""
    g = rand_su3(orc, 20, 18)
    h = rand_su3(orc, 20, 19)
    for k in range(20):
        a, b = flat(g[k]), flat(g[k])
        orc.lib.orc_overrelax_step(a, flat(h[k]), omega, 3)
        orc.lib.orc_su3_normalize(a)
        orc.lib.orc_overrelax_step_short(b, flat(h[k]), omega)
        assert np.abs(a - b).max() < 1e-14
        gp = g[k] * (1 + 1e-8 * np.cos(np.arange(9)).reshape(3, 3))
        a, b = flat(gp), flat(gp)
        orc.lib.orc_overrelax_step(a, flat(h[k]), omega, 3)
        orc.lib.orc_su3_normalize(a)
        orc.lib.orc_overrelax_step_short(b, flat(h[k]), omega)
        if omega != 1.0:
            assert np.abs(a - b).max() > 1e-11


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (4, 6, 4, 2)])
def test_short_sweep_equals_reference_sweep(orc, dims):
    """A half sweep in the short form (nterms = -3: the series in its short
    form on h as the relaxation leaves it, not normalised) equals the
    reference's half sweep (h normalised, overrelax.cu:411-417) to rounding
    on unitary links and g, both parities in turn; with g off the group by
    1e-6 (the short form relies on the unitarity of g) the two differ by
    more than rounding, so the comparison can tell them apart."""
    V = orc.volume(dims)
    u2 = orc.su3_to_2v(orc.random_init_gauge(V, 4401))
    g0 = orc.su3_to_2v(orc.random_init_trafo(V, 4402))
    for off, lo, hi in ((0.0, 0.0, 1e-13), (1e-6, 1e-9, 1.0)):
        g = g0 * (1.0 + off * np.cos(np.arange(g0.size)))
        a, b = g.copy(), g.copy()
        for parity in (0, 1):
            orc.halfsweep(u2, a, dims, parity, 4, 1.7)
            orc.halfsweep(u2, b, dims, parity, 4, 1.7, nterms=-3)
        d = np.abs(a - b).max()
        assert lo <= d < hi, (off, d)


@pytest.mark.parametrize("omega", [1.0, 2.0, 1.68, 1.7])
def test_overrelax_series_closed_form(orc, omega):
    """overrelax.cu:149-177 with N = 3: g -> normalize(1 + w R + w(w-1)/2 R^2) g,
    R = gn g^+ - 1.  w = 1 gives gn, w = 2 gives (gn g^+)^2 g."""
    g = rand_su3(orc, 10, 8)
    gn = rand_su3(orc, 10, 9)
    for k in range(10):
        ga = flat(g[k])
        orc.lib.orc_overrelax_step(ga, flat(gn[k]), omega, 3)
        P = gn[k] @ g[k].conj().T
        R = P - np.eye(3)
        S = np.eye(3) + omega * R + omega * (omega - 1) / 2 * R @ R
        # Gram-Schmidt of rows 0,1, row 2 = conj cross product
        a = S[0] / np.linalg.norm(S[0])
        b = S[1] - np.vdot(a, S[1]) * a
        b /= np.linalg.norm(b)
        Sn = np.array([a, b, np.conj(np.cross(a, b))])
        want = Sn @ g[k]
        assert np.abs(mats(ga)[0] - want).max() < 1e-13
        if omega == 1.0:
            assert np.abs(mats(ga)[0] - gn[k]).max() < 1e-13
        if omega == 2.0:
            assert np.abs(mats(ga)[0] - P @ P @ g[k]).max() < 1e-12


def test_halfsweep_brute_force_2x4(orc):
    """A 2^4 lattice (every neighbour pair is doubly connected) against a
    direct numpy evaluation of the star, CM relaxation and overrelaxation."""
    dims = (2, 2, 2, 2)
    V = 16
    u = orc.random_init_gauge(V, 21)
    g = orc.random_init_trafo(V, 22)
    U = mats(orc.from_2v(orc.su3_to_2v(u))).reshape(V, 4, 3, 3)
    G = mats(orc.from_2v(orc.su3_to_2v(g)))
    g2 = orc.su3_to_2v(g)
    orc.halfsweep(orc.su3_to_2v(u), g2, dims, 0, 4, 1.68)
    got = mats(orc.from_2v(g2))

    def coords(s):
        return [s % 2, (s // 2) % 2, (s // 4) % 2, s // 8]

    def site(x):
        return x[0] + 2 * (x[1] + 2 * (x[2] + 2 * x[3]))

    for s in range(V):
        x = coords(s)
        if sum(x) % 2:
            assert np.array_equal(got[s], G[s])  # odd sites untouched
            continue
        K = np.zeros((3, 3), complex)
        for mu in range(4):
            xp = list(x)
            xp[mu] = (xp[mu] + 1) % 2
            xm = list(x)
            xm[mu] = (xm[mu] - 1) % 2
            K += U[s, mu] @ G[site(xp)].conj().T + U[site(xm), mu].conj().T @ G[site(xm)].conj().T
        h = G[s].copy()
        for a, b in ((0, 1), (0, 2), (1, 2)):
            X = h @ K
            w = np.array([X[a, a].real + X[b, b].real, -X[a, b].imag - X[b, a].imag,
                          -X[a, b].real + X[b, a].real, -X[a, a].imag + X[b, b].imag])
            w /= np.linalg.norm(w)
            A = np.array([[w[0] + 1j * w[3], w[2] + 1j * w[1]], [-w[2] + 1j * w[1], w[0] - 1j * w[3]]])
            h[[a, b]] = A @ h[[a, b]]

        def gs(M):
            a0 = M[0] / np.linalg.norm(M[0])
            b0 = M[1] - np.vdot(a0, M[1]) * a0
            b0 /= np.linalg.norm(b0)
            return np.array([a0, b0, np.conj(np.cross(a0, b0))])

        h = gs(h)
        R = h @ G[s].conj().T - np.eye(3)
        S = gs(np.eye(3) + 1.68 * R + 1.68 * 0.68 / 2 * R @ R)
        want = gs(S @ G[s])
        assert np.abs(got[s] - want).max() < 1e-13


def test_functional_of_cold_and_pure_gauge(orc):
    dims = (4, 4, 4, 4)
    V = orc.volume(dims)
    unit = np.zeros(18 * 4 * V)
    unit.reshape(-1, 18)[:, [0, 8, 16]] = 1.0
    gu = np.zeros(18 * V)
    gu.reshape(-1, 18)[:, [0, 8, 16]] = 1.0
    F, D, M = orc.functional(orc.su3_to_2v(unit), orc.su3_to_2v(gu), dims)
    assert F == 3.0 and D == 0.0 and M == 0.0
    # pure gauge U = h h^+(x+mu) transformed back with g = h^+ gives F = 3
    h = orc.random_init_trafo(V, 5)
    u = orc.g_trafo(unit, h, dims)
    hd = flat(np.conj(np.swapaxes(mats(h), -1, -2)))
    F, D, M = orc.functional(orc.su3_to_2v(u), orc.su3_to_2v(hd), dims)
    assert abs(F - 3.0) < 1e-14 and M < 1e-25


def np_site_theta(ut, dims, ndir=4, half=0.5, traceless=True):
    """Per-site Re tr(Delta Delta^+) of host-layout links ut, restated in
    numpy from observables.c:70-165 independently of oracle.c:
    A(U) = (U - U^+)/(2i) made traceless, Delta(x) = sum_mu A(U_mu(x-mu)) -
    A(U_mu(x)).  `half` and `traceless` exist only to plant mistakes."""
    LX, LY, LZ, LT = dims
    U = mats(ut).reshape(LT, LZ, LY, LX, 4, 3, 3)

    def A(M):
        h = (M - np.conj(np.swapaxes(M, -1, -2))) * (-1j * half)
        if traceless:
            h = h - (np.trace(h, axis1=-2, axis2=-1) / 3.0)[..., None, None] * np.eye(3)
        return h

    D = np.zeros(U.shape[:4] + (3, 3), dtype=complex)
    for mu in range(ndir):
        Um = U[..., mu, :, :]
        # out[x] = Um[x - mu]: x is the last lattice axis, t the first
        D += A(np.roll(Um, 1, axis=3 - mu)) - A(Um)
    return (np.abs(D) ** 2).sum(axis=(-2, -1))


def test_theta_pinned_on_unfixed_fields(orc):
    """The stop quantity of the sweeps and of calc_functional (dAdA and max
    dAdA of orc.functional: Delta(x) = -A(g(x) K(x)) from the star, the form
    of overrelax.cu:482-605 that the fused checks use) equals the reference's
    host dAdA (observables.c:70-165) of the transformed links on random,
    not gauge-fixed fields: against the golden values written by the
    reference build (4^4), then against the pinned host dAdA and a numpy
    restatement on other shapes, where oracle/_ref is present also against
    the live reference routine.  A dropped 1/2 in A or a dropped trace
    subtraction changes the values far beyond the tolerance."""
    gd = _gold()
    dims = tuple(int(x) for x in gd["dims"])
    F, D, M = orc.functional(orc.su3_to_2v(gd["gauge"]), orc.su3_to_2v(gd["trafo"]), dims)
    assert abs(D - gd["ut_dAdA"]) <= 1e-13 * gd["ut_dAdA"]
    assert abs(M - gd["ut_maxdAdA"]) <= 1e-13 * gd["ut_maxdAdA"]
    assert abs(F - gd["ut_F"]) <= 1e-13
    th = np_site_theta(gd["gauge_transformed"], dims)
    assert abs(th.mean() - gd["ut_dAdA"]) <= 1e-13 * gd["ut_dAdA"]
    assert abs(th.max() - gd["ut_maxdAdA"]) <= 1e-13 * gd["ut_maxdAdA"]
    # the planted mistakes are seen
    assert abs(np_site_theta(gd["gauge_transformed"], dims, half=1.0).mean() / gd["ut_dAdA"] - 4) < 1e-12
    assert abs(np_site_theta(gd["gauge_transformed"], dims, traceless=False).mean() / gd["ut_dAdA"] - 1) > 1e-3
    ref = orc.ref_lib()
    for dims, seed in (((4, 4, 4, 4), 6101), ((4, 6, 4, 8), 6102), ((6, 4, 8, 4), 6103),
                       ((4, 6, 4, 8), 6104)):
        V = orc.volume(dims)
        u = orc.random_init_gauge(V, seed)
        g = orc.random_init_trafo(V, seed + 500)
        ut = orc.g_trafo(u, g, dims)
        F, D, M = orc.functional(orc.su3_to_2v(u), orc.su3_to_2v(g), dims)
        d_host, m_host = orc.dAdA(ut, dims)
        assert D > 1.0 and M > D
        assert abs(D - d_host) <= 1e-13 * d_host and abs(M - m_host) <= 1e-13 * m_host
        assert abs(F - orc.gauge_functional(ut, dims)) <= 1e-13
        th = np_site_theta(ut, dims)
        assert abs(th.mean() - d_host) <= 1e-13 * d_host and abs(th.max() - m_host) <= 1e-13 * m_host
        if ref is not None:
            import ctypes as C
            for n, v in zip(["LX", "LY", "LZ", "T", "VOLUME"], list(dims) + [V]):
                C.c_int.in_dll(ref, n).value = v
            nn = orc.neighbours(dims)
            C.c_void_p.in_dll(ref, "nn").value = nn.ctypes.data
            ref.dAdA.restype = C.c_double
            ref.dAdA.argtypes = [C.c_void_p]
            d_ref = ref.dAdA(ut.ctypes.data_as(C.c_void_p))
            m_ref = C.c_double.in_dll(ref, "maxDADA").value
            assert abs(D - d_ref) <= 1e-13 * d_ref and abs(M - m_ref) <= 1e-13 * m_ref


def test_theta_slices_pinned(orc):
    """Per-slice F, dAdA and max dAdA of the Coulomb checks
    (orc.functional_slices over the 3 spatial directions) against the numpy
    restatement of the spatial Delta on the transformed links, per time
    slice, on random fields that are not gauge fixed."""
    for dims, seed in (((4, 4, 4, 6), 6201), ((6, 4, 4, 4), 6202)):
        V = orc.volume(dims)
        u = orc.random_init_gauge(V, seed)
        g = orc.random_init_trafo(V, seed + 500)
        ut = orc.g_trafo(u, g, dims)
        F, D, M = orc.functional_slices(orc.su3_to_2v(u), orc.su3_to_2v(g), dims, 3)
        th = np_site_theta(ut, dims, ndir=3).reshape(dims[3], -1)
        Ut = mats(ut).reshape(dims[3], -1, 4, 3, 3)
        f_np = np.trace(Ut[:, :, :3], axis1=-2, axis2=-1).real.sum(axis=(1, 2)) / (3 * th.shape[1])
        assert np.all(D > 0.5)
        assert np.allclose(D, th.mean(axis=1), rtol=1e-13, atol=0)
        assert np.allclose(M, th.max(axis=1), rtol=1e-13, atol=0)
        assert np.allclose(F, f_np, rtol=0, atol=1e-14)
        # the 4-direction form differs: the test sees the direction count
        assert not np.allclose(D, np_site_theta(ut, dims, ndir=4).reshape(dims[3], -1).mean(axis=1), rtol=1e-3)


def test_stop_rule_quirks(orc):
    """overrelax.cu:775-822: the stop is checked only when i % ck == 0, so the
    stop iteration is a multiple of ck; maxit = 0 returns 0."""
    dims = (4, 4, 4, 4)
    V = orc.volume(dims)
    unit = np.zeros(18 * 4 * V)
    unit.reshape(-1, 18)[:, [0, 8, 16]] = 1.0
    u = orc.su3_to_2v(orc.g_trafo(unit, orc.random_init_trafo(V, 5), dims))
    g = orc.su3_to_2v(orc.random_init_trafo(V, 6))
    it1, h1 = orc.overrelax_gauge(u, g.copy(), dims, 5000, 1e-12, 1)
    it7, h7 = orc.overrelax_gauge(u, g.copy(), dims, 5000, 1e-12, 7)
    assert it1 > 0 and it7 % 7 == 0 and it1 <= it7 < it1 + 7
    n = len(h1[::7])
    assert np.array_equal(h1[::7], h7[:n])  # same trajectory, sampled
    it0, _ = orc.overrelax_gauge(u, g.copy(), dims, 0, 1e-12, 1)
    assert it0 == 0
    itf, _ = orc.overrelax_gauge(u, g.copy(), dims, 3, 1e-30, 1)
    assert itf == -1


def test_functional_monotone_for_omega_one(orc):
    dims = (4, 4, 4, 4)
    V = orc.volume(dims)
    u = orc.su3_to_2v(orc.random_init_gauge(V, 31))
    g = orc.su3_to_2v(orc.random_init_trafo(V, 32))
    it, hist = orc.overrelax_gauge(u, g, dims, 40, 1e-30, 1, 1.0)
    assert np.all(np.diff(hist[:, 0]) >= -1e-14)


def test_plaquette_gauge_invariant(orc):
    dims = (4, 4, 4, 4)
    V = orc.volume(dims)
    u = orc.random_init_gauge(V, 41)
    ut = orc.g_trafo(u, orc.random_init_trafo(V, 42), dims)
    assert abs(orc.mean_plaq(u, dims) - orc.mean_plaq(ut, dims)) < 1e-14
    assert abs(orc.plaquette_2v(orc.su3_to_2v(u), dims) - orc.mean_plaq(u, dims)) < 1e-14


def test_coulomb_slices_are_independent(orc):
    """Coulomb gauge uses only spatial links: fixing a 4^3 x 4 lattice slice by
    slice equals fixing each slice alone; every slice reaches the bound."""
    dims = (4, 4, 4, 4)
    V = orc.volume(dims)
    u = orc.su3_to_2v(orc.random_init_gauge(V, 51))
    g = orc.su3_to_2v(orc.random_init_trafo(V, 52))
    rc, iters = orc.coulomb_gauge(u, g, dims, 4000, 1e-12, 1)
    assert rc == 0 and np.all(iters > 0)
    F, D, M = orc.functional_slices(u, g, dims, 3)
    assert np.all(M < 1e-12)


def test_reference_sa_rule_density(orc):
    """Pin of the oracle's restatement of the reference SA sampler
    (simulated_annealing.cu:1-73): y = (g^2 - ln(1-u0))/kb with g normal and
    u0 uniform, accepted when 2 u1^2 <= 2 - y, a0 = 1 - y.  S = g^2 + E
    (chi-square(1) plus Exp(1)) has density exp(-s) erfi(sqrt(s/2)), so the
    accepted a0 follow exp(-kb y) erfi(sqrt(kb y/2)) sqrt(1 - y/2), y = 1-a0
    (not the Kennedy-Pendleton density: the reference squares g where KP
    halves it); rejected draws return w."""
    from scipy import integrate, special, stats
    rng = np.random.default_rng(5)
    n, beta, wnorm = 200000, 0.9, 2.0
    gauss = rng.standard_normal(n).astype(np.float32)
    unif = (1.0 - rng.random(4 * n)).astype(np.float32)  # (0, 1] like the twisters
    a0 = orc.sa_sample_ref(gauss, unif, beta, wnorm, n)
    acc = a0[a0 != wnorm]
    kb = wnorm * wnorm * beta
    grid = np.linspace(-1.0, 1.0, 8001)
    y = 1.0 - grid
    pdf = np.exp(-kb * y) * special.erfi(np.sqrt(kb * y / 2.0)) * np.sqrt(np.clip(1 - y / 2, 0, None))

    def cdf_of(p):
        c = integrate.cumulative_trapezoid(p, grid, initial=0.0)
        return lambda x: np.interp(x, grid, c / c[-1])

    assert stats.kstest(acc, cdf_of(pdf)).pvalue > 1e-3
    # the acceptance rate is the integral of the same density over y in [0, 2]
    rate = integrate.quad(lambda t: np.exp(-t) * special.erfi(np.sqrt(t / 2)) * np.sqrt(max(0.0, 1 - t / (2 * kb))), 0, 2 * kb)[0]
    assert abs(acc.size / n - rate) < 0.005
    # a planted mistake (the Kennedy-Pendleton density with a = kb/2) is caught
    kp = np.sqrt(np.clip(1 - grid ** 2, 0, None)) * np.exp(kb / 2 * (grid - 1))
    assert stats.kstest(acc, cdf_of(kp)).pvalue < 1e-6


def _sphere_ks(n3, phi):
    """Kolmogorov-Smirnov p-values of a direction's z component against the
    uniform law on [-1, 1] and of its azimuth against the uniform law on
    (-pi, pi]: together they say the direction is uniform on the sphere."""
    from scipy import stats
    return (stats.kstest(n3, stats.uniform(loc=-1.0, scale=2.0).cdf).pvalue,
            stats.kstest(phi, stats.uniform(loc=-np.pi, scale=2.0 * np.pi).cdf).pvalue)


def test_reference_sa_direction_isotropic(orc):
    """The direction part of the reference SA sampler's alpha
    (simulated_annealing.cu:47-55) as the oracle restates it: for an
    accepted draw alpha is a unit quaternion, and (alpha_1, alpha_2,
    alpha_3)/|.| is uniform on the sphere and independent of alpha_0, so that
    with the a0 law pinned above the draw follows the SU(2) conditional
    exp(kb Re tr alpha) of the Haar measure (which is what gives the heatbath
    detailed balance).  A planted mistake (cos_theta = u2 instead of 2 u2 - 1)
    is caught."""
    from scipy import stats
    rng = np.random.default_rng(6)
    n, beta, wnorm = 200000, 0.9, 2.0
    gauss = rng.standard_normal(n).astype(np.float32)
    unif = (1.0 - rng.random(4 * n)).astype(np.float32)
    al = orc.sa_sample_ref(gauss, unif, beta, wnorm, n, components=4)
    assert np.array_equal(al[:, 0], orc.sa_sample_ref(gauss, unif, beta, wnorm, n))
    acc = al[al[:, 0] != wnorm]
    assert np.abs(np.sum(acc * acc, axis=1) - 1.0).max() < 1e-12
    vec = acc[:, 1:] / np.linalg.norm(acc[:, 1:], axis=1)[:, None]
    p3, pphi = _sphere_ks(vec[:, 2], np.arctan2(vec[:, 1], vec[:, 0]))
    assert p3 > 1e-3 and pphi > 1e-3, (p3, pphi)
    assert abs(stats.pearsonr(acc[:, 0], vec[:, 2])[0]) < 0.01
    # planted: cos_theta drawn from [0, 1) only
    bad_ct = unif.reshape(n, 4)[:, 2].astype(np.float64)[al[:, 0] != wnorm]
    assert _sphere_ks(bad_ct, np.arctan2(vec[:, 1], vec[:, 0]))[0] < 1e-6


def packed_round(u2, lo=True):
    """The packed links of the mixed Landau run restated in numpy: each real
    u of the 2-row links as hi = rint(u 2^23) clamped to 24 bits and, with
    lo, lo = rint((u - hi 2^-23) 2^38) clamped to 16 bits; the value read
    back is hi 2^-23 (+ lo 2^-38)."""
    hi = np.clip(np.rint(u2 * 2.0 ** 23), -8388607, 8388607)
    if not lo:
        return hi * 2.0 ** -23
    q = np.clip(np.rint((u2 - hi * 2.0 ** -23) * 2.0 ** 38), -32767, 32767)
    return hi * 2.0 ** -23 + q * 2.0 ** -38


def p21_round(a):
    """The 21-bit links of the FP32 phase of mixed runs restated in numpy:
    each real u as q = rint(u 2^20) clamped to +-(2^20 - 1), read back as
    q 2^-20."""
    return np.clip(np.rint(a * 2.0 ** 20), -1048575, 1048575) * 2.0 ** -20


def test_p21_round_bound(orc):
    """|p21(u) - u| <= 2^-21 for the reals of SU(3) matrices (2^-20 at
    |u| = 1, where q clamps); and the worst effect of that rounding on a
    site's sqrt(theta), 83.2 2^-21 = 4e-5, is far below the 1e-3 at which the
    FP32 phase ends by default (switch theta = 1e-6): measured on random
    links and g it stays below that bound."""
    u2 = orc.su3_to_2v(orc.random_init_gauge(4096, 4394))
    err = np.abs(p21_round(u2) - u2)
    assert err.max() <= 2.0 ** -21 and err.max() > 2.0 ** -25
    one = np.array([1.0, -1.0, 0.0, 0.5])
    assert np.abs(p21_round(one) - one).max() == 2.0 ** -20
    dims = (4, 4, 4, 4)
    V = orc.volume(dims)
    u = orc.su3_to_2v(orc.random_init_gauge(V, 4395))
    g = orc.su3_to_2v(orc.random_init_trafo(V, 4396))
    d = np.abs(np.sqrt(np_star_theta(p21_round(u), g, dims)) - np.sqrt(np_star_theta(u, g, dims)))
    assert 1e-3 * 2.0 ** -21 < d.max() < 83.2 * 2.0 ** -21


def fx_round(a):
    """The 32-bit fixed-point form of the mixed Landau run's polish
    restated in numpy: each real u as q = rint(u 2^31) clamped to
    +-(2^31 - 1), read back as q 2^-31."""
    return np.clip(np.rint(a * 2.0 ** 31), -2147483647, 2147483647) * 2.0 ** -31


def test_fx_round_bound(orc):
    """|fx(u) - u| <= 2^-32 for the reals of SU(3) matrices (2^-31 at |u| =
    1, where q clamps); the rounding is seen on random links."""
    u2 = orc.su3_to_2v(orc.random_init_gauge(4096, 4393))
    err = np.abs(fx_round(u2) - u2)
    assert err.max() <= 2.0 ** -32 and err.max() > 2.0 ** -36
    one = np.array([1.0, -1.0, 0.0, 0.5])
    assert np.abs(fx_round(one) - one).max() == 2.0 ** -31
    assert fx_round(np.array([0.5]))[0] == 0.5


def test_packing_bound_on_theta(orc):
    """The stop rule of runs on packed links takes the packing's worst
    effect on Delta as 83.2 e (e = 2^-38 for high + low parts, 2^-23 for the
    high parts alone): the max dAdA of the same g on packed-rounded and on
    exact links differ by less than that in sqrt, on random links and g."""
    dims = (4, 4, 4, 4)
    V = orc.volume(dims)
    for seed in (5101, 5102, 5103):
        u2 = orc.su3_to_2v(orc.random_init_gauge(V, seed))
        g2 = orc.su3_to_2v(orc.random_init_trafo(V, seed + 50))
        m = orc.functional(u2, g2, dims)[2]
        for lo, e in ((True, 2.0 ** -38), (False, 2.0 ** -23)):
            mp = orc.functional(packed_round(u2, lo), g2, dims)[2]
            d = abs(np.sqrt(mp) - np.sqrt(m))
            assert d < 83.2 * e, (lo, d / e)
            assert d > 1e-3 * e  # the rounding is seen


def test_packed_round_bound(orc):
    """|packed(u) - u| <= 2^-39 for SU(3) links (2^-38 at |u| = 1, where hi
    and lo both clamp), the high parts alone within 2^-24 (2^-23 at |u| = 1),
    and the low part never clamps elsewhere."""
    u2 = orc.su3_to_2v(orc.random_init_gauge(4096, 4391))
    hi = np.rint(u2 * 2.0 ** 23)
    assert np.abs(hi).max() <= 8388607
    assert np.abs(np.rint((u2 - hi * 2.0 ** -23) * 2.0 ** 38)).max() <= 2 ** 14
    err = np.abs(packed_round(u2) - u2)
    assert err.max() <= 2.0 ** -39 and err.max() > 2.0 ** -43
    err_hi = np.abs(packed_round(u2, lo=False) - u2)
    assert err_hi.max() <= 2.0 ** -24 and err_hi.max() > 2.0 ** -28
    # the unit link: 1 is not a 24-bit multiple of 2^-23, both parts clamp
    one = np.array([1.0, -1.0, 0.0])
    assert np.abs(packed_round(one) - one).max() == 2.0 ** -38
    assert np.abs(packed_round(one, lo=False) - one).max() == 2.0 ** -23


def np_short_theta(u2, g2, dims):
    """Per-site theta = |A(g(x) K(x))|_F^2 with the star K formed in the short
    form of the batch sweeps and of the FP32 sweeps (star_terms_short in the
    lattice header): each term U_mu(x) g^+(x+mu) and (g(x-mu) U_mu(x-mu))^+
    from rows 0, 1 of the product with row 2 rebuilt as conj(row0 x row1),
    link and g rows 0, 1 as stored (2-row layout), their row 2 rebuilt the
    same way.  numpy, independent of oracle.c."""
    LX, LY, LZ, LT = dims

    def full(a2):
        r = a2.reshape(-1, 2, 3, 2)
        m = r[..., 0] + 1j * r[..., 1]
        return np.concatenate([m, np.conj(np.cross(m[:, 0], m[:, 1]))[:, None]], axis=1)

    def rebuild(p01):
        return np.concatenate([p01, np.conj(np.cross(p01[..., 0, :], p01[..., 1, :]))[..., None, :]], axis=-2)

    U = full(u2).reshape(LT, LZ, LY, LX, 4, 3, 3)
    G = full(g2).reshape(LT, LZ, LY, LX, 3, 3)
    K = np.zeros(G.shape, dtype=complex)
    for mu in range(4):
        ax = 3 - mu
        Um = U[..., mu, :, :]
        gf = np.roll(G, -1, axis=ax)                  # g(x + mu)
        K += rebuild(Um[..., :2, :] @ np.conj(np.swapaxes(gf, -1, -2)))
        gb = np.roll(G, 1, axis=ax)                   # g(x - mu)
        Ub = np.roll(Um, 1, axis=ax)                  # U_mu(x - mu)
        K += np.conj(np.swapaxes(rebuild(gb[..., :2, :] @ Ub), -1, -2))
    M = G @ K
    A = (M - np.conj(np.swapaxes(M, -1, -2))) * (-0.5j)
    A = A - (np.trace(A, axis1=-2, axis2=-1) / 3.0)[..., None, None] * np.eye(3)
    return (np.abs(A) ** 2).sum(axis=(-2, -1)).reshape(-1)


def test_short_star_packing_bound(orc):
    """The stop rule of batch lattices on packed links takes the packing's
    worst effect on Delta with the short star as 139.2 e: the short-form
    theta on packed-rounded links and the exact theta of the same g differ
    by less than that in sqrt at every site, on random links and g
    (e = 2^-38 for high + low parts; 2^-23 for the high parts alone).  The
    short form on the exact links equals the reference form to rounding."""
    dims = (4, 4, 4, 6)
    V = orc.volume(dims)
    for seed in (5201, 5202):
        u = orc.random_init_gauge(V, seed)
        g = orc.random_init_trafo(V, seed + 50)
        u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
        exact = np_site_theta(orc.g_trafo(u, g, dims), dims).reshape(-1)
        sh = np_short_theta(u2, g2, dims)
        assert np.abs(np.sqrt(sh) - np.sqrt(exact)).max() < 1e-12
        assert abs(sh.max() - orc.functional(u2, g2, dims)[2]) < 1e-12 * sh.max()
        for lo, e in ((True, 2.0 ** -38), (False, 2.0 ** -23)):
            shp = np_short_theta(packed_round(u2, lo), g2, dims)
            d = np.abs(np.sqrt(shp) - np.sqrt(exact)).max()
            assert d < 139.2 * e, (lo, d / e)
            assert d > 1e-2 * e  # the rounding is seen


def test_fixed_point_bound_on_theta(orc):
    """The stop rule of Coulomb slices polished on 32-bit fixed-point links
    and g takes the rounding's worst effect on Delta as 600 e (e = 2^-32): the
    check computes Delta = -A(G K) from the rounded rows of g (row 2 rebuilt)
    and the rounded links (reference-order star), while the slice's result
    is g' = normalize(G) on the exact links.  Per site, sqrt(theta) of the
    two differ by less than 600 e on random links and g (both with 3 and 4
    directions).  The bound: links 83.2 e (as for the packed links), plus
    |normalize(G) - G|_F <= 31.2 e entering G K once through G and once
    through the 8 star terms (16 x 31.2 e)."""
    dims = (4, 4, 4, 6)
    V = orc.volume(dims)
    e = 2.0 ** -32
    for seed in (5301, 5302):
        u = orc.random_init_gauge(V, seed)
        g = orc.random_init_trafo(V, seed + 50)
        u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
        ux, gx = fx_round(u2), fx_round(g2)
        # g' = normalize(G) with G = rounded rows, row 2 rebuilt
        gn = orc.from_2v(gx).reshape(-1, 18)
        for k in range(V):
            orc.lib.orc_su3_normalize(gn[k])
        gn = gn.ravel()
        for ndir in (4, 3):
            # the star form equals the per-link form on exact fields
            ex = np_site_theta(orc.g_trafo(u, g, dims), dims, ndir=ndir).reshape(-1)
            assert np.abs(np.sqrt(np_star_theta(u2, g2, dims, ndir)) - np.sqrt(ex)).max() < 1e-12
            # the check's Delta: reference-order star on the rounded fields
            th_c = np_star_theta(ux, gx, dims, ndir)
            th_t = np_site_theta(orc.g_trafo(u, gn, dims), dims, ndir=ndir).reshape(-1)
            d = np.abs(np.sqrt(th_c) - np.sqrt(th_t)).max()
            assert d < 600 * e, (ndir, d / e)
            assert d > 1e-2 * e


def np_star_theta(u2, g2, dims, ndir=4):
    """Per-site |A(G(x) K(x))|_F^2 with the star K = sum_mu U_mu(x) G^+(x+mu)
    + U_mu^+(x-mu) G^+(x-mu) formed from full matrices whose row 2 is
    rebuilt from the stored rows 0, 1 (the reference-order star of the
    sweeps and their fused checks); numpy."""
    LX, LY, LZ, LT = dims

    def full(a2):
        r = a2.reshape(-1, 2, 3, 2)
        m = r[..., 0] + 1j * r[..., 1]
        return np.concatenate([m, np.conj(np.cross(m[:, 0], m[:, 1]))[:, None]], axis=1)

    U = full(u2).reshape(LT, LZ, LY, LX, 4, 3, 3)
    G = full(g2).reshape(LT, LZ, LY, LX, 3, 3)
    dag = lambda m: np.conj(np.swapaxes(m, -1, -2))
    K = np.zeros(G.shape, dtype=complex)
    for mu in range(ndir):
        ax = 3 - mu
        Um = U[..., mu, :, :]
        K += Um @ dag(np.roll(G, -1, axis=ax))
        K += dag(np.roll(Um, 1, axis=ax)) @ dag(np.roll(G, 1, axis=ax))
    M = G @ K
    A = (M - dag(M)) * (-0.5j)
    A = A - (np.trace(A, axis1=-2, axis2=-1) / 3.0)[..., None, None] * np.eye(3)
    return (np.abs(A) ** 2).sum(axis=(-2, -1)).reshape(-1)
