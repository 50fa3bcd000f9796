"""FP64 parity on a heatbath-thermalized lattice: an 8^4 configuration after
200 SU(3) heatbath sweeps at beta = 6.0 from a cold start (twisters seeded
with srand(4357)), initial g from random_init_trafo with RandSeed 351, fixed
to Landau gauge with checkinterval 1 to max dAdA < (float)1e-12 at
omega = 1.68 (the reference's value, overrelax.cu:416) and 1.7.  The GPU
path and the oracle (overrelax.cu:677-823 restated) must stop within 2
iterations of each other, give fixed links (observables.c:17-47) within
1e-10 max-abs and F within 1e-12."""
import numpy as np
import pytest

from test_gpu_or import EPS12, _setup
from test_gpu_stoch import synthetic_params

pytestmark = pytest.mark.gpu

DIMS = (8, 8, 8, 8)


@pytest.fixture(scope="module")
def thermalized(gfx, orc):
    gfx.set_lattice(*DIMS)
    u0 = gfx.unit_gauge()
    gfx.init_thermalization(u0)
    gfx.set_mt_params(synthetic_params(4357).ravel())
    gfx.set_mt_seed(4357)
    gfx.heatbath_sweeps(200, 6.0)
    plaq = gfx.calc_plaquette(0)
    u = gfx.get_gauge()
    gfx.finalize_MT()
    gfx.finalize_thermalization()
    print(f"8^4 beta=6.0 after 200 heatbath sweeps: plaquette {plaq:.6f}", flush=True)
    assert 0.55 < plaq < 0.65
    assert abs(orc.mean_plaq(u, DIMS) - plaq) < 1e-13
    return u


@pytest.mark.parametrize("omega", [1.68, 1.7])
def test_thermalized_8x4_matches_oracle(gfx, orc, thermalized, omega):
    u = thermalized
    V = orc.volume(DIMS)
    g = orc.random_init_trafo(V, 351)
    _setup(gfx, DIMS, u, g, omega)
    it_gpu = gfx.overrelax_gauge(20000, EPS12, 1)
    F_gpu, _, M_gpu = gfx.results()
    g_gpu = gfx.get_trafo()
    fixed_gpu = gfx.apply_trafo(u)
    gfx.finalize_gaugefixing()
    u2, g2 = orc.su3_to_2v(u), orc.su3_to_2v(g)
    it_cpu, hist = orc.overrelax_gauge(u2, g2, DIMS, 20000, EPS12, 1, omega)
    g_cpu = orc.from_2v(g2)
    fixed_cpu = orc.g_trafo(u, g_cpu, DIMS)
    print(f"omega {omega}: stop gpu {it_gpu} oracle {it_cpu}, F gpu {F_gpu:.15f} oracle "
          f"{hist[-1, 0]:.15f}, max|g| diff {np.abs(g_gpu - g_cpu).max():.3e}, fixed links diff "
          f"{np.abs(fixed_gpu - fixed_cpu).max():.3e}", flush=True)
    assert it_cpu > 10 and it_gpu > 10
    assert abs(it_gpu - it_cpu) <= 2
    assert M_gpu < EPS12 and hist[-1, 2] < EPS12
    assert np.abs(g_gpu - g_cpu).max() < 1e-10
    assert np.abs(fixed_gpu - fixed_cpu).max() < 1e-10
    assert abs(F_gpu - hist[-1, 0]) < 1e-12
    # the fixed field satisfies the gauge condition by the host functional
    d, m = orc.dAdA(fixed_cpu, DIMS)
    assert m < EPS12 * (1 + 1e-6)
