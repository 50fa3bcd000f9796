"""End-to-end tests of the cudagaugefix program (cudagaugefix.cu:599-828 as a
drop-in): thermalize a configuration into an ILDG file, gauge fix it from an
input file in the reference grammar, and check the files it writes."""
import os
import subprocess

import numpy as np
import pytest

from test_gpu_or import EPS12
from test_gpu_stoch import synthetic_params

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cudagfx_amd", "bin", "cudagaugefix")


def _run(args, cwd):
    r = subprocess.run([EXE] + args, cwd=cwd, capture_output=True, text=True, timeout=300)
    return r.returncode, r.stdout + r.stderr


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)


@pytest.fixture(scope="module")
def thermalized(tmp_path_factory):
    d = tmp_path_factory.mktemp("driver")
    # the twister table the program reads from its working directory
    synthetic_params(2024).astype("<u4").tofile(str(d / "MersenneTwister.dat"))
    _write(str(d / "therm.in"), "T=8\nL=8\n\nRandSeed = 351\nMTSeed = 4357\n\n"
           "BeginTHERMInit\n  Nsweep = 80\n  beta = 6.0\n  ckinterval = 20\n  start = cold\n"
           "EndTHERMInit\n")
    rc, log = _run(["-i", "therm.in", "-f", "conf.ildg"], str(d))
    assert rc == 0, log[-2000:]
    assert os.path.exists(str(d / "conf.ildg"))
    return d


def test_driver_landau(gfx, orc, thermalized):
    d = thermalized
    _write(str(d / "or.in"), "T=8\nL=8\n\nRandSeed = 351\n\nBeginORXInit\n  eps = 1.0e-12\n"
           "  ckinterval = 10\n  maxit = 20000\n  omega = 1.7\nEndORXInit\n")
    rc, log = _run(["-i", "or.in", "-f", "conf.ildg"], str(d))
    assert rc == 0, log[-2000:]
    assert "CONVERGENCE!" in log and "Final HOST values:" in log
    dims = (8, 8, 8, 8)
    gfx.set_lattice(*dims)
    rc_u, u = gfx.read_gf_ildg(str(d / "conf.ildg"))
    rc_f, fixed = gfx.read_gf_ildg(str(d / "landau_conf.ildg"))
    assert rc_u == 0 and rc_f == 0
    _, mx = orc.dAdA(fixed, dims)
    assert mx < 2 * EPS12
    # gauge invariance of the plaquette, unitarity of the fixed links
    assert abs(orc.mean_plaq(fixed, dims) - orc.mean_plaq(u, dims)) < 1e-12
    m = fixed.reshape(-1, 3, 3, 2)
    m = m[..., 0] + 1j * m[..., 1]
    assert np.abs(m @ np.conj(np.swapaxes(m, -1, -2)) - np.eye(3)).max() < 1e-12
    # a second run refuses to overwrite the output (exit -1 as the reference)
    rc2, _ = _run(["-i", "or.in", "-f", "conf.ildg"], str(d))
    assert rc2 == 255


def test_driver_coulomb_mixed(gfx, orc, thermalized):
    d = thermalized
    _write(str(d / "coul.in"), "T=8\nL=8\n\nRandSeed = 352\n\nBeginORXInit\n  eps = 1.0e-12\n"
           "  ckinterval = 5\n  maxit = 20000\n  omega = 1.7\n  precision = mixed\n"
           "  gauge = coulomb\nEndORXInit\n")
    rc, log = _run(["-i", "coul.in", "-f", "conf.ildg"], str(d))
    assert rc == 0, log[-2000:]
    dims = (8, 8, 8, 8)
    gfx.set_lattice(*dims)
    rc_f, fixed = gfx.read_gf_ildg(str(d / "coulomb_conf.ildg"))
    assert rc_f == 0
    unit = np.zeros(18 * orc.volume(dims))
    unit.reshape(-1, 18)[:, [0, 8, 16]] = 1.0
    _, _, M = orc.functional_slices(orc.su3_to_2v(fixed), orc.su3_to_2v(unit), dims, 3)
    assert np.all(M < 2 * EPS12)
