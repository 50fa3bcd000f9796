"""End-to-end tests of the cudagaugefix program (cudagaugefix.cu:599-828 as a
drop-in): thermalize a configuration into an ILDG file, gauge fix it from an
input file in the reference grammar, and check the files it writes."""
import os
import shutil
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


def _host_check(gfx, orc, path, dims):
    gfx.set_lattice(*dims)
    rc, fixed = gfx.read_gf_ildg(str(path))
    assert rc == 0
    _, mx = orc.dAdA(fixed, dims)
    return fixed, mx


def test_driver_gribov_copies(gfx, orc, thermalized):
    """copies = 4 with a BeginSAInit block: every start is annealed and fixed
    in one batch, the copy with the largest F is written."""
    d = thermalized
    shutil.copy(str(d / "conf.ildg"), str(d / "gribov.ildg"))
    _write(str(d / "gribov.in"), "T=8\nL=8\n\nRandSeed = 353\nMTSeed = 4359\n\n"
           "BeginSAInit\n  Tmin = 0.01\n  Tmax = 0.45\n  Nstep = 30\n  expo = -1.0\n  micro = 2\n"
           "EndSAInit\n\nBeginORXInit\n  eps = 1.0e-12\n  ckinterval = 10\n  maxit = 20000\n"
           "  omega = 1.7\n  precision = mixed\n  copies = 4\nEndORXInit\n")
    rc, log = _run(["-i", "gribov.in", "-f", "gribov.ildg"], str(d))
    assert rc == 0, log[-3000:]
    Fs = [float(l.split("F = ")[1].split()[0]) for l in log.splitlines()
          if l.startswith("gribov.ildg copy ")]
    assert len(Fs) == 4
    best = [l for l in log.splitlines() if "best copy" in l]
    assert len(best) == 1 and float(best[0].split("F = ")[1]) == max(Fs)
    fixed, mx = _host_check(gfx, orc, d / "landau_gribov.ildg", (8, 8, 8, 8))
    assert mx < 2 * EPS12
    assert abs(orc.gauge_functional(fixed, (8, 8, 8, 8)) - max(Fs)) < 1e-10


def test_driver_uneven_gaugefiles(gfx, orc, thermalized):
    """-f with a gaugefiles list of another size: both configurations are
    fixed in one batch (extents from each file), each written; with one start
    in FP64 the 8^4 result equals the single run's (same start, same stop)."""
    d = thermalized
    _write(str(d / "therm4.in"), "T=4\nL=8\n\nRandSeed = 361\nMTSeed = 4361\n\n"
           "BeginTHERMInit\n  Nsweep = 60\n  beta = 6.0\n  ckinterval = 20\n  start = cold\n"
           "EndTHERMInit\n")
    rc, log = _run(["-i", "therm4.in", "-f", "small.ildg"], str(d))
    assert rc == 0, log[-2000:]
    shutil.copy(str(d / "conf.ildg"), str(d / "listed.ildg"))
    _write(str(d / "list.in"), "T=8\nL=8\n\nRandSeed = 351\n\nBeginORXInit\n  eps = 1.0e-12\n"
           "  ckinterval = 10\n  maxit = 20000\n  omega = 1.7\n"
           "  gaugefiles = small.ildg  # another size\nEndORXInit\n")
    rc, log = _run(["-i", "list.in", "-f", "listed.ildg"], str(d))
    assert rc == 0, log[-3000:]
    big, mx_big = _host_check(gfx, orc, d / "landau_listed.ildg", (8, 8, 8, 8))
    small, mx_small = _host_check(gfx, orc, d / "landau_small.ildg", (8, 8, 8, 4))
    assert mx_big < 2 * EPS12 and mx_small < 2 * EPS12
    # the single run of the same file and start (test_driver_landau's input)
    single_path = d / "landau_conf.ildg"
    if not single_path.exists():
        _write(str(d / "or.in"), "T=8\nL=8\n\nRandSeed = 351\n\nBeginORXInit\n  eps = 1.0e-12\n"
               "  ckinterval = 10\n  maxit = 20000\n  omega = 1.7\nEndORXInit\n")
        assert _run(["-i", "or.in", "-f", "conf.ildg"], str(d))[0] == 0
    single, _ = _host_check(gfx, orc, single_path, (8, 8, 8, 8))
    assert np.abs(big - single).max() < 1e-10
    # an existing output stops the run before any fixing (exit -1)
    assert _run(["-i", "list.in", "-f", "listed.ildg"], str(d))[0] == 255
