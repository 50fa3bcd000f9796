"""Host services of libcudagfx.so (no GPU needed): the input grammar, ILDG
I/O, the Lehmer RNG, random initialisation and host observables, each
against the reference's own host C (golden vectors) or the file format."""
import json
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_lehmer_and_random_init_bitexact(gfx):
    gold = json.load(open(os.path.join(GOLD, "lehmer.json")))
    lib = gfx.load()
    lib.PlantSeeds(gold["seed"])
    assert [lib.Random() for _ in range(20)] == gold["draws"]
    gd = np.load(os.path.join(GOLD, "ref_fields.npz"))
    gfx.set_lattice(*gd["dims"])
    gfx.plant_seeds(351)
    assert np.array_equal(gfx.random_init_trafo_host(), gd["trafo"])
    gfx.plant_seeds(4357)
    assert np.array_equal(gfx.random_init_gauge_host(), gd["gauge"])


def test_host_observables_bitexact(gfx):
    gd = np.load(os.path.join(GOLD, "ref_fields.npz"))
    gfx.set_lattice(*gd["dims"])
    u, g = gd["gauge"], gd["trafo"]
    ut = gfx.host_g_trafo(u, g)
    assert np.array_equal(ut, gd["gauge_transformed"])
    for tag, a in (("u", u), ("ut", ut)):
        assert gfx.host_gauge_functional(a) == gd[tag + "_F"]
        d, m = gfx.host_dAdA(a)
        assert d == gd[tag + "_dAdA"] and m == gd[tag + "_maxdAdA"]
        assert gfx.host_mean_plaq(a) == gd[tag + "_plaq"]


_PARSE_CHILD = r"""
import sys, json
sys.path.insert(0, sys.argv[1])
import cudagfx_amd as g, ctypes as C
lib = g.load()
rc = g.read_input(sys.argv[2])
gi = lambda n: C.c_int.in_dll(lib, n).value
gd = lambda n: C.c_double.in_dll(lib, n).value
sa, th = g.saparam(), g.thermparam()
print("JSON" + json.dumps(dict(rc=rc, T=gi("T"), L=gi("L"), LX=gi("LX"), VOLUME=gi("VOLUME"),
      randseed=gi("randseed"), orxeps=gd("orxeps"), orxmaxit=gi("orxmaxit"),
      orxcheckinterval=gi("orxcheckinterval"), orxflag=gi("orxflag"), saflag=gi("saflag"),
      thermflag=gi("thermflag"), sa=[sa.Tmin, sa.Tmax, sa.expo, sa.N, sa.checkint],
      therm=[th.Nsweep, th.Nsave, th.beta, th.checkint, th.startcond])))
"""

_CASES = json.load(open(os.path.join(GOLD, "read_input.json")))


@pytest.mark.parametrize("name", sorted(_CASES))
def test_read_input_matches_reference(tmp_path, name):
    """read_input.l behaviour, one process per input as the grammar keeps state."""
    case = _CASES[name]
    path = tmp_path / (name + ".in")
    path.write_text(case["text"])
    r = subprocess.run([sys.executable, "-c", _PARSE_CHILD, ROOT, str(path)],
                       capture_output=True, text=True, timeout=120)
    got = None
    for line in r.stdout.splitlines():
        if line.startswith("JSON"):
            got = json.loads(line[4:])
    want = case["result"]
    assert got is not None, r.stderr
    if want["exit"] != 0:
        # the reference exits(1) on a parsing error; the library returns 1
        assert got["rc"] == 1
        return
    for k in ("T", "L", "LX", "VOLUME", "randseed", "orxeps", "orxmaxit", "orxcheckinterval",
              "orxflag", "saflag", "thermflag", "sa", "therm"):
        assert got[k] == want[k], k


def _lime_records(path):
    recs = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        magic, ver, flags, n = struct.unpack(">IHHQ", data[pos:pos + 16])
        typ = data[pos + 16:pos + 144].split(b"\0")[0].decode()
        assert magic == 0x456789AB and ver == 1
        body = data[pos + 144:pos + 144 + n]
        recs.append((typ, flags, body))
        pos += 144 + n + (8 - n % 8) % 8
    assert pos == len(data)
    return recs


@pytest.mark.parametrize("prec", [64, 32])
def test_ildg_roundtrip_and_layout(gfx, tmp_path, prec):
    dims = (4, 4, 2, 6)
    gfx.set_lattice(*dims)
    gfx.plant_seeds(7)
    u = gfx.random_init_gauge_host()
    lib = gfx.load()
    C = __import__("ctypes")
    C.c_double.in_dll(lib, "PLAQ").value = 0.5937
    C.c_double.in_dll(lib, "BETA").value = 6.0
    path = tmp_path / "conf.ildg"
    assert gfx.write_gf_ildg(u, path, prec) == 0
    recs = _lime_records(path)
    assert [r[0] for r in recs] == ["ildg-format", "xlf-info", "ildg-binary-data"]
    assert all(r[1] == 0xC000 for r in recs)  # MB and ME on every record
    xml = recs[0][2].decode()
    for tag, v in (("precision", prec), ("lx", 4), ("ly", 4), ("lz", 2), ("lt", 6)):
        assert f"<{tag}> {v} </{tag}>" in xml
    assert "plaquette = 0.593700" in recs[1][2].decode()
    raw = np.frombuffer(recs[2][2], dtype=">f8" if prec == 64 else ">f4")
    want = u if prec == 64 else u.astype(np.float32).astype(np.float64)
    assert np.array_equal(raw.astype(np.float64), want)  # big-endian, s-major, mu-minor
    rc, back = gfx.read_gf_ildg(path)
    assert rc == 0 and np.array_equal(back, want)
    assert abs(C.c_double.in_dll(lib, "BETA").value - 6.0) < 1e-12


def test_ildg_errors(gfx, tmp_path):
    dims = (4, 4, 2, 6)
    gfx.set_lattice(*dims)
    u = gfx.unit_gauge()
    path = tmp_path / "c.ildg"
    assert gfx.write_gf_ildg(u, path, 64) == 0
    gfx.set_lattice(4, 4, 2, 8)  # dimension mismatch (gauge_io.c:314-342)
    assert gfx.read_gf_ildg(path)[0] == 120
    gfx.set_lattice(*dims)
    assert gfx.read_gf_ildg(tmp_path / "missing.ildg")[0] == -1
    assert gfx.write_gf_ildg(u, tmp_path / "x.ildg", 16) == 500
    # precision entry that does not match the data block (gauge_io.c:425-428)
    data = bytearray(open(path, "rb").read())
    i = data.find(b"<precision> 64 </precision>")
    data[i:i + 27] = b"<precision> 32 </precision>"
    bad = tmp_path / "bad.ildg"
    bad.write_bytes(bytes(data))
    assert gfx.read_gf_ildg(bad)[0] == 500


def test_sa_schedule(gfx):
    """set_sa_temperature (simulated_annealing.cu:265-299)."""
    N, Tmin, Tmax = 11, 0.01, 0.45
    lin = [gfx.sa_temperature(i, Tmin, Tmax, -1.0, N) for i in range(N)]
    assert abs(lin[0] - Tmax) < 1e-15 and abs(lin[-1] - Tmin) < 1e-15
    assert np.allclose(np.diff(lin), (Tmin - Tmax) / (N - 1))
    geo = [gfx.sa_temperature(i, Tmin, Tmax, 0.0, N) for i in range(N)]
    assert np.allclose(geo, Tmax * (Tmin / Tmax) ** (np.arange(N) / (N - 1)))
    pw = [gfx.sa_temperature(i, Tmin, Tmax, 2.0, N) for i in range(N)]
    assert abs(pw[0] - Tmax) < 1e-12 and abs(pw[-1] - Tmin) < 1e-12
    assert gfx.sa_temperature(3, 0.2, 0.2, -1.0, N) == 0.2


def test_ildg_dims_without_reading_the_field(gfx, tmp_path):
    """read_ildg_dims takes the extents from the ildg-format record, whatever
    LX..T are set to (the batch driver reads files of different sizes)."""
    gfx.set_lattice(6, 4, 2, 8)
    path = tmp_path / "d.ildg"
    assert gfx.write_gf_ildg(gfx.unit_gauge(), path, 32) == 0
    gfx.set_lattice(4, 4, 4, 4)
    assert gfx.ildg_dims(path) == (6, 4, 2, 8)
    assert gfx.ildg_dims(tmp_path / "missing.ildg") is None
    (tmp_path / "junk.ildg").write_bytes(b"\0" * 300)
    assert gfx.ildg_dims(tmp_path / "junk.ildg") is None


_BATCH_KEYS_CHILD = r"""
import sys, json
sys.path.insert(0, sys.argv[1])
import cudagfx_amd as g, ctypes as C
lib = g.load()
rc = g.read_input(sys.argv[2])
n = C.c_int.in_dll(lib, "norxfiles").value
files = (C.c_char * 1024 * 64).in_dll(lib, "orxfiles")
print("JSON" + json.dumps(dict(rc=rc, copies=C.c_int.in_dll(lib, "orxcopies").value,
      files=[bytes(files[k]).split(b"\0")[0].decode() for k in range(n)],
      orxmaxit=C.c_int.in_dll(lib, "orxmaxit").value, orxflag=C.c_int.in_dll(lib, "orxflag").value)))
"""


@pytest.mark.parametrize("text,want", [
    ("BeginORXInit\n  maxit = 100\n  copies = 8\n  gaugefiles = a.ildg b.ildg,c.ildg  # three\nEndORXInit\n",
     dict(rc=0, copies=8, files=["a.ildg", "b.ildg", "c.ildg"], orxmaxit=100, orxflag=1)),
    ("BeginORXInit\n  gaugefiles = one.ildg\n  maxit = 7\nEndORXInit\n",
     dict(rc=0, copies=1, files=["one.ildg"], orxmaxit=7, orxflag=1)),
    ("BeginORXInit\n  gaugefiles =\nEndORXInit\n", dict(rc=1)),
    ("BeginORXInit\n  copies = x\nEndORXInit\n", dict(rc=1)),
])
def test_batch_keys_of_the_grammar(tmp_path, text, want):
    """The batch keys of BeginORXInit: copies = <int>, gaugefiles = names to
    the end of the line (blanks or commas between, '#' ends the list); an
    empty list or a non-integer count is a parsing error."""
    path = tmp_path / "batch.in"
    path.write_text("T=8\nL=8\n\n" + text)
    r = subprocess.run([sys.executable, "-c", _BATCH_KEYS_CHILD, ROOT, str(path)],
                       capture_output=True, text=True, timeout=120)
    got = next(json.loads(l[4:]) for l in r.stdout.splitlines() if l.startswith("JSON"))
    assert got["rc"] == want["rc"]
    if want["rc"] == 0:
        assert {k: got[k] for k in want} == want
