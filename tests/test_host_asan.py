"""The host parsers of untrusted input -- the input grammar (read_input) and
the ILDG/LIME reader (read_gf_ildg, read_ildg_dims) -- under AddressSanitizer
and UndefinedBehaviorSanitizer, on the CPU.  `make -C cudagfx_amd asan`
builds the host services alone (no HIP) with a small driver,
cudagfx_amd/host/gfx_host_check.c; a sanitizer finding aborts the driver
with a non-zero exit code, so every case asserts a clean exit.  Inputs: the
golden input texts of the reference grammar, valid ILDG files, and truncated
or corrupt files, among them LIME headers whose 64-bit lengths point far past
the end of the file (gauge_io.c:245-443 is the reader these replace)."""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "cudagfx_amd", "bin", "gfx_host_check_asan")
DIMS = ("4", "4", "2", "6")


@pytest.fixture(scope="module")
def exe():
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "cudagfx_amd"), "asan"],
                       capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        pytest.skip("sanitizer build not available: " + r.stderr[-500:])
    return EXE


def run(exe, *args, cwd=None):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True,
                       errors="replace", timeout=30, cwd=cwd, env=env)
    assert r.returncode == 0, (args, r.stderr[-3000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-3000:]
    out = {}
    for line in r.stdout.splitlines():
        for tok in line.split():
            if "=" in tok:
                k, v = tok.split("=", 1)
                out[k] = v
    assert "rc" in out, r.stdout[-2000:]
    return out, r.stdout


def lime(typ, body, length=None):
    n = len(body) if length is None else length
    h = struct.pack(">IHHQ", 0x456789AB, 1, 0xC000, n) + typ.encode().ljust(128, b"\0")
    return h + body + b"\0" * ((8 - len(body) % 8) % 8)


def fmt_xml(prec=64, dims=(4, 4, 2, 6)):
    return ("<?xml version=\"1.0\"?><ildgFormat><precision> %d </precision><lx> %d </lx>"
            "<ly> %d </ly><lz> %d </lz><lt> %d </lt></ildgFormat>" % ((prec,) + dims)).encode()


def test_input_files(exe, tmp_path):
    cases = json.load(open(os.path.join(GOLD, "read_input.json")))
    for name, case in sorted(cases.items()):
        p = tmp_path / (name + ".in")
        text = case["text"].encode()
        p.write_bytes(text)
        out, _ = run(exe, "input", p)
        want = case["result"]
        if want["exit"] == 0:
            assert int(out["rc"]) == 0 and int(out["T"]) == want["T"] and int(out["L"]) == want["L"]
        else:
            assert int(out["rc"]) == 1
        # every prefix length of the text, and garbage spliced in
        for cut in range(0, len(text), max(1, len(text) // 25)):
            p.write_bytes(text[:cut])
            run(exe, "input", p)
        rng = np.random.default_rng(len(text))
        for _ in range(5):
            b = bytearray(text)
            for k in rng.integers(0, len(b), 6):
                b[k] = int(rng.integers(1, 256))
            p.write_bytes(bytes(b))
            run(exe, "input", p)
    # long lines, no newline at the end, NUL bytes, a huge number, a file name list
    for text in (b"T=" + b"9" * 5000, b"L = 8\nT = 4", b"T=4\0L=4\n", b"#" * 100000,
                 b"BeginORXInit\n gaugefiles = " + b"a" * 5000 + b" b" * 200 + b"\nEndORXInit\n",
                 b"BeginORXInit\n eps = 1." + b"0" * 4000 + b"\nEndORXInit\n"):
        p = tmp_path / "x.in"
        p.write_bytes(text)
        run(exe, "input", p)
    run(exe, "input", tmp_path / "missing.in")


@pytest.mark.parametrize("prec", [64, 32])
def test_valid_ildg_files(exe, tmp_path, prec):
    path = tmp_path / "c.ildg"
    out, _ = run(exe, "write", path, *DIMS, prec, 7)
    assert int(out["rc"]) == 0
    out, _ = run(exe, "read", path, *DIMS)
    assert int(out["rc"]) == 0
    data = open(path, "rb").read()
    # the binary record is the last one: its data are the last 4V*18 reals
    n = 4 * 4 * 4 * 2 * 6 * 18
    w = 8 if prec == 64 else 4
    body = data[len(data) - n * w - (8 - (n * w) % 8) % 8:][: n * w]
    vals = np.frombuffer(body, dtype=">f8" if prec == 64 else ">f4").astype(np.float64)
    assert abs(float(out["sum"]) - vals.sum()) < 1e-9
    out, _ = run(exe, "dims", path)
    assert int(out["rc"]) == 0
    # truncated at many points (headers, metadata, data)
    cuts = sorted(set(list(range(0, 800, 13)) + list(range(800, len(data), max(1, len(data) // 40)))))
    for cut in cuts:
        p = tmp_path / "t.ildg"
        p.write_bytes(data[:cut])
        out, _ = run(exe, "read", p, *DIMS)
        assert int(out["rc"]) != 0
        run(exe, "dims", p)
    # random corruption of the record headers and metadata
    rng = np.random.default_rng(prec)
    for _ in range(60):
        b = bytearray(data)
        for k in rng.integers(0, 700, 3):
            b[k] = int(rng.integers(0, 256))
        p = tmp_path / "r.ildg"
        p.write_bytes(bytes(b))
        run(exe, "read", p, *DIMS)
        run(exe, "dims", p)


def test_crafted_lime_lengths(exe, tmp_path):
    """Lengths that wrap calloc(len + 1) or that would seek back onto the
    same header: each read ends with an error status, no overflow, no loop."""
    n = 4 * 4 * 4 * 2 * 6 * 18
    data = lime("ildg-binary-data", b"\0" * (8 * n))
    cases = {
        # calloc(len + 1) wraps to 0, then fread(len) would write the heap
        "wrap": lime("ildg-format", fmt_xml()[:64], length=2 ** 64 - 1) + data,
        "xlf_wrap": lime("ildg-format", fmt_xml()) + lime("xlf-info", b"x", length=2 ** 64 - 1) + data,
        # padded length 2^64 - 144: seeking by it goes back onto this header
        "loop": lime("ildg-format", fmt_xml()) + lime("other", b"", length=2 ** 64 - 144) + data,
        "loop_first": lime("other", b"", length=2 ** 64 - 144) + lime("ildg-format", fmt_xml()) + data,
        "neg_seek": lime("other", b"", length=2 ** 63 + 8) + data,
        # a metadata record larger than the cap, and one longer than the file
        "huge_meta": lime("ildg-format", b"<lx>" + b" " * (2 ** 20 + 8) + b"</lx>") + data,
        "past_end": lime("ildg-format", fmt_xml(), length=10 ** 6) + data,
        "data_past_end": lime("ildg-format", fmt_xml()) + lime("ildg-binary-data", b"\0" * 64, length=8 * n),
        "no_format": data,
        "bad_magic": b"\x12\x34\x56\x78" + lime("ildg-format", fmt_xml())[4:] + data,
    }
    for name, blob in cases.items():
        p = tmp_path / (name + ".ildg")
        p.write_bytes(blob)
        out, _ = run(exe, "read", p, *DIMS)
        assert int(out["rc"]) != 0, name
        out, _ = run(exe, "dims", p)
        if name in ("wrap", "huge_meta", "past_end", "loop_first", "neg_seek", "no_format", "bad_magic"):
            assert int(out["rc"]) == -1, name
    # a well-formed file built the same way reads back (the crafted cases
    # fail for their lengths, not for the layout)
    p = tmp_path / "good.ildg"
    p.write_bytes(lime("ildg-format", fmt_xml()) + lime("xlf-info", b"plaquette = 0.5 beta = 6.0") + data)
    out, _ = run(exe, "read", p, *DIMS)
    assert int(out["rc"]) == 0 and float(out["sum"]) == 0.0
    out, _ = run(exe, "dims", p)
    assert int(out["rc"]) == 0 and out["dims"] == "4"


def test_host_observables_under_sanitizers(exe, orc):
    """g_trafo, F, dAdA and the plaquette run clean and give the oracle's
    host values (which are pinned bit for bit to the reference's)."""
    dims = (4, 4, 4, 4)
    out, _ = run(exe, "observe", *dims, 5)
    V = orc.volume(dims)
    orc.lib.orc_lehmer_plant_seeds(5)
    u = np.zeros(18 * 4 * V)
    orc.lib.orc_random_init_gauge(u, V)
    g = np.zeros(18 * V)
    orc.lib.orc_random_init_trafo(g, V)
    ut = orc.g_trafo(u, g, dims)
    d, m = orc.dAdA(ut, dims)
    assert abs(float(out["F"]) - orc.gauge_functional(ut, dims)) < 1e-14
    assert abs(float(out["dAdA"]) - d) < 1e-13 * d and abs(float(out["max"]) - m) < 1e-13 * m
    assert abs(float(out["plaq"]) - orc.mean_plaq(ut, dims)) < 1e-14
