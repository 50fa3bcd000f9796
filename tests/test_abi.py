"""The C-ABI library loads and exports every function and global declared in
include/*.h (no compute calls: these run without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    funcs, globs = set(), set()
    for h in ("cudagfx.h", "cudagfx_host.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w]+\s*\*?\s*(\w+)\s*\(", text, flags=re.M):
            funcs.add(m.group(1))
        for m in re.finditer(r"^extern\s+[\w ]+?\s+([\w ,]+);", text, flags=re.M):
            for name in m.group(1).split(","):
                globs.add(name.strip())
    funcs.discard("defined")
    return funcs, globs


def test_headers_declare_the_reference_entry_points():
    funcs, globs = _declared()
    for name in ("init_gaugefixing", "finalize_gaugefixing", "init_thermalization",
                 "finalize_thermalization", "overrelax_gauge", "calc_functional",
                 "calc_plaquette", "simannealing_gauge", "thermalize_gauge", "init_MT",
                 "update_MT", "finalize_MT", "read_input", "read_gf_ildg", "write_gf_ildg",
                 "PlantSeeds", "random_init_trafo", "g_trafo", "dAdA", "mean_plaq"):
        assert name in funcs, name
    for name in ("T", "LX", "VOLUME", "FUNC", "DADA", "maxDADA", "orxeps", "saparam",
                 "thermparam"):
        assert name in globs, name


def test_library_exports_every_declared_symbol(gfx):
    lib = C.CDLL(gfx.LIB_PATH)
    funcs, globs = _declared()
    missing = [n for n in sorted(funcs | globs) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_covers_the_abi(gfx):
    funcs, _ = _declared()
    assert set(gfx.ABI_FUNCTIONS) <= funcs


def test_compute_without_init_is_an_error(gfx):
    """Calls before init_* return GFX_ERR_STATE instead of touching the device."""
    lib = gfx.load()
    lib.finalize_gaugefixing()
    assert lib.get_trafo(None) == 600
    assert lib.gfx_halfsweep(0, 4) == 600
    assert lib.overrelax_gauge(1, 1e-12, 1) == -1


def test_setting_validation(gfx):
    """Setters check their arguments on the host (no device work)."""
    lib = gfx.load()
    assert lib.gfx_set_packed_links(2) == 600 and lib.gfx_set_packed_links(-1) == 600
    assert lib.gfx_set_packed_links(0) == 0 and lib.gfx_set_packed_links(1) == 0
    assert lib.gfx_set_precision(7, 1e-6) == 600
    assert lib.gfx_set_fp32_stall(-1, 1e-8) == 600 and lib.gfx_set_fp32_stall(300, 0.0) == 600
    assert lib.gfx_set_fp32_stall(300, 1e-8) == 0


def test_geometry_is_checked_before_device_work(gfx):
    """Odd or missing extents and lattices above 2^29 sites (the kernels
    address a parity's plane elements with 32-bit byte offsets) are refused
    with GFX_ERR_GEOMETRY before any device call, by every entry point that
    takes a lattice."""
    lib = gfx.load()
    keep = gfx.lattice()
    try:
        for dims in [(5, 4, 4, 4), (4, 4, 4, 3), (0, 4, 4, 4), (1024, 1024, 32, 32)]:
            gfx.set_lattice(*dims)
            assert lib.init_gaugefixing(None, None) == 100, dims
            assert lib.init_thermalization(None) == 100, dims
            assert lib.gfx_batch_add(*dims, None, None) == -100, dims
    finally:
        gfx.set_lattice(*keep)


def test_driver_binary_built():
    assert os.access(os.path.join(ROOT, "cudagfx_amd", "bin", "cudagaugefix"), os.X_OK)


def test_product_does_not_import_the_oracle():
    """The shipped package and C sources never reference the oracle."""
    for d in ("cudagfx_amd",):
        for dirpath, _, files in os.walk(os.path.join(ROOT, d)):
            for f in files:
                if f.endswith((".py", ".c", ".h", ".hip", "Makefile")):
                    text = open(os.path.join(dirpath, f), errors="ignore").read()
                    assert "import oracle" not in text and "liboracle" not in text, f
