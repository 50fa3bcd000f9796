"""Writes the golden vectors under tests/golden/ from the reference's own host
C routines (rngs.c, su3manip.c, observables.c, read_input.c), compiled from
the reference checkout into oracle/_ref/libcudagfx_ref.so by
``make -C oracle ref``.  Run in the build container:

    python tests/golden/make_golden.py

Outputs (small, committed):
  lehmer.json          PlantSeeds(351) stream-0 draws; TestRandom constants
  ref_fields.npz       random_init_trafo(seed 351), random_init_gauge(seed
                       4357) on 4x4x4x2; g_trafo of that pair; gauge_functional,
                       dAdA, maxDADA, mean_plaq of the gauge and of the
                       transformed field
  read_input.json      parsed globals (or the exit status) for each input text
                       in INPUTS below, one reference process per input
"""
import ctypes as C
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "libcudagfx_ref.so")
DIMS = (4, 4, 4, 2)

# input texts for the grammar (read_input.l); the first reproduces the
# structure of the reference's example input
INPUTS = {
    "example": ("# input file for GPU gauge fixing\n\n\n#T = 16 \n#L = 8\n\nT=48\nL=24\n\n"
                "RandSeed = 351\n\n\n#BeginTHERMInit \n#  Nsweep = 20000\n#EndTHERMInit\n\n\n"
                "BeginORXInit \n  eps = 1.0e-15\n  ckinterval = 50\n  maxit = 20000\nEndORXInit\n\n\n"
                "BeginSAInit \n  Tmin = 1.0e-4\n  Tmax = 0.5\n  ckinterval = 100\n  Nstep = 5000\n"
                "  expo = -1.0\nEndSAInit\n\n\n"),
    "therm": ("T = 8\nL = 8\nRandSeed = 4357\nBeginTHERMInit\n  Nsweep = 200\n  Nsave = 10\n"
              "  beta = 6.0\n  ckinterval = 1\n  start = cold\nEndTHERMInit\n"),
    "hot": ("T=4\nL=4\nBeginTHERMInit\n  Nsweep = 3\n  beta = 5.7\n  start = hot\nEndTHERMInit\n"),
    "eps12": ("T = 32\nL = 32\nRandSeed = 1\nBeginORXInit\n  eps = 1.0e-12\n  maxit = 5000\n"
              "  ckinterval = 1\nEndORXInit\n"),
    "comments": ("T = 16 # time extent\nL = 8\n# full line\nBeginORXInit # block\n"
                 "  eps = 0.5e-10 # inline\n  maxit = 7\nEndORXInit\n"),
    "two_on_line": ("T=4\nL=4\nBeginORXInit\n  eps = 1.0e-8  maxit = 3\nEndORXInit\n"),
    "expo_geom": ("T=4\nL=4\nBeginSAInit\n  Tmin = 0.01\n  Tmax = 0.45\n  Nstep = 4000\n"
                  "  expo = 0.0\n  ckinterval = 10\nEndSAInit\n"),
    "err_noexp_dot": ("T=4\nL=4\nBeginORXInit\n  eps = 1e-12\nEndORXInit\n"),
    "err_trailing_space": ("T=4 \nL=4\n"),
    "err_unknown_key": ("T=4\nL=4\nfoo = 3\n"),
    "err_in_block": ("T=4\nL=4\nBeginORXInit\n  omega = x\nEndORXInit\n"),
}

_CHILD = r"""
import ctypes as C, json, sys
ref = C.CDLL(sys.argv[1])
class SA(C.Structure):
    _fields_ = [("Tmin", C.c_double), ("Tmax", C.c_double), ("expo", C.c_double), ("N", C.c_int), ("checkint", C.c_int)]
class TH(C.Structure):
    _fields_ = [("Nsweep", C.c_int), ("Nsave", C.c_int), ("beta", C.c_double), ("checkint", C.c_int), ("startcond", C.c_int)]
ref.read_input.argtypes = [C.c_char_p]
rc = ref.read_input(sys.argv[2].encode())
gi = lambda n: C.c_int.in_dll(ref, n).value
gd = lambda n: C.c_double.in_dll(ref, n).value
sa = SA.in_dll(ref, "saparam"); th = TH.in_dll(ref, "thermparam")
out = dict(rc=rc, T=gi("T"), L=gi("L"), LX=gi("LX"), VOLUME=gi("VOLUME"), randseed=gi("randseed"),
           orxeps=gd("orxeps"), orxmaxit=gi("orxmaxit"), orxcheckinterval=gi("orxcheckinterval"),
           orxflag=gi("orxflag"), saflag=gi("saflag"), thermflag=gi("thermflag"),
           sa=[sa.Tmin, sa.Tmax, sa.expo, sa.N, sa.checkint],
           therm=[th.Nsweep, th.Nsave, th.beta, th.checkint, th.startcond])
sys.stdout.flush()
print("JSON" + json.dumps(out))
"""


def lehmer(ref):
    ref.PlantSeeds.argtypes = [C.c_long]
    ref.Random.restype = C.c_double
    ref.GetSeed.argtypes = [C.POINTER(C.c_long)]
    ref.SelectStream.argtypes = [C.c_int]
    ref.PutSeed.argtypes = [C.c_long]
    ref.PlantSeeds(351)
    draws = [ref.Random() for _ in range(20)]
    # TestRandom (rngs.c:153-179)
    ref.SelectStream(0)
    ref.PutSeed(1)
    for _ in range(10000):
        ref.Random()
    x = C.c_long()
    ref.GetSeed(C.byref(x))
    check = x.value
    ref.SelectStream(1)
    ref.PlantSeeds(1)
    ref.GetSeed(C.byref(x))
    stream1 = x.value
    ref.SelectStream(0)
    return dict(seed=351, draws=draws, check_after_10000=check, stream1_after_plant1=stream1)


def fields(ref):
    LX, LY, LZ, LT = DIMS
    V = LX * LY * LZ * LT
    for n, v in zip(["LX", "LY", "LZ", "T", "VOLUME"], [LX, LY, LZ, LT, V]):
        C.c_int.in_dll(ref, n).value = v
    sys.path.insert(0, ROOT)
    import oracle  # neighbour table only (cudagaugefix.cu:72-118 layout)

    nn = oracle.neighbours(DIMS)
    C.c_void_p.in_dll(ref, "nn").value = nn.ctypes.data
    vp = C.c_void_p
    ref.random_init_trafo.argtypes = [vp]
    ref.random_init_gauge.argtypes = [vp]
    ref.g_trafo.argtypes = [vp, vp]
    for f in ("gauge_functional", "dAdA", "mean_plaq"):
        getattr(ref, f).restype = C.c_double
        getattr(ref, f).argtypes = [vp]
    ref.PlantSeeds(351)
    g = np.zeros(18 * V)
    ref.random_init_trafo(g.ctypes.data_as(vp))
    ref.PlantSeeds(4357)
    u = np.zeros(18 * 4 * V)
    ref.random_init_gauge(u.ctypes.data_as(vp))
    obs = {}

    def observe(tag, a):
        obs[tag + "_F"] = ref.gauge_functional(a.ctypes.data_as(vp))
        obs[tag + "_dAdA"] = ref.dAdA(a.ctypes.data_as(vp))
        obs[tag + "_maxdAdA"] = C.c_double.in_dll(ref, "maxDADA").value
        obs[tag + "_plaq"] = ref.mean_plaq(a.ctypes.data_as(vp))

    observe("u", u)
    ut = u.copy()
    ref.g_trafo(ut.ctypes.data_as(vp), g.ctypes.data_as(vp))
    observe("ut", ut)
    return g, u, ut, obs


def read_inputs():
    import tempfile

    out = {}
    with tempfile.TemporaryDirectory() as d:
        for name, text in INPUTS.items():
            path = os.path.join(d, name + ".in")
            with open(path, "w") as f:
                f.write(text)
            r = subprocess.run([sys.executable, "-c", _CHILD, REF, path], capture_output=True,
                               text=True)
            rec = {"exit": r.returncode}
            for line in r.stdout.splitlines():
                if line.startswith("JSON"):
                    rec.update(json.loads(line[4:]))
            out[name] = {"text": text, "result": rec}
    return out


def main():
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/libcudagfx_ref.so missing: run make -C oracle ref")
    ref = C.CDLL(REF)
    with open(os.path.join(HERE, "lehmer.json"), "w") as f:
        json.dump(lehmer(ref), f, indent=1)
    g, u, ut, obs = fields(ref)
    np.savez_compressed(os.path.join(HERE, "ref_fields.npz"), dims=np.array(DIMS), trafo=g,
                        gauge=u, gauge_transformed=ut, **{k: np.float64(v) for k, v in obs.items()})
    with open(os.path.join(HERE, "read_input.json"), "w") as f:
        json.dump(read_inputs(), f, indent=1)
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
