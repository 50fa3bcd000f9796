"""Writes tests/golden/mt_dat_words.json: twister output words of the
oracle's MT607 (MersenneTwister.cu:142-199 semantics) for the reference's own
parameter table MersenneTwister.dat, read here from the reference checkout.

Twisters 0..7 and 4088..4095 are seeded as seedMTGPU does
(MersenneTwister.cu:98-114: srand(seed), then one rand() per twister in
order), for srand(4357) and srand(351).  Each generates 40 numbers from its
seed (initialized = 0), then 19 more after a restart (initialized = 1, the
19 stored words reloaded with the buffer index back at 0).  The fixture keeps
the 16 parameter records used (3 words each) as the inputs of the known
answers, the 16 seeds, and the expected words; the table itself is not
copied.

Run from the repository root:  python3 tests/golden/make_mt_golden.py
"""
import ctypes as C
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

DAT = "/root/reference/MersenneTwister.dat"
IDX = list(range(8)) + list(range(4088, 4096))
NFIRST, NRESTART = 40, 19


def words_for(records, seeds):
    n = len(records)
    params = (oracle.MTParams * n)()
    for i, ((a, b, c), s) in enumerate(zip(records, seeds)):
        params[i].matrix_a, params[i].mask_b, params[i].mask_c, params[i].seed = a, b, c, s
    state = np.zeros(19 * n, dtype=np.uint32)
    _, w1 = oracle.mt_generate(params, state, NFIRST, 0, want_words=True)
    _, w2 = oracle.mt_generate(params, state, NRESTART, 1, want_words=True)
    return w1.reshape(NFIRST, n).T.tolist(), w2.reshape(NRESTART, n).T.tolist()


def glibc_seeds(seed, idx):
    libc = C.CDLL("libc.so.6")
    libc.srand(seed)
    r = [libc.rand() & 0xFFFFFFFF for _ in range(4096)]
    return [r[i] for i in idx]


def main():
    raw = open(DAT, "rb").read()
    p = np.frombuffer(raw, dtype="<u4").reshape(4096, 4)
    records = [[int(v) for v in p[i, :3]] for i in IDX]
    out = {"source": "MersenneTwister.dat of the reference (sha256 %s), records %s" %
                     (hashlib.sha256(raw).hexdigest(), IDX),
           "indices": IDX, "records": records, "n_first": NFIRST, "n_restart": NRESTART,
           "seeds": {}}
    for s in (4357, 351):
        seeds = glibc_seeds(s, IDX)
        w1, w2 = words_for(records, seeds)
        out["seeds"][str(s)] = {"twister_seeds": seeds, "words": w1, "words_after_restart": w2}
    json.dump(out, open(os.path.join(ROOT, "tests", "golden", "mt_dat_words.json"), "w"))


if __name__ == "__main__":
    main()
