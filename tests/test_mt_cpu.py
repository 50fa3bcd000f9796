"""Pins of the oracle's Mersenne Twister (MersenneTwister.cu:142-199 semantics)."""
import os

import numpy as np
import pytest

DAT = "/root/reference/MersenneTwister.dat"


def test_generic_engine_known_answer(orc):
    """The engine (in-place circular buffer) with the MT19937 parameters: the
    10000th output for the default seed 5489 is 4123659995 (C++11
    std::mt19937 requirement)."""
    assert orc.lib.orc_mt19937_nth(5489, 10000) == 4123659995
    assert orc.lib.orc_mt19937_nth(5489, 1) == 3499211612


def _mt607_reference_form(a, b, c, seed, n):
    """x_{k+19} = x_{k+9} ^ ((x_k & ~1 | x_{k+1} & 1) >> 1) ^ (lsb ? a : 0) on
    a growing list (no circular buffer), then tempering: an independent
    formulation of the published recurrence."""
    x = [seed & 0xFFFFFFFF]
    for j in range(1, 19):
        x.append((1812433253 * (x[-1] ^ (x[-1] >> 30)) + j) & 0xFFFFFFFF)
    out = []
    for k in range(n):
        y = (x[k] & 0xFFFFFFFE) | (x[k + 1] & 1)
        v = x[k + 9] ^ (y >> 1) ^ (a if y & 1 else 0)
        x.append(v)
        t = v
        t ^= t >> 12
        t ^= (t << 7) & b
        t ^= (t << 15) & c
        t ^= t >> 18
        out.append(t & 0xFFFFFFFF)
    return out, x


def test_mt607_recurrence_and_output_transform(orc):
    rng = np.random.default_rng(5)
    count, nper = 8, 45
    params = (orc.MTParams * count)()
    raw = []
    for i in range(count):
        a = int(rng.integers(0, 2 ** 32)) | 0x80000000
        b, c, s = (int(v) for v in rng.integers(0, 2 ** 32, 3))
        params[i].matrix_a, params[i].mask_b, params[i].mask_c, params[i].seed = a, b, c, s
        raw.append((a, b, c, s))
    state = np.zeros(count * 19, dtype=np.uint32)
    f, words = orc.mt_generate(params, state, nper, 0, want_words=True)
    for i, (a, b, c, s) in enumerate(raw):
        want, _ = _mt607_reference_form(a, b, c, s, nper)
        got = words.reshape(nper, count)[:, i]
        assert list(got) == want
        # ((float)x + 1.0f) / 2^32 in float arithmetic, lane-strided layout
        wf = ((np.array(want, dtype=np.uint32).astype(np.float32) + np.float32(1.0))
              / np.float32(4294967296.0))
        assert np.array_equal(f.reshape(nper, count)[:, i], wf)
        # after the call params.seed = state[0] (MersenneTwister.cu:196)
        assert params[i].seed == state[19 * i]


def test_restart_semantics(orc):
    """A second call with initialized=1 reloads the 19 words as stored and
    restarts at buffer index 0 (MersenneTwister.cu:164-168): for a first call
    of 19*k numbers this is exactly the continuation of the sequence."""
    a, b, c, s = 0x9908B0DF, 0x9D2C5680, 0xEFC60000, 12345
    for first in (19 * 2, 23):
        params = (orc.MTParams * 1)()
        params[0].matrix_a, params[0].mask_b, params[0].mask_c, params[0].seed = a, b, c, s
        state = np.zeros(19, dtype=np.uint32)
        _, w1 = orc.mt_generate(params, state, first, 0, want_words=True)
        _, w2 = orc.mt_generate(params, state, 10, 1, want_words=True)
        full, _ = _mt607_reference_form(a, b, c, s, first + 10)
        assert list(w1) == full[:first]
        if first % 19 == 0:
            assert list(w2) == full[first:]
        else:
            assert list(w2) != full[first:]  # the restart is not a continuation


@pytest.mark.skipif(not os.path.exists(DAT), reason="reference twister table not present")
def test_reference_parameter_table_drives_the_same_recurrence(orc):
    """The reference's parameter records (4096 x {matrix_a, mask_b, mask_c,
    seed}) through both formulations."""
    p = np.fromfile(DAT, dtype="<u4").reshape(4096, 4)
    assert len(np.unique(p[:, 0])) == 4096 and np.all(p[:, 3] == 0)
    params = (orc.MTParams * 16)()
    for i in range(16):
        params[i].matrix_a, params[i].mask_b, params[i].mask_c = (int(v) for v in p[i, :3])
        params[i].seed = 4357 + i
    state = np.zeros(16 * 19, dtype=np.uint32)
    _, words = orc.mt_generate(params, state, 40, 0, want_words=True)
    for i in range(16):
        want, _ = _mt607_reference_form(*(int(v) for v in p[i, :3]), 4357 + i, 40)
        assert list(words.reshape(40, 16)[:, i]) == want


def test_real_table_words_fixture(orc):
    """Known answers for the reference's own parameter table: the oracle's
    words for 16 of its records (0..7, 4088..4095), seeded as seedMTGPU
    seeds them (srand(4357) and srand(351), one rand() per twister,
    MersenneTwister.cu:98-114), 40 numbers and then 19 after a restart.
    The fixture (tests/golden/make_mt_golden.py) holds the records it used;
    where the reference checkout is present they are checked against the
    table as well."""
    import ctypes as C
    import json
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "mt_dat_words.json")))
    idx, records = fx["indices"], fx["records"]
    libc = C.CDLL("libc.so.6")
    for seed, e in fx["seeds"].items():
        libc.srand(int(seed))
        r = [libc.rand() & 0xFFFFFFFF for _ in range(4096)]
        assert [r[i] for i in idx] == e["twister_seeds"]
        n = len(idx)
        params = (orc.MTParams * n)()
        for k, ((a, b, c), s) in enumerate(zip(records, e["twister_seeds"])):
            params[k].matrix_a, params[k].mask_b, params[k].mask_c, params[k].seed = a, b, c, s
        state = np.zeros(19 * n, dtype=np.uint32)
        _, w1 = orc.mt_generate(params, state, fx["n_first"], 0, want_words=True)
        _, w2 = orc.mt_generate(params, state, fx["n_restart"], 1, want_words=True)
        assert w1.reshape(fx["n_first"], n).T.tolist() == e["words"]
        assert w2.reshape(fx["n_restart"], n).T.tolist() == e["words_after_restart"]
        for k in range(n):
            want, _ = _mt607_reference_form(*records[k], e["twister_seeds"][k], fx["n_first"])
            assert e["words"][k] == want
    if os.path.exists(DAT):
        p = np.fromfile(DAT, dtype="<u4").reshape(4096, 4)
        assert [[int(v) for v in p[i, :3]] for i in idx] == records
