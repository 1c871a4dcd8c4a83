"""The committed golden fixtures (tests/golden/) against the CPU oracle and
against what the mathematics fixes, without a GPU.

The fixtures were written by tests/golden/make_golden.py from the oracle;
this test pins them: the oracle must still regenerate every input bit for bit
and sort it to the committed output, numpy's sort must agree, and the cases
with closed-form answers must have them."""
import hashlib
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SMALL_N = [1, 2, 31, 32, 33, 63, 64, 65, 129]
ARRAYS = 2


@pytest.fixture(scope="module")
def small():
    with np.load(os.path.join(GOLDEN, "small.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def hashes():
    with open(os.path.join(GOLDEN, "hashes.json")) as f:
        return json.load(f)["sha256"]


def keys_of(a):
    return a["key"] if a.dtype.names else a


@pytest.mark.parametrize("dtname", ["I32", "F32", "REC_I32_I32"])
def test_small_fixtures_match_oracle(orc, small, dtname):
    dt = getattr(orc, dtname)
    for dist in orc.NINE:
        for n in SMALL_N:
            k = f"{dtname}_{orc.DIST_NAMES[dist]}_{n}"
            inp, want = small[k + "_in"], small[k + "_out"]
            assert inp.tobytes() == orc.generate(dt, dist, n, n * ARRAYS, 42).tobytes(), k
            got = keys_of(orc.sort_batched(dt, inp, n))
            assert got.tobytes() == want.tobytes(), k


def test_small_fixtures_match_numpy(orc, small):
    for dist in orc.NINE:
        for n in SMALL_N:
            k = f"I32_{orc.DIST_NAMES[dist]}_{n}"
            ref = np.sort(small[k + "_in"].reshape(ARRAYS, n), axis=1).ravel()
            assert ref.tobytes() == small[k + "_out"].tobytes(), k
            # fp32 keys ordered by IEEE totalOrder on their bit patterns
            kf = f"F32_{orc.DIST_NAMES[dist]}_{n}"
            bits = small[kf + "_in"].astype(np.uint32)
            u = bits ^ np.where(bits >> 31, np.uint32(0xFFFFFFFF), np.uint32(0x80000000))
            u = np.sort(u.reshape(ARRAYS, n), axis=1).ravel()
            back = u ^ np.where(u >> 31, np.uint32(0x80000000), np.uint32(0xFFFFFFFF))
            assert back.astype(np.uint32).tobytes() == small[kf + "_out"].tobytes(), kf


def test_small_fixtures_closed_forms(orc, small):
    for n in SMALL_N:
        for dtname in ("I32", "REC_I32_I32"):
            srt = small[f"{dtname}_SORTED_{n}_out"].reshape(ARRAYS, n)
            rev = small[f"{dtname}_REVERSED_{n}_out"].reshape(ARRAYS, n)
            # SORTED and REVERSED hold each of 0..n-1 once per array
            assert (srt == np.arange(n)).all() and (rev == np.arange(n)).all()
            assert (small[f"{dtname}_CONSTANT_{n}_out"] == 99).all()
            alt = small[f"{dtname}_ALT_EXTREMES_{n}_out"].reshape(ARRAYS, n)
            lo = (n + 1) // 2   # alternating INT_MIN / INT_MAX, starting with INT_MIN
            assert (alt[:, :lo] == np.iinfo(np.int32).min).all() and (alt[:, lo:] == np.iinfo(np.int32).max).all()
        # records keep their keys: the key multiset of every array is unchanged
        rin = small[f"REC_I32_I32_DUP50_{n}_in"]["key"].reshape(ARRAYS, n)
        assert (np.sort(rin, axis=1).ravel() == small[f"REC_I32_I32_DUP50_{n}_out"]).all()


def test_hashes_match_oracle(orc, hashes):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    assert set(hashes) == {c[0] for c in mg.HASH_CASES}
    for name, dtname, dist, n, arrays in mg.HASH_CASES:
        dt = getattr(orc, dtname)
        _, out = mg.sorted_case(dt, dist, n, arrays)
        assert mg.digest(dt, out) == hashes[name], name
    # SORTED and REVERSED inputs sort to the same sequence (0..n-1 per array)
    assert hashes["batched I32 SORTED 2x1000"] == hashes["batched I32 REVERSED 2x1000"]
    want = hashlib.sha256(np.tile(np.arange(1000, dtype="<i4"), 2).tobytes()).hexdigest()
    assert hashes["batched I32 SORTED 2x1000"] == want
