"""Pins the CPU oracle (oracle/) before anything is checked against it.

Each check is chosen so that a plausible mistake in the oracle (a wrong
constant, a dropped payload check, a comparator that reads the payload, a
sign error in the totalOrder transform) fails at least one of them; none of
them re-types the oracle's own formula.
"""
import itertools

import numpy as np
import pytest


def test_splitmix64_published_vectors(orc):
    # Vigna's splitmix64.c: first output for state 0, and the first three
    # outputs of a generator seeded 1234567 (state advances by the golden gamma).
    L = orc.lib()
    assert L.orc_splitmix64(0) == 0xE220A8397B1DCDAF
    gamma = 0x9E3779B97F4A7C15
    s = 1234567
    outs = []
    for _ in range(3):
        outs.append(L.orc_splitmix64(s))
        s = (s + gamma) % 2**64
    assert outs == [6457827717110365317, 3203168211198807973, 9817491932198370423]


def test_mixbits_inverts_with_published_inverse(orc):
    # lowbias32 (bench.cu:152-155) and Wellons' published inverse lowbias32_r;
    # a wrong multiplier or shift breaks the round trip.
    def inv(x):
        M = 0xFFFFFFFF
        x ^= x >> 16
        x = (x * 0x43021123) & M
        x ^= (x >> 15) ^ (x >> 30)
        x = (x * 0x1D69E2A5) & M
        x ^= x >> 16
        return x
    L = orc.lib()
    rng = np.random.default_rng(1)
    for x in [0, 1, 2, 0xFFFFFFFF, 0x80000000, *rng.integers(0, 2**32, 200).tolist()]:
        assert inv(L.orc_mixbits(int(x))) == int(x)
    assert L.orc_mixbits(0) == 0


@pytest.mark.parametrize("n", [1, 2, 7, 32, 129])
def test_structured_distributions_closed_forms(orc, n):
    arrays = 3
    c = n * arrays
    loc = np.arange(c) % n
    assert (orc.generate(orc.I32, orc.D_SORTED, n, c) == loc).all()
    assert (orc.generate(orc.I32, orc.D_REVERSED, n, c) == n - 1 - loc).all()
    assert (orc.generate(orc.I32, orc.D_CONSTANT, n, c) == 99).all()
    assert (orc.generate(orc.I32, orc.D_BINARY, n, c) == loc % 2).all()
    assert (orc.generate(orc.I32, orc.D_ORGAN_PIPE, n, c) == np.minimum(loc, n - 1 - loc)).all()
    alt = orc.generate(orc.I32, orc.D_ALT_EXTREMES, n, c)
    assert (alt == np.where(loc % 2 == 1, 2**31 - 1, -2**31)).all()
    ext = orc.generate(orc.I32, orc.D_EXTREMES6, n, c)
    assert (ext == np.array([2**31 - 1, -2**31, 0, 2**31 - 1, -1, 1])[loc % 6]).all()
    per = orc.generate(orc.I32, orc.D_PERTURBED, n, c).astype(np.int64)
    assert (np.abs(per - 4 * loc) <= 8).all()
    d16 = orc.generate(orc.I32, orc.D_DISTINCT16, n, 4096)
    assert len(set(d16.tolist())) <= 16
    rec = orc.generate(orc.REC_I32_I32, orc.D_UNIFORM, n, c)
    assert (rec["payload"] == np.arange(c)).all()              # bench.cu:189


def test_uniform_and_dup50_statistics(orc):
    u = orc.generate(orc.I32, orc.D_UNIFORM, 1024, 1 << 20, seed=42)
    assert u.min() < -2**31 + 2**24 and u.max() > 2**31 - 2**24
    assert abs(float(u.astype(np.float64).mean())) < 2**31 * 0.01
    d = orc.generate(orc.I32, orc.D_DUP50, 1000, 1000, seed=7)
    # second half only repeats first-half keys
    assert set(d[500:].tolist()) <= set(d[:500].tolist())
    f = orc.generate(orc.F32, orc.D_NORMAL_SPECIAL, 1 << 20, 1 << 20).view(np.float32)
    finite = f[np.isfinite(f)]
    assert abs(finite.mean()) < 0.01 and abs(finite.std() - 1.0) < 0.01
    special = ~np.isfinite(f) | ((f == 0) & np.signbit(f))
    assert 0.005 < special.mean() < 0.012


@pytest.mark.parametrize("dtype", ["I32", "U64", "F32", "REC_I32_I32", "REC_I64_I32", "F64"])
def test_sort_agrees_with_numpy(orc, dtype):
    dt = getattr(orc, dtype)
    for dist in orc.NINE + [orc.D_EXTREMES6, orc.D_RANDOM1M]:
        n = 129
        data = orc.generate(dt, dist, n, n * 40, seed=dist + 3)
        got = orc.sort_batched(dt, data, n)
        keys = got["key"] if dt in (orc.REC_I32_I32, orc.REC_I64_I32) else got
        want = orc.sort_keys_numpy(dt, data, n)
        assert (keys.reshape(-1) == want).all(), orc.DIST_NAMES[dist]
        assert orc.validate_batched(dt, data, got, n)[0] == 0


def test_sort_brute_force_tiny(orc):
    # every permutation of a small multiset: sorted output is the unique
    # non-decreasing arrangement
    for ms in ([3, 1, 2, 1], [5, -2**31, 2**31 - 1, 0], [7, 7, 7, 1]):
        want = None
        for perm in itertools.permutations(ms):
            a = np.array(perm, dtype=np.int32)
            s = orc.sort_batched(orc.I32, a, a.size)
            if want is None:
                want = s
            assert (s == want).all()
            assert all(s[i] <= s[i + 1] for i in range(s.size - 1))


def test_total_order_special_values(orc):
    L = orc.lib()
    order = [0xFFC00001, 0xFFC00000, 0xFF800000, 0xBF800000, 0x80000001, 0x80000000,
             0x00000000, 0x00000001, 0x3F800000, 0x7F800000, 0x7F800001, 0x7FC00000]
    keys = [L.orc_f32_to_ordered(b) for b in order]
    assert keys == sorted(keys) and len(set(keys)) == len(keys)
    for b in order:
        assert L.orc_ordered_to_f32(L.orc_f32_to_ordered(b)) == b
    data = np.array(order[::-1] * 3, dtype=np.uint32)
    got = orc.sort_batched(orc.F32, data, data.size)
    assert [int(x) for x in got[::3]] == order


def test_total_order_special_values_f64(orc):
    # IEEE-754 totalOrder: -NaN (larger payload first) < -inf < negatives <
    # -0 < +0 < positives < +inf < +NaN (smaller payload first)
    order = [0xFFFFFFFFFFFFFFFF, 0xFFF8000000000000, 0xFFF0000000000001, 0xFFF0000000000000,
             0xBFF0000000000000, 0x8000000000000001, 0x8000000000000000, 0x0000000000000000,
             0x0000000000000001, 0x3FF0000000000000, 0x7FEFFFFFFFFFFFFF, 0x7FF0000000000000,
             0x7FF0000000000001, 0x7FF8000000000000]
    data = np.array(order[::-1] * 3, dtype=np.uint64)
    got = orc.sort_batched(orc.F64, data, data.size)
    assert [int(x) for x in got[::3]] == order
    # sorted N(0,1)+specials: non-NaN values ascend as doubles, NaNs sit at the ends by sign
    n = 1 << 12
    d = orc.generate(orc.F64, orc.D_NORMAL_SPECIAL, n, n * 8, seed=11)
    s = orc.sort_batched(orc.F64, d, n).reshape(-1, n)
    for row in s:
        v = row.view(np.float64)
        nan = np.isnan(v)
        neg_nan = nan & np.signbit(v)
        k = int(neg_nan.sum())
        assert neg_nan[:k].all()
        body = v[k:k + int((~nan).sum())]
        assert not np.isnan(body).any() and (body[1:] >= body[:-1]).all()
        assert nan[k + body.size:].all() and not np.signbit(v[k + body.size:]).any()
        zero_sign = np.signbit(body[body == 0]).astype(np.int8)
        assert (np.diff(zero_sign) <= 0).all()   # -0 before +0
    assert orc.validate_batched(orc.F64, d, s.reshape(-1), n)[0] == 0


def test_validator_catches_planted_mistakes(orc):
    n = 64
    data = orc.generate(orc.REC_I32_I32, orc.D_DISTINCT16, n, n * 4, seed=5)
    good = orc.sort_batched(orc.REC_I32_I32, data, n)
    assert orc.validate_batched(orc.REC_I32_I32, data, good, n)[0] == 0

    def bad(mutate):
        b = good.copy()
        mutate(b)
        return orc.validate_batched(orc.REC_I32_I32, data, b, n)[0]

    # key order broken
    def swap(b):
        b[3], b[40] = b[40].copy(), b[3].copy()
    assert bad(swap) != 0
    # record duplicated (same key twice, one payload lost)
    i = next(j for j in range(n - 1) if good["key"][j] == good["key"][j + 1])
    assert bad(lambda b: b.__setitem__(i + 1, b[i])) != 0
    # payload points at an input record with another payload
    assert bad(lambda b: b.__setitem__((0,), (b["key"][0], (b["payload"][0] + 1) % (4 * n)))) != 0
    # record moved across arrays (two equal keys exchanged between arrays)
    keys0 = {int(k): j for j, k in enumerate(good["key"][:n])}
    cross = [(j, keys0[int(k)]) for j, k in enumerate(good["key"][n:2 * n], start=n) if int(k) in keys0]
    assert cross
    a, b_ = cross[0]

    def xchg(b):
        b[a], b[b_] = b[b_].copy(), b[a].copy()
    assert bad(xchg) != 0
    # scalar keys: a single changed key
    d = orc.generate(orc.I32, orc.D_UNIFORM, n, n * 2)
    s = orc.sort_batched(orc.I32, d, n)
    s[5] += 1
    assert orc.validate_batched(orc.I32, d, s, n)[0] != 0


def test_segmented_and_large(orc):
    rng = np.random.default_rng(3)
    lens = rng.integers(0, 50, 200)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    data = orc.generate(orc.I32, orc.D_UNIFORM, int(off[-1]), int(off[-1]), seed=9)
    got = orc.sort_segmented(orc.I32, data, off)
    for s in range(200):
        assert (got[off[s]:off[s + 1]] == np.sort(data[off[s]:off[s + 1]])).all()
    assert orc.validate_segmented(orc.I32, data, got, off)[0] == 0
    big = orc.generate(orc.U64, orc.D_PRESORTED_RUNS, 1 << 20, 1 << 20)
    assert (orc.sort_large(orc.U64, big) == np.sort(big)).all()
