"""GPU parity of the structure-of-arrays pair sort (cs_batched_sort_pairs).

The pair sort is the record sort with keys and payloads in separate buffers,
so its output must be bit-identical to sorting the same records with
cs_batched_sort (CS_REC_I32_I32 / CS_REC_I64_I32) and splitting them; the
record result itself is checked against the oracle's validator (host
std::sort key sequence, payload membership / association / uniqueness,
bench.cu:209-228).  Inputs carry the payload = global index of the record
generator (bench.cu:189), with the tie-heavy distributions among them.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 3, 31, 32, 33, 100, 255, 256, 257, 1000, 1024, 2049, 4095, 4096, 5000, 8192]


def run_pairs(cs, orc, rt, key_dtype, dist, n, arrays, inplace=False, seed=7):
    rec = cs.REC_I32_I32 if key_dtype == cs.I32 else cs.REC_I64_I32
    count = n * arrays
    host = orc.generate(rec, dist, n, count, seed)
    keys = np.ascontiguousarray(host["key"])
    vals = np.ascontiguousarray(host["payload"]).astype(np.int32)
    # the record sort, split afterwards: the expected pairs
    rbuf = cs.from_numpy(host)
    cs.batched_sort(rec, rbuf.ptr, rbuf.ptr, arrays, n)
    rt.synchronize()
    want = cs.to_numpy(rbuf, rec, count)
    rc, msg = orc.validate_batched(rec, host, want, n)
    where = f"key_dtype={key_dtype} dist={orc.DIST_NAMES[dist]} n={n} arrays={arrays} inplace={inplace}"
    assert rc == 0, f"{where}: record sort: {msg}"
    kin, vin = cs.from_numpy(keys), cs.from_numpy(vals)
    if inplace:
        kout, vout = kin, vin
    else:
        kout, vout = cs.from_numpy(np.zeros_like(keys)), cs.from_numpy(np.zeros_like(vals))
    cs.batched_sort_pairs(key_dtype, kin.ptr, vin.ptr, kout.ptr, vout.ptr, arrays, n)
    rt.synchronize()
    got_k = kout.download(keys.dtype, count)
    got_v = vout.download(np.int32, count)
    assert got_k.tobytes() == np.ascontiguousarray(want["key"]).tobytes(), f"{where}: keys differ"
    assert got_v.tobytes() == np.ascontiguousarray(want["payload"]).astype(np.int32).tobytes(), \
        f"{where}: payloads differ"
    if not inplace:
        assert kin.download(keys.dtype, count).tobytes() == keys.tobytes(), f"{where}: input keys modified"


@pytest.mark.parametrize("key_dtype", ["I32", "I64"])
@pytest.mark.parametrize("n", SIZES)
def test_pairs_match_records(cs, orc, gpu, key_dtype, n):
    kd = getattr(cs, key_dtype)
    rec = cs.REC_I32_I32 if kd == cs.I32 else cs.REC_I64_I32
    if n > cs.batched_max_n(rec):
        pytest.skip("beyond capacity")
    arrays = max(3, min(2000, (1 << 17) // n))
    for dist in (orc.D_UNIFORM, orc.D_DUP50, orc.D_DISTINCT16, orc.D_CONSTANT, orc.D_REVERSED):
        run_pairs(cs, orc, gpu, kd, dist, n, arrays)


@pytest.mark.parametrize("key_dtype", ["I32", "I64"])
def test_pairs_in_place(cs, orc, gpu, key_dtype):
    kd = getattr(cs, key_dtype)
    for n in (33, 1000, 4096):
        run_pairs(cs, orc, gpu, kd, orc.D_DUP50, n, 257, inplace=True)


def pair_arrays(cs, orc, key_dtype, dist, count, seed):
    """Records from the oracle's generator (payload = global index) and the
    same pairs as two arrays."""
    rec = cs.REC_I32_I32 if key_dtype == cs.I32 else cs.REC_I64_I32
    host = orc.generate(rec, dist, max(count, 1), count, seed)
    return rec, host, np.ascontiguousarray(host["key"]), np.ascontiguousarray(host["payload"]).astype(np.int32)


def check_pairs_vs_records(cs, orc, rt, rec, host, keys, vals, got_k, got_v, want, off, where, same_ties=True):
    """The pairs, put back into records, pass the oracle's record checks; the
    keys equal the record sort's; with same_ties the payloads do too (a
    deterministic path: the partition scatter ranks equal keys with atomics,
    so its tie order differs from run to run, which the checks allow)."""
    rc, msg = orc.validate_segmented(rec, host, want, off)
    assert rc == 0, f"{where}: record sort: {msg}"
    got = np.zeros(got_k.size, dtype=host.dtype)
    got["key"] = got_k
    got["payload"] = got_v
    rc, msg = orc.validate_segmented(rec, host, got, off)
    assert rc == 0, f"{where}: pair sort: {msg}"
    assert got_k.tobytes() == np.ascontiguousarray(want["key"]).tobytes(), f"{where}: keys differ"
    if same_ties:
        assert got_v.tobytes() == np.ascontiguousarray(want["payload"]).astype(np.int32).tobytes(), \
            f"{where}: payloads differ"


@pytest.mark.parametrize("key_dtype", ["I32", "I64"])
def test_ragged_pairs_match_records(cs, orc, gpu, key_dtype):
    """cs_segmented_sort_pairs: ragged SoA pairs through a record-type plan,
    bit-identical to the plan's record sort of the same pairs, out of place
    and in place, with segments up to the plan's longest, empty and
    one-element ones, and a first offset past zero."""
    kd = getattr(cs, key_dtype)
    rng = np.random.default_rng(11)
    for dist, segs, max_len, first in ((orc.D_UNIFORM, 500, 4096, 0), (orc.D_DUP50, 300, 1500, 17),
                                       (orc.D_DISTINCT16, 1000, 300, 0), (orc.D_REVERSED, 64, 4096, 5)):
        lens = rng.integers(0, max_len + 1, segs)
        lens[:: 7] = 0
        lens[1:: 11] = 1
        off = np.concatenate([[first], first + np.cumsum(lens)]).astype(np.int64)
        count = int(off[-1])
        rec, host, keys, vals = pair_arrays(cs, orc, kd, dist, count, int(segs))
        plan = cs.SegPlan(rec, off)
        doff = cs.from_numpy(off)
        rbuf = cs.from_numpy(host)
        plan.sort(rbuf.ptr, rbuf.ptr, doff.ptr)
        gpu.synchronize()
        want = cs.to_numpy(rbuf, rec, count)
        where = f"key_dtype={key_dtype} dist={orc.DIST_NAMES[dist]} segs={segs} first={first}"
        for inplace in (False, True):
            kin, vin = cs.from_numpy(keys), cs.from_numpy(vals)
            kout, vout = (kin, vin) if inplace else (cs.from_numpy(keys.copy()), cs.from_numpy(vals.copy()))
            plan.sort_pairs(kin.ptr, vin.ptr, kout.ptr, vout.ptr, doff.ptr)
            gpu.synchronize()
            check_pairs_vs_records(cs, orc, gpu, rec, host, keys, vals, kout.download(keys.dtype, count),
                                   vout.download(np.int32, count), want, off, f"{where} inplace={inplace}")


@pytest.mark.parametrize("key_dtype", ["I32", "I64"])
def test_large_pairs_match_records(cs, orc, gpu, key_dtype):
    """cs_sort_pairs: one array of SoA pairs through the hybrid record sort
    (merge below the record type's threshold, partition above), with and
    without a workspace: the oracle's record checks on the pairs, keys equal
    to the hybrid sort of the records, and below the threshold (the merge
    path, deterministic) the payloads too."""
    kd = getattr(cs, key_dtype)
    rec = cs.REC_I32_I32 if kd == cs.I32 else cs.REC_I64_I32
    thr = cs.hybrid_threshold(rec)
    for n, dist in ((100003, orc.D_UNIFORM), (thr - 5, orc.D_DUP50), (thr + 77, orc.D_UNIFORM)):
        _, host, keys, vals = pair_arrays(cs, orc, kd, dist, n, n)
        rbuf = cs.from_numpy(host)
        cs.sort_hybrid(rec, rbuf.ptr, n)
        gpu.synchronize()
        want = cs.to_numpy(rbuf, rec, n)
        off = np.array([0, n], dtype=np.int64)
        for ws in (None, cs.Workspace(rec, n)):
            kbuf, vbuf = cs.from_numpy(keys), cs.from_numpy(vals)
            cs.sort_pairs(kd, kbuf.ptr, vbuf.ptr, n, ws)
            gpu.synchronize()
            check_pairs_vs_records(cs, orc, gpu, rec, host, keys, vals, kbuf.download(keys.dtype, n),
                                   vbuf.download(np.int32, n), want, off,
                                   f"key_dtype={key_dtype} n={n} dist={orc.DIST_NAMES[dist]} ws={ws is not None}",
                                   same_ties=n < thr)


def test_pairs_arguments(cs, gpu):
    buf = cs.device_buffer(cs.I32, 8)
    with pytest.raises(cs.CsError):
        cs.sort_pairs(cs.U64, buf.ptr, buf.ptr, 8)                   # keys must be int32 or int64
    plan = cs.SegPlan(cs.I32, np.array([0, 4, 8], dtype=np.int64))
    doff = cs.from_numpy(np.array([0, 4, 8], dtype=np.int64))
    with pytest.raises(cs.CsError):
        plan.sort_pairs(buf.ptr, buf.ptr, buf.ptr, buf.ptr, doff.ptr)   # plan of a key-only type
