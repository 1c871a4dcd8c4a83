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
