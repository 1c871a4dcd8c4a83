"""GPU regression suite with the shape of the reference's historical tests
(test.cu:10-602): single arrays of 32*2^k and 32*3^k keys, every
multiple-of-32 thread count for v3, every combination of the TestIfSorted /
FullSM / ASync switches, int and IntAndData, and the large-array entry
points (Sort = partition, SortMerge, HybridSort) up to 2^19 keys.  The
reference fills its arrays with curand (valgen.hpp:38-96, not reproducible
here); these use the oracle's seeded distributions instead, and every result
is checked against the oracle's std::sort (test.cu does the same check,
test.cu:75-92)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FLAG_SETS = list(range(8))   # TESTSORTED=1 | FULLSM=2 | ASYNC=4, all combinations


def sizes_upto(limit):
    s = set()
    v = 32
    while v <= limit:
        s.add(v)
        v *= 2
    v = 32
    while v <= limit:
        s.add(v)
        v *= 3
    return sorted(s)


def check(orc, dtype, host_in, got, n, where):
    rc, msg = orc.validate_batched(dtype, host_in, got, n)
    assert rc == 0, f"{where}: {msg}"


def sweep_one_array(cs, orc, gpu, dtype, n, cases, dists):
    """cases: iterable of (method, threads, flags); one array of n keys."""
    for dist in dists:
        host_in = orc.generate(dtype, dist, n, n, 1000 + n)
        src = cs.from_numpy(host_in)
        dst = cs.device_buffer(dtype, n)
        for method, threads, flags in cases:
            print(f"historical case dtype={dtype} n={n} dist={orc.DIST_NAMES[dist]} method={cs.METHODS[method]} "
                  f"threads={threads} flags={flags}", flush=True)
            cs.batched_sort(dtype, src.ptr, dst.ptr, 1, n, method, threads, flags)
            gpu.synchronize()
            got = cs.to_numpy(dst, dtype, n)
            check(orc, dtype, host_in, got, n,
                  f"n={n} dist={orc.DIST_NAMES[dist]} method={cs.METHODS[method]} threads={threads} flags={flags}")


@pytest.mark.parametrize("dtname", ["I32", "REC_I32_I32"])
def test_small(cs, orc, gpu, dtname):
    """testSmall: warp for n <= 32, v1 and v2 with threads = W up to 512
    (test.cu:10-300); IntAndData runs with FullSM only (test.cu:614-618)."""
    dt = getattr(cs, dtname)
    flag_sets = FLAG_SETS if dtname == "I32" else [2, 6]
    dists = [orc.D_UNIFORM, orc.D_DUP50, orc.D_REVERSED]
    for n in [1, 2, 5, 17, 31] + sizes_upto(512):
        w = max(32, 1 << (n - 1).bit_length())
        if n <= 32:
            cases = [(cs.WARP, 32, f) for f in flag_sets]
        else:
            cases = [(m, w, f) for m in (cs.V1, cs.V2) for f in flag_sets]
        sweep_one_array(cs, orc, gpu, dt, n, cases, dists)


@pytest.mark.parametrize("dtname", ["I32", "REC_I32_I32"])
def test_v3_every_thread_count(cs, orc, gpu, dtname):
    """testV3: threads 32..512 step 32 at every size (test.cu:302-441)."""
    dt = getattr(cs, dtname)
    limit = 8192 if dtname == "I32" else 4096
    flag_sets = FLAG_SETS if dtname == "I32" else [2]
    for n in sizes_upto(limit):
        cases = [(cs.V3, t, f) for t in range(32, 513, 32) for f in flag_sets]
        sweep_one_array(cs, orc, gpu, dt, n, cases, [orc.D_UNIFORM])
    # a second distribution on a thinner grid
    for n in sizes_upto(limit):
        cases = [(cs.V3, t, 0) for t in (32, 96, 256, 512)]
        sweep_one_array(cs, orc, gpu, dt, n, cases, [orc.D_DUP50, orc.D_SORTED])


@pytest.mark.parametrize("dtname", ["I32", "REC_I32_I32"])
@pytest.mark.parametrize("entry", ["partition", "merge", "hybrid"])
def test_large_entry_points(cs, orc, gpu, dtname, entry):
    """testLarge / testLargeMerge / testLargeHybrid: Sort, SortMerge and
    HybridSort on one array of 32*2^k and 32*3^k keys below 2^20
    (test.cu:443-602)."""
    dt = getattr(cs, dtname)
    for n in sizes_upto((1 << 19)):
        host_in = orc.generate(dt, orc.D_UNIFORM, n, n, 7 * n + 1)
        buf = cs.from_numpy(host_in)
        print(f"historical large case dtype={dtname} entry={entry} n={n}", flush=True)
        if entry == "partition":
            cs.sort_partition(dt, buf.ptr, n)
        elif entry == "merge":
            cs.sort_merge(dt, buf.ptr, n)
        else:
            cs.sort_hybrid(dt, buf.ptr, n)
        gpu.synchronize()
        got = cs.to_numpy(buf, dt, n)
        rc, msg = orc.validate_segmented(dt, host_in, got, np.array([0, n], dtype=np.int64))
        assert rc == 0, f"{entry} n={n}: {msg}"
