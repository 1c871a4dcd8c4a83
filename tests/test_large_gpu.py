"""GPU parity of the large-array sorts (merge, partition, hybrid) against
the CPU oracle (std::sort / __gnu_parallel::sort), including the full-size
2^28 int32 and 2^27 uint64 configurations."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run_large(cs, orc, rt, dtype, dist, n, method, seed=42, **kw):
    host = orc.generate(dtype, dist, max(n, 1), n, seed)
    buf = cs.from_numpy(host) if n else cs.device_buffer(dtype, 1)
    if method == "merge":
        cs.sort_merge(dtype, buf.ptr, n, **kw)
    elif method == "merge_out":
        out = cs.device_buffer(dtype, n)
        cs.sort_merge(dtype, buf.ptr, n, ext=out.ptr, ext_elems=n, in_place=False)
        buf = out
    elif method == "merge_ws":
        ws = cs.Workspace(dtype, n)
        cs.sort_merge_ws(dtype, buf.ptr, n, ws)
    elif method == "partition":
        cs.sort_partition(dtype, buf.ptr, n, **kw)
    elif method == "partition_ws":
        ws = kw.pop("ws", None) or cs.Workspace(dtype, n)
        cs.sort_partition_ws(dtype, buf.ptr, n, ws, **kw)
    elif method == "hybrid":
        cs.sort_hybrid(dtype, buf.ptr, n, kw.get("ws"))
    elif method == "hybrid_ws":
        ws = cs.Workspace(dtype, n)
        cs.sort_hybrid(dtype, buf.ptr, n, ws)
    rt.synchronize()
    got = cs.to_numpy(buf, dtype, n)
    where = f"dtype={dtype} dist={orc.DIST_NAMES[dist]} n={n} method={method}"
    if dtype in (cs.REC_I32_I32, cs.REC_I64_I32) or n <= (1 << 22):
        rc, msg = orc.validate_segmented(dtype, host, got, np.array([0, n], dtype=np.int64))
        assert rc == 0, f"{where}: {msg}"
    else:
        want = orc.sort_large(dtype, host)
        assert got.tobytes() == want.tobytes(), where


SIZES = [0, 1, 2, 33, 1000, 4096, 4097, 8192, 8193, 100003, (1 << 20) + 3, 3 << 20]
DISTS = ["D_UNIFORM", "D_SORTED", "D_REVERSED", "D_CONSTANT", "D_DUP50", "D_EXTREMES6", "D_DISTINCT16",
         "D_PRESORTED_RUNS"]


@pytest.mark.parametrize("method", ["merge", "merge_out", "merge_ws", "partition", "partition_ws", "hybrid",
                                    "hybrid_ws"])
@pytest.mark.parametrize("dtype", ["I32", "U64", "F32", "REC_I32_I32", "REC_I64_I32", "F64"])
def test_large_methods(cs, orc, gpu, dtype, method):
    dt = getattr(cs, dtype)
    for n in SIZES:
        dists = DISTS if n in (100003, (1 << 20) + 3) else ["D_UNIFORM", "D_DUP50"]
        for d in dists:
            run_large(cs, orc, gpu, dt, getattr(orc, d), n, method, seed=n + 7)


def test_partition_part_counts(cs, orc, gpu):
    n = (1 << 21) + 5
    for parts in (2, 3, 32, 1024, 2048):
        run_large(cs, orc, gpu, cs.I32, orc.D_UNIFORM, n, "partition", max_parts=parts)
    run_large(cs, orc, gpu, cs.U64, orc.D_DISTINCT16, n, "partition", max_parts=64)


def test_merge_arguments(cs, gpu):
    buf = cs.device_buffer(cs.I32, 16)
    with pytest.raises(cs.CsError):
        cs.sort_merge(cs.I32, buf.ptr, 16, in_place=False)          # out of place needs ext
    with pytest.raises(cs.CsError):
        cs.sort_merge(cs.I32, buf.ptr, 16, tile=100)                 # not a power of two
    ws = cs.Workspace(cs.I32, 8)
    with pytest.raises(cs.CsError):
        cs.sort_merge_ws(cs.I32, buf.ptr, 16, ws)                    # workspace too small
    ws.reserve(16)
    assert ws.capacity() == 16
    with pytest.raises(cs.CsError):
        cs.sort_hybrid(cs.U64, buf.ptr, 2, ws)                       # wrong element type


@pytest.mark.parametrize("dtype,n", [("I32", 1 << 28), ("U64", 1 << 27)])
@pytest.mark.parametrize("dist", ["D_UNIFORM", "D_PRESORTED_RUNS"])
def test_config4_full(cs, orc, gpu, dtype, n, dist):
    dt = getattr(cs, dtype)
    ws = cs.Workspace(dt, n)
    run_large(cs, orc, gpu, dt, getattr(orc, dist), n, "hybrid", ws=ws)
    run_large(cs, orc, gpu, dt, getattr(orc, dist), n, "merge")
    run_large(cs, orc, gpu, dt, getattr(orc, dist), n, "partition")


@pytest.mark.parametrize("dtype", ["I32", "F32", "U64", "F64", "REC_I64_I32"])
def test_partition_level_boundaries(cs, orc, gpu, dtype):
    """Sizes around the one-level / two-level switch of the partition path
    (256 parts of 1536 elements) and small part bounds that force buckets
    longer than the chunk onto the long-bucket chunk sort + merge passes."""
    dt = getattr(cs, dtype)
    target = 1536
    for n in (256 * target - 1, 256 * target, 256 * target + 1, 300 * target + 17):
        run_large(cs, orc, gpu, dt, orc.D_UNIFORM, n, "partition", seed=n)
    run_large(cs, orc, gpu, dt, orc.D_DUP50, 300 * target + 17, "partition", max_parts=4)
    run_large(cs, orc, gpu, dt, orc.D_UNIFORM, 70000, "partition", min_parts=4000)


@pytest.mark.parametrize("dtype", ["I32", "U64", "F32", "REC_I32_I32", "REC_I64_I32"])
def test_partition_split_buckets(cs, orc, gpu, dtype):
    """Final buckets a little longer than the single-wave tile (2048
    elements, 1024 for 16-byte records) are sorted as a head and a tail and
    merged: part bounds that put most buckets there, with tails from one
    element to a full head, heavy duplicates (equality buckets stay out of
    it) and presorted runs, one level and two, in place and through a
    workspace."""
    dt = getattr(cs, dtype)
    head = 1024 if cs.DTYPE_SIZE[dt] == 16 else 2048
    for n, parts in (((1 << 21) + 3, ((1 << 21) + 3) // (head * 3 // 2)),   # two levels, buckets ~1.5 heads
                     (200 * head + 1, 200),                                   # one level, buckets ~1 head
                     (150 * head + 77, 100)):                                # one level, buckets ~1.5 heads
        for d in ("D_UNIFORM", "D_DUP50", "D_PRESORTED_RUNS", "D_DISTINCT16"):
            run_large(cs, orc, gpu, dt, getattr(orc, d), n, "partition", seed=n + len(d), max_parts=parts)
        run_large(cs, orc, gpu, dt, orc.D_UNIFORM, n, "partition_ws", seed=n + 5, max_parts=parts)


def test_workspace_holds_every_smaller_sort(cs, orc, gpu):
    """A workspace reserved for n elements serves every smaller partition
    sort: below 2^26 8-byte elements the count / scatter workgroups shrink
    (so there can be more of them), the two-level part counts do not grow
    evenly with n, and 256 parts or fewer take one level with more samples
    per part."""
    cap = 1 << 26
    ws = cs.Workspace(cs.U64, cap)
    for n in (cap, cap - 3, (1 << 25) + 1, 256 * 1536, 256 * 1536 + 1, 100003):
        run_large(cs, orc, gpu, cs.U64, orc.D_UNIFORM, n, "partition_ws", seed=n, ws=ws)


@pytest.mark.parametrize("dtype", ["I32", "U64", "REC_I32_I32", "REC_I64_I32"])
def test_partition_long_buckets_two_levels(cs, orc, gpu, dtype):
    """A part bound above the one-level limit but far below n / target: two
    levels laid out on the device whose final buckets are several tiles
    long, so the cooperative chunk sort + merge passes run after the
    persistent bucket sort (in place and through a workspace)."""
    dt = getattr(cs, dtype)
    n = (1 << 23) + 11 if cs.DTYPE_SIZE[dt] <= 8 else (1 << 22) + 11
    for d in ("D_UNIFORM", "D_DUP50", "D_PRESORTED_RUNS"):
        run_large(cs, orc, gpu, dt, getattr(orc, d), n, "partition", seed=n + 1, max_parts=300)
    run_large(cs, orc, gpu, dt, orc.D_UNIFORM, n, "partition_ws", seed=n + 2, max_parts=300)
    run_large(cs, orc, gpu, dt, orc.D_CONSTANT, n, "partition", seed=n + 3, max_parts=300)


@pytest.mark.parametrize("dtype", ["F32", "F64"])
@pytest.mark.parametrize("method", ["partition", "merge"])
def test_large_special_floats(cs, orc, gpu, dtype, method):
    """NaNs of both signs, infinities and signed zeros (1 % of the keys)
    through the large paths: the equality buckets of repeated special values
    and the totalOrder transform in every classification and merge."""
    dt = getattr(cs, dtype)
    for n in (100003, (1 << 21) + 7):
        run_large(cs, orc, gpu, dt, orc.D_NORMAL_SPECIAL, n, method, seed=n)


def test_repeated_sorts_keep_memory_bounded(cs, orc, gpu):
    """The stream-ordered temporaries of the partition and merge paths come
    from a pool that keeps freed memory cached: repeated sorts reuse it, so
    device memory in use settles after the first calls instead of growing."""
    n = (1 << 22) + 13
    host = orc.generate(cs.I32, orc.D_UNIFORM, n, n, 3)
    buf = cs.from_numpy(host)
    for _ in range(2):
        cs.sort_partition(cs.I32, buf.ptr, n)
        cs.sort_merge(cs.I32, buf.ptr, n)
    gpu.synchronize()
    free0, _ = gpu.mem_info()
    for _ in range(10):
        buf.upload(host)
        cs.sort_partition(cs.I32, buf.ptr, n)
        cs.sort_merge(cs.I32, buf.ptr, n)
    gpu.synchronize()
    free1, _ = gpu.mem_info()
    assert free0 - free1 < (64 << 20), (free0, free1)
    got = cs.to_numpy(buf, cs.I32, n)
    assert got.tobytes() == np.sort(host).tobytes()


@pytest.mark.parametrize("dtype,n", [("U64", (1 << 24) + 3), ("I32", (1 << 25) + 5), ("I32", (1 << 22) + 9),
                                     ("REC_I64_I32", (1 << 22) + 1)])
def test_workspace_sorts_allocate_nothing(cs, orc, gpu, dtype, n):
    """The reference's allocation-free path (CuSortWorkspace, cusort.hpp:67):
    with a reserved workspace the hybrid sort, above the threshold (the
    partition path) and below it (the merge path), takes no device memory
    from the library's pool, call after call, and still sorts correctly."""
    dt = getattr(cs, dtype)
    host = orc.generate(dt, orc.D_UNIFORM, n, n, 11)
    buf = cs.from_numpy(host)
    ws = cs.Workspace(dt, n)
    gpu.synchronize()
    before = cs.memory_stats()
    for _ in range(3):
        buf.upload(host)
        cs.sort_hybrid(dt, buf.ptr, n, ws)
    gpu.synchronize()
    assert cs.memory_stats() == before
    merges, parts = ws.calls()
    assert (merges, parts) == ((0, 3) if n >= cs.hybrid_threshold(dt) else (3, 0))
    got = cs.to_numpy(buf, dt, n)
    if dt == cs.REC_I64_I32:
        rc, msg = orc.validate_segmented(dt, host, got, np.array([0, n], dtype=np.int64))
        assert rc == 0, msg
    else:
        assert got.tobytes() == orc.sort_large(dt, host).tobytes()


def test_workspace_too_small_is_refused(cs, gpu):
    n = 1 << 20
    buf = cs.device_buffer(cs.I32, n)
    ws = cs.Workspace(cs.I32, n // 2)
    with pytest.raises(cs.CsError):
        cs.sort_partition_ws(cs.I32, buf.ptr, n, ws)
    ws.reserve(n)
    cs.sort_partition_ws(cs.I32, buf.ptr, n, ws, max_parts=2048)
