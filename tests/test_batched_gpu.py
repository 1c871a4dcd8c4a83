"""GPU parity of the batched sort (cs_batched_sort) against the CPU oracle.

Every case: generate the input on the device (cs_generate), check it is
bit-identical to the oracle's host generator, sort through the C ABI, copy
back and run the oracle's validator (host std::sort per array, bit-exact key
sequence, and for records the payload range / membership / association /
uniqueness checks of reference bench.cu:209-228).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run_case(cs, orc, rt, dtype, dist, n, arrays, method=0, threads=0, flags=0, inplace=False,
             seed=42, offset_elems=0):
    count = n * arrays
    esz = cs.DTYPE_SIZE[dtype]
    base = cs.device_buffer(dtype, count + offset_elems + 1)
    src_ptr = base.ptr + offset_elems * esz
    cs.generate(dtype, dist, n, count, seed, src_ptr)
    host_in = orc.generate(dtype, dist, n, count, seed)
    rt.synchronize()
    dev_in = cs.to_numpy(base, dtype, count, offset_elems)
    assert dev_in.tobytes() == host_in.tobytes(), "device generator differs from the oracle's"
    if inplace:
        out, dst_ptr = base, src_ptr
    else:
        out = cs.device_buffer(dtype, count + offset_elems + 1)
        dst_ptr = out.ptr + offset_elems * esz
    cs.batched_sort(dtype, src_ptr, dst_ptr, arrays, n, method, threads, flags)
    rt.synchronize()
    res = cs.to_numpy(out, dtype, count, offset_elems)
    rc, msg = orc.validate_batched(dtype, host_in, res, n)
    where = (f"dtype={dtype} dist={orc.DIST_NAMES[dist]} n={n} arrays={arrays} method={cs.METHODS[method]} "
             f"threads={threads} flags={flags} inplace={inplace}")
    assert rc == 0, f"{where}: {msg}"
    if not inplace:
        again = cs.to_numpy(base, dtype, count, offset_elems)
        assert again.tobytes() == host_in.tobytes(), f"{where}: input modified by an out-of-place sort"


def arrays_for(n, budget=1 << 18, lo=3):
    return max(lo, min(4096, budget // max(n, 1)))


SIZES = list(range(1, 34)) + [63, 64, 65, 100, 127, 128, 129, 255, 256, 257, 511, 512, 513, 1000, 1023, 1024,
                              1025, 2047, 2048, 2049, 4095, 4096, 4097, 8191, 8192, 8193, 16383, 16384]


@pytest.mark.parametrize("n", SIZES)
def test_auto_i32_sizes(cs, orc, gpu, n):
    for dist in (orc.D_UNIFORM, orc.D_EXTREMES6, orc.D_ALT_EXTREMES, orc.D_DUP50):
        run_case(cs, orc, gpu, cs.I32, dist, n, arrays_for(n))


@pytest.mark.parametrize("dtype", ["I32", "REC_I32_I32", "F32", "U64", "REC_I64_I32", "U32", "I64", "F64"])
@pytest.mark.parametrize("n", [1, 5, 32, 33, 129, 1000, 2048, 4096])
def test_auto_dtypes_nine_distributions(cs, orc, gpu, dtype, n):
    dt = getattr(cs, dtype)
    if n > cs.batched_max_n(dt):
        pytest.skip("beyond capacity")
    dists = orc.NINE + [orc.D_EXTREMES6, orc.D_RANDOM1M, orc.D_BINARY]
    if dt in (cs.F32, cs.F64):
        dists = dists + [orc.D_NORMAL_SPECIAL]
    for dist in dists:
        run_case(cs, orc, gpu, dt, dist, n, arrays_for(n, 1 << 16), seed=1000 + dist)


@pytest.mark.parametrize("dtype", ["I32", "REC_I32_I32", "U64", "REC_I64_I32"])
def test_capacity_limit(cs, orc, gpu, dtype):
    dt = getattr(cs, dtype)
    n = cs.batched_max_n(dt)
    for dist in (orc.D_UNIFORM, orc.D_EXTREMES6, orc.D_ALT_EXTREMES):
        run_case(cs, orc, gpu, dt, dist, n, 5)


def method_domain(cs, n):
    """(method, threads) pairs of the reference's sweep (bench.cu:523-552)."""
    w = max(32, 1 << (n - 1).bit_length())
    tmax = min(w, 1024)
    out = []
    if n <= 32:
        out.append((cs.WARP, 32))
    if n < 32:
        return out
    if w <= 1024:
        out += [(cs.V1, w), (cs.V2, w)]
    mult = sorted({32, 64, 96, 160, tmax, max(32, tmax // 2)} & set(range(32, tmax + 1, 32)))
    out += [(cs.V3, t) for t in mult] + [(cs.V4, t) for t in mult]
    pow2 = [t for t in (32, 64, 128, 256, 512, 1024) if t <= tmax]
    out += [(cs.V5, t) for t in pow2]
    out += [(cs.V6, t) for t in pow2 if w // t <= 16]
    out += [(cs.AUTO, 0)]
    return out


@pytest.mark.parametrize("dtype", ["I32", "REC_I32_I32"])
@pytest.mark.parametrize("n", [7, 32, 33, 100, 129, 1000, 1024, 2048, 4096])
def test_every_method_and_thread_count(cs, orc, gpu, dtype, n):
    dt = getattr(cs, dtype)
    for method, threads in method_domain(cs, n):
        for flags in (0, cs.FLAG_TESTSORTED, cs.FLAG_FULLSM, cs.FLAG_ASYNC,
                      cs.FLAG_TESTSORTED | cs.FLAG_FULLSM | cs.FLAG_ASYNC):
            for dist in (orc.D_UNIFORM, orc.D_EXTREMES6, orc.D_SORTED):
                run_case(cs, orc, gpu, dt, dist, n, 7, method, threads, flags)


@pytest.mark.parametrize("dtype", ["I32", "U32", "F32", "I64", "U64", "F64", "REC_I32_I32", "REC_I64_I32"])
@pytest.mark.parametrize("n", [32, 33, 100, 255, 1000, 1024])
def test_async_direct_lds_loads(cs, orc, gpu, dtype, n):
    """ASYNC on the LDS network (v1, v2, and v3 / v4 at thread counts that
    are not a power of two): key-only arrays reach LDS by direct
    global-to-LDS loads, fp keys are converted in place, records keep the
    staged loads; every case against the oracle, including odd lengths,
    partial last waves and fp special values."""
    dt = getattr(cs, dtype)
    w = max(32, 1 << (n - 1).bit_length())
    dists = [orc.D_UNIFORM, orc.D_DUP50, orc.D_SORTED, orc.D_EXTREMES6]
    if dt in (cs.F32, cs.F64):
        dists.append(orc.D_NORMAL_SPECIAL)
    cases = [(cs.V1, w), (cs.V2, w), (cs.V3, 96), (cs.V4, 160)]
    for method, threads in cases:
        if threads > w:
            continue
        for flags in (cs.FLAG_ASYNC, cs.FLAG_ASYNC | cs.FLAG_TESTSORTED):
            for dist in dists:
                run_case(cs, orc, gpu, dt, dist, n, 9, method, threads, flags, seed=n + dist)
        run_case(cs, orc, gpu, dt, orc.D_UNIFORM, n, 5, method, threads, cs.FLAG_ASYNC, inplace=True, offset_elems=1)


@pytest.mark.parametrize("dtype", ["I32", "F32", "REC_I32_I32", "U64", "REC_I64_I32"])
@pytest.mark.parametrize("n", [32, 100, 1024, 2048, 4096])
def test_testsorted_mixed_arrays(cs, orc, gpu, dtype, n):
    """TESTSORTED inside the register engines: an array found sorted right
    after the load skips the network.  Already-sorted (host std::sort) and
    unsorted arrays alternate, so one workgroup holds both kinds, with
    several arrays per wave and, for the long records, several waves per
    array (the skip must then be taken by the whole workgroup or not at
    all).  Every method whose domain holds n runs with the flag."""
    dt = getattr(cs, dtype)
    if n > cs.batched_max_n(dt):
        pytest.skip("beyond capacity")
    arrays = 37
    raw = orc.generate(dt, orc.D_UNIFORM, n, n * arrays, 5)
    presorted = orc.sort_batched(dt, raw, n)
    host = raw.copy()
    for a in range(0, arrays, 2):
        host[a * n:(a + 1) * n] = presorted[a * n:(a + 1) * n]
    if host.dtype.names:   # records: the payload is the record's input position (the validator's convention)
        host["payload"] = np.arange(host.size, dtype=np.int32)
    src = cs.from_numpy(host)
    for method, threads in method_domain(cs, n):
        for flags in (cs.FLAG_TESTSORTED, cs.FLAG_TESTSORTED | cs.FLAG_FULLSM):
            out = cs.device_buffer(dt, n * arrays)
            cs.batched_sort(dt, src.ptr, out.ptr, arrays, n, method, threads, flags)
            gpu.synchronize()
            res = cs.to_numpy(out, dt, n * arrays)
            rc, msg = orc.validate_batched(dt, host, res, n)
            assert rc == 0, f"dtype={dtype} n={n} method={cs.METHODS[method]} threads={threads} flags={flags}: {msg}"
            for a in range(0, arrays, 2):   # a sorted array leaves unchanged, records included
                assert res[a * n:(a + 1) * n].tobytes() == host[a * n:(a + 1) * n].tobytes()


@pytest.mark.parametrize("inplace", [False, True])
def test_config1_full_workload(cs, orc, gpu, inplace):
    """4096 x 1024 int32, uniform full range, seed 42 (the headline config)."""
    run_case(cs, orc, gpu, cs.I32, orc.D_UNIFORM, 1024, 4096, inplace=inplace, seed=42)


@pytest.mark.parametrize("n", [32, 129, 1000, 2048, 4096])
def test_config2_records_full_batches(cs, orc, gpu, n):
    """IntAndData records, 4,194,304 valid elements per batch, nine distributions."""
    arrays = 4194304 // n
    for dist in orc.NINE:
        run_case(cs, orc, gpu, cs.REC_I32_I32, dist, n, arrays, seed=42)


def test_unaligned_and_in_place(cs, orc, gpu):
    for n in (4, 16, 64, 1024, 4096):
        for off in (1, 3):
            run_case(cs, orc, gpu, cs.I32, orc.D_UNIFORM, n, 9, offset_elems=off)
            run_case(cs, orc, gpu, cs.I32, orc.D_UNIFORM, n, 9, offset_elems=off, inplace=True)
        run_case(cs, orc, gpu, cs.REC_I32_I32, orc.D_DUP50, n, 9, offset_elems=1, inplace=True)


def test_many_tiny_arrays(cs, orc, gpu):
    run_case(cs, orc, gpu, cs.I32, orc.D_UNIFORM, 2, 1 << 20)
    run_case(cs, orc, gpu, cs.I32, orc.D_UNIFORM, 3, 100003)
    run_case(cs, orc, gpu, cs.REC_I32_I32, orc.D_UNIFORM, 1, 1000)


def test_empty_is_noop(cs, gpu):
    buf = cs.device_buffer(cs.I32, 16)
    cs.batched_sort(cs.I32, buf.ptr, buf.ptr, 0, 16)
    cs.batched_sort(cs.I32, buf.ptr, buf.ptr, 4, 0)
    gpu.synchronize()


def test_describe_reports_resources(cs, gpu):
    d = cs.describe(cs.I32, 1024)
    assert d["items_per_thread"] == 32 and d["block_threads"] == 256 and d["arrays_per_block"] == 8
    assert 0 < d["vgprs"] <= 128
    d = cs.describe(cs.I32, 1000, cs.V3, 96)
    assert d["block_threads"] == 96 and d["arrays_per_block"] == 1


def test_stream_hold_delays_queued_work(cs, gpu):
    """cs_stream_hold keeps its stream busy for about the time asked (the
    benchmark queues its timed launches behind it), then exits."""
    stream = gpu.Stream()
    start, mid, end = gpu.Event(), gpu.Event(), gpu.Event()
    start.record(stream)
    cs.stream_hold(2000.0, stream.handle)
    mid.record(stream)
    cs.stream_hold(0.0, stream.handle)
    end.record(stream)
    end.synchronize()
    held = start.elapsed_ms(mid)
    assert 1.9 <= held <= 50.0, held
    assert mid.elapsed_ms(end) < 1.0
