"""Python binding of libcusort_hip.so (the C ABI in include/cusort_hip.h).

The product is the HIP library; this module is the thin host-side mirror the
tests and the benchmark call through.  Device memory, streams and events come
from the HIP runtime the library links (hiprt.py).  There is no CPU fallback: if the library is
missing or no GPU is present the calls fail loudly.
"""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcusort_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "cusort_hip.h")

# cs_dtype
I32, REC_I32_I32, F32, U64, REC_I64_I32, U32, I64, F64 = range(8)
DTYPE_NAMES = {"i32": I32, "rec32": REC_I32_I32, "f32": F32, "u64": U64, "rec64": REC_I64_I32,
               "u32": U32, "i64": I64, "f64": F64}
DTYPE_SIZE = {I32: 4, REC_I32_I32: 8, F32: 4, U64: 8, REC_I64_I32: 16, U32: 4, I64: 8, F64: 8}

# cs_method (bench.cu:129-131 names)
METHODS = ["auto", "bitonic_warp", "bitonic_v1", "bitonic_v2", "bitonic_v3", "bitonic_v4_xor",
           "bitonic_v5_static", "bitonic_v6_register"]
AUTO, WARP, V1, V2, V3, V4, V5, V6 = range(8)

FLAG_TESTSORTED, FLAG_FULLSM, FLAG_ASYNC = 1, 2, 4

OK = 0
E_INVALID_VALUE, E_UNSUPPORTED, E_CAPACITY, E_NO_DEVICE, E_BUDGET, E_AGAIN, E_WORKSPACE = (
    -1, -2, -3, -4, -5, -6, -7)

(D_UNIFORM, D_SORTED, D_REVERSED, D_CONSTANT, D_DISTINCT16, D_PERTURBED, D_ORGAN_PIPE,
 D_ALT_EXTREMES, D_DUP50, D_EXTREMES6, D_RANDOM1M, D_BINARY, D_NORMAL_SPECIAL,
 D_PRESORTED_RUNS) = range(14)


class CsError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        super().__init__(f"{what}: {strerror(status)} ({status})")


_lib = None


def build(jobs: int = 8) -> None:
    """Compile the HIP sources for gfx950 into libcusort_hip.so (make)."""
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run build() (make -C cuda-sort_amd)")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, u32, u64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint64
        P32 = ctypes.POINTER(ctypes.c_int32)
        L.cs_strerror.restype = ctypes.c_char_p
        L.cs_strerror.argtypes = [ctypes.c_int]
        L.cs_version.restype = ctypes.c_int
        L.cs_dtype_size.restype = ctypes.c_size_t
        L.cs_dtype_size.argtypes = [ctypes.c_int]
        L.cs_method_from_name.restype = ctypes.c_int
        L.cs_method_from_name.argtypes = [ctypes.c_char_p]
        L.cs_batched_sort.restype = ctypes.c_int
        L.cs_batched_sort.argtypes = [ctypes.c_int, vp, vp, i64, i32, ctypes.c_int, i32, u32, vp]
        L.cs_batched_sort_pairs.restype = ctypes.c_int
        L.cs_batched_sort_pairs.argtypes = [ctypes.c_int, vp, vp, vp, vp, i64, i32, vp]
        L.cs_batched_max_n.restype = i32
        L.cs_batched_max_n.argtypes = [ctypes.c_int]
        L.cs_batched_describe.restype = ctypes.c_int
        L.cs_batched_describe.argtypes = [ctypes.c_int, i32, ctypes.c_int, i32, u32, P32, P32, P32, P32, P32]
        P64 = ctypes.POINTER(ctypes.c_int64)
        L.cs_segmented_plan_create.restype = ctypes.c_int
        L.cs_segmented_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, vp, i64, u32]
        L.cs_segmented_sort.restype = ctypes.c_int
        L.cs_segmented_sort.argtypes = [vp, vp, vp, vp, vp]
        L.cs_segmented_sort_pairs.restype = ctypes.c_int
        L.cs_segmented_sort_pairs.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.cs_sort_pairs.restype = ctypes.c_int
        L.cs_sort_pairs.argtypes = [ctypes.c_int, vp, vp, i64, vp, vp]
        L.cs_segmented_plan_destroy.restype = ctypes.c_int
        L.cs_segmented_plan_destroy.argtypes = [vp]
        L.cs_segmented_plan_info.restype = ctypes.c_int
        L.cs_segmented_plan_info.argtypes = [vp, P32, P32, P64, P32, P64, i32]
        L.cs_segmented_max_len.restype = i32
        L.cs_segmented_max_len.argtypes = [ctypes.c_int]
        L.cs_segmented_sort_plain.restype = ctypes.c_int
        L.cs_segmented_sort_plain.argtypes = [ctypes.c_int, vp, vp, vp, i64, i32, vp]
        L.cs_segmented_tune.restype = ctypes.c_int
        L.cs_segmented_tune.argtypes = [ctypes.c_int, i32, vp]
        L.cs_segmented_get_variant.restype = ctypes.c_int
        L.cs_segmented_get_variant.argtypes = [ctypes.c_int, i32, P32]
        L.cs_segmented_set_variant.restype = ctypes.c_int
        L.cs_segmented_set_variant.argtypes = [ctypes.c_int, i32, i32]
        sz = ctypes.c_size_t
        L.cs_workspace_create.restype = ctypes.c_int
        L.cs_workspace_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, sz, i32, i32]
        L.cs_workspace_reserve.restype = ctypes.c_int
        L.cs_workspace_reserve.argtypes = [vp, sz]
        L.cs_workspace_capacity.restype = ctypes.c_int
        L.cs_workspace_capacity.argtypes = [vp, ctypes.POINTER(sz)]
        L.cs_workspace_destroy.restype = ctypes.c_int
        L.cs_workspace_destroy.argtypes = [vp]
        L.cs_workspace_calls.restype = ctypes.c_int
        L.cs_workspace_calls.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
        L.cs_memory_stats.restype = ctypes.c_int
        L.cs_memory_stats.argtypes = [ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
        L.cs_sort_merge_static.restype = ctypes.c_int
        L.cs_sort_merge_static.argtypes = [ctypes.c_int, vp, ctypes.c_int64, vp]
        L.cs_sort_merge_static_max_bytes.restype = ctypes.c_size_t
        L.cs_sort_merge_static_max_bytes.argtypes = []
        L.cs_sort_partition_ws.restype = ctypes.c_int
        L.cs_sort_partition_ws.argtypes = [ctypes.c_int, vp, ctypes.c_int64, i32, i32, ctypes.c_uint32, vp, vp]
        L.cs_sort_merge.restype = ctypes.c_int
        L.cs_sort_merge.argtypes = [ctypes.c_int, vp, i64, i32, i32, vp, sz, ctypes.c_int, vp]
        L.cs_sort_merge_ws.restype = ctypes.c_int
        L.cs_sort_merge_ws.argtypes = [ctypes.c_int, vp, i64, vp, vp]
        L.cs_sort_partition.restype = ctypes.c_int
        L.cs_sort_partition.argtypes = [ctypes.c_int, vp, i64, i32, i32, u32, vp]
        L.cs_sort_hybrid.restype = ctypes.c_int
        L.cs_sort_hybrid.argtypes = [ctypes.c_int, vp, i64, vp, vp]
        L.cs_hybrid_threshold.restype = i64
        L.cs_hybrid_threshold.argtypes = [ctypes.c_int]
        L.cs_server_now.restype = ctypes.c_double
        L.cs_server_create.restype = ctypes.c_int
        L.cs_server_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, sz, i32, i64, i32]
        L.cs_server_submit.restype = ctypes.c_int
        L.cs_server_submit.argtypes = [vp, vp, vp, i64, ctypes.c_double, ctypes.POINTER(u64)]
        L.cs_server_poll.restype = ctypes.c_int
        L.cs_server_poll.argtypes = [vp, u64, P32, P32]
        L.cs_server_wait.restype = ctypes.c_int
        L.cs_server_wait.argtypes = [vp, u64, ctypes.c_double]
        L.cs_server_latency.restype = ctypes.c_int
        L.cs_server_latency.argtypes = [vp, u64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.cs_server_release.restype = ctypes.c_int
        L.cs_server_release.argtypes = [vp, u64]
        L.cs_server_drain.restype = ctypes.c_int
        L.cs_server_drain.argtypes = [vp]
        L.cs_server_get_stats.restype = ctypes.c_int
        L.cs_server_get_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
        L.cs_server_destroy.restype = ctypes.c_int
        L.cs_server_destroy.argtypes = [vp]
        L.cs_server_inject_launch_failure.restype = ctypes.c_int
        L.cs_server_inject_launch_failure.argtypes = [vp, i32]
        L.cs_server_pool_in_use.restype = ctypes.c_int
        L.cs_server_pool_in_use.argtypes = [vp, P64]
        L.cs_device_copy.restype = ctypes.c_int
        L.cs_device_copy.argtypes = [vp, vp, ctypes.c_size_t, vp]
        L.cs_stream_hold.restype = ctypes.c_int
        L.cs_stream_hold.argtypes = [ctypes.c_double, vp]
        L.cs_generate.restype = ctypes.c_int
        L.cs_generate.argtypes = [ctypes.c_int, ctypes.c_int, i64, i64, u64, vp, vp]
        _lib = L
    return _lib


def header_symbols(path: str = HEADER_PATH) -> list[str]:
    """Function names declared in the C header (for export checks)."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cs_[a-z0-9_]+)\s*\(", text)))


def strerror(status: int) -> str:
    return lib().cs_strerror(status).decode()


def check(status: int, what: str) -> None:
    if status != OK:
        raise CsError(status, what)


def batched_sort(dtype: int, src: int, dst: int, arrays: int, n: int, method: int = AUTO,
                 threads: int = 0, flags: int = 0, stream: int = 0) -> None:
    """cs_batched_sort on raw device pointers (ints)."""
    check(lib().cs_batched_sort(dtype, src, dst, arrays, n, method, threads, flags, stream),
          f"cs_batched_sort(dtype={dtype}, n={n}, method={METHODS[method] if 0 <= method < 8 else method}, "
          f"threads={threads}, flags={flags})")


def batched_sort_pairs(key_dtype: int, keys_in: int, vals_in: int, keys_out: int, vals_out: int, arrays: int,
                       n: int, stream: int = 0) -> None:
    """cs_batched_sort_pairs: keys (I32 or I64) and int32 payloads in two arrays."""
    check(lib().cs_batched_sort_pairs(key_dtype, keys_in, vals_in, keys_out, vals_out, arrays, n, stream),
          f"cs_batched_sort_pairs(key_dtype={key_dtype}, n={n})")


def batched_status(dtype: int, src: int, dst: int, arrays: int, n: int, method: int = AUTO,
                   threads: int = 0, flags: int = 0, stream: int = 0) -> int:
    return int(lib().cs_batched_sort(dtype, src, dst, arrays, n, method, threads, flags, stream))


def batched_max_n(dtype: int) -> int:
    return int(lib().cs_batched_max_n(dtype))


def describe(dtype: int, n: int, method: int = AUTO, threads: int = 0, flags: int = 0) -> dict:
    vals = [ctypes.c_int32(0) for _ in range(5)]
    check(lib().cs_batched_describe(dtype, n, method, threads, flags, *[ctypes.byref(v) for v in vals]),
          "cs_batched_describe")
    keys = ["block_threads", "items_per_thread", "lds_bytes", "vgprs", "arrays_per_block"]
    return {k: int(v.value) for k, v in zip(keys, vals)}


def generate(dtype: int, dist: int, n: int, count: int, seed: int, dst: int, stream: int = 0) -> None:
    check(lib().cs_generate(dtype, dist, n, count, seed & (2**64 - 1), dst, stream), "cs_generate")


def device_copy(dst: int, src: int, nbytes: int, stream: int = 0) -> None:
    check(lib().cs_device_copy(dst, src, nbytes, stream), "cs_device_copy")


def stream_hold(microseconds: float, stream: int = 0) -> None:
    check(lib().cs_stream_hold(float(microseconds), stream), "cs_stream_hold")


SEG_NO_GRAPH = 1


class ServerStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_int64), ("rejected", ctypes.c_int64), ("batches", ctypes.c_int64),
                ("keys", ctypes.c_int64), ("p50_s", ctypes.c_double), ("p99_s", ctypes.c_double),
                ("max_s", ctypes.c_double), ("mean_s", ctypes.c_double), ("keys_per_s", ctypes.c_double),
                ("mean_batch_requests", ctypes.c_double), ("peak_pool_bytes", ctypes.c_int64),
                ("max_inflight_batches", ctypes.c_int64), ("bypasses", ctypes.c_int64)]


class Server:
    """cs_server: continuous batching of host sort requests."""

    def __init__(self, dtype: int, budget_bytes: int, slots: int = 4, max_batch_elems: int = 0,
                 max_batch_requests: int = 0):
        h = ctypes.c_void_p()
        check(lib().cs_server_create(ctypes.byref(h), dtype, budget_bytes, slots, max_batch_elems,
                                     max_batch_requests), "cs_server_create")
        self.handle = h.value
        self.dtype = dtype
        self._keep = {}

    def submit(self, arr, out=None, arrival_s: float = -1.0) -> int:
        """Submit a host numpy array; returns the ticket.  The result lands in
        `out` (a new array by default), available from result(ticket)."""
        import numpy as np
        a = np.ascontiguousarray(arr)
        o = np.empty_like(a) if out is None else out
        t = ctypes.c_uint64(0)
        check(lib().cs_server_submit(self.handle, a.ctypes.data if a.size else None,
                                     o.ctypes.data if o.size else None, a.size, arrival_s, ctypes.byref(t)),
              "cs_server_submit")
        self._keep[t.value] = (a, o)
        return int(t.value)

    def submit_status(self, arr) -> int:
        import numpy as np
        a = np.ascontiguousarray(arr)
        o = np.empty_like(a)
        t = ctypes.c_uint64(0)
        return int(lib().cs_server_submit(self.handle, a.ctypes.data, o.ctypes.data, a.size, -1.0, ctypes.byref(t)))

    def wait(self, ticket: int, timeout_s: float = -1.0) -> int:
        return int(lib().cs_server_wait(self.handle, ticket, timeout_s))

    def result(self, ticket: int):
        rc = self.wait(ticket)
        check(rc, "server request")
        return self._keep[ticket][1]

    def latency(self, ticket: int) -> float:
        a, d = ctypes.c_double(0), ctypes.c_double(0)
        check(lib().cs_server_latency(self.handle, ticket, ctypes.byref(a), ctypes.byref(d)), "cs_server_latency")
        return d.value - a.value

    def release(self, ticket: int) -> None:
        check(lib().cs_server_release(self.handle, ticket), "cs_server_release")
        self._keep.pop(ticket, None)

    def drain(self) -> None:
        check(lib().cs_server_drain(self.handle), "cs_server_drain")

    def done_time(self, ticket: int) -> float:
        a, d = ctypes.c_double(0), ctypes.c_double(0)
        check(lib().cs_server_latency(self.handle, ticket, ctypes.byref(a), ctypes.byref(d)), "cs_server_latency")
        return d.value

    def inject_launch_failure(self, status: int) -> None:
        check(lib().cs_server_inject_launch_failure(self.handle, status), "cs_server_inject_launch_failure")

    def pool_in_use(self) -> int:
        b = ctypes.c_int64(0)
        check(lib().cs_server_pool_in_use(self.handle, ctypes.byref(b)), "cs_server_pool_in_use")
        return int(b.value)

    def stats(self) -> dict:
        st = ServerStats()
        check(lib().cs_server_get_stats(self.handle, ctypes.byref(st)), "cs_server_get_stats")
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self) -> None:
        if getattr(self, "handle", None) and _lib is not None:
            _lib.cs_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


def server_now() -> float:
    return float(lib().cs_server_now())


class Workspace:
    """cs_workspace: reusable scratch for the large-array sorts."""

    def __init__(self, dtype: int, capacity: int = 0, blocks_per_cu: int = 4, n_streams: int = 8):
        h = ctypes.c_void_p()
        check(lib().cs_workspace_create(ctypes.byref(h), dtype, capacity, blocks_per_cu, n_streams),
              "cs_workspace_create")
        self.handle = h.value
        self.dtype = dtype

    def reserve(self, capacity: int) -> None:
        check(lib().cs_workspace_reserve(self.handle, capacity), "cs_workspace_reserve")

    def capacity(self) -> int:
        c = ctypes.c_size_t(0)
        check(lib().cs_workspace_capacity(self.handle, ctypes.byref(c)), "cs_workspace_capacity")
        return int(c.value)

    def calls(self) -> tuple:
        """(merge sorts, partition sorts) run through this workspace."""
        m, p = ctypes.c_int64(0), ctypes.c_int64(0)
        check(lib().cs_workspace_calls(self.handle, ctypes.byref(m), ctypes.byref(p)), "cs_workspace_calls")
        return int(m.value), int(p.value)

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.cs_workspace_destroy(self.handle)
            self.handle = None


def memory_stats() -> tuple:
    """(allocations, bytes) the library has taken from its device pool."""
    a, b = ctypes.c_int64(0), ctypes.c_int64(0)
    check(lib().cs_memory_stats(ctypes.byref(a), ctypes.byref(b)), "cs_memory_stats")
    return int(a.value), int(b.value)


def sort_merge_static(dtype: int, data: int, n: int, stream: int = 0) -> None:
    """cs_sort_merge_static: the merge sort with the library's fixed scratch."""
    check(lib().cs_sort_merge_static(dtype, data, n, stream), "cs_sort_merge_static")


def sort_merge_static_status(dtype: int, data: int, n: int, stream: int = 0) -> int:
    return int(lib().cs_sort_merge_static(dtype, data, n, stream))


def sort_partition_ws(dtype: int, data: int, n: int, ws: "Workspace", min_parts: int = 0, max_parts: int = 0,
                      flags: int = 0, stream: int = 0) -> None:
    check(lib().cs_sort_partition_ws(dtype, data, n, min_parts, max_parts, flags, ws.handle, stream),
          "cs_sort_partition_ws")


def sort_merge(dtype: int, data: int, n: int, tile: int = 0, merge_threads: int = 0, ext: int = 0,
               ext_elems: int = 0, in_place: bool = True, stream: int = 0) -> None:
    check(lib().cs_sort_merge(dtype, data, n, tile, merge_threads, ext or None, ext_elems, 1 if in_place else 0,
                              stream), "cs_sort_merge")


def sort_merge_ws(dtype: int, data: int, n: int, ws: "Workspace", stream: int = 0) -> None:
    check(lib().cs_sort_merge_ws(dtype, data, n, ws.handle, stream), "cs_sort_merge_ws")


def sort_partition(dtype: int, data: int, n: int, min_parts: int = 0, max_parts: int = 0, flags: int = 0,
                   stream: int = 0) -> None:
    check(lib().cs_sort_partition(dtype, data, n, min_parts, max_parts, flags, stream), "cs_sort_partition")


def sort_pairs(key_dtype: int, keys: int, vals: int, n: int, ws: "Workspace | None" = None, stream: int = 0) -> None:
    """cs_sort_pairs: one array of SoA key / int32 payload pairs, in place."""
    check(lib().cs_sort_pairs(key_dtype, keys, vals, n, ws.handle if ws else None, stream), "cs_sort_pairs")


def sort_hybrid(dtype: int, data: int, n: int, ws: "Workspace | None" = None, stream: int = 0) -> None:
    check(lib().cs_sort_hybrid(dtype, data, n, ws.handle if ws else None, stream), "cs_sort_hybrid")


def hybrid_threshold(dtype: int) -> int:
    return int(lib().cs_hybrid_threshold(dtype))


class SegPlan:
    """A ragged-batch plan (cs_segplan) built from host CSR offsets."""

    def __init__(self, dtype: int, offsets, flags: int = 0):
        import numpy as np
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        self.dtype = dtype
        self.segments = off.size - 1
        h = ctypes.c_void_p()
        check(lib().cs_segmented_plan_create(ctypes.byref(h), dtype, off.ctypes.data, self.segments, flags),
              "cs_segmented_plan_create")
        self.handle = h.value

    def sort(self, src: int, dst: int, d_offsets: int, stream: int = 0) -> None:
        check(lib().cs_segmented_sort(self.handle, src, dst, d_offsets, stream), "cs_segmented_sort")

    def sort_pairs(self, keys_in: int, vals_in: int, keys_out: int, vals_out: int, d_offsets: int,
                   stream: int = 0) -> None:
        """cs_segmented_sort_pairs: SoA keys / int32 payloads through a record-type plan."""
        check(lib().cs_segmented_sort_pairs(self.handle, keys_in, vals_in, keys_out, vals_out, d_offsets, stream),
              "cs_segmented_sort_pairs")

    def info(self) -> dict:
        items, grid = ctypes.c_int32(0), ctypes.c_int32(0)
        mx = ctypes.c_int64(0)
        var = (ctypes.c_int32 * 16)()
        cnt = (ctypes.c_int64 * 16)()
        check(lib().cs_segmented_plan_info(self.handle, ctypes.byref(items), ctypes.byref(grid), ctypes.byref(mx),
                                           var, cnt, 16), "cs_segmented_plan_info")
        return {"work_items": items.value, "grid": grid.value, "max_len": mx.value,
                "variants": list(var), "bucket_segments": list(cnt)}

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.cs_segmented_plan_destroy(self.handle)
            self.handle = None


def segmented_sort_plain(dtype: int, src: int, dst: int, d_offsets: int, segments: int, max_len: int,
                         stream: int = 0) -> None:
    check(lib().cs_segmented_sort_plain(dtype, src, dst, d_offsets, segments, max_len, stream),
          "cs_segmented_sort_plain")


def segmented_max_len(dtype: int) -> int:
    return int(lib().cs_segmented_max_len(dtype))


def segmented_tune(dtype: int, trials: int = 3, stream: int = 0) -> None:
    check(lib().cs_segmented_tune(dtype, trials, stream), "cs_segmented_tune")


def segmented_variant(dtype: int, bucket: int) -> int:
    v = ctypes.c_int32(0)
    check(lib().cs_segmented_get_variant(dtype, bucket, ctypes.byref(v)), "cs_segmented_get_variant")
    return int(v.value)


def segmented_set_variant(dtype: int, bucket: int, variant: int) -> int:
    return int(lib().cs_segmented_set_variant(dtype, bucket, variant))


def ragged_lengths(segments: int = 100000, max_len: int = 8192, seed: int = 42):
    """Lengths floor(exp(U[0, ln(max_len+1)))) clipped to [1, max_len] (log-uniform)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    lens = np.floor(np.exp(rng.uniform(0.0, np.log(max_len + 1.0), segments))).astype(np.int64)
    return np.clip(lens, 1, max_len)


def offsets_from_lengths(lens):
    import numpy as np
    return np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)


# ------------------------------------------------------------ buffers
def _np_dtype(dtype: int):
    import numpy as np
    return {I32: np.dtype("<i4"), F32: np.dtype("<u4"), U32: np.dtype("<u4"), U64: np.dtype("<u8"),
            I64: np.dtype("<i8"), F64: np.dtype("<u8"),
            REC_I32_I32: np.dtype([("key", "<i4"), ("payload", "<i4")]),
            REC_I64_I32: np.dtype([("key", "<i8"), ("payload", "<i4"), ("pad", "<i4")])}[dtype]


def device_buffer(dtype: int, count: int):
    """Device allocation for `count` elements of `dtype` (hiprt.Buffer)."""
    import hiprt
    return hiprt.Buffer(count * DTYPE_SIZE[dtype])


def to_numpy(buf, dtype: int, count: int, offset_elems: int = 0):
    """Copy `count` elements starting at element `offset_elems` to the host."""
    return buf.download(_np_dtype(dtype), count, offset_elems * DTYPE_SIZE[dtype])


def from_numpy(arr):
    import hiprt
    b = hiprt.Buffer(arr.nbytes)
    b.upload(arr)
    return b
