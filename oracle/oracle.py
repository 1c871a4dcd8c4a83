"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py for its cpu_baseline leg.  The product path
(cuda-sort_amd/) never imports this module.

The oracle restates the reference's correctness check: host std::sort per
array (reference bench.cu:300-301) plus the record invariants of
validateBatch (bench.cu:209-228).  See oracle.h for the pinning of each
function.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

I32, REC_I32_I32, F32, U64, REC_I64_I32, U32, I64, F64 = range(8)

(D_UNIFORM, D_SORTED, D_REVERSED, D_CONSTANT, D_DISTINCT16, D_PERTURBED, D_ORGAN_PIPE,
 D_ALT_EXTREMES, D_DUP50, D_EXTREMES6, D_RANDOM1M, D_BINARY, D_NORMAL_SPECIAL,
 D_PRESORTED_RUNS) = range(14)

#: the nine seeded distributions of the record configuration, in order
NINE = [D_UNIFORM, D_SORTED, D_REVERSED, D_CONSTANT, D_DISTINCT16, D_PERTURBED,
        D_ORGAN_PIPE, D_ALT_EXTREMES, D_DUP50]
DIST_NAMES = ["UNIFORM", "SORTED", "REVERSED", "CONSTANT", "DISTINCT16", "PERTURBED",
              "ORGAN_PIPE", "ALT_EXTREMES", "DUP50", "EXTREMES6", "RANDOM1M", "BINARY",
              "NORMAL_SPECIAL", "PRESORTED_RUNS"]

REC_I32 = np.dtype([("key", "<i4"), ("payload", "<i4")])
REC_I64 = np.dtype([("key", "<i8"), ("payload", "<i4"), ("pad", "<i4")])

NP_DTYPE = {I32: np.dtype("<i4"), REC_I32_I32: REC_I32, F32: np.dtype("<u4"),
            U64: np.dtype("<u8"), REC_I64_I32: REC_I64, U32: np.dtype("<u4"),
            I64: np.dtype("<i8"), F64: np.dtype("<u8")}

_lib = None


def build() -> None:
    """Compile liboracle.so from oracle.cpp (g++, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i64, i32, u64 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_uint64
        L.orc_splitmix64.restype = u64
        L.orc_splitmix64.argtypes = [u64]
        L.orc_mixbits.restype = ctypes.c_uint32
        L.orc_mixbits.argtypes = [ctypes.c_uint32]
        L.orc_f32_to_ordered.restype = ctypes.c_uint32
        L.orc_f32_to_ordered.argtypes = [ctypes.c_uint32]
        L.orc_ordered_to_f32.restype = ctypes.c_uint32
        L.orc_ordered_to_f32.argtypes = [ctypes.c_uint32]
        L.orc_generate.argtypes = [i32, i32, i64, i64, u64, vp]
        L.orc_sort_batched.argtypes = [i32, vp, vp, i64, i64, i32]
        L.orc_sort_segmented.argtypes = [i32, vp, vp, vp, i64, i32]
        L.orc_sort_large.argtypes = [i32, vp, vp, i64, i32]
        L.orc_validate_batched.restype = i64
        L.orc_validate_batched.argtypes = [i32, vp, vp, i64, i64, ctypes.c_char_p, i32]
        L.orc_validate_segmented.restype = i64
        L.orc_validate_segmented.argtypes = [i32, vp, vp, vp, i64, ctypes.c_char_p, i32]
        L.orc_time_batched.restype = ctypes.c_double
        L.orc_time_batched.argtypes = [i32, vp, i64, i64, i32, i32, ctypes.POINTER(i32)]
        L.orc_time_segmented.restype = ctypes.c_double
        L.orc_time_segmented.argtypes = [i32, vp, vp, i64, i32, i32, ctypes.POINTER(i32)]
        L.orc_time_large.restype = ctypes.c_double
        L.orc_time_large.argtypes = [i32, vp, i64, i32, i32, ctypes.POINTER(i32)]
        L.orc_host_cores.restype = i32
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    assert a.flags.c_contiguous
    return a.ctypes.data


def generate(dtype: int, dist: int, n: int, count: int, seed: int = 42) -> np.ndarray:
    out = np.empty(count, dtype=NP_DTYPE[dtype])
    rc = lib().orc_generate(dtype, dist, n, count, seed & (2**64 - 1), _ptr(out))
    if rc != 0:
        raise ValueError("orc_generate rejected its arguments")
    return out


def sort_batched(dtype: int, data: np.ndarray, n: int, threads: int = 0) -> np.ndarray:
    out = np.empty_like(data)
    arrays = data.size // n
    if lib().orc_sort_batched(dtype, _ptr(data), _ptr(out), n, arrays, threads) != 0:
        raise ValueError("orc_sort_batched rejected its arguments")
    return out


def sort_segmented(dtype: int, data: np.ndarray, offsets: np.ndarray, threads: int = 0) -> np.ndarray:
    out = np.empty_like(data)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    if lib().orc_sort_segmented(dtype, _ptr(data), _ptr(out), _ptr(off), off.size - 1, threads) != 0:
        raise ValueError("orc_sort_segmented rejected its arguments")
    return out


def sort_large(dtype: int, data: np.ndarray, threads: int = 0) -> np.ndarray:
    out = np.empty_like(data)
    if lib().orc_sort_large(dtype, _ptr(data), _ptr(out), data.size, threads) != 0:
        raise ValueError("orc_sort_large rejected its arguments")
    return out


def validate_batched(dtype: int, inp: np.ndarray, res: np.ndarray, n: int) -> tuple[int, str]:
    """0 if `res` is a correct per-array sort of `inp`, else 1-based bad index."""
    msg = ctypes.create_string_buffer(256)
    rc = lib().orc_validate_batched(dtype, _ptr(inp), _ptr(res), n, inp.size // n, msg, 256)
    return int(rc), msg.value.decode()


def validate_segmented(dtype: int, inp: np.ndarray, res: np.ndarray, offsets: np.ndarray) -> tuple[int, str]:
    msg = ctypes.create_string_buffer(256)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    rc = lib().orc_validate_segmented(dtype, _ptr(inp), _ptr(res), _ptr(off), off.size - 1, msg, 256)
    return int(rc), msg.value.decode()


def time_batched(dtype: int, data: np.ndarray, n: int, threads: int = 0, trials: int = 3) -> tuple[float, int]:
    used = ctypes.c_int(0)
    s = lib().orc_time_batched(dtype, _ptr(data), n, data.size // n, threads, trials, ctypes.byref(used))
    return float(s), int(used.value)


def time_segmented(dtype: int, data: np.ndarray, offsets: np.ndarray, threads: int = 0,
                   trials: int = 3) -> tuple[float, int]:
    used = ctypes.c_int(0)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    s = lib().orc_time_segmented(dtype, _ptr(data), _ptr(off), off.size - 1, threads, trials, ctypes.byref(used))
    return float(s), int(used.value)


def time_large(dtype: int, data: np.ndarray, threads: int = 0, trials: int = 3) -> tuple[float, int]:
    used = ctypes.c_int(0)
    s = lib().orc_time_large(dtype, _ptr(data), data.size, threads, trials, ctypes.byref(used))
    return float(s), int(used.value)


def host_cores() -> int:
    return int(lib().orc_host_cores())


def sort_keys_numpy(dtype: int, data: np.ndarray, n: int) -> np.ndarray:
    """Independent check: per-array key sort with numpy (for pinning only)."""
    if dtype in (REC_I32_I32, REC_I64_I32):
        keys = data["key"].reshape(-1, n)
    elif dtype == F32:
        b = data.astype(np.uint32)
        keys = np.where(b >> 31 == 1, b ^ np.uint32(0xFFFFFFFF), b ^ np.uint32(0x80000000)).reshape(-1, n)
        s = np.sort(keys, axis=1, kind="stable").reshape(-1)
        return np.where(s >> 31 == 1, s ^ np.uint32(0x80000000), s ^ np.uint32(0xFFFFFFFF)).astype(np.uint32)
    elif dtype == F64:
        # totalOrder from the float value itself: NaNs by sign to the ends
        # (payload order among them from the bit pattern), -0 before +0.
        v = data.view(np.float64).reshape(-1, n)
        b = data.reshape(-1, n)
        nan = np.isnan(v)
        neg = np.signbit(v)
        cls = np.where(nan & neg, 0, np.where(nan, 2, 1))
        tie = np.where(nan & neg, ~b, np.where(nan, b, neg.astype(np.uint64) ^ np.uint64(1)))
        val = np.where(nan, 0.0, v)
        order = np.lexsort((tie, val, cls), axis=1)
        return np.take_along_axis(b, order, axis=1).reshape(-1)
    else:
        keys = data.reshape(-1, n)
    return np.sort(keys, axis=1, kind="stable").reshape(-1)
