"""C-ABI checks that need no GPU: every declared symbol is exported, and the
argument checks of the reference's (method, n, threads) domain reject bad
input before any device work is enqueued."""
import ctypes
import os
import subprocess

import pytest


def test_library_exports_every_declared_symbol(cs):
    names = cs.header_symbols()
    assert "cs_batched_sort" in names and len(names) >= 8
    L = ctypes.CDLL(cs.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_header_compiles_as_c(cs, tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "cusort_hip.h"\nint main(void){return cs_version()>0?0:1;}\n')
    inc = os.path.dirname(cs.HEADER_PATH)
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-c", str(src), "-I", inc, "-o",
                    str(tmp_path / "t.o")], check=True)


def test_strings_and_sizes(cs):
    L = cs.lib()
    assert L.cs_version() == 100
    for name, idx in zip(cs.METHODS, range(8)):
        assert L.cs_method_from_name(name.encode()) == idx
    assert L.cs_method_from_name(b"cub_block_merge") == -1
    assert [L.cs_dtype_size(d) for d in range(8)] == [4, 8, 4, 8, 16, 4, 8, 8]
    assert cs.batched_max_n(cs.I32) == 16384 and cs.batched_max_n(cs.REC_I32_I32) == 8192
    assert cs.batched_max_n(cs.REC_I64_I32) == 4096 and cs.batched_max_n(99) == 0
    assert "domain" in cs.strerror(cs.E_UNSUPPORTED)


FAKE = 1 << 40   # never dereferenced: every case below is rejected first


@pytest.mark.parametrize("args, want", [
    # dtype, arrays, n, method, threads, flags
    ((9, 4, 64, 0, 0, 0), -1),                      # unknown dtype
    ((0, -1, 64, 0, 0, 0), -1),                     # negative arrays
    ((0, 4, -5, 0, 0, 0), -1),                      # negative n
    ((0, 4, 64, 8, 0, 0), -1),                      # unknown method
    ((0, 4, 64, 0, 0, 8), -1),                      # unknown flag
    ((0, 4, 16385, 0, 0, 0), -3),                   # beyond 160 KB LDS capacity
    ((1, 4, 8193, 0, 0, 0), -3),
    ((4, 4, 4097, 0, 0, 0), -3),
    ((0, 4, 33, 1, 0, 0), -2),                      # bitonic_warp needs n <= 32
    ((0, 4, 20, 1, 64, 0), -2),                     # ... and 32 threads
    ((0, 4, 20, 2, 0, 0), -2),                      # v1 needs n >= 32 (bench.cu:536)
    ((0, 4, 100, 2, 64, 0), -2),                    # v1 needs threads == W
    ((0, 4, 2000, 3, 0, 0), -2),                    # v2 needs W <= 1024
    ((0, 4, 100, 4, 48, 0), -2),                    # v3 needs a multiple of 32
    ((0, 4, 100, 4, 1056, 0), -2),                  # v3 threads <= 1024
    ((0, 4, 100, 5, 256, 0), -2),                   # v4 threads <= W
    ((0, 4, 100, 6, 96, 0), -2),                    # v5 needs a power of two
    ((0, 4, 4096, 7, 128, 0), -2),                  # v6 needs W/threads <= 16
    ((0, 4, 4096, 7, 2048, 0), -2),                 # ... and threads <= 1024
])
def test_domain_rejections(cs, args, want):
    dt, arrays, n, method, threads, flags = args
    assert cs.batched_status(dt, FAKE, FAKE, arrays, n, method, threads, flags) == want


def test_stream_hold_rejects_out_of_range(cs):
    L = cs.lib()
    for us in (-1.0, 1e6 + 1, float("nan")):
        assert L.cs_stream_hold(us, None) == cs.E_INVALID_VALUE


def test_null_and_overlap(cs):
    assert cs.batched_status(0, 0, FAKE, 4, 64) == cs.E_INVALID_VALUE
    # partially overlapping in/out ranges are rejected; identical = in place
    assert cs.batched_status(0, FAKE, FAKE + 4, 4, 64) == cs.E_INVALID_VALUE
    # empty work is a no-op even with null pointers
    assert cs.batched_status(0, 0, 0, 0, 64) == cs.OK
    assert cs.batched_status(0, 0, 0, 4, 0) == cs.OK


def test_pairs_arguments(cs):
    L = cs.lib()
    assert L.cs_batched_sort_pairs(cs.F32, FAKE, FAKE, FAKE, FAKE, 1, 8, None) == cs.E_INVALID_VALUE
    assert L.cs_batched_sort_pairs(cs.I32, FAKE, FAKE, FAKE, FAKE, 1, -1, None) == cs.E_INVALID_VALUE
    assert L.cs_batched_sort_pairs(cs.I32, FAKE, FAKE, FAKE + 4, FAKE, 4, 64, None) == cs.E_INVALID_VALUE
    assert L.cs_batched_sort_pairs(cs.I32, FAKE, FAKE, FAKE, FAKE, -1, 8, None) == cs.E_INVALID_VALUE
    assert L.cs_batched_sort_pairs(cs.I32, FAKE, FAKE, FAKE, FAKE, 1, 8193, None) == cs.E_CAPACITY
    assert L.cs_batched_sort_pairs(cs.I64, FAKE, FAKE, FAKE, FAKE, 1, 4097, None) == cs.E_CAPACITY
    assert L.cs_batched_sort_pairs(cs.I32, None, FAKE, FAKE, FAKE, 1, 8, None) == cs.E_INVALID_VALUE
    assert L.cs_batched_sort_pairs(cs.I32, None, None, None, None, 0, 8, None) == cs.OK
    assert L.cs_batched_sort_pairs(cs.I64, None, None, None, None, 4, 0, None) == cs.OK


def test_hybrid_thresholds(cs):
    # measured merge / partition crossovers: 2^23 for 4- and 8-byte, 2^21 for
    # 16-byte elements; 0 for an unknown element type
    want = {cs.I32: 23, cs.U32: 23, cs.F32: 23, cs.U64: 23, cs.I64: 23, cs.F64: 23, cs.REC_I32_I32: 23,
            cs.REC_I64_I32: 21}
    for dt, lg in want.items():
        assert cs.hybrid_threshold(dt) == 1 << lg, dt
    assert cs.hybrid_threshold(99) == 0
