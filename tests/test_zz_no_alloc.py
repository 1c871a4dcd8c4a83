"""cs_sort_merge_static, the allocation-free variant of cs_sort_merge: on the
GPU its output equals cs_sort_merge's bit for bit on two calls in a row (the
second smaller, over the first call's leftovers in the fixed scratch), it
takes no memory from the library's pool, and a size beyond its fixed scratch
is refused before anything runs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def sort_both(cs, orc, gpu, dtype, n, seed):
    host = orc.generate(dtype, orc.D_UNIFORM, n, n, seed)
    a = cs.from_numpy(host)
    b = cs.from_numpy(host)
    gpu.synchronize()
    before = cs.memory_stats()
    cs.sort_merge_static(dtype, a.ptr, n)
    gpu.synchronize()
    assert cs.memory_stats() == before, "the static entry point allocated"
    cs.sort_merge(dtype, b.ptr, n)
    gpu.synchronize()
    got, want = cs.to_numpy(a, dtype, n), cs.to_numpy(b, dtype, n)
    assert got.tobytes() == want.tobytes(), f"dtype={dtype} n={n} seed={seed}"
    return got


@pytest.mark.parametrize("dtype", ["I32", "U64", "REC_I32_I32", "F32"])
def test_static_matches_merge_twice(cs, orc, gpu, dtype):
    dt = getattr(cs, dtype)
    cap = cs.lib().cs_sort_merge_static_max_bytes()
    esz = {cs.I32: 4, cs.F32: 4, cs.U64: 8, cs.REC_I32_I32: 8}[dt]
    n1 = min(3 << 20, cap // esz // 2) + 17
    n2 = 100003
    got1 = sort_both(cs, orc, gpu, dt, n1, 5)
    got2 = sort_both(cs, orc, gpu, dt, n2, 6)
    if dt in (cs.I32, cs.U64):
        assert got1.tobytes() == np.sort(orc.generate(dt, orc.D_UNIFORM, n1, n1, 5)).tobytes()
        assert got2.tobytes() == np.sort(orc.generate(dt, orc.D_UNIFORM, n2, n2, 6)).tobytes()


def test_static_refuses_beyond_capacity(cs, gpu):
    cap = cs.lib().cs_sort_merge_static_max_bytes()
    n = cap // 4 + 1
    buf = cs.device_buffer(cs.I32, 16)
    before = cs.memory_stats()
    assert cs.sort_merge_static_status(cs.I32, buf.ptr, n) == cs.E_CAPACITY
    assert cs.memory_stats() == before
