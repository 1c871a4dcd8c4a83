"""GPU parity of the ragged (CSR-segmented) sort against the CPU oracle:
persistent bucketed kernel (graph and direct launch), the plain baseline,
every tuner variant, pointer updates of a captured graph, graph replays,
and the full-size ragged configuration (int32 and fp32 totalOrder)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def upload(cs, arr):
    return cs.from_numpy(arr)


def check_seg(cs, orc, rt, dtype, data, off, plan=None, flags=0, plain=False, inplace=False):
    src = upload(cs, data)
    doff = upload(cs, off)
    dst = src if inplace else cs.device_buffer(dtype, data.size)
    if plain:
        lens = np.diff(off)
        cs.segmented_sort_plain(dtype, src.ptr, dst.ptr, doff.ptr, off.size - 1, int(lens.max(initial=0)))
    else:
        p = plan or cs.SegPlan(dtype, off, flags)
        p.sort(src.ptr, dst.ptr, doff.ptr)
    rt.synchronize()
    res = cs.to_numpy(dst, dtype, data.size)
    rc, msg = orc.validate_segmented(dtype, data, res, off)
    assert rc == 0, msg
    return res


def random_offsets(rng, segs, max_len):
    lens = rng.integers(0, max_len + 1, segs)
    lens[rng.integers(0, segs, max(1, segs // 10))] = 0
    lens[rng.integers(0, segs, max(1, segs // 10))] = 1
    return np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)


@pytest.mark.parametrize("dtype", ["I32", "F32", "REC_I32_I32", "U64", "REC_I64_I32", "U32", "I64", "F64"])
def test_random_ragged_all_variants(cs, orc, gpu, dtype):
    dt = getattr(cs, dtype)
    rng = np.random.default_rng(7 + dt)
    mx = cs.segmented_max_len(dt)
    off = random_offsets(rng, 600, mx)
    off = np.concatenate([off, off[-1] + np.array([mx, 2 * mx, 2 * mx + 3])])   # full-length segments too
    for dist in (orc.D_UNIFORM, orc.D_EXTREMES6, orc.D_DUP50):
        data = orc.generate(dt, dist, int(off[-1]), int(off[-1]), seed=dist)
        check_seg(cs, orc, gpu, dt, data, off)
        check_seg(cs, orc, gpu, dt, data, off, flags=cs.SEG_NO_GRAPH)
        check_seg(cs, orc, gpu, dt, data, off, plain=True)
    data = orc.generate(dt, orc.D_UNIFORM, int(off[-1]), int(off[-1]), seed=3)
    for variant in (0, 1, 2, 3):   # 3: cross-lane engine, 4-byte keys without payload, buckets 5..11
        for b in range(16):
            cs.segmented_set_variant(dt, b, variant)
        try:
            check_seg(cs, orc, gpu, dt, data, off)
        finally:
            for b in range(16):
                cs.segmented_set_variant(dt, b, -1)


def test_graph_replays_and_pointer_updates(cs, orc, gpu):
    rng = np.random.default_rng(11)
    off = random_offsets(rng, 2000, 3000)
    n = int(off[-1])
    plan = cs.SegPlan(cs.I32, off)
    datas = [orc.generate(cs.I32, orc.D_UNIFORM, n, n, seed=s) for s in (1, 2)]
    srcs = [upload(cs, d) for d in datas]
    doff = upload(cs, off)
    outs = [cs.device_buffer(cs.I32, n) for _ in range(2)]
    want = [orc.sort_segmented(cs.I32, d, off) for d in datas]
    for rep in range(6):                  # replay, switching buffers every call
        i = rep % 2
        plan.sort(srcs[i].ptr, outs[i].ptr, doff.ptr)
        gpu.synchronize()
        got = cs.to_numpy(outs[i], cs.I32, n)
        assert (got == want[i]).all(), f"replay {rep}"
    # in place on the same plan
    plan.sort(srcs[0].ptr, srcs[0].ptr, doff.ptr)
    gpu.synchronize()
    assert (cs.to_numpy(srcs[0], cs.I32, n) == want[0]).all()


def test_empty_and_degenerate(cs, orc, gpu):
    off = np.zeros(5, dtype=np.int64)             # four empty segments
    plan = cs.SegPlan(cs.I32, off)
    assert plan.info()["work_items"] >= 1
    buf = cs.device_buffer(cs.I32, 4)
    doff = upload(cs, off)
    plan.sort(buf.ptr, buf.ptr, doff.ptr)
    gpu.synchronize()
    plan0 = cs.SegPlan(cs.I32, np.zeros(1, dtype=np.int64))   # no segments at all
    plan0.sort(buf.ptr, buf.ptr, doff.ptr)
    with pytest.raises(cs.CsError):
        cs.SegPlan(cs.I32, np.array([0, cs.segmented_max_len(cs.I32) + 1], dtype=np.int64))
    with pytest.raises(cs.CsError):
        cs.SegPlan(cs.I32, np.array([0, 5, 3], dtype=np.int64))


def test_tuner_picks_a_valid_variant(cs, orc, gpu):
    cs.segmented_tune(cs.I32, 2)
    for b in range(1, 14):
        v = cs.segmented_variant(cs.I32, b)
        assert v in (0, 1, 2, 3)
        assert v != 3 or 5 <= b <= 11      # the cross-lane engine covers W = 32..2048
    rng = np.random.default_rng(5)
    off = random_offsets(rng, 3000, 8192)
    data = orc.generate(cs.I32, orc.D_UNIFORM, int(off[-1]), int(off[-1]), seed=9)
    check_seg(cs, orc, gpu, cs.I32, data, off)


@pytest.mark.parametrize("dtype,dist", [("I32", "D_UNIFORM"), ("F32", "D_NORMAL_SPECIAL")])
def test_config3_full_ragged(cs, orc, gpu, dtype, dist):
    """100,000 segments, lengths log-uniform over 1..8192 (seed 42): the
    persistent path and the plain per-array launch both match std::sort."""
    dt = getattr(cs, dtype)
    off = cs.offsets_from_lengths(cs.ragged_lengths(100000, 8192, 42))
    n = int(off[-1])
    data = orc.generate(dt, getattr(orc, dist), n, n, seed=42)
    a = check_seg(cs, orc, gpu, dt, data, off)
    b = check_seg(cs, orc, gpu, dt, data, off, plain=True)
    assert a.tobytes() == b.tobytes()
