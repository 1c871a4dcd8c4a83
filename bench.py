#!/usr/bin/env python3
"""Benchmark of the MI355X sorting path.

Default workload (the headline configuration): one batched sort of 4096
independent int32 arrays x 1024 keys, keys uniform over the full int32 range
from seed 42, one lane group per array, through the C ABI
(cs_batched_sort, method auto).  A step is one sort of the whole batch from
the pristine input buffer into the output buffer (one HBM read + one write,
the reference's in-place sort moves the same bytes); inputs are resident in
HBM before the timed region starts.

Prints one JSON line.  `value` is Gkeys/s of the whole job; `roofline`
prices the batched-sort kernel against MI355X's 8 TB/s HBM with the
algorithmic 8 bytes per int32 key (read + write); `cpu_baseline` is the
oracle's per-array std::sort over all host cores on the same input.

Other workloads (--workload) print their own lines for the other
configurations; the default run is the one the headline metric is quoted on.
"""
from __future__ import annotations

import argparse
import json
import time
import os
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cuda-sort_amd"))

METRIC = "Gkeys/s & %HBM roofline, batched int32 (fixed vs ragged); serve p99 latency"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="batched_i32",
                    choices=["batched_i32", "batched_rec32", "batched_pairs64", "copy", "ragged_i32",
                             "ragged_f32", "large_i32", "large_u64", "serve"])
    ap.add_argument("--rate", type=float, default=20000.0)
    ap.add_argument("--seconds", type=float, default=10.0)
    ap.add_argument("--budget-gb", type=float, default=32.0)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--large-method", default="hybrid", choices=["hybrid", "merge", "partition"])
    ap.add_argument("--large-log2", type=int, default=0, help="log2 of the array length (0: 28 for int32, 27 for uint64)")
    ap.add_argument("--large-dist", default="uniform", choices=["uniform", "presorted"])
    ap.add_argument("--segments", type=int, default=100000)
    ap.add_argument("--max-len", type=int, default=8192)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--arrays", type=int, default=4096)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--method", default="auto")
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-trials", type=int, default=7)
    return ap.parse_args()


def load_traffic(name: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by a copy-kernel calibration), committed under
    profiles/ as JSON; None when absent."""
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def warm_up(step, count, min_seconds=0.1):
    """At least `count` untimed steps and at least `min_seconds` of them: a
    short batch of kernels on a cold GPU runs at a lower clock (measured on
    MI355X: 18.3 us per 4096x1024 batch over the first 200 launches, 13.2 us
    once the clock has ramped)."""
    import hiprt
    t0 = time.perf_counter()
    done = 0
    while done < count or time.perf_counter() - t0 < min_seconds:
        step()
        done += 1
        if done % 64 == 0:
            hiprt.synchronize()   # keep the host from queueing far ahead of the GPU
    hiprt.synchronize()
    return done


def time_launches(step, steps, stream, hold_us=200.0):
    """Milliseconds per step over `steps` calls of `step` on `stream`, from HIP
    events recorded on that stream.  A short bounded hold kernel
    (cs_stream_hold) keeps the stream busy while the host enqueues the first
    launches, so that the run starts with launches already queued: a ctypes
    call and launch take ~10 us, close to one 4096x1024 sort, and without the
    hold the first launch of a short run lands after a host gap.  The hold is
    kept short because an idle GPU lowers its clock within milliseconds
    (measured on MI355X: a 17 ms hold before 400 sorts made them 10 %
    slower).  Warns when the host enqueued more slowly than the device ran,
    in which case host gaps may be part of the time."""
    import cusort_hip as cs
    import hiprt
    cs.stream_hold(hold_us, stream.handle)
    start, end = hiprt.Event(), hiprt.Event()
    start.record(stream)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    end.record(stream)
    enqueue_us = (time.perf_counter() - t0) * 1e6
    end.synchronize()
    ms = start.elapsed_ms(end) / steps
    if enqueue_us > hold_us + ms * 1e3 * steps:
        print(f"bench: enqueueing {steps} steps took {enqueue_us:.0f} us, longer than the hold and the device time "
              f"({hold_us:.0f} + {ms * 1e3 * steps:.0f} us): host gaps may be part of the time", file=sys.stderr)
    return ms


def bench_batched(args, dtype_name: str):
    import cusort_hip as cs
    import hiprt

    dt = cs.DTYPE_NAMES[dtype_name]
    method = cs.METHODS.index(args.method)
    n, arrays = args.n, args.arrays
    count = n * arrays
    hiprt.set_device(0)
    stream = hiprt.Stream()                      # default priority
    src = cs.device_buffer(dt, count)
    dst = cs.device_buffer(dt, count)
    seed = 42
    cs.generate(dt, cs.D_UNIFORM, n, count, seed, src.ptr, stream.handle)
    stream.synchronize()

    def step():
        cs.batched_sort(dt, src.ptr, dst.ptr, arrays, n, method, args.threads, 0, stream.handle)

    warm_up(step, args.warmup)
    # HIP events on the stream the kernels run on bracket exactly K launches
    ms = time_launches(step, args.steps, stream)
    hiprt.synchronize()

    # correctness of the measured output (outside the timed region)
    res = cs.to_numpy(dst, dt, count)
    ok = None
    cpu = None
    bytes_per_key = 8 if dt in (cs.I32, cs.U32, cs.F32) else 16
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        host_in = oracle.generate(dt, oracle.D_UNIFORM, n, count, seed)
        ok = oracle.validate_batched(dt, host_in, res, n)[0] == 0
        secs, threads = oracle.time_batched(dt, host_in, n, threads=0, trials=args.cpu_trials)
        cpu = {"value": count / secs / 1e9, "unit": "Gkeys/s", "cores": threads, "kind": "port",
               "sample": f"the full {arrays}x{n} batch, per-array std::sort over {threads} host threads, "
                         f"median of {args.cpu_trials} trials"}
    keys_per_s = count / (ms * 1e-3)
    alg_bytes = count * bytes_per_key
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    name = f"batched_{dtype_name}_{arrays}x{n}"
    line = {
        "metric": METRIC,
        "value": keys_per_s / 1e9,
        "unit": "Gkeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32" if dt == cs.I32 else dtype_name,
        "data": "synthetic",
        "config": {"workload": f"batched sort, {arrays} arrays x {n} {dtype_name} keys, uniform full range, "
                               f"seed {seed}, method {args.method}",
                   "arrays": arrays, "n": n, "method": args.method, "parallelism": "one GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic(name)},
        "cpu_baseline": cpu,
        "validated": ok,
    }
    return line


def bench_pairs(args):
    """Key / payload pairs in two arrays (int64 keys, int32 payloads), the
    SoA form of the {int64 key, int32 payload} record: --arrays x --n pairs
    sorted by cs_batched_sort_pairs, timed like the batched sort, next to the
    padded 16-byte record sort of the same data (cs_batched_sort,
    CS_REC_I64_I32)."""
    import numpy as np
    import cusort_hip as cs
    import hiprt

    n, arrays = args.n, args.arrays
    count = n * arrays
    hiprt.set_device(0)
    stream = hiprt.Stream()
    rec = cs.device_buffer(cs.REC_I64_I32, count)
    rec_out = cs.device_buffer(cs.REC_I64_I32, count)
    seed = 42
    cs.generate(cs.REC_I64_I32, cs.D_UNIFORM, n, count, seed, rec.ptr, stream.handle)
    stream.synchronize()
    host = cs.to_numpy(rec, cs.REC_I64_I32, count)
    keys = np.ascontiguousarray(host["key"])
    vals = np.ascontiguousarray(host["payload"]).astype(np.int32)
    kin, vin = cs.from_numpy(keys), cs.from_numpy(vals)
    kout, vout = cs.from_numpy(np.zeros_like(keys)), cs.from_numpy(np.zeros_like(vals))

    def timed(step):
        warm_up(step, args.warmup)
        return time_launches(step, args.steps, stream)

    ms = timed(lambda: cs.batched_sort_pairs(cs.I64, kin.ptr, vin.ptr, kout.ptr, vout.ptr, arrays, n,
                                             stream.handle))
    ms_rec = timed(lambda: cs.batched_sort(cs.REC_I64_I32, rec.ptr, rec_out.ptr, arrays, n, 0, 0, 0,
                                           stream.handle))
    got_k, got_v = kout.download(np.int64, count), vout.download(np.int32, count)
    want = cs.to_numpy(rec_out, cs.REC_I64_I32, count)
    same = got_k.tobytes() == np.ascontiguousarray(want["key"]).tobytes() and \
        got_v.tobytes() == np.ascontiguousarray(want["payload"]).astype(np.int32).tobytes()
    ok, cpu = None, None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        ok = same and oracle.validate_batched(cs.REC_I64_I32, host, want, n)[0] == 0
        secs, threads = oracle.time_batched(cs.REC_I64_I32, host, n, threads=0, trials=args.cpu_trials)
        cpu = {"value": count / secs / 1e9, "unit": "Gkeys/s", "cores": threads, "kind": "port",
               "sample": f"the full {arrays}x{n} batch as records, per-array std::sort over {threads} host "
                         f"threads, median of {args.cpu_trials} trials"}
    alg_bytes = count * 24   # 12 bytes per pair, read and written
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    return {"metric": METRIC, "value": count / (ms * 1e-3) / 1e9, "unit": "Gkeys/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "int64+int32", "data": "synthetic",
            "config": {"workload": f"batched pair sort, {arrays} arrays x {n} (int64 key, int32 payload) pairs in "
                                   f"two arrays, uniform keys, seed {seed}", "arrays": arrays, "n": n},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None},
            "record_sort_ms": ms_rec, "same_as_record_sort": same,
            "cpu_baseline": cpu, "validated": ok}


def bench_copy(args):
    """Plain device copy (cs_device_copy) of --copy-mib MiB: the measured HBM
    copy bandwidth, and the known-traffic kernel that calibrates the
    FETCH_SIZE / WRITE_SIZE counters."""
    import cusort_hip as cs
    import hiprt

    hiprt.set_device(0)
    stream = hiprt.Stream()
    nbytes = args.copy_mib << 20
    src, dst = hiprt.Buffer(nbytes), hiprt.Buffer(nbytes)
    src.fill(1)
    copy = lambda: cs.device_copy(dst.ptr, src.ptr, nbytes, stream.handle)
    warm_up(copy, args.warmup)
    ms = time_launches(copy, args.steps, stream)
    gbs = 2 * nbytes / (ms * 1e-3) / 1e9
    return {"metric": "device copy bandwidth (read + write)", "value": gbs, "unit": "GB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
            "bytes_per_launch": 2 * nbytes, "config": {"workload": f"copy {args.copy_mib} MiB"}}


def bench_ragged(args, dtype_name: str):
    """Ragged batch: --segments arrays, lengths log-uniform over 1..--max-len
    (seed 42), CSR offsets; persistent bucketed graph launch vs the plain
    per-array launch (one lane group per segment, longest-segment tiling)."""
    import numpy as np
    import cusort_hip as cs
    import hiprt

    dt = cs.DTYPE_NAMES[dtype_name]
    hiprt.set_device(0)
    stream = hiprt.Stream()
    lens = cs.ragged_lengths(args.segments, args.max_len, 42)
    off = cs.offsets_from_lengths(lens)
    total = int(off[-1])
    dist = cs.D_UNIFORM if dt != cs.F32 else cs.D_NORMAL_SPECIAL
    src, dst = cs.device_buffer(dt, total), cs.device_buffer(dt, total)
    doff = cs.from_numpy(off)
    cs.generate(dt, dist, total, total, 42, src.ptr, stream.handle)
    if not args.no_tune:
        cs.segmented_tune(dt, 3, stream.handle)
    plan = cs.SegPlan(dt, off)
    stream.synchronize()

    def timed(fn):
        warm_up(fn, args.warmup)
        return time_launches(fn, args.steps, stream)

    ms_p = timed(lambda: plan.sort(src.ptr, dst.ptr, doff.ptr, stream.handle))
    res_p = cs.to_numpy(dst, dt, total)
    ms_b = timed(lambda: cs.segmented_sort_plain(dt, src.ptr, dst.ptr, doff.ptr, args.segments, args.max_len,
                                                 stream.handle))
    res_b = cs.to_numpy(dst, dt, total)
    same = res_p.tobytes() == res_b.tobytes()
    ok, cpu = None, None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        host_in = oracle.generate(dt, dist, total, total, 42)
        ok = oracle.validate_segmented(dt, host_in, res_p, off)[0] == 0
        secs, threads = oracle.time_segmented(dt, host_in, off, threads=0, trials=3)
        cpu = {"value": total / secs / 1e9, "unit": "Gkeys/s", "cores": threads, "kind": "port",
               "sample": f"the full ragged batch ({total} keys), per-segment std::sort over {threads} host "
                         "threads balanced by n log n, median of 3 trials"}
    bytes_per_key = cs.DTYPE_SIZE[dt] * 2
    alg = total * bytes_per_key + off.nbytes
    info = plan.info()
    return {"metric": METRIC, "value": total / (ms_p * 1e-3) / 1e9, "unit": "Gkeys/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_p, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "int32" if dt == cs.I32 else "fp32 (totalOrder)",
            "data": "synthetic",
            "config": {"workload": f"ragged batch, {args.segments} segments, lengths log-uniform 1..{args.max_len} "
                                   f"(seed 42), {total} {dtype_name} keys, persistent bucketed hipGraph launch",
                       "segments": args.segments, "total_keys": total, "work_items": info["work_items"],
                       "grid": info["grid"], "variants": info["variants"][:14]},
            "plain_per_array": {"ms_per_step": ms_b, "value": total / (ms_b * 1e-3) / 1e9,
                                "speedup_persistent_vs_plain": ms_b / ms_p, "identical_output": same},
            "roofline": {"bound": "hbm", "achieved": alg / (ms_p * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": alg / (ms_p * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "traffic": load_traffic(f"ragged_{dtype_name}_{args.segments}x{args.max_len}")},
            "cpu_baseline": cpu, "validated": ok}


def bench_large(args, dtype_name: str):
    """One array of 2^28 int32 or 2^27 uint64 keys (uniform, or 10% presorted
    4096-runs), sorted in place by the hybrid / merge / partition path.  Each
    step restores the pristine input (untimed, as bench.cu:335) and times the
    sort alone with events on the sort's stream."""
    import cusort_hip as cs
    import hiprt

    dt = cs.DTYPE_NAMES[dtype_name]
    log2n = args.large_log2 or (28 if dt in (cs.I32, cs.U32, cs.F32) else 27)
    n = 1 << log2n
    dist = cs.D_UNIFORM if args.large_dist == "uniform" else cs.D_PRESORTED_RUNS
    hiprt.set_device(0)
    stream = hiprt.Stream()
    pristine, work = cs.device_buffer(dt, n), cs.device_buffer(dt, n)
    cs.generate(dt, dist, n, n, 42, pristine.ptr, stream.handle)
    ws = cs.Workspace(dt, n)
    nbytes = n * cs.DTYPE_SIZE[dt]

    def sort():
        if args.large_method == "hybrid":
            cs.sort_hybrid(dt, work.ptr, n, ws, stream.handle)
        elif args.large_method == "merge":
            cs.sort_merge_ws(dt, work.ptr, n, ws, stream.handle)
        else:
            cs.sort_partition(dt, work.ptr, n, stream=stream.handle)

    def restore_and_sort():
        cs.device_copy(work.ptr, pristine.ptr, nbytes, stream.handle)
        sort()

    warm_up(restore_and_sort, args.warmup)
    total = 0.0
    for _ in range(args.steps):
        cs.device_copy(work.ptr, pristine.ptr, nbytes, stream.handle)
        total += time_launches(sort, 1, stream, hold_us=300.0)
    ms = total / args.steps
    ok, cpu = None, None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        host_in = oracle.generate(dt, dist, n, n, 42)
        got = cs.to_numpy(work, dt, n)
        ok = got.tobytes() == oracle.sort_large(dt, host_in).tobytes()
        secs, threads = oracle.time_large(dt, host_in, threads=0, trials=1)
        cpu = {"value": n / secs / 1e9, "unit": "Gkeys/s", "cores": threads, "kind": "port",
               "sample": f"the full array (2^{log2n} keys), __gnu_parallel::sort over {threads} host threads, 1 trial"}
    # Algorithmic HBM bytes: the merge path streams the whole array once for
    # the chunk sort and once per merge pass (read + write every key); the
    # partition path is given the whole-sort lower bound (one read + one
    # write), since its pass count depends on the bucket sizes.
    merge_like = args.large_method == "merge" or (args.large_method == "hybrid" and n < cs.hybrid_threshold(dt))
    chunk = 8192 if cs.DTYPE_SIZE[dt] == 4 else 4096
    passes = max(0, (n - 1).bit_length() - (chunk - 1).bit_length()) if merge_like else 0
    alg = (passes + 1) * 2 * nbytes if merge_like else 2 * nbytes
    alg_note = (f"chunk sort + {passes} merge passes, each reading and writing all {n} keys" if merge_like
                else "one read and one write of every key (whole-sort lower bound)")
    return {"metric": METRIC, "value": n / (ms * 1e-3) / 1e9, "unit": "Gkeys/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "int32" if dt == cs.I32 else "uint64", "data": "synthetic",
            "config": {"workload": f"large-array sort, 2^{log2n} {dtype_name} keys, {args.large_dist}, seed 42, "
                                   f"{args.large_method} path", "n": n, "method": args.large_method},
            "roofline": {"bound": "hbm", "achieved": alg / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "traffic": load_traffic(f"large_{dtype_name}_2p{log2n}_{'merge' if merge_like else 'partition'}"),
                         "algorithmic_bytes": alg, "basis": alg_note},
            "cpu_baseline": cpu, "validated": ok}


def bench_serve(args):
    """Serving loop: Poisson arrivals at --rate req/s for --seconds, sizes
    log-uniform over 16..65536 elements, {int64 key, int32 payload} records,
    device pool of --budget-gb GB (decimal), --slots in-flight batches.  The
    open-loop generator is cuda-sort_amd/cusort_serve_bench (C++)."""
    import subprocess
    exe = os.path.join(ROOT, "cuda-sort_amd", "cusort_serve_bench")
    cmd = [exe, "--rate", str(args.rate), "--seconds", str(args.seconds), "--budget-gb", str(args.budget_gb),
           "--slots", str(args.slots)]
    r = subprocess.run(cmd, capture_output=True, text=True, check=False)
    if r.returncode != 0 and not r.stdout.strip():
        raise RuntimeError(f"{exe} failed: {r.stderr[-2000:]}")
    res = json.loads(r.stdout.strip().splitlines()[-1])
    return {"metric": METRIC, "value": res["p99_ms"], "unit": "ms (p99 latency)", "n_gpus": 1,
            "steps": res["requests"], "warmup": 0, "ms_per_step": res["mean_ms"], "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "int64 key + int32 payload", "data": "synthetic",
            "config": {"workload": f"serving loop, Poisson {args.rate:g} req/s for {args.seconds:g} s, sizes "
                                   f"log-uniform 16..65536, {args.budget_gb:g} GB HBM budget, {args.slots} slots"},
            "serve": res, "validated": res["failures"] == 0}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(every array is independent: N GPUs would each run this workload on their own share)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched_i32":
        line = bench_batched(args, "i32")
    elif args.workload == "copy":
        line = bench_copy(args)
    elif args.workload == "batched_pairs64":
        line = bench_pairs(args)
    elif args.workload == "serve":
        line = bench_serve(args)
    elif args.workload.startswith("large"):
        line = bench_large(args, args.workload.split("_")[1])
    elif args.workload.startswith("ragged"):
        line = bench_ragged(args, args.workload.split("_")[1])
    else:
        line = bench_batched(args, "rec32")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
