"""cuda-sort_amd/cusort_bench, the reference-compatible benchmark driver
(bench.cu:44-136): option handling on the CPU; on the GPU a check-only sweep
and a timed run whose CSV has exactly the reference's columns (bench.cu:240-271)."""
import csv
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cuda-sort_amd", "cusort_bench")
COLUMNS = ("method,distribution,N,arrays,total_values,working_values,threads,items_per_thread,"
           "testsorted,fullsm,asynccpy,shared_bytes,trial,E_s,X_s,E_wall_s,"
           "wall_X_s,time_per_array_s,time_per_value_s,validated,cuda_runtime,thrust_version,cub_version,"
           "registers_per_thread,active_blocks_per_sm,local_bytes_per_thread").split(",")


def run(*args, timeout=600):
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=timeout)


@pytest.fixture(scope="module")
def exe():
    if not os.path.exists(EXE):
        pytest.skip("cusort_bench is not built (make -C cuda-sort_amd)")
    return EXE


def test_help(exe):
    r = run("--help")
    assert r.returncode == 0 and "--max-threads" in r.stdout and "--check-only" in r.stdout


@pytest.mark.parametrize("args,msg", [
    (["--mode", "tiny"], "Invalid --mode"),
    (["--type", "float"], "Invalid --type"),
    (["--max-threads", "48"], "multiple of 32"),
    (["--min-size", "9", "--max-size", "8"], "exceeds --max-size"),
    (["--methods", "bitonic_v9"], "Unknown --methods entry"),
    (["--methods", "bitonic_v3,"], "Empty name"),
    (["--arrays", "0"], "Invalid positive integer"),
    (["--repetitions"], "Missing value"),
    (["--frobnicate", "3"], "Unknown option"),
])
def test_option_errors(exe, args, msg):
    r = run(*args)
    assert r.returncode == 1 and msg in r.stderr, r.stderr


@pytest.mark.gpu
def test_check_only_sweep(exe, tmp_path):
    r = run("--check-only", "--type", "both", "--max-size", "130", "--output", str(tmp_path / "chk"))
    assert r.returncode == 0, r.stderr[-2000:]
    passes = [l for l in r.stdout.splitlines() if l.startswith("PASS ")]
    methods = {l.split()[1] for l in passes}
    assert {"bitonic_warp", "bitonic_v1", "bitonic_v2", "bitonic_v3", "bitonic_v4_xor", "bitonic_v5_static",
            "bitonic_v6_register", "cub_block_merge", "thrust_segment_loop"} <= methods
    assert any(" EXTREMES" in l for l in passes)
    for t in ("int", "record"):
        with open(tmp_path / f"chk_{t}_small.csv") as f:
            assert f.readline().strip().split(",") == COLUMNS


@pytest.mark.gpu
def test_timed_csv(exe, tmp_path):
    r = run("--type", "int", "--min-size", "1000", "--max-size", "1025", "--repetitions", "2",
            "--methods", "bitonic_v3,bitonic_v6_register,cub_block_merge", "--output", str(tmp_path / "t"))
    assert r.returncode == 0, r.stderr[-2000:]
    with open(tmp_path / "t_int_small.csv") as f:
        rows = list(csv.DictReader(f))
    assert rows and list(rows[0].keys()) == COLUMNS
    assert {row["method"] for row in rows} == {"bitonic_v3", "bitonic_v6_register", "cub_block_merge"}
    for row in rows:
        assert float(row["E_s"]) > 0 and row["validated"] == "1"
        n, a = int(row["N"]), int(row["arrays"])
        assert abs(float(row["time_per_value_s"]) - float(row["E_s"]) / (n * a)) < 1e-12


@pytest.mark.gpu
def test_large_mode(exe, tmp_path):
    r = run("--mode", "large", "--type", "both", "--max-size", str(1 << 17), "--repetitions", "1",
            "--methods", "large_partition,large_merge,large_hybrid", "--output", str(tmp_path / "l"))
    assert r.returncode == 0, r.stderr[-2000:]
    with open(tmp_path / "l_record_large.csv") as f:
        rows = list(csv.DictReader(f))
    assert {row["method"] for row in rows} == {"large_partition", "large_merge", "large_hybrid"}
