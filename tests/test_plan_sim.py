"""Executes the register-tile schedule (cuda-sort_amd/csrc/plan.hpp) on the
host with the kernel's own layout/flip/LDS-address helpers (tests/sim/
plan_sim.cpp) and checks it is a sorting network by the 0-1 principle
(exhaustive for W <= 16), plus the record and padding invariants."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def sim(tmp_path_factory):
    exe = tmp_path_factory.mktemp("sim") / "plan_sim"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", os.path.join(HERE, "sim", "plan_sim.cpp"),
                    "-o", str(exe)], check=True)
    return str(exe)


# every (log2 W, log2 E) tiling the library instantiates for 4-byte keys,
# plus a few wider ones
CASES = [(w, e) for w in range(1, 15) for e in range(1, 6) if e <= w and w - e <= 10]


@pytest.mark.parametrize("lanesort", [0, 1])
@pytest.mark.parametrize("w,e", CASES)
def test_schedule_sorts(sim, w, e, lanesort):
    r = subprocess.run([sim, str(w), str(e), str(lanesort)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
