"""Executes the cross-lane register bitonic engine's schedule
(cuda-sort_amd/csrc/bitonic_xlane.hpp, xlane_plan.hpp) on the host, lane by
lane with the hardware's DPP / permlane-swap semantics and the kernel's own
LDS addresses (tests/sim/xlane_sim.cpp).  The in-lane odd-even sorter is
checked over all 2^32 binary inputs, every merge stage over every pair of
sorted binary halves, W <= 128 over every binary state, and random integer
inputs against std::sort."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def sim(tmp_path_factory):
    exe = tmp_path_factory.mktemp("xsim") / "xlane_sim"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", os.path.join(HERE, "sim", "xlane_sim.cpp"),
                    "-o", str(exe)], check=True)
    return str(exe)


@pytest.mark.parametrize("order", ["oem", "oemdf"])   # engine's level order, register tile's depth-first order
def test_lane_sorter_zero_one_exhaustive(sim, order):
    r = subprocess.run([sim, "5", "11", order], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr


# (log2 W, first stage through an LDS window): the library's instantiations
# (windows from stage 9) and the all-register / all-window variants of W = 1024
CASES = [(w, 9) for w in range(5, 12)] + [(10, 10), (10, 11), (9, 11), (8, 6), (11, 11), (11, 10)]


@pytest.mark.parametrize("logw,ldsk", CASES)
def test_schedule_sorts(sim, logw, ldsk):
    r = subprocess.run([sim, str(logw), str(ldsk)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
