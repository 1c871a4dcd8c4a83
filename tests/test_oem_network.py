"""The odd-even merge network that replaces bitonic stages 1..e inside each
lane of key-only register tiles (cuda-sort_amd/csrc/plan.hpp, make_oem):
exhaustive 0-1 principle (every 0/1 input of n wires comes out sorted,
bit-sliced 64 inputs at a time), Batcher's comparator counts, and a planted
mistake (one comparator removed) that the check must catch."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = tmp_path_factory.mktemp("oem") / "oem_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", os.path.join(HERE, "sim", "oem_check.cpp"),
                    "-o", str(exe)], check=True)
    return str(exe)


@pytest.mark.parametrize("n", [2, 4, 8, 16, 32])
def test_network_sorts_every_01_input(checker, n):
    r = subprocess.run([checker, str(n)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.parametrize("drop", [0, 10, 40, 62])
def test_planted_mistake_is_caught(checker, drop):
    r = subprocess.run([checker, "16", str(drop)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "unsorted" in r.stdout, r.stdout
