"""The serving layer's batch admission (bounded bypass of a request that does
not fit), checked on the host by fixed cases and a randomised discrete-event
run of a server (tests/sim/admission_test.cpp)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_admission_policy(tmp_path):
    exe = tmp_path / "admission_test"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", os.path.join(HERE, "sim", "admission_test.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
