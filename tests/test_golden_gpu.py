"""The HIP library against the committed golden fixtures (tests/golden/):
inputs uploaded from the fixtures, sorted through the C ABI, compared with
the committed outputs; larger cases generated on the device and compared by
the SHA-256 of the sorted keys."""
import hashlib
import importlib.util
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SMALL_N = [1, 2, 31, 32, 33, 63, 64, 65, 129]
ARRAYS = 2


def _make_golden():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg


@pytest.fixture(scope="module")
def small():
    with np.load(os.path.join(GOLDEN, "small.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("dtname", ["I32", "F32", "REC_I32_I32"])
@pytest.mark.parametrize("inplace", [False, True])
def test_small_fixtures(cs, gpu, small, dtname, inplace):
    dt = getattr(cs, dtname)
    for dist_name in ("UNIFORM", "SORTED", "REVERSED", "CONSTANT", "DISTINCT16", "PERTURBED", "ORGAN_PIPE",
                      "ALT_EXTREMES", "DUP50"):
        for n in SMALL_N:
            k = f"{dtname}_{dist_name}_{n}"
            inp, want = small[k + "_in"], small[k + "_out"]
            print(f"golden case {k} inplace={inplace}", flush=True)
            src = cs.from_numpy(inp)
            dst = src if inplace else cs.device_buffer(dt, inp.size)
            cs.batched_sort(dt, src.ptr, dst.ptr, ARRAYS, n)
            gpu.synchronize()
            got = cs.to_numpy(dst, dt, inp.size)
            keys = got["key"] if got.dtype.names else got
            assert keys.tobytes() == want.tobytes(), k
            if got.dtype.names:   # every record of each array kept, none duplicated
                for a in range(ARRAYS):
                    s = slice(a * n, (a + 1) * n)
                    assert sorted(got[s].tolist()) == sorted(inp[s].tolist()), k


def test_hashes(cs, gpu):
    mg = _make_golden()
    with open(os.path.join(GOLDEN, "hashes.json")) as f:
        hashes = json.load(f)["sha256"]
    for name, dtname, dist, n, arrays in mg.HASH_CASES:
        dt = getattr(cs, dtname)
        count = n * (arrays or 1)
        print(f"golden hash case {name}", flush=True)
        buf = cs.device_buffer(dt, count)
        cs.generate(dt, dist, n, count, 42, buf.ptr)
        if arrays:
            cs.batched_sort(dt, buf.ptr, buf.ptr, arrays, n)
        else:
            cs.sort_hybrid(dt, buf.ptr, n)
        gpu.synchronize()
        got = cs.to_numpy(buf, dt, count)
        keys = np.ascontiguousarray(got["key"] if got.dtype.names else got)
        assert hashlib.sha256(keys.tobytes()).hexdigest() == hashes[name], name
