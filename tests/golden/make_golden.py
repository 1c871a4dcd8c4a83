"""Writes the committed golden fixtures of the sorting path from the CPU oracle.

The reference ships no data files: every expected output in its tests is
computed at run time by host std::sort (reference bench.cu:300-301,
test.cu:46-48).  These fixtures freeze that oracle's answers for the seeded
inputs so that the GPU tests can also compare against committed data:

  small.npz    inputs and expected outputs of the batched sort for the nine
               seeded distributions x n in {1, 2, 31, 32, 33, 63, 64, 65, 129},
               two arrays per case, for int32 keys, fp32 keys (IEEE totalOrder)
               and int32+int32 records (expected output: the key sequence;
               the order of equal keys is free, as in the reference).
  hashes.json  SHA-256 of the oracle's sorted output (keys only for records)
               for larger cases: n = 1000, the 4096 x 1024 batches and two
               single large arrays.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402

SEED = 42
SMALL_N = [1, 2, 31, 32, 33, 63, 64, 65, 129]
SMALL_ARRAYS = 2
SMALL_DTYPES = {"I32": oracle.I32, "F32": oracle.F32, "REC_I32_I32": oracle.REC_I32_I32}

# (name, dtype, dist, n, arrays); arrays == 0 means one large array of n keys
HASH_CASES = (
    [(f"batched I32 {oracle.DIST_NAMES[d]} 2x1000", "I32", d, 1000, 2) for d in oracle.NINE]
    + [
        ("batched I32 UNIFORM 4096x1024", "I32", oracle.D_UNIFORM, 1024, 4096),
        ("batched REC_I32_I32 UNIFORM 4096x1024", "REC_I32_I32", oracle.D_UNIFORM, 1024, 4096),
        ("batched F32 NORMAL_SPECIAL 512x1024", "F32", oracle.D_NORMAL_SPECIAL, 1024, 512),
        ("batched U64 DUP50 1024x2048", "U64", oracle.D_DUP50, 2048, 1024),
        ("large I32 UNIFORM 2^22", "I32", oracle.D_UNIFORM, 1 << 22, 0),
        ("large U64 UNIFORM 2^21", "U64", oracle.D_UNIFORM, 1 << 21, 0),
    ]
)


def keys_of(dtype: int, arr: np.ndarray) -> np.ndarray:
    return arr["key"] if arr.dtype.names else arr


def small_key(dtname: str, dist: int, n: int) -> str:
    return f"{dtname}_{oracle.DIST_NAMES[dist]}_{n}"


def sorted_case(dtype: int, dist: int, n: int, arrays: int) -> tuple[np.ndarray, np.ndarray]:
    if arrays == 0:
        inp = oracle.generate(dtype, dist, n, n, SEED)
        return inp, oracle.sort_large(dtype, inp)
    inp = oracle.generate(dtype, dist, n, n * arrays, SEED)
    return inp, oracle.sort_batched(dtype, inp, n)


def digest(dtype: int, out: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(keys_of(dtype, out)).tobytes()).hexdigest()


def build_small() -> dict[str, np.ndarray]:
    data = {}
    for dtname, dt in SMALL_DTYPES.items():
        for dist in oracle.NINE:
            for n in SMALL_N:
                inp, out = sorted_case(dt, dist, n, SMALL_ARRAYS)
                k = small_key(dtname, dist, n)
                data[k + "_in"] = inp
                data[k + "_out"] = np.ascontiguousarray(keys_of(dt, out))
    return data


def build_hashes() -> dict[str, str]:
    res = {}
    for name, dtname, dist, n, arrays in HASH_CASES:
        dt = getattr(oracle, dtname)
        _, out = sorted_case(dt, dist, n, arrays)
        res[name] = digest(dt, out)
    return res


def main() -> None:
    oracle.lib()
    np.savez_compressed(os.path.join(HERE, "small.npz"), **build_small())
    with open(os.path.join(HERE, "hashes.json"), "w") as f:
        json.dump({"seed": SEED, "sha256": build_hashes()}, f, indent=1, sort_keys=True)
        f.write("\n")
    print("wrote", os.path.join(HERE, "small.npz"), "and hashes.json")


if __name__ == "__main__":
    main()
