"""Writes tests/golden/ntt_golden.npz and tests/golden/ntt_checksums.json.

The reference holds no golden vectors (its tests draw unseeded random inputs
and compare against the boojum CPU FFT at run time, src/ntt.rs:391-611).
These fixtures are produced by this repo's oracle and, before being written,
every small case is checked against tests/pyref.py (a direct O(n^2)
statement of the semantics in Python big ints), so the committed values are
pinned by the mathematics, not by the oracle alone.

Inputs are raw u64 draws (non-canonical values included), as the reference's
tests use (src/ntt.rs:424).  Run:  python tests/golden/gen_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import oracle  # noqa: E402
import pyref  # noqa: E402

LOGS = [1, 2, 3, 4, 6, 9]
LDE = [(0, 0), (1, 0), (1, 1), (2, 3), (3, 0), (3, 7)]
COLS = 3
SEED = 20240601
PYREF_MAX_LOG = 6  # O(n^2) big-int check up to here; sampled Horner beyond

CHECKSUM_CASES = [  # (log_n, num_cols, seed)
    (12, 2, 11), (16, 1, 12), (20, 1, 13),
]


def case_key(log_n, bitrev, inverse, k, c):
    return f"out_L{log_n}_b{int(bitrev)}_i{int(inverse)}_k{k}_c{c}"


def checksum(arr):
    return hashlib.sha256(oracle.canon(arr).astype("<u8").tobytes()).hexdigest()


def main():
    data = {}
    for log_n in LOGS:
        n = 1 << log_n
        inp = np.stack([oracle.fill_raw(n, SEED + log_n, c) for c in range(COLS)])
        data[f"in_L{log_n}"] = inp
        for bitrev in (False, True):
            for inverse in (False, True):
                for k, c in LDE:
                    outs = np.stack([oracle.ntt_column(inp[col], log_n, bitrev, inverse, k, c) for col in range(COLS)])
                    if log_n <= PYREF_MAX_LOG:
                        for col in range(COLS):
                            ref = pyref.transform(inp[col].tolist(), log_n, bitrev, inverse, k, c)
                            assert [int(v) for v in outs[col]] == ref, (log_n, bitrev, inverse, k, c, col)
                    data[case_key(log_n, bitrev, inverse, k, c)] = outs
    np.savez_compressed(os.path.join(HERE, "ntt_golden.npz"), **data)

    sums = []
    for log_n, cols, seed in CHECKSUM_CASES:
        n = 1 << log_n
        inp = np.concatenate([oracle.fill_uniform(n, seed, c) for c in range(cols)])
        for bitrev in (False, True):
            for inverse in (False, True):
                out = oracle.batch_ntt(inp, log_n, cols, n, n, bitrev, inverse, threads=8)
                sums.append({"log_n": log_n, "num_cols": cols, "seed": seed, "bitrev_inputs": bitrev,
                             "inverse": inverse, "sha256": checksum(out)})
    with open(os.path.join(HERE, "ntt_checksums.json"), "w") as f:
        json.dump(sums, f, indent=1)
    print("wrote", len(data), "arrays and", len(sums), "checksums")


if __name__ == "__main__":
    main()
