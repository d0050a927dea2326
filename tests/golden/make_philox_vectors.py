"""Writes tests/golden/philox4x64.json: raw Philox4x64-10 outputs from numpy's independent
implementation (np.random.Philox), for counters and keys of the form the library uses
(counter = (step, atom, 0, 0), key = (seed, 0)) plus a few wide-bit cases.

numpy increments the 256-bit counter before each block, so a block generated from numpy state
counter c is Philox(c + 1, key).
"""
import json
import os

import numpy as np

CASES = [
    ((1, 0, 0, 0), (13, 0)),
    ((1, 647, 0, 0), (13, 0)),
    ((1000, 98303, 0, 0), (13, 0)),
    ((7, 5, 0, 0), (0xFFFFFFFFFFFFFFFF, 0)),
    ((0xDEADBEEF, 0x12345678, 0, 0), (0x0123456789ABCDEF, 0)),
    ((2, 3, 5, 7), (11, 13)),
    ((0xFFFFFFFFFFFFFFFE, 0xFFFFFFFFFFFFFFFF, 0xFFFFFFFFFFFFFFFF, 0xFFFFFFFFFFFFFFFF), (1, 2)),
]


def block(ctr, key):
    c = list(ctr)
    c[0] = (c[0] - 1) & 0xFFFFFFFFFFFFFFFF  # numpy pre-increments (no borrow needed for c[0] >= 1)
    assert ctr[0] >= 1
    bg = np.random.Philox(counter=np.array(c, dtype=np.uint64), key=np.array(key, dtype=np.uint64))
    return [int(v) for v in bg.random_raw(4)]


def main():
    out = [{"counter": [hex(v) for v in ctr], "key": [hex(v) for v in key],
            "out": [hex(v) for v in block(ctr, key)]} for ctr, key in CASES]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "philox4x64.json")
    with open(path, "w") as f:
        json.dump({"generator": f"numpy {np.__version__} np.random.Philox (Philox4x64-10)",
                   "cases": out}, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
