"""Write tests/golden/rng_vectors.json.

glibc rand() words come from the C library itself (ctypes, srand/rand); the
xorshift128+ floats come from the independent pure-Python statement in
tests/helpers.py.  Both are then frozen so later changes to either the oracle
or the kernel are checked against fixed numbers.
"""

import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from helpers import xorshift128p_floats  # noqa: E402


def main():
    libc = ctypes.CDLL("libc.so.6")
    out = {"glibc_rand": {}, "xorshift128p": []}
    for seed in (1648335518, 29012392, 1):
        libc.srand(seed)
        out["glibc_rand"][str(seed)] = [int(libc.rand()) for _ in range(16)]
    for seed4 in ([1648335518, 1, 2, 3], [0x12345678, 0x9ABCDEF0, 0x0F1E2D3C, 0x4B5A6978], [7, 0, 0, 1]):
        f = xorshift128p_floats(seed4, 16)
        out["xorshift128p"].append({"seed4": seed4, "bits": [int(b) for b in f.view(np.uint32)]})
    with open(os.path.join(HERE, "rng_vectors.json"), "w") as fp:
        json.dump(out, fp, indent=1)


if __name__ == "__main__":
    main()
