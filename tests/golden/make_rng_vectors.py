"""Write tests/golden/rng_vectors.json: known answers for the reference's RNG.

An independent numpy restatement of trace.slang:195-202 (LCG step, then the
xorshift-multiply output permutation, 23 mantissa bits -> [0, 1)) and of the
per-(pixel, sample) seeding at trace.slang:736-737, in vectorised uint32
arithmetic. Run: python tests/golden/make_rng_vectors.py
"""
import json
import os

import numpy as np

M = np.uint32


def seed(px, py, sample):
    with np.errstate(over="ignore"):
        s = (M(px) * M(0x9E3779B9)) ^ (M(py) * M(0x85EBCA6B)) ^ (M(0) * M(0xC2B2AE35)) ^ \
            (M(sample) * M(0x27D4EB2D)) ^ M(0xA511E9B3)
    return s | M(1)


def draws(state, n):
    out = []
    s = np.array(state, dtype=np.uint32)
    with np.errstate(over="ignore"):
        for _ in range(n):
            s = s * M(747796405) + M(2891336453)
            word = ((s >> ((s >> M(28)) + M(4))) ^ s) * M(277803737)
            bits = M(0x3F800000) | (((word >> M(22)) ^ word) >> M(9))
            f = bits.view(np.float32) - np.float32(1.0)
            out.append(int(f.view(np.uint32)))
    return int(s), out


if __name__ == "__main__":
    cases = []
    for px, py, smp in [(0, 0, 1), (1, 0, 1), (0, 1, 1), (255, 255, 16), (1919, 1079, 8), (3839, 2159, 64),
                        (123, 456, 7), (65535, 65535, 4294967295)]:
        st = seed(px, py, smp)
        final, vals = draws(st, 48)
        cases.append({"px": px, "py": py, "sample": smp, "seed": int(st), "draw_bits": vals, "state_after": final})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_vectors.json")
    with open(path, "w") as f:
        json.dump({"source": "numpy restatement of trace.slang:195-202, 736-737", "cases": cases}, f, indent=1)
    print("wrote", path)
