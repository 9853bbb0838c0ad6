"""Writes tests/golden/pyrdown_kat.npz: the pyrDown known-answer vectors.

Inputs are the reference test's pattern (x*13 + y*29 + (x*y)/7) & 0xff on the
sizes {64, 97} x {48, 33} (tests/matcher_test.cpp:105-110); outputs are computed
here with a separable [1 4 6 4 1] correlation, reflect-101 borders, even
samples and (sum + 128) >> 8 -- the algorithm of the reference's CPU
reference pyrDownRef (tests/matcher_test.cpp:74-101), written independently of
the oracle.  Run: python tests/golden/make_golden.py
"""
import os

import numpy as np


def pyrdown(src):
    k = np.array([1, 4, 6, 4, 1], np.int64)
    p = np.pad(src.astype(np.int64), 2, mode="reflect")
    h, w = src.shape
    rows = sum(k[i] * p[i:i + h, :] for i in range(5))
    full = sum(k[i] * rows[:, i:i + w] for i in range(5))
    return ((full[::2, ::2] + 128) >> 8).astype(np.uint8)


def main():
    out = {}
    for w in (64, 97):
        for h in (48, 33):
            xs, ys = np.meshgrid(np.arange(w), np.arange(h))
            src = ((xs * 13 + ys * 29 + (xs * ys) // 7) & 0xFF).astype(np.uint8)
            out[f"in_{w}x{h}"] = src
            out[f"out_{w}x{h}"] = pyrdown(src)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pyrdown_kat.npz")
    np.savez_compressed(path, **out)
    print(path)


if __name__ == "__main__":
    main()
