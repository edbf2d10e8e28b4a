"""Writes tests/golden/lb_vectors.npz from the reference's own hilbert2d and
indexx (src/loadbalance.f90:251-400, built into oracle/_ref/libwxref_lb.so
by oracle/ref/Makefile): Hilbert keys of seeded coordinates, including bits
above the key length, and index tables of seeded arrays with many ties.
Inputs and outputs are both stored (int32).  Run from the repo root:
python tests/golden/make_lb_golden.py"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
_ip = ctypes.POINTER(ctypes.c_int32)


def main():
    R = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libwxref_lb.so"))
    rng = np.random.default_rng(2024)
    out = {}
    for bits in (1, 2, 4, 6, 9):
        x = rng.integers(0, 1 << (bits + 2), 400).astype(np.int32)
        y = rng.integers(0, 1 << (bits + 2), 400).astype(np.int32)
        k = np.zeros(400, dtype=np.int32)
        R.wxr_lb_hilbert2d(x.ctypes.data_as(_ip), y.ctypes.data_as(_ip), k.ctypes.data_as(_ip), bits, 400)
        out[f"hilbert_{bits}_x"], out[f"hilbert_{bits}_y"], out[f"hilbert_{bits}_key"] = x, y, k
    for n in (1, 5, 6, 7, 40, 777, 4000):
        for hi in (3, 50, 1 << 24):
            a = rng.integers(0, hi, n).astype(np.int32)
            ix = np.zeros(n, dtype=np.int32)
            R.wxr_lb_indexx(n, a.ctypes.data_as(_ip), ix.ctypes.data_as(_ip))
            out[f"indexx_{n}_{hi}_arr"], out[f"indexx_{n}_{hi}_idx"] = a, ix
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "lb_vectors.npz"), **out)


if __name__ == "__main__":
    main()
