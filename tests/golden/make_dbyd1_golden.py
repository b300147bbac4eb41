"""Writes tests/golden/dbyd1_ref.npz: inputs and outputs of the reference's
own DBYD1 (dbyd1.f + tridag.f, compiled from /root/reference with amdflang
into oracle/_ref by ``make -C oracle ref``) on seeded random rows, for ibc
0..4 and N1 in {4, 12, 64, 256}.  Run in the build container:

    make -C oracle ref && python tests/golden/make_dbyd1_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402


def main():
    if O.ref_dbyd1() is None:
        raise SystemExit("oracle/_ref/libref_dbyd1.so missing: run make -C oracle ref")
    rng = np.random.default_rng(1994 + 1)
    out = {}
    for n1 in (4, 12, 64, 256):
        for ibc in range(5):
            B = rng.standard_normal((n1, 3))
            A0 = rng.standard_normal((n1, 3))  # caller boundary values (ibc 2,3,4)
            A = O.call_ref_dbyd1(A0.copy(), B, ibc)
            out[f"B_{n1}_{ibc}"] = B
            out[f"A0_{n1}_{ibc}"] = A0
            out[f"A_{n1}_{ibc}"] = A
    np.savez_compressed(os.path.join(os.path.dirname(__file__), "dbyd1_ref.npz"), **out)
    print("wrote dbyd1_ref.npz with", len(out) // 3, "cases")


if __name__ == "__main__":
    main()
