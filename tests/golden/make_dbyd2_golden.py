"""Writes tests/golden/dbyd2_ref.npz: inputs and outputs of the reference's
own DBYD2 (dbyd2.f90 + mtridag.f90, compiled from /root/reference with amdflang
into oracle/_ref by ``make -C oracle ref``) on seeded random columns, for
every ibc 0..5 and N2 in {12, 100, 150, 300}.  Run in the build container:

    make -C oracle ref && python tests/golden/make_dbyd2_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402


def main():
    if O.ref_dbyd2() is None:
        raise SystemExit("oracle/_ref/libref_dbyd2.so missing: run make -C oracle ref")
    rng = np.random.default_rng(1994)
    out = {}
    for n2 in (12, 100, 150, 300):
        for ibc in range(6):
            B = rng.standard_normal((2, n2))
            A0 = rng.standard_normal((2, n2))  # caller boundary values (ibc 2,3,4)
            A = O.call_ref_dbyd2(A0.copy(), B, ibc)
            out[f"B_{n2}_{ibc}"] = B
            out[f"A0_{n2}_{ibc}"] = A0
            out[f"A_{n2}_{ibc}"] = A
    np.savez_compressed(os.path.join(os.path.dirname(__file__), "dbyd2_ref.npz"), **out)
    print("wrote dbyd2_ref.npz with", len(out) // 3, "cases")


if __name__ == "__main__":
    main()
