"""Precision report of the FP32 path against the FP64 path on one workload.

    python -m sparc_amd.validate [--workload box2] [--maxtime 1999]

Runs the full single-source solve (forward with snapshots at the output
cadence, travel times against synthetic observations of a +2 % Gaussian
sound-speed bump of radius 5 Mm at 3 Mm depth, adjoint, kernels) once in
FP64 and once in FP32 with the compensated state, receivers on a 10 Mm ring
of 8 pixels, and prints one JSON line with the error of each FP32 output:
receiver traces, travel times, misfit, and the sound-speed kernel both over
the model layers (outside the PML) and over the whole box, end-to-end and
with the FP64 adjoint source fed to the FP32 adjoint (which isolates the
FP32 adjoint and kernel path from the travel-time differences).  The FP64
engine agrees with the FP64 CPU oracle to rounding on the grids the oracle
can run (the GPU parity tests); at this size it stands in for it.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from dataclasses import replace

import numpy as np

from . import abi
from . import driver as D
from . import measure as M
from . import pipeline as PL


def _rl2(a, b) -> float:
    den = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / (den if den > 0 else 1.0)


def report(workload: str = "box2", maxtime: int = 1999, ring_mm: float = 10.0) -> dict:
    W = D.workload(workload)
    P = W.params
    S0 = D.setup_for(W)
    rec = M.ring_pixels(P.nx, P.ny, P.xlength, P.ylength, 0.0, 0.0, ring_mm, 8)

    def spec(prec, comp=True):
        W2 = D.Workload(W.name, replace(P, compensated=comp), W.tab, W.nsteps)
        return PL.SolveSpec(W=W2, receivers=rec, maxtime=maxtime, precision=prec)

    t0 = time.time()
    obs = PL.observe(spec(abi.FP64), D.setup_for(W, c2_3d=PL.bump_c2(S0, W)))
    r64, meas = PL.full_solve(spec(abi.FP64), S0, obs)
    print(f"[validate] FP64 solve done ({time.time() - t0:.0f} s)", file=sys.stderr, flush=True)
    inner = slice(P.npmlbot, P.nz - P.npmltop)
    out = {"workload": f"{W.name} {P.nx}x{P.ny}x{P.nz}, {maxtime + 1} steps", "receivers": len(rec),
           "tau_fp64_s": [round(float(t), 6) for t in r64.tau], "misfit_fp64": r64.misfit}
    for comp in (True, False):
        r32, _ = PL.full_solve(spec(abi.FP32, comp), S0, obs, meas=meas)
        sp = spec(abi.FP32, comp)
        fe, ae = PL.make_engines(sp, S0)
        try:
            PL.forward_solve(fe, sp, store=True)
            PL.copy_snapshots(ae, fe, sp.maxtime // sp.steps + 1)
            PL.adjoint_solve(ae, sp, PL.adjoint_vr(sp, r64.adj, S0), S0.g_raw)
            K_same = PL.kernel_lnc(ae.kernel_get(abi.KERN_C2), S0, sp)
        finally:
            fe.close()
            ae.close()
        key = "fp32_compensated" if comp else "fp32_plain"
        out[key] = {
            "traces_rel_l2": _rl2(r32.traces, r64.traces),
            "tau_max_abs_s": float(np.max(np.abs(r32.tau - r64.tau))),
            "misfit_rel": abs(r32.misfit - r64.misfit) / r64.misfit,
            "K_lnc_rel_l2_model_layers": _rl2(r32.K_lnc[:, :, inner], r64.K_lnc[:, :, inner]),
            "K_lnc_rel_l2_whole_box": _rl2(r32.K_lnc, r64.K_lnc),
            "K_lnc_same_adjoint_source_model_layers": _rl2(K_same[:, :, inner],
                                                           r64.K_lnc[:, :, inner]),
            "K_lnc_same_adjoint_source_whole_box": _rl2(K_same, r64.K_lnc),
            "K_rho_rel_l2_model_layers": _rl2(r32.kernel_raw[abi.KERN_RHO][:, :, inner],
                                              r64.kernel_raw[abi.KERN_RHO][:, :, inner]),
            "hessian_rel_l2": _rl2(r32.kernel_raw[abi.KERN_HESSIAN],
                                   r64.kernel_raw[abi.KERN_HESSIAN]),
        }
        print(f"[validate] {key} done ({time.time() - t0:.0f} s)", file=sys.stderr, flush=True)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="box2")
    ap.add_argument("--maxtime", type=int, default=1999)
    a = ap.parse_args()
    print(json.dumps(report(a.workload, a.maxtime)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
