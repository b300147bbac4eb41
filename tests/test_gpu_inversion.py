"""One iteration of the inversion around the GPU solver (data_forward.py,
full.py, grad.py, linesearch.py and misfit.py behaviour): observations of two
sources in a model with a sound-speed bump, the gradient pass in the smooth
model, the preconditioned and smoothed gradient, and a line search whose
first steps along the descent direction lower the summed misfit."""
import dataclasses

import numpy as np
import pytest

from sparc_amd import abi
from sparc_amd import driver as D
from sparc_amd import inversion as INV
from sparc_amd import pipeline as PL

pytestmark = pytest.mark.gpu


def test_one_iteration_lowers_the_misfit(tmp_path):
    W = D.workload("tiny")
    S0 = D.setup_for(W)
    rec = np.array([[16, 16], [18, 16], [16, 13]])
    sources = [(0.0, 0.0), (2.0, -1.0)]

    def spec_for(s):
        return PL.SolveSpec(W=dataclasses.replace(W, src_xy_mm=s), receivers=rec,
                            maxtime=W.params.steps * 20, precision=abi.FP64)

    true = PL.bump_c2(S0, W, amp=0.05, radius_mm=3.0, depth_mm=2.0)
    obs = INV.observe_all(spec_for, D.setup_for(W, c2_3d=true), sources)
    res, meas = INV.gradient_pass(spec_for, S0, sources, obs, kernel_dir=str(tmp_path))
    chi0 = INV.total_misfit([r.misfit for r in res])
    assert chi0 > 0
    # the kernel files are the reference's kernel_c (a fixed multiple of K_lnc)
    kc, _, hs = INV.read_gradient_inputs(str(tmp_path), len(sources), psi=False)
    for r, k, h in zip(res, kc, hs):
        ratio = k[r.K_lnc != 0] / r.K_lnc[r.K_lnc != 0]
        assert np.allclose(ratio, ratio[0], rtol=1e-12)
        assert np.array_equal(h, r.kernel_raw[abi.KERN_HESSIAN])
    grad, _ = INV.gradients([r.K_lnc for r in res], None,
                            [r.kernel_raw[abi.KERN_HESSIAN] for r in res], W.tab[:, 2])
    assert np.all(np.isfinite(grad))
    P = W.params
    mask = np.zeros(P.nz)
    mask[P.npmlbot:P.nz - P.npmltop] = 1.0
    d = -INV.update_direction(grad) * mask[None, None, :]
    d /= np.abs(d).max()
    steps = [0.005 * i for i in range(1, 6)]
    setups = [D.setup_for(W, c2_3d=S0.c2 * np.exp(2.0 * a * d)) for a in steps]
    per_job = INV.linesearch_pass(spec_for, setups, sources, obs, meas)
    totals = INV.linesearch_totals(per_job, len(sources), len(steps))
    print("misfit", chi0, "line search", totals)
    assert totals[0] < chi0
    assert min(totals) < 0.9 * chi0
