"""The differentiation configuration at its own size: torch.autograd over the
256 x 256 x 300 forward + adjoint solve in FP32 (compensated state, kernel
sums in FP64) with forward-wavefield checkpoints every 100 steps recomputed
in the backward pass, mapping a 3-D d ln c to the travel-time misfit at 8
receivers on a 10 Mm ring (observed data from a +2 % Gaussian sound-speed
bump, radius 5 Mm, 3 Mm deep).

* The FP32 checkpointed gradient against the FP64 engine's gradient (which
  equals the oracle, tests/test_gpu_pipeline.py and test_gpu_shapes.py):
  within 5e-3 relative L2 (the travel-time residuals carry the FP32 traces'
  error, amplified; the kernels themselves meet 1e-4 on a common adjoint
  source).
* Both against central differences of the misfit along 16 random Gaussian
  bumps in ln c (numpy default_rng(4), step 1e-4 relative), in four tests of
  four directions so that each finishes in about two minutes.  The
  differences are taken with the FP64 solve: an FP32 misfit carries
  rounding noise of ~4e-5 relative, against misfit changes of ~1e-3
  relative at this step, which moved the FP32 differences' agreement by
  several hundredths of the 0.1 bar between builds with different FP32
  roundings (0.092 and 0.125); the FP64 differences are free of it.
* Peak device memory of the checkpointed solve is reported.

Measured: misfits 7.570765 (FP64) and 7.570786 (FP32), FP32 gradient 6.8e-5
relative L2 from the FP64 one; adjoint against the
FP64 differences 8.9e-2 relative L2 over the 16 directions for the FP64 and
the FP32 adjoint alike, the large ones within 0.79-1.18 (the kernel formula
is the adjoint of the continuous equations, so the agreement is at
discretisation level; the outliers are the directions whose misfit change is
under 1 % of the largest); peak device memory 59.2 GiB (a 100-step segment
of per-step snapshots plus the checkpoints).

1401 steps (16.3 min of solar time) instead of the configuration's 2000
keep the 32 finite-difference solves within the test budget; the ring
receivers' arrivals from the Ricker source (peaked at 400 s) come at 10-13
min, inside the record rather than at its end (arrivals cut by the end of
the record leave the measurement at the start of the adjoint run, where the
7-point Lagrange source interpolation has fewer points).  Kernel sums are taken every step (the exact
time-reversal pairing of the autograd problem) from per-segment snapshots.
"""
import numpy as np
import pytest
import torch

from sparc_amd import abi
from sparc_amd import background as B
from sparc_amd import driver as D
from sparc_amd import pipeline as PL
from sparc_amd.autograd import Problem, misfit
from sparc_amd.measure import ring_pixels
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

STEPS = 1401
EPS = 1e-4


def _bump(W, S, cx, cy, cz, r):
    P = W.params
    x = ((np.arange(P.nx) / P.nx) - 0.5) * P.xlength / 1e8
    y = ((np.arange(P.ny) / P.ny) - 0.5) * P.ylength / 1e8
    zmm = (S.z - 1.0) * B.RSUN_R / 1e8
    return np.exp(-((x[:, None, None] - cx) ** 2 + (y[None, :, None] - cy) ** 2
                    + (zmm[None, None, :] - cz) ** 2) / r ** 2)


class Case:
    def __init__(self):
        W = D.workload("box2", precision=abi.FP32)
        P = W.params
        S0 = D.setup_for(W)
        self.W, self.S0 = W, S0
        rec = ring_pixels(P.nx, P.ny, P.xlength, P.ylength, 0.0, 0.0, 10.0, 8)
        maxtime = STEPS - 1
        spec = PL.SolveSpec(W=W, receivers=rec, maxtime=maxtime, precision=abi.FP64)
        obs = PL.observe(spec, D.setup_for(W, c2_3d=PL.bump_c2(S0, W)))
        self.p64 = Problem.build(W, rec, maxtime, obs, precision=abi.FP64, checkpoint_every=100,
                                 kstride=1)
        self.p32 = Problem.build(W, rec, maxtime, obs, precision=abi.FP32, checkpoint_every=100,
                                 kstride=1)
        shape = (P.nx, P.ny, P.nz)
        g = {}
        for name, pr, dt in (("fp64", self.p64, torch.float64), ("fp32", self.p32, torch.float32)):
            dlnc = torch.zeros(shape, dtype=dt, requires_grad=True)
            chi = misfit(dlnc, pr)
            chi.backward()
            g[name] = dlnc.grad.numpy().astype(np.float64)
            setattr(self, f"chi_{name}", float(chi.detach()))
        self.g = g
        self.peak32 = self.p32.last_peak_bytes
        rng = np.random.default_rng(4)
        self.dirs = [_bump(W, S0, rng.uniform(-12, 12), rng.uniform(-12, 12), rng.uniform(-9, -2),
                           rng.uniform(2.0, 4.0)) for _ in range(16)]
        self.fd = {}


@pytest.fixture(scope="module")
def case():
    return Case()


@pytest.mark.timeout(900)
def test_box2_fp32_checkpointed_gradient_matches_fp64(case):
    C = case
    assert C.chi_fp64 > 0
    err = rel_l2(C.g["fp32"], C.g["fp64"])
    print(f"\n  misfit FP64 {C.chi_fp64:.6e} FP32 {C.chi_fp32:.6e}; gradient rel L2 {err:.3e}; "
          f"peak device memory of the checkpointed FP32 solve {C.peak32 / 2**30:.2f} GiB")
    assert abs(C.chi_fp32 - C.chi_fp64) < 1e-3 * C.chi_fp64
    assert err < 5e-3


@pytest.mark.timeout(600)
@pytest.mark.parametrize("part", range(4))
def test_box2_gradient_matches_finite_differences(case, part):
    C = case
    idx = range(4 * part, 4 * part + 4)
    for i in idx:
        h = C.dirs[i]
        C.fd[i] = (C.p64.evaluate(EPS * h) - C.p64.evaluate(-EPS * h)) / (2 * EPS)
    print(f"\n  FD directions {list(idx)}: {[f'{C.fd[i]:.4e}' for i in idx]}")
    if part < 3:
        return
    fd = np.array([C.fd[i] for i in range(16)])
    for name in ("fp64", "fp32"):
        ad = np.array([float(np.sum(C.g[name] * h)) for h in C.dirs])
        err = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
        print(f"  {name}: adjoint / FD {np.round(ad / fd, 3)}; rel L2 {err:.3e}")
        assert err < 0.1, name
        big = np.abs(fd) > 0.05 * np.max(np.abs(fd))
        assert np.all(np.sign(ad[big]) == np.sign(fd[big])), name
