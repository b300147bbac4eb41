"""torch.autograd wrapper of the forward + adjoint solve: 3-D sound-speed
and density perturbations dlnc = d ln c, dlnrho = d ln rho on the grid ->
the travel-time misfit 0.5 * sum tau^2 at a set of receivers (MISFIT_ALL,
driver.f90:1702-2076), with the gradients from the adjoint sensitivity
kernels (kernels.f90:530,537): d misfit / d dlnc_j = K_lnc(x_j) dV_j and
d misfit / d dlnrho_j = K_lnrho(x_j) dV_j (density at fixed sound speed,
hydrostatic background: background.perturb_density).

PyTorch is used for the autograd plumbing only; the solves run on the HIP
engine through the C ABI, the set-up and measurement in FP64 on the host.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import abi
from . import background as bgm
from . import driver as D
from . import pipeline as PL
from .measure import RingMeasurement


@dataclass
class Problem:
    """A fixed inversion problem: workload, receivers, observed traces and
    the time windows (fixed by the reference model's synthetics)."""
    spec: PL.SolveSpec
    S0: bgm.Setup
    obs: np.ndarray
    meas: RingMeasurement
    checkpoint_every: int | None = None  # steps between restart states (None: keep all snapshots)
    kstride: int | None = None           # steps between kernel snapshots (None: output cadence)
    aligned: bool = True                 # exact time-reversal pairing of snapshots
    exact_adjoint: bool = True           # derivative of the discrete travel-time estimator
    density_hydrostatic: bool = True     # density perturbations keep p0' = -rho0 g
    last_peak_bytes: int = 0

    @classmethod
    def build(cls, W: D.Workload, receivers: np.ndarray, maxtime: int, obs: np.ndarray,
              precision: int = abi.FP32, window_width: float = 120.0, **kw) -> "Problem":
        spec = PL.SolveSpec(W=W, receivers=np.asarray(receivers), maxtime=maxtime,
                            precision=precision, window_width=window_width)
        S0 = D.setup_for(W)
        syn = PL.observe(spec, S0)
        meas = RingMeasurement(syn, spec.cad, window_width)
        return cls(spec=spec, S0=S0, obs=obs, meas=meas, **kw)

    def model_mask(self) -> np.ndarray:
        """(nz,) 1 on the layers the model lives on, 0 in the PML layers.  The
        kernel formula (kernels.f90:537) is the adjoint of the interior
        equations only: the PML damping profiles are built from c at the
        boundaries (physics.f90:254-290) and the PML equations are not
        self-adjoint, so d ln c is held at zero there."""
        P = self.spec.W.params
        m = np.zeros(P.nz)
        m[P.npmlbot:P.nz - P.npmltop] = 1.0
        return m

    def setup(self, dlnc: np.ndarray, dlnrho: np.ndarray | None = None) -> bgm.Setup:
        W = self.spec.W
        m = self.model_mask()[None, None, :]
        S = D.setup_for(W, c2_3d=self.S0.c2 * np.exp(2.0 * dlnc * m))
        if dlnrho is not None:
            S = bgm.perturb_density(S, dlnrho * m, hydrostatic=self.density_hydrostatic)
        return S

    def evaluate(self, dlnc: np.ndarray, dlnrho: np.ndarray | None = None) -> float:
        """Misfit only (no snapshots kept)."""
        S = self.setup(dlnc, dlnrho)
        fe, ae = PL.make_engines(self.spec, S)
        try:
            syn = PL.forward_solve(fe, self.spec, store=False)
        finally:
            fe.close()
            ae.close()
        return self.meas.misfit(self.meas.travel_times(self.obs, syn))

    def forward(self, dlnc: np.ndarray, dlnrho: np.ndarray | None = None):
        """Forward solve; returns (misfit, state for backward)."""
        S = self.setup(dlnc, dlnrho)
        fe, ae = PL.make_engines(self.spec, S)
        try:
            if self.checkpoint_every:
                syn = PL.forward_checkpointed(fe, self.spec, self.checkpoint_every)
            else:
                syn = PL.forward_solve(fe, self.spec, store=True, kstride=self.kstride)
        except Exception:
            fe.close()
            ae.close()
            raise
        tau = self.meas.travel_times(self.obs, syn)
        return self.meas.misfit(tau), (S, fe, ae, syn, tau)

    def backward(self, saved) -> np.ndarray:
        """Gradient with respect to dlnc."""
        return self.backward_both(saved)[0]

    def backward_both(self, saved) -> tuple[np.ndarray, np.ndarray]:
        """Gradients with respect to dlnc and dlnrho (one adjoint solve)."""
        S, fe, ae, syn, tau = saved
        try:
            adj = self.meas.adjoint_source(syn, tau, obs=self.obs, exact=self.exact_adjoint)
            vr = PL.adjoint_vr(self.spec, adj, S)
            if self.checkpoint_every:
                PL.adjoint_recompute(fe, ae, self.spec, vr, S.g_raw, self.checkpoint_every,
                                     kstride=self.kstride, aligned=self.aligned)
            else:
                PL.adjoint_solve(ae, self.spec, vr, S.g_raw, kstride=self.kstride,
                                 aligned=self.aligned, src=fe)
            self.last_peak_bytes = fe.device_bytes + ae.device_bytes
            w = (PL.cell_volume(S) * self.model_mask())[None, None, :]
            Kc = PL.kernel_lnc(ae.kernel_get(abi.KERN_C2), S, self.spec, kstride=self.kstride)
            Kr = PL.kernel_lnrho(ae.kernel_get(abi.KERN_RHO), S, self.spec,
                                 kstride=self.kstride)
            return Kc * w, Kr * w
        finally:
            fe.close()
            ae.close()


class SoundSpeedMisfit(torch.autograd.Function):
    """misfit = f(dlnc); backward runs the adjoint solve."""

    @staticmethod
    def forward(ctx, dlnc: torch.Tensor, problem: Problem) -> torch.Tensor:
        d = dlnc.detach().to("cpu", torch.float64).numpy()
        chi, saved = problem.forward(d)
        ctx.problem = problem
        ctx.saved = saved
        ctx.in_device = dlnc.device
        ctx.in_dtype = dlnc.dtype
        return torch.tensor(chi, dtype=torch.float64, device=dlnc.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        g = ctx.problem.backward(ctx.saved)
        ctx.saved = None
        grad = torch.from_numpy(g).to(ctx.in_device, ctx.in_dtype) * grad_out.to(ctx.in_dtype)
        return grad, None


def misfit(dlnc: torch.Tensor, problem: Problem) -> torch.Tensor:
    return SoundSpeedMisfit.apply(dlnc, problem)


class ModelMisfit(torch.autograd.Function):
    """misfit = f(dlnc, dlnrho); backward runs one adjoint solve for both."""

    @staticmethod
    def forward(ctx, dlnc: torch.Tensor, dlnrho: torch.Tensor, problem: Problem) -> torch.Tensor:
        c = dlnc.detach().to("cpu", torch.float64).numpy()
        r = dlnrho.detach().to("cpu", torch.float64).numpy()
        chi, saved = problem.forward(c, r)
        ctx.problem, ctx.saved = problem, saved
        ctx.meta = (dlnc.device, dlnc.dtype, dlnrho.device, dlnrho.dtype)
        return torch.tensor(chi, dtype=torch.float64, device=dlnc.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        gc, gr = ctx.problem.backward_both(ctx.saved)
        ctx.saved = None
        dc, tc, dr, tr = ctx.meta
        return (torch.from_numpy(gc).to(dc, tc) * grad_out.to(tc),
                torch.from_numpy(gr).to(dr, tr) * grad_out.to(tr), None)


def misfit_c_rho(dlnc: torch.Tensor, dlnrho: torch.Tensor, problem: Problem) -> torch.Tensor:
    return ModelMisfit.apply(dlnc, dlnrho, problem)
