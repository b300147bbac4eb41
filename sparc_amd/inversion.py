"""The inversion around the solver: the behaviour of the reference's Python
drivers (data_forward.py, full.py, linesearch.py, grad.py, misfit.py),
rewritten over this package's pipeline instead of `sparc` processes and FITS
files.

* observations: one forward solve per source in the true model
  (data_forward.py:48-52);
* gradient pass: forward + adjoint + kernel per source, kernels summed over
  sources (full.py:48-97, grad.py:158-164);
* gradient post-processing: hessian preconditioning with rho0 and a 5e-3
  floor, Gaussian x smoothing (sigma 4 px), z smoothing, antisymmetrisation
  of the stream-function kernel (grad.py:44-107, 166-186);
* update direction: steepest descent or Polak-Ribiere conjugate gradient
  (grad.py:192-239);
* line search: trial models (grad.py:293-311), one forward solve per
  (step, source) job with the job number mapped as linesearch.py:30-31,
  misfits summed over sources (misfit.py:24-26, lsmisfit.py:24-26).

Arrays are (nx, ny, nz) like the rest of the package (grad.py transposes the
FITS (nz, ny, nx) order to the same).
"""
from __future__ import annotations

import numpy as np

# --------------------------------------------------------------------------
# gradient post-processing (grad.py:44-107)
# --------------------------------------------------------------------------
SMOOTH_COEFFS = (0.75390625, 0.41015625, -0.234375, 0.087890625, -0.01953125, 0.001953125)


def filterx(kern: np.ndarray, sigma_px: float = 4.0) -> np.ndarray:
    """Gaussian smoothing along x by an FFT product with the transform of a
    periodic Gaussian of sigma 4 pixels (grad.py:44-59), in place."""
    nx = kern.shape[0]
    x = np.fft.fftfreq(nx) * nx
    g = np.exp(-x ** 2.0 / (2.0 * sigma_px ** 2.0))
    filt = np.fft.rfft(g)[:, None, None]
    kern[:] = np.fft.irfft(np.fft.rfft(kern, axis=0) * filt, n=nx, axis=0).real
    return kern


def filterz(arr: np.ndarray, algo: str = "spline", sp: float = 1.0) -> np.ndarray:
    """z smoothing (grad.py:61-100), in place: 'smooth' = the 11-point filter
    on planes 6..nz-5 (1-based) with two-point averages at 2..4 and
    nz-4..nz-1; planes 5 and nz (1-based) are zeroed, because the reference
    copies its temporary back over them without filling them (grad.py:86-91).
    'gaussian' = gaussian_filter1d(sigma=sp); 'spline' = a smoothing spline
    per column with s=sp on the column scaled by its max."""
    nx, ny, nz = arr.shape
    if algo == "smooth":
        c = SMOOTH_COEFFS
        temp = np.zeros_like(arr)
        for k in range(5, nz - 5):
            temp[:, :, k] = (c[0] * arr[:, :, k]
                             + sum(0.5 * c[j] * (arr[:, :, k - j] + arr[:, :, k + j])
                                   for j in range(1, 6)))
        temp[:, :, nz - 5:nz - 1] = 0.5 * (arr[:, :, nz - 4:] + arr[:, :, nz - 6:nz - 2])
        temp[:, :, 1:4] = 0.5 * (arr[:, :, :3] + arr[:, :, 2:5])
        arr[:, :, 1:nz] = temp[:, :, 1:nz]
    elif algo == "gaussian":
        from scipy.ndimage import gaussian_filter1d
        arr[:] = gaussian_filter1d(arr, sigma=sp, axis=-1)
    elif algo == "spline":
        from scipy.interpolate import UnivariateSpline
        z = np.arange(nz)
        for i in range(nx):
            for j in range(ny):
                col = arr[i, j]
                m = col.max()
                s = UnivariateSpline(z, col / m, s=sp)
                arr[i, j] = s(z) * m
    else:
        raise ValueError(f"unknown filter {algo}")
    return arr


def antisymmetrize(arr: np.ndarray) -> np.ndarray:
    """0.5 (f(x) - f(-x)) about the box centre along x (grad.py:102), in place."""
    arr[:] = 0.5 * (arr - arr[::-1])
    return arr


def rms(arr: np.ndarray) -> float:
    return float(np.sqrt(np.sum(arr ** 2) / arr.size))


def precondition(hessians: list, rho0_profile: np.ndarray, floor: float = 5e-3) -> np.ndarray:
    """sum_src |hessian| * rho0(z), normalised to max 1, floored at 5e-3
    (grad.py:158-174)."""
    h = np.zeros_like(hessians[0])
    for x in hessians:
        h += np.abs(x)
    h = h * np.asarray(rho0_profile)[None, None, :]
    h = h / h.max()
    h[h < floor] = floor
    return h


def gradients(kernels_c: list, kernels_psi: list | None, hessians: list,
              rho0_profile: np.ndarray):
    """Summed, preconditioned and smoothed gradients (grad.py:158-186):
    sound speed (x Gaussian, z 'smooth'), stream function (x Gaussian,
    antisymmetrised, z Gaussian sigma 2).  Returns (grad_c, grad_psi)."""
    h = precondition(hessians, rho0_profile)
    gc = sum(kernels_c) / h
    filterx(gc)
    filterz(gc, algo="smooth", sp=0.3)
    gpsi = None
    if kernels_psi is not None:
        gpsi = sum(kernels_psi) / h
        filterx(gpsi)
        antisymmetrize(gpsi)
        filterz(gpsi, algo="gaussian", sp=2.0)
    return gc, gpsi


def cg_coefficient(grad_psi: np.ndarray, last_grad_psi: np.ndarray) -> float:
    """Polak-Ribiere coefficient from the stream-function gradients, clipped
    at zero (grad.py:224-230); it also scales the sound-speed update."""
    con = float(np.sum(grad_psi * (grad_psi - last_grad_psi)) / np.sum(last_grad_psi ** 2.0))
    return max(con, 0.0)


def update_direction(grad: np.ndarray, last_update: np.ndarray | None = None,
                     con: float = 0.0) -> np.ndarray:
    """grad (steepest descent, and the first iteration), or grad + con *
    last_update (conjugate gradient) (grad.py:192-239).  The reference's
    L-BFGS branch is switched off by its own flags and names undefined
    variables (grad.py:134-136, 241-290), so it has no counterpart here."""
    if last_update is None:
        return grad.copy()
    return grad + con * last_update


def linesearch_models(last_psi: np.ndarray, update_psi: np.ndarray,
                      last_c: np.ndarray | None = None, n: int = 5, eps: float = 1e-7):
    """Trial models i = 1..n (grad.py:293-311): psi_i = last_psi + eps * i *
    rms(last_psi) * update_psi / max|update_psi|; the sound-speed model is
    passed on unchanged (the reference mutes its update).  Returns a list of
    (c_i, psi_i)."""
    u = update_psi / np.abs(update_psi).max()
    scale = rms(last_psi)
    return [(last_c, last_psi + eps * i * scale * u) for i in range(1, n + 1)]


def linesearch_job(procno: int, nsources: int, nsteps: int = 5):
    """(line-search step, source) of job procno, both 1-based, or None past the
    last job (linesearch.py:22-31)."""
    if procno >= nsources * nsteps:
        return None
    return procno // nsources + 1, procno % nsources + 1


def linesearch_totals(per_job: np.ndarray, nsources: int, nsteps: int = 5) -> list:
    """Total misfit of each trial step from the per-job misfits in job order
    (lsmisfit.py:24-26)."""
    per_job = np.asarray(per_job, dtype=np.float64)
    return [float(np.sum(per_job[i * nsources:(i + 1) * nsources])) for i in range(nsteps)]


def total_misfit(per_source: list) -> float:
    """Sum of the per-source misfits of the master pixels (misfit.py:24-26)."""
    return float(np.sum(per_source))


# --------------------------------------------------------------------------
# the solves (data_forward.py, full.py, linesearch.py over the pipeline)
# --------------------------------------------------------------------------
def observe_all(spec_for_source, setup, sources: list, eng_factory=None) -> list:
    """data_forward.py:48-52: the 'observed' traces of every source in the
    true model, one forward solve each."""
    from . import pipeline as PL
    return [PL.observe(spec_for_source(s), setup, eng_factory=eng_factory) for s in sources]


def gradient_pass(spec_for_source, setup, sources: list, observations: list,
                  eng_factory=None, kernel_dir: str | None = None, **kw):
    """full.py:48-97: forward, travel times, adjoint and kernels for every
    source.  Returns the per-source SolveResults (traces, misfit, kernel
    sums) and RingMeasurements (windows reused by the line search).  With
    kernel_dir, each source's kernel files and misfit line are also written
    there as PRODUCE_KERNELS and MISFIT_ALL name them (kernel_c_SS.fits,
    hessian_SS.fits, ..., misfit_SS_00)."""
    from . import pipeline as PL
    results, meas = [], []
    for i, (s, obs) in enumerate(zip(sources, observations)):
        spec = spec_for_source(s)
        r, m = PL.full_solve(spec, setup, obs, eng_factory=eng_factory, **kw)
        results.append(r)
        meas.append(m)
        if kernel_dir is not None:
            write_solve_files(kernel_dir, i + 1, r, setup, spec)
    return results, meas


def write_solve_files(directory: str, contrib: int, res, setup, spec) -> list:
    """One source's kernel files and misfit line in the reference's layout."""
    from . import abi
    from . import formats as FM
    files = FM.kernel_files(res.kernel_raw, setup.rho0, setup.c2, stepskern=spec.steps,
                            timestep=spec.W.params.timestep,
                            magnetic=abi.KERN_BX in res.kernel_raw)
    return FM.write_kernel_files(directory, contrib, files, misfit=res.misfit,
                                 nmeas=len(spec.receivers))


def read_gradient_inputs(directory: str, nsources: int, psi: bool = True):
    """The per-source files grad.py reads (grad.py:158-164): kernel_c_SS,
    kernel_vectorpsi_SS (when psi) and hessian_SS.  Returns (kernels_c,
    kernels_psi or None, hessians)."""
    import os
    from . import formats as FM

    def rd(name, s):
        return FM.fits_read(os.path.join(directory, "%s_%02d.fits" % (name, s)))
    srcs = range(1, nsources + 1)
    kc = [rd("kernel_c", s) for s in srcs]
    kp = [rd("kernel_vectorpsi", s) for s in srcs] if psi else None
    hs = [rd("hessian", s) for s in srcs]
    return kc, kp, hs


def linesearch_pass(spec_for_source, setups: list, sources: list, observations: list,
                    measurements: list, eng_factory=None) -> np.ndarray:
    """linesearch.py: one forward solve per (trial model, source) job, in the
    job order of linesearch_job, with the misfit of each job measured in the
    windows of the gradient pass.  Returns the per-job misfits; their
    per-step totals are linesearch_totals(...)."""
    from . import pipeline as PL
    nsrc, nls = len(sources), len(setups)
    out = np.zeros(nsrc * nls)
    for procno in range(nsrc * nls):
        ls, src = linesearch_job(procno, nsrc, nls)
        spec = spec_for_source(sources[src - 1])
        syn = PL.observe(spec, setups[ls - 1], eng_factory=eng_factory)
        m = measurements[src - 1]
        out[procno] = m.misfit(m.travel_times(observations[src - 1], syn))
    return out
