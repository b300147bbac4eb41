"""Travel times and adjoint sources (product, FP64 host) against the
oracle's restatement of COMPUTE_TT and against finite differences."""
import numpy as np
import pytest

from oracle import oracle as O
from sparc_amd import measure as M


def _wavelet(t, t0, f=3e-3, sig=150.0):
    return np.cos(2 * np.pi * f * (t - t0)) * np.exp(-0.5 * ((t - t0) / sig) ** 2)


@pytest.mark.parametrize("tau", [-7.3, -1.1, 0.0, 0.4, 5.9])
def test_compute_tt_matches_oracle_and_recovers_shift(tau):
    dt = 30.0
    t = np.arange(80) * dt
    u0 = _wavelet(t, 1000.0)
    u = _wavelet(t, 1000.0 + tau)
    got = M.compute_tt(u0, u, dt)
    assert abs(got - O.compute_tt(u0, u, dt)) < 1e-9
    assert abs(got - tau) < 0.5  # parabola through 30 s samples of a 3 mHz cc peak


@pytest.mark.parametrize("exact,dt,tol", [(True, 30.0, 1e-6), (True, 2.0, 1e-6),
                                          (False, 2.0, 0.03)])
def test_adjoint_source_is_the_misfit_gradient(exact, dt, tol):
    """a = d(0.5 tau^2)/du against central differences of the windowed
    measurement on trace samples.  exact: the estimator's own derivative, to
    FD accuracy at any cadence.  The reference's linearised form matches it
    when the correlation is finely sampled (the parabola through 30 s samples
    of a 3 mHz correlation peak biases it by ~25%)."""
    t = np.arange(int(2700 / dt)) * dt
    rng = np.random.default_rng(0)
    syn = np.stack([_wavelet(t, 900.0 + 40 * r) for r in range(3)])
    obs = np.stack([_wavelet(t, 900.0 + 40 * r + 2.0 + r) for r in range(3)])
    meas = M.RingMeasurement(syn, dt, 150.0)
    tau = meas.travel_times(obs, syn)
    adj = meas.adjoint_source(syn, tau, obs=obs, exact=exact)
    for _ in range(4):
        du = rng.standard_normal(syn.shape) * np.exp(-0.5 * ((t - 1000) / 200) ** 2) * 1e-3
        if dt < 10:  # smooth in time, like a wavefield perturbation
            du = np.stack([np.convolve(d, np.ones(15) / 15, mode="same") for d in du])
        eps = 1e-3
        chi_p = meas.misfit(meas.travel_times(obs, syn + eps * du))
        chi_m = meas.misfit(meas.travel_times(obs, syn - eps * du))
        fd = (chi_p - chi_m) / (2 * eps)
        lin = float(np.sum(adj * du) * dt)
        assert abs(lin - fd) < tol * abs(fd) + 1e-12, (lin, fd)


def test_ring_pixels():
    pts = M.ring_pixels(256, 256, 200e8, 200e8, 0.0, 0.0, 10.0, 8)
    dx = 200.0 / 256
    r = np.hypot((pts[:, 0] - 128) * dx, (pts[:, 1] - 128) * dx)
    assert pts.shape == (8, 2) and np.all(np.abs(r - 10.0) < dx)
