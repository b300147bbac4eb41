"""Travel-time measurement and adjoint sources (FP64, host).

Restates COMPUTE_TT (driver.f90:921-960) with its 3x3 parabola fit
(inverse, driver.f90:964-1042) and the adjoint-source construction of
ADJOINT_SOURCE_COMP (driver.f90:729-917), generalised from the reference's
x-line of receivers (ny = 1) to receiver pixels anywhere on the observation
plane (the ring of the 3-D configurations), with a Gaussian time window
around each synthetic arrival in place of the reference's box window.
"""
from __future__ import annotations

import math

import numpy as np


def _tt_parts(u0: np.ndarray, u: np.ndarray, dt: float):
    nt = u0.size
    # cc(i) for i = -(nt-1)..nt-1 with cc(i) = sum_q u0(q) u(q+i)
    cc = np.correlate(u, u0, mode="full") * dt
    t = np.arange(-nt + 1, nt) * dt
    loc = int(np.argmax(cc))  # first maximum, like maxloc
    loc = min(max(loc, 1), cc.size - 2)
    times = t[loc - 1:loc + 2]
    mat = np.stack([np.ones(3), times, times ** 2.0], axis=1)
    inv = np.linalg.inv(mat)
    p1 = inv[1] @ cc[loc - 1:loc + 2]
    p2 = inv[2] @ cc[loc - 1:loc + 2]
    return loc - (nt - 1), inv, p1, p2


def compute_tt(u0: np.ndarray, u: np.ndarray, dt: float) -> float:
    """Travel-time shift of u relative to u0 (COMPUTE_TT): cc(lag) =
    sum u0(t) u(t+lag) dt over all lags, parabola through the peak and its
    neighbours, tau = -p1/(2 p2).  Positive tau: u arrives later."""
    _, _, p1, p2 = _tt_parts(u0, u, dt)
    return float(-p1 * 0.5 / p2)


def compute_tt_grad(u0: np.ndarray, u: np.ndarray, dt: float) -> np.ndarray:
    """Exact derivative of compute_tt with respect to the samples of u
    (valid away from ties of the correlation peak): d tau/d cc_i through
    the parabola vertex, d cc(lag)/d u(m) = u0(m - lag) dt."""
    lag0, inv, p1, p2 = _tt_parts(u0, u, dt)
    dtau_dcc = -0.5 * inv[1] / p2 + 0.5 * p1 * inv[2] / p2 ** 2
    nt = u0.size
    g = np.zeros(nt)
    for i, lag in enumerate((lag0 - 1, lag0, lag0 + 1)):
        m = np.arange(nt)
        q = m - lag
        ok = (q >= 0) & (q < nt)
        g[m[ok]] += dtau_dcc[i] * u0[q[ok]] * dt
    return g


def gaussian_window(nt: int, dt: float, center: float, width: float) -> np.ndarray:
    t = np.arange(nt) * dt
    return np.exp(-0.5 * ((t - center) / width) ** 2)


def arrival_center(u: np.ndarray, dt: float) -> float:
    """Time of the envelope peak of a trace (the reference picks the window
    centre as the maximum of the filtered synthetic, driver.f90:857-859)."""
    return float(np.argmax(np.abs(u))) * dt


def time_derivative(u: np.ndarray, dt: float) -> np.ndarray:
    """Spectral time derivative of a trace, as the reference multiplies the
    temporal spectrum by i*2*pi*nu (driver.f90:825-830)."""
    n = u.size
    nu = np.fft.fftfreq(n, d=dt)
    return np.real(np.fft.ifft(np.fft.fft(u) * (2j * np.pi * nu)))


class RingMeasurement:
    """Windowed cross-correlation travel times at a set of receiver pixels.

    obs, syn: (nrec, nt) traces of v_z at the observation plane sampled every
    ``dt`` seconds.  Windows are fixed by the synthetic traces (their arrival
    centres) and reused for every model of an inversion, like the cached
    windows of the line search (driver.f90:1427-1437).
    """

    def __init__(self, syn: np.ndarray, dt: float, width: float):
        self.dt = dt
        self.width = width
        self.centers = np.array([arrival_center(u, dt) for u in syn])
        nt = syn.shape[1]
        self.windows = np.stack([gaussian_window(nt, dt, c, width) for c in self.centers])

    def travel_times(self, obs: np.ndarray, syn: np.ndarray) -> np.ndarray:
        return np.array([compute_tt(w * o, w * s, self.dt)
                         for w, o, s in zip(self.windows, obs, syn)])

    def misfit(self, tau: np.ndarray) -> float:
        """0.5 * sum tau^2 (driver.f90:908, MISFIT_ALL)."""
        return 0.5 * float(np.sum(tau ** 2))

    def adjoint_source(self, syn: np.ndarray, tau: np.ndarray, obs: np.ndarray | None = None,
                       exact: bool = False) -> np.ndarray:
        """d(misfit)/d(u_i(t)) per receiver as a density in time (the misfit
        changes by sum_m a(t_m) du(t_m) dt), in forward time order.

        Default: the reference's linearised form (ADJOINT_SOURCE_COMP,
        driver.f90:870-877): tau * W(t), W = -w d(wu)/dt / (sum (d(wu)/dt)^2 dt)
        (with the reference's box window w in {0,1} this is its
        -ccdot/(sum ccdot^2 dt) over the window).
        exact=True: the derivative of the discrete estimator itself
        (compute_tt_grad on the windowed traces; needs obs), which the
        parabola fit through cadence-spaced correlation samples makes differ
        from the linearised form by the fit's bias."""
        out = np.zeros_like(syn)
        for r, (w, s) in enumerate(zip(self.windows, syn)):
            if exact:
                g = compute_tt_grad(w * obs[r], w * s, self.dt)
                out[r] = tau[r] * w * g / self.dt
                continue
            ds = time_derivative(w * s, self.dt)
            den = float(np.sum(ds ** 2) * self.dt)
            if den > 0.0:
                out[r] = -tau[r] / den * w * ds
        return out


def ring_pixels(nx: int, ny: int, xlength: float, ylength: float, x0_mm: float, y0_mm: float,
                radius_mm: float, n: int = 8) -> np.ndarray:
    """(n, 2) 0-based (i, j) grid points nearest to n points on a ring around
    (x0, y0) Mm, coordinates measured from the box centre like (x-0.5)*xlength."""
    pts = []
    dxm = xlength / 1e8 / nx
    dym = ylength / 1e8 / ny
    for q in range(n):
        ang = 2.0 * math.pi * q / n
        xm = x0_mm + radius_mm * math.cos(ang)
        ym = y0_mm + radius_mm * math.sin(ang)
        i = int(round(xm / dxm + nx / 2)) % nx
        j = int(round(ym / dym + ny / 2)) % ny
        pts.append((i, j))
    return np.array(pts, dtype=np.int64)
