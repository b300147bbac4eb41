"""Filtered travel-time measurement and its adjoint source (FP64, host):
ADJOINT_SOURCE_FILT (driver.f90:1262-1686) with its ridge and band filters
FMODE_FILTER, PMODE_FILTER, HIGHPMODE_FILTER, FREQ_FILTER and PHASE_FILTER
(driver.f90:1074-1258), generalised from the reference's x-line of
receivers (ny = 1) to the observation plane of a 3-D box: the filters act
on |k_h| = (kx^2 + ky^2)^(1/2), receivers are the pixels whose distance
from the source lies in (mindist, maxdist) (the reference's rule) or an
explicit pixel list (a ring), and the point source of each measurement's
adjoint term is exp(-i k.x_i) in the (kx, ky) spectrum.

Conventions kept from the reference: transforms are unnormalised in both
directions (FFTW's), the measurement works in minutes (dt = cadence/60, tau
in minutes, misfit = 0.5 sum tau^2 in min^2), window indices are 1-based
and inclusive, ccdot is the spectral time derivative of the filtered
synthetic (i 2 pi nu, nu in Hz, the signed wavenumber list of distmat with
+n/2 at Nyquist), con = -tau / (sum ccdot^2 dt), and the source is
time-reversed and scaled by nt/60 at the end.  The adjoint terms of all
receivers are summed in the spectral domain before one inverse transform
(the reference transforms each one; the sum is the same up to rounding).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .background import PI_R, sp
from .inputs import RidgeParams

EYE = 1j
# default-real literals of the reference are single precision (no -r8,
# makefile:27): 10.**(-8.), the filter constants and pi (initialize.f90:38)
E8 = sp(1e-8)


def nint(x: float) -> int:
    """Fortran NINT: nearest integer, halves away from zero."""
    return int(math.copysign(math.floor(abs(x) + 0.5), x))


def distmat(n: int) -> np.ndarray:
    """Signed integer wavenumbers 0, 1, ..., n/2, -(n/2 - 1), ..., -1 with
    +n/2 at Nyquist (distmat(n, 1), initialize.f90:1124-1141)."""
    i = np.arange(1, n + 1)
    i2 = np.minimum(i - 1, n - i + 1).astype(np.float64)
    sig = np.where((i - 1) > (n - i + 1), -1.0, 1.0)
    return i2 * sig


def _fwd(a: np.ndarray) -> np.ndarray:
    """FFTW forward (sign -1), unnormalised."""
    return np.fft.fftn(a)


def _bwd(a: np.ndarray) -> np.ndarray:
    """FFTW backward (sign +1), unnormalised."""
    return np.fft.ifftn(a) * a.size


def horizontal_k(nx: int, ny: int, xlength: float, ylength: float, ridge: bool = True):
    """|k_h| in Mm^-1 on the (kx, ky) grid.  The ridge filters divide by
    L n/(n-1) (driver.f90:1091-1092), the phase filter by L (:1249)."""
    def axis(n, L):
        if n == 1:
            return np.zeros(1)
        Lm = L * E8 * n / (n - 1.0) if ridge else L * E8
        return np.abs(distmat(n)) * 2.0 * PI_R / Lm
    kx = axis(nx, xlength)
    ky = axis(ny, ylength)
    return np.sqrt(kx[:, None] ** 2 + ky[None, :] ** 2)


def freq_mhz(nt: int, dt_s: float) -> np.ndarray:
    """|nu| in mHz of the nt-point transform (w/(2 pi)*1e3, driver.f90:1094-1097)."""
    w = np.abs(distmat(nt)) * 2.0 * PI_R / (nt * dt_s)
    return w / (2.0 * PI_R) * 1e3


def _low_taper(filt: np.ndarray, f: np.ndarray, f_low: float, df: float, zero_first: bool):
    """The low-frequency taper of the ridge filters: below f_low + df times
    0.5 (1 + cos(pi (f - f_low - df)/df)), zero below f_low; FMODE_FILTER
    applies the taper first, the p-mode filters zero first (the order only
    matters at f = f_low exactly)."""
    tap = np.where(f < f_low + df, 0.5 * (1.0 + np.cos(PI_R * (f - (f_low + df)) / df)), 1.0)
    low = f < f_low
    if zero_first:
        filt[..., low] = 0.0
        filt *= tap
    else:
        filt *= tap
        filt[..., low] = 0.0
    return filt


def _raised_band(k, f, f0, f1):
    """0.5 (1 + cos(pi (2 d/delta - 1))) for 0 < d = f - f0 < delta = f1 - f0."""
    delta = (f1 - f0)[..., None]
    d = f[None, None, :] - f0[..., None]
    inside = (d < delta) & (d > 0)
    return np.where(inside, 0.5 * (1.0 + np.cos(PI_R * (2.0 * d / np.where(inside, delta, 1.0)
                                                         - 1.0))), 0.0)


def fmode_filter(kh: np.ndarray, nt: int, dt_s: float) -> np.ndarray:
    """FMODE_FILTER (driver.f90:1074-1113): the band between
    f0 = 2.056565 sqrt(k) and f1 = 1 + 3 k - 0.6 k^2 mHz, raised-cosine
    across it, tapered below 1.6 mHz and zero below 1.1 mHz."""
    f = freq_mhz(nt, dt_s)
    f0 = sp(2.056565) * kh ** 0.5
    f1 = 1.0 + 3.0 * kh + sp(-0.6) * kh ** 2
    return _low_taper(_raised_band(kh, f, f0, f1), f, sp(1.1), 0.5, zero_first=False)


def pmode_filter(kh: np.ndarray, nt: int, dt_s: float) -> np.ndarray:
    """PMODE_FILTER (driver.f90:1162-1208): between f0 = 3.96 sqrt(k) and
    f1 = 1.1 + 5.2 k - 1.3 k^2 mHz, zero below 1.6 mHz, tapered to 2.1."""
    f = freq_mhz(nt, dt_s)
    f0 = sp(3.96) * np.abs(kh) ** 0.5
    f1 = sp(1.1) + sp(5.2) * kh + sp(-1.3) * kh ** 2
    return _low_taper(_raised_band(kh, f, f0, f1), f, sp(1.6), 0.5, zero_first=True)


def highpmode_filter(kh: np.ndarray, nt: int, dt_s: float) -> np.ndarray:
    """HIGHPMODE_FILTER (driver.f90:1117-1158): flat within half the p-band
    width of f0 = 2 sqrt(254e-6 k)/(2 pi) kHz, raised-cosine edges over the
    outer half."""
    f = freq_mhz(nt, dt_s)
    k = np.abs(kh)
    f0 = 2.0 * (sp(254e-6) * k) ** 0.5 / (2 * PI_R) * 1e3
    f1 = sp(1.85) * (sp(0.0080513021) + sp(0.02369524) * k + sp(-0.00460597) * k ** 2) \
        / (2 * PI_R) * 1e3
    delta = ((f1 - f0) * 0.5)[..., None]
    d = np.abs(f[None, None, :] - f0[..., None])
    ad = np.abs(delta)
    out = np.where(d < delta, 1.0, 0.0)
    edge = (d < delta) & (d > delta * 0.5)
    out = np.where(edge, 0.5 * (1.0 + np.cos(PI_R * (d - ad * 0.5) / np.where(edge, ad * 0.5, 1.0))),
                   out)
    return _low_taper(out, f, sp(1.6), 0.5, zero_first=True)


def freq_filter(f1: float, f2: float, nt: int, dt_s: float) -> np.ndarray:
    """FREQ_FILTER (driver.f90:1213-1227): 1/(1 + e^((f1 - f)/wid)) -
    1/(1 + e^((f2 - f)/wid)), wid = 0.4 of the frequency step (mHz)."""
    w = np.abs(distmat(nt)) * 1e3 / (nt * dt_s)
    wid = sp(0.4) * (w[1] - w[0])
    with np.errstate(over="ignore"):
        return 1.0 / (1.0 + np.exp((f1 - w) / wid)) - 1.0 / (1.0 + np.exp((f2 - w) / wid))


def phase_filter(speed: float, var: float, nx: int, ny: int, xlength: float, ylength: float,
                 nt: int, dt_s: float) -> np.ndarray:
    """PHASE_FILTER (driver.f90:1231-1256): a Gaussian in phase speed
    w/(k + 0.01) about `speed` (km/s with w in rad/s x 1e3 and k in Mm^-1)
    times a high-pass at 2 mHz."""
    w = np.abs(distmat(nt)) * 2.0 * PI_R / (nt * dt_s)
    dw = w[1] - w[0]
    kay = horizontal_k(nx, ny, xlength, ylength, ridge=False)
    with np.errstate(over="ignore"):
        hp = 1.0 / (1.0 + np.exp((sp(0.002) * 2.0 * PI_R - w) / (dw * 0.5)))
    return np.exp(-(w[None, None, :] * 1e3 / (kay[..., None] + sp(0.01)) - speed) ** 2
                  / (2.0 * var ** 2)) * hp[None, None, :]


def compute_tt(u0: np.ndarray, u: np.ndarray, dt: float) -> float:
    """COMPUTE_TT (driver.f90:921-960) on one window (see measure.compute_tt)."""
    from .measure import compute_tt as ctt
    return ctt(u0, u, dt)


# window picking of the phase-speed measurement (pord = 2): arrival time in
# minutes against distance in Mm, pcoef(:, 1) of driver.f90:1540-1546
PCOEF_P2 = (sp(13.4090), sp(0.3421), sp(-0.0005), 0.0, 0.0)
RIDGE_SPEED = (0.75, 1.00)  # Mm per minute: f mode, p1 mode (driver.f90:1463-1464)


@dataclass
class Measurement:
    pord: int
    pixel: tuple        # 0-based (i, j)
    tau_min: float
    lef: int            # 1-based inclusive window
    rig: int
    loc: int


@dataclass
class FiltResult:
    adj: np.ndarray                 # (nx, ny, nt) adjoint source, reversed time, x nt/60
    misfit: float                   # 0.5 sum tau^2 (min^2)
    nmeas: int
    measurements: list = field(default_factory=list)
    windows: dict = field(default_factory=dict)   # pord -> [(lef, rig)] in pixel order


def _timeline_t1(nt: int, steps: int, timestep: float, solartime_h: float) -> float:
    """t(1) in minutes as ADJOINT_SOURCE_FILT forms it from the forward
    timeline file (driver.f90:1313-1318): the file holds (time, timeline)
    for time = 0, steps, ..., timeline = -solartime*3600 + (time+1)*timestep
    (driver.f90:453-457); it is read into t in reverse and t = -t/60."""
    last = -solartime_h * 3600.0 + ((nt - 1) * steps + 1) * timestep
    return -last / 60.0


def adjoint_source_filt(acc: np.ndarray, dat: np.ndarray, cadence_s: float,
                        xlength: float, ylength: float, src_mm: tuple,
                        ridge: dict, bands: dict, t1_min: float | None = None,
                        receivers: np.ndarray | None = None,
                        windows: dict | None = None) -> FiltResult:
    """ADJOINT_SOURCE_FILT.  acc, dat: (nx, ny, nt) v_z at the observation
    plane (synthetic and observed) every cadence_s seconds, forward time
    order.  ridge: pord -> RidgeParams (params.0 f mode, params.1 p1 mode,
    params.2 the band-only measurement); bands: pord -> (f1, f2) mHz
    (filter.params.0 / .1).  receivers: optional (n, 2) 0-based pixels
    replacing the distance-annulus rule.  windows: pord -> [(lef, rig)] of an
    earlier run to reuse (the line search's windows.0/.1, driver.f90:
    1325-1336); otherwise they are picked on the synthetic."""
    acc = np.asarray(acc, dtype=np.float64)
    dat = np.asarray(dat, dtype=np.float64)
    nx, ny, nt = acc.shape
    dt = cadence_s / 60.0  # minutes
    x = np.arange(nx) / nx
    y = np.arange(ny) / ny if ny > 1 else np.zeros(1)
    xm = (x - 0.5) * xlength * E8 - src_mm[0]
    ym = (y - 0.5) * ylength * E8 - src_mm[1] if ny > 1 else np.zeros(1)
    dist = np.sqrt(xm[:, None] ** 2 + ym[None, :] ** 2)
    freqnu = distmat(nt) / (nt * dt * 60.0)  # Hz, signed
    kx, ky = distmat(nx), distmat(ny)
    filtout = _fwd(acc.astype(np.complex128))
    filtdat = _fwd(dat.astype(np.complex128))
    kh = horizontal_k(nx, ny, xlength, ylength, ridge=True)
    fmode = fmode_filter(kh, nt, cadence_s)
    pmode = pmode_filter(kh, nt, cadence_s)
    filt = np.ones(nt)
    if 0 in bands:
        filt = freq_filter(bands[0][0], bands[0][1], nt, cadence_s)
    fmode = fmode * filt[None, None, :]
    if 1 in bands:
        filt = freq_filter(bands[1][0], bands[1][1], nt, cadence_s)
    pmode = pmode * filt[None, None, :]
    highp = np.broadcast_to(filt[None, None, :], (nx, ny, nt))
    if receivers is None:
        pix = None
    else:
        pix = [tuple(int(v) for v in p) for p in np.asarray(receivers)]
    spec_sum = np.zeros((nx, ny, nt), dtype=np.complex128)
    res = FiltResult(adj=np.zeros((nx, ny, nt)), misfit=0.0, nmeas=0)
    sq = 0.0
    for pord in (0, 1, 2):
        if pord not in ridge:
            continue
        rp: RidgeParams = ridge[pord]
        halftime = nint(rp.window / (2.0 * dt))
        flt = fmode if pord == 0 else (pmode if pord == 1 else highp)
        acc_f = _bwd(filtout * flt)
        dat_f = _bwd(filtdat * flt)
        ccdot = _bwd(filtout * flt * (EYE * freqnu * 2.0 * PI_R)[None, None, :])
        inside = [(i, j) for i in range(nx) for j in range(ny)
                  if rp.mindist < dist[i, j] < rp.maxdist]
        chosen = inside if pix is None else [p for p in pix if p in set(inside)]
        reuse = windows.get(pord) if windows else None
        picked = []
        for n, (i, j) in enumerate(chosen):
            d = dist[i, j]
            if reuse is not None:
                lef, rig = reuse[n]
                loc = (lef + rig) // 2
            elif pord < 2:
                # the reference indexes acc(timest:timefin) unchecked; the
                # window start is kept inside the trace here
                timest = min(max(int(math.floor(d / RIDGE_SPEED[pord] * 1.0 / dt)), 1), nt)
                timefin = min(timest + 40, nt)
                seg = np.abs(acc_f[i, j, timest - 1:timefin].real)
                loc = int(np.argmax(seg)) + timest
                lef, rig = loc - halftime, loc + halftime
            else:
                if t1_min is None:
                    raise ValueError("the phase-speed measurement needs t(1) (t1_min)")
                if d < 250.0:
                    c = PCOEF_P2
                    tm = nint((c[0] - t1_min + c[1] * d + c[2] * d ** 2 + c[3] * d ** 3
                               + c[4] * d ** 4) / dt) + 1
                    seg = acc_f[i, j, tm - 7:tm + 6].real
                    loc = int(np.argmax(seg)) + tm - 6
                else:
                    loc = int(np.argmax(acc_f[i, j, :].real)) + 1
                lef, rig = loc - halftime, loc + halftime
            picked.append((lef, rig))
            # 0-based slice of the inclusive window; the reference reads past
            # the trace when a window overhangs it, here it is cut at the ends
            s0, s1 = max(lef - 1, 0), min(rig, nt)
            tau = compute_tt(acc_f[i, j, s0:s1].real, dat_f[i, j, s0:s1].real, dt)
            win = np.zeros(nt)
            win[s0:s1] = 1.0
            ccdotone = np.fft.fft(ccdot[i, j, :] * win)
            con = -tau / (np.sum(ccdot[i, j, s0:s1] ** 2) * dt)
            phase = np.exp(-EYE * 2.0 * PI_R * (kx[:, None] * x[i] + ky[None, :] * y[j]))
            spec_sum += (con * phase)[:, :, None] * ccdotone[None, None, :]
            sq += tau ** 2
            res.nmeas += 1
            res.measurements.append(Measurement(pord, (i, j), tau, lef, rig, loc))
        res.windows[pord] = picked
        if chosen:
            filtex = _bwd(flt * spec_sum)
            res.adj += filtex[:, :, ::-1].real
            spec_sum[:] = 0.0
    res.adj *= nt / 60.0
    res.misfit = 0.5 * sq
    return res


def write_windows(path: str, wins: list) -> None:
    """windows.0 / windows.1 of the first measurement (driver.f90:1474):
    one 'lef rig' line per measured pixel."""
    with open(path, "w") as f:
        for lef, rig in wins:
            f.write(f"{lef:12d}{rig:12d}\n")


def read_windows(path: str) -> list:
    out = []
    with open(path) as f:
        for ln in f:
            if ln.strip():
                a, b = ln.split()[:2]
                out.append((int(a), int(b)))
    return out
