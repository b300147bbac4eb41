"""The filtered travel-time measurement (sparc_amd/measure_filt.py,
ADJOINT_SOURCE_FILT and its filters, driver.f90:1074-1686), pinned by
closed forms and by finite differences:

* distmat's wavenumber order (initialize.f90:1124-1141);
* FREQ_FILTER is a logistic band: one half at each corner, flat inside,
  zero outside, area the band width;
* the f-mode ridge filter passes the f-mode dispersion relation
  omega^2 = g k (g = 274 m/s^2) and the p-mode filter rejects it;
* the adjoint source is the gradient of the misfit 0.5 sum tau^2 with
  respect to the synthetic wavefield at fixed windows: for the unnormalised
  transforms and the reference's scalings, d chi = -(dt/nt) <adj reversed,
  d acc> with dt in minutes; checked by central differences in 2-D (the
  reference's x-line of receivers) and 3-D (an annulus and a ring).
"""
import math

import numpy as np
import pytest

from sparc_amd import measure_filt as F
from sparc_amd.inputs import RidgeParams


def test_distmat_order():
    assert np.array_equal(F.distmat(8), [0, 1, 2, 3, 4, -3, -2, -1])
    assert np.array_equal(F.distmat(7), [0, 1, 2, 3, -3, -2, -1])
    assert np.array_equal(F.distmat(1), [0])


def test_freq_filter_is_a_logistic_band():
    nt, dt = 1024, 30.0
    filt = F.freq_filter(2.0, 3.5, nt, dt)
    f = np.abs(F.distmat(nt)) * 1e3 / (nt * dt)
    df = f[1] - f[0]
    inside = (f > 2.3) & (f < 3.2)
    outside = (f < 1.6) | (f > 3.9)
    assert np.all(filt[inside] > 0.99) and np.all(np.abs(filt[outside]) < 1e-3)
    # odd about each corner: filt(f1 + d) + filt(f1 - d) = 1 away from the other corner
    for fc in (2.0, 3.5):
        g = lambda v: np.interp(v, f[:nt // 2], filt[:nt // 2])
        for d in (0.5 * df, 1.3 * df):
            s = g(fc + d) + g(fc - d)
            assert abs(s - 1.0) < 0.02, (fc, d, s)
    # area of the positive-frequency band = f2 - f1
    assert abs(np.sum(filt[:nt // 2]) * df - 1.5) < 0.01


def test_fmode_filter_passes_the_fmode_ridge():
    nx, nt, dt = 256, 1024, 30.0
    L = 400e8
    kh = F.horizontal_k(nx, 1, L, L)
    fm = F.fmode_filter(kh, nt, dt)[:, 0, :]
    pm = F.pmode_filter(kh, nt, dt)[:, 0, :]
    f = F.freq_mhz(nt, dt)
    g = 274.0  # m/s^2
    hits = 0
    for ik in range(1, nx // 2):
        k = kh[ik, 0]
        if not 0.5 < k < 1.5:
            continue
        ff = math.sqrt(g * k * 1e-6) / (2 * math.pi) * 1e3  # mHz, omega^2 = g k
        j = int(np.argmin(np.abs(f[:nt // 2] - ff)))
        assert fm[ik, j] > 0.5, (k, ff, fm[ik, j])
        assert pm[ik, j] == 0.0, (k, ff)
        assert fm[ik, np.argmin(np.abs(f[:nt // 2] - 1.0))] == 0.0  # below f_low = 1.1 mHz
        hits += 1
    assert hits > 20
    assert np.all((fm >= 0) & (fm <= 1)) and np.all((pm >= 0) & (pm <= 1))


def _packets(nx, ny, L, src, nt, cad, shift_s=0.0, amp=1.0):
    """Outgoing wave packets at 3 mHz with group and phase speed 0.75 Mm/min,
    Gaussian envelope of 300 s, as a ridge-filtered f mode looks."""
    x = np.arange(nx) / nx
    y = np.arange(ny) / ny if ny > 1 else np.zeros(1)
    xm = (x - 0.5) * L * 1e-8 - src[0]
    ym = (y - 0.5) * L * 1e-8 - src[1] if ny > 1 else np.zeros(1)
    r = np.sqrt(xm[:, None] ** 2 + ym[None, :] ** 2)
    t = np.arange(nt) * cad
    v = 0.75 / 60.0  # Mm/s
    tt = t[None, None, :] - r[..., None] / v - shift_s
    return amp * np.exp(-(tt / 300.0) ** 2) * np.cos(2 * np.pi * 3e-3 * tt)


@pytest.mark.parametrize("case", ["2d_line", "3d_annulus", "3d_ring"])
def test_adjoint_source_is_the_misfit_gradient(case):
    cad = 10.0
    nt = 512
    L = 100e8
    nx, ny = (64, 1) if case == "2d_line" else (48, 48)
    src = (-5.0, 0.0) if case == "2d_line" else (2.0, -3.0)
    acc = _packets(nx, ny, L, src, nt, cad)
    dat = _packets(nx, ny, L, src, nt, cad, shift_s=12.0, amp=0.9)
    ridge = {0: RidgeParams(10.0, 30.0, 20.0)}
    bands = {0: (1.0, 4.0)}
    rec = None
    if case == "3d_ring":
        rec = np.array([[int(round((src[0] + 15 * math.cos(a)) / (L * 1e-8 / nx) + nx / 2)),
                         int(round((src[1] + 15 * math.sin(a)) / (L * 1e-8 / ny) + ny / 2))]
                        for a in np.linspace(0, 2 * math.pi, 8, endpoint=False)])
    res = F.adjoint_source_filt(acc, dat, cad, L, L, src, ridge, bands, receivers=rec)
    assert res.nmeas == (8 if case == "3d_ring" else res.nmeas) and res.nmeas > 0
    taus = np.array([m.tau_min for m in res.measurements])
    # the data arrive 12 s (0.2 min) late; the box window of 20 min cuts the
    # 5-min packets, which biases the correlation by ~5 %
    assert np.all(np.abs(taus - 0.2) < 0.02), taus
    assert abs(res.misfit - 0.5 * np.sum(taus ** 2)) <= 1e-14 * res.misfit
    dt_min = cad / 60.0
    rng = np.random.default_rng(11)
    adj_fwd = res.adj[:, :, ::-1]
    fds, lins = [], []
    for _ in range(6):
        d = _packets(nx, ny, L, (src[0] + rng.uniform(-1, 1), src[1] + rng.uniform(-1, 1)),
                     nt, cad, shift_s=rng.uniform(-60, 60), amp=1e-3)
        eps = 1e-2
        chi = [F.adjoint_source_filt(acc + s * eps * d, dat, cad, L, L, src, ridge, bands,
                                     receivers=rec, windows=res.windows).misfit
               for s in (1, -1)]
        fds.append((chi[0] - chi[1]) / (2 * eps))
        lins.append(-dt_min / nt * float(np.sum(adj_fwd * d)))
    fds, lins = np.array(fds), np.array(lins)
    # the reference's source is the Born-linearised gradient and the misfit
    # uses the parabola estimator: over the six directions they agree to a
    # few per cent (a wrong normalisation would be off by factors of nt, 60
    # or the transform sizes)
    err = np.linalg.norm(lins - fds) / np.linalg.norm(fds)
    assert err < 0.05, (err, lins / fds)


def test_windows_reuse_and_files(tmp_path):
    cad, nt, L = 30.0, 256, 200e8
    acc = _packets(64, 1, L, (-5.0, 0.0), nt, cad)
    dat = _packets(64, 1, L, (-5.0, 0.0), nt, cad, shift_s=20.0)
    ridge = {0: RidgeParams(10.0, 80.0, 10.0), 1: RidgeParams(20.0, 100.0, 10.0)}
    bands = {0: (1.0, 4.0), 1: (2.0, 5.0)}
    a = F.adjoint_source_filt(acc, dat, cad, L, L, (-5.0, 0.0), ridge, bands)
    for p, w in a.windows.items():
        F.write_windows(str(tmp_path / f"windows.{p}"), w)
    wins = {p: F.read_windows(str(tmp_path / f"windows.{p}")) for p in a.windows}
    assert wins == a.windows
    b = F.adjoint_source_filt(acc, dat, cad, L, L, (-5.0, 0.0), ridge, bands, windows=wins)
    assert np.array_equal(a.adj, b.adj) and a.misfit == b.misfit
    for m in a.measurements:
        assert m.rig - m.lef == 2 * F.nint(10.0 / (2 * cad / 60.0))
