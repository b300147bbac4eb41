"""Pins the FP64 CPU oracle: against the reference's own compiled DBYD2
(golden vectors), numpy's FFT, and closed forms of each operator."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from sparc_amd import abi
from sparc_amd import background as B
from tests.helpers import rel_err

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(GOLD, "dbyd2_ref.npz"))


@pytest.mark.parametrize("n2", [12, 100, 150, 300])
@pytest.mark.parametrize("ibc", range(6))
def test_dbyd2_matches_reference_build(gold, n2, ibc):
    Bm, A0, A = gold[f"B_{n2}_{ibc}"], gold[f"A0_{n2}_{ibc}"], gold[f"A_{n2}_{ibc}"]
    out = O.dbyd2(A0.copy(), Bm, ibc)
    assert np.array_equal(out, A), np.abs(out - A).max()


@pytest.mark.parametrize("n2", [12, 100, 300])
@pytest.mark.parametrize("ibc", range(5))
def test_product_setup_dbyd2_matches_reference_build(gold, n2, ibc):
    # the product's numpy DBYD2 (used for set-up gradients) against the reference build
    Bm, A0, A = gold[f"B_{n2}_{ibc}"], gold[f"A0_{n2}_{ibc}"], gold[f"A_{n2}_{ibc}"]
    out = B.dbyd2(Bm, ibc, A0.copy())
    assert rel_err(out, A) < 1e-13


def test_dbyd2_live_reference_build_when_present():
    if O.ref_dbyd2() is None:
        pytest.skip("oracle/_ref not built (no reference tree here)")
    rng = np.random.default_rng(7)
    for ibc in range(6):
        Bm = rng.standard_normal((4, 77))
        A0 = rng.standard_normal((4, 77))
        assert np.array_equal(O.dbyd2(A0.copy(), Bm, ibc), O.call_ref_dbyd2(A0.copy(), Bm, ibc))


def test_dbyd2_polynomials():
    # compact scheme + 4th/5th order closures: cubics are differentiated to the
    # precision of the reference's single-precision closure constants
    # (-43./96., -10/3. ... are default reals, dbyd2.f90:104-125): ~1e-7 * n2
    # at the boundary rows, decaying into the interior through the solve
    n2 = 60
    s = np.linspace(0.0, 1.0, n2)
    for p in range(0, 4):
        f = s ** p
        d = O.dbyd2(np.zeros((1, n2)), f[None, :], 1)[0]
        exact = p * s ** (p - 1) if p > 0 else np.zeros(n2)
        assert np.max(np.abs(d - exact)) < 1e-4 * max(1, p)
        # alpha = 1./3. is a default real too: ~1e-8 relative in the interior
        assert np.max(np.abs(d - exact)[12:-12]) < 1e-7 * max(1, p)


def test_dbyd2_matches_dense_solve_interior():
    # interior rows: (1/3) u_{j-1} + u_j + (1/3) u_{j+1} = 7/9 d (B_{j+1}-B_{j-1}) + d/36 (B_{j+2}-B_{j-2})
    rng = np.random.default_rng(3)
    n2 = 50
    Bv = rng.standard_normal(n2)
    u = O.dbyd2(np.zeros((1, n2)), Bv[None, :], 1)[0]
    d = n2 - 1.0
    a = np.float32(1.0 / 3.0)
    j = np.arange(3, n2 - 3)
    lhs = a * u[j - 1] + u[j] + a * u[j + 1]
    rhs = d * 7 / 9 * (Bv[j + 1] - Bv[j - 1]) + d / 36 * (Bv[j + 2] - Bv[j - 2])
    assert np.max(np.abs(lhs - rhs)) < 1e-10 * np.max(np.abs(rhs))


@pytest.mark.parametrize("n", [1, 6, 64, 96, 128, 256, 384])
def test_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel_err(O.fft(x, -1), np.fft.fft(x)) < 1e-13
    assert rel_err(O.fft(x, +1), np.fft.ifft(x) * n) < 1e-13


@pytest.mark.parametrize("nx", [32, 48, 64])
def test_ddx_spectral_exact(nx):
    ny, nz = 4, 3
    x = np.arange(nx) / nx
    st = 7.5
    f = np.zeros((nx, ny, nz))
    exact = np.zeros_like(f)
    for m in (1, 3, nx // 2 - 1):
        f += np.cos(2 * np.pi * m * x)[:, None, None]
        exact += (-2 * B.PI_R * m * st * np.sin(2 * np.pi * m * x))[:, None, None]
    # the Nyquist mode is dropped by c2r (derivatives.f90:36-46)
    f += np.cos(np.pi * nx * x)[:, None, None]
    assert rel_err(O.ddx(f, st), exact) < 1e-12
    g = np.transpose(f, (1, 0, 2)).copy()
    assert rel_err(O.ddy(g, st), np.transpose(exact, (1, 0, 2))) < 1e-12


def test_zfilter_closed_forms():
    nx, ny, nz = 3, 2, 30
    const = np.full((nx, ny, nz), 2.5)
    assert rel_err(O.zfilter(const, 2), const) < 1e-15
    alt = np.broadcast_to(((-1.0) ** np.arange(nz)), (nx, ny, nz)).copy()
    out = O.zfilter(alt, 2)
    # the 11-point stencil annihilates the Nyquist mode on interior planes 6..nz-5
    assert np.max(np.abs(out[:, :, 5:nz - 5])) < 1e-15
    # 2-point averages near the ends, untouched planes 1 and nz
    assert np.allclose(out[:, :, 1:5], -alt[:, :, 1:5])
    assert np.array_equal(out[:, :, 0], alt[:, :, 0])
    assert np.array_equal(out[:, :, -1], alt[:, :, -1])
    # kst = 6 (FILTER_VARS_2D) leaves planes 2..5 untouched
    out6 = O.zfilter(alt, 6)
    assert np.array_equal(out6[:, :, 1:5], alt[:, :, 1:5])


def test_zfilter_stencil_brute_force():
    rng = np.random.default_rng(11)
    f = rng.standard_normal((2, 2, 40))
    out = O.zfilter(f, 2)
    c = [0.75390625, 0.41015625, -0.234375, 0.087890625, -0.01953125, 0.001953125]
    for k in range(5, 35):
        want = c[0] * f[:, :, k] + sum(0.5 * c[m] * (f[:, :, k - m] + f[:, :, k + m]) for m in range(1, 6))
        assert np.allclose(out[:, :, k], want, rtol=0, atol=1e-14)
    # the filter's transfer function at the Nyquist wavenumber is zero and at k=0 is one
    H = c[0] + sum(c[m] * np.cos(m * np.pi) for m in range(1, 6))
    assert abs(H) < 1e-15 and abs(sum(c) - 1.0) < 1e-15


def test_lagrange_exact_on_degree6():
    cad, timestep = 30.0, 2.0
    nt = 40
    coef = np.array([0.3, -1.1, 0.7, 0.05, -0.002, 1e-4, -3e-6])
    poly = lambda t: sum(c * (t / 100.0) ** q for q, c in enumerate(coef))
    tsamp = np.arange(nt) * cad  # sample t-1 (1-based t) at (t-1)*cad
    vr = np.array([poly(tsamp)])
    for time in (60, 61, 77, 150, 300):
        t = time * timestep
        got = O.lagrange(vr, time, timestep, cad, num_steps=nt)[0]
        assert abs(got - poly(t)) < 1e-9 * max(1.0, abs(poly(t))), (time, got, poly(t))


def test_inverse3_matches_numpy():
    rng = np.random.default_rng(5)
    a = rng.standard_normal((3, 3)) + 3 * np.eye(3)
    assert rel_err(O.inverse3(a), np.linalg.inv(a)) < 1e-13


def test_compute_tt_recovers_shift():
    dt = 0.5
    t = np.arange(400) * dt
    sig, t0 = 12.0, 90.0
    for tau in (-3.3, -0.7, 0.0, 0.21, 2.6):
        u0 = np.exp(-(t - t0) ** 2 / (2 * sig ** 2))
        u = np.exp(-(t - t0 - tau) ** 2 / (2 * sig ** 2))
        est = O.compute_tt(u0, u, dt)
        assert abs(est - tau) < 0.02 * dt, (tau, est)


def test_lddrk_plane_wave_amplification():
    """One step of a horizontal acoustic mode in a uniform, unstratified box
    equals the LDDRK stability polynomial R(L dt) = sum o_j (L dt)^j applied to
    the 2x2 system xi' = v, v' = -(c k)^2 xi (step.f90:24-118)."""
    nx, ny, nz = 32, 4, 40
    cfg = abi.SparcConfig()
    cfg.abi_version = 1
    cfg.nx, cfg.ny, cfg.nz = nx, ny, nz
    cfg.option = 4
    cfg.deltat = 2e-3
    cfg.stretchx = cfg.stretchy = 1.0
    cfg.e_rad, cfg.o_rad = 5, 6
    o = O.Oracle(cfg)
    c2 = 1.7
    ones = np.ones(nz)
    o.set_background(abi.BG_RHO0, ones)
    o.set_background(abi.BG_RHOINV, ones)
    o.set_background(abi.BG_C2RHO0, c2 * ones)
    for b in (abi.BG_GRADP0_Z, abi.BG_GRADRHO0_Z, abi.BG_AZ, abi.BG_BZPML):
        o.set_background(b, 0 * ones)
    o.set_profile(abi.PROF_STRETCH, ones)
    o.set_profile(abi.PROF_G, 0 * ones)
    m = 3
    x = np.arange(nx) / nx
    k = 2 * B.PI_R * m
    xi0 = np.broadcast_to(np.cos(2 * np.pi * m * x)[:, None, None], (nx, ny, nz)).copy()
    o.put(abi.ARR_A, abi.XI_X, xi0)
    o.step(1, 1)  # time 1: no spectral filter (freq_filtering = 0)
    dt = cfg.deltat
    opt = [1.0, 1.0, 0.5, float(np.float32(0.1665579725791151184)),
           float(np.float32(0.03950410250586729987)), float(np.float32(0.007810706393337838236))]
    L = np.array([[0.0, 1.0], [-c2 * k * k, 0.0]]) * dt
    R = sum(opt[j] * np.linalg.matrix_power(L, j) for j in range(6))
    want_xi = R[0, 0] * xi0
    want_v = R[1, 0] * xi0
    mid = slice(15, 25)
    assert rel_err(o.get(abi.ARR_A, abi.XI_X)[:, :, mid], want_xi[:, :, mid]) < 1e-9
    assert rel_err(o.get(abi.ARR_A, abi.V_X)[:, :, mid], want_v[:, :, mid]) < 1e-9


def test_damping_closed_form():
    """DAMP of a single horizontal Fourier mode is rate(k) * mode * nx * ny
    (damping_rates folds normx*normy for the unnormalized FFT pair)."""
    nx, ny, nz = 32, 16, 14
    cfg = abi.SparcConfig()
    cfg.abi_version = 1
    cfg.nx, cfg.ny, cfg.nz = nx, ny, nz
    cfg.option = 4
    cfg.deltat = 1e-3
    o = O.Oracle(cfg)
    L = 50e8
    rates = B.damping_rates(nx, ny, L, L, False)
    o.set_damping(rates)
    x = np.arange(nx) / nx
    y = np.arange(ny) / ny
    mx, my = 5, 3
    v = np.broadcast_to((np.cos(2 * np.pi * (mx * x[:, None] + my * y[None, :])))[:, :, None],
                        (nx, ny, nz)).copy()
    out = o.damp_field(v, np.zeros_like(v))
    assert rel_err(-out, rates[mx, my] * nx * ny * v) < 1e-12
