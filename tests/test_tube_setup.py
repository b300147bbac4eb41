"""Flux-tube background and the reference's magnetic set-up
(physics.f90:113-160, 204-316) on the 32 x 32 x 100 tube case."""
import numpy as np
import pytest

from sparc_amd import abi
from sparc_amd import background as B
from sparc_amd import driver as D


@pytest.fixture(scope="module")
def tube():
    W = D.workload("tube")
    P = W.params
    f = B.flux_tube_vector_potential(W.tab, P.nx, P.ny, P.xlength, P.ylength, **W.extra["tube"])
    S = D.setup_for(W)
    return W, f, S


def test_field_is_divergence_free_on_the_grid(tube):
    W, f, S = tube
    P = W.params
    sx, sy = B.RSUN_R / P.xlength, B.RSUN_R / P.ylength
    stretch = 1.0 / B.dbyd2(W.tab[:, 0][None, :], 1)[0]
    div = B.ddh(f["bx"], 0, sx) + B.ddh(f["by"], 1, sy) + B.ddz(f["bz"], stretch, 1)
    dbz = B.ddz(f["bz"], stretch, 1)
    # zero up to the plane-mean mode and the modes whose derivatives the
    # spectral operators drop (Nyquist in both x and y)
    resid = div - div.mean(axis=(0, 1), keepdims=True)
    assert np.abs(resid).max() < 1e-7 * np.abs(dbz).max()
    assert np.abs(div).max() < 2e-3 * np.abs(dbz).max()


def test_tube_profile_and_pressure_balance(tube):
    W, f, S = tube
    P = W.params
    c = P.nx // 2
    z = f["z"]
    assert abs(np.interp(0.0, z, f["bz"][c, c, :]) - 3000.0) < 1.0
    # flux pi R^2 B0 is the same at every height (Gaussian integral, box >> R)
    dA = (P.xlength / P.nx) * (P.ylength / P.ny)
    flux = f["bz"].sum(axis=(0, 1)) * dA
    assert np.allclose(flux, flux[-1], rtol=1e-3)
    # lateral balance at the outside temperature
    p0 = W.tab[:, 3][None, None, :]
    assert np.allclose(f["p"] + f["bz"] ** 2 / (8 * np.pi), p0 + 0 * f["p"], rtol=1e-12)
    assert f["p"].min() > 0
    assert np.allclose(f["rho"] / f["p"], (W.tab[:, 2] / W.tab[:, 3])[None, None, :], rtol=1e-12)


def test_magnetic_setup_restates_the_reference(tube):
    W, f, S = tube
    P = W.params
    rho0 = f["rho"] / B.DIMRHO
    c2 = (f["c"] / B.DIMC) ** 2
    braw = [f[k] / B.DIMB for k in ("bx", "by", "bz")]
    cq = B.solar_data(W.tab)["c"]
    red = W.extra["cap2"] / (W.extra["cap2"] + sum(b ** 2 for b in braw) / (rho0 * cq ** 2))
    nz = P.nz
    for k1 in range(nz - 10, nz + 1):
        red[:, :, k1 - 1] *= ((k1 - nz) / 10.0) ** 4
    m = S.mhd
    for bid, b in zip((abi.BG_BX, abi.BG_BY, abi.BG_BZ), braw):
        assert np.allclose(m[bid], b * np.sqrt(red), rtol=1e-13, atol=0)
    assert np.all(m[abi.BG_BZ][:, :, -1] == 0)          # top plane tapered to zero
    va2 = sum(m[b] ** 2 for b in (abi.BG_BX, abi.BG_BY, abi.BG_BZ)) / rho0
    assert np.all(va2 <= W.extra["cap2"] * cq ** 2 * (1 + 1e-12))  # v_A <= 4 c_q
    sx, sy = B.RSUN_R / P.xlength, B.RSUN_R / P.ylength
    assert np.allclose(m[abi.BG_CURLBX], B.ddh(m[abi.BG_BZ], 1, sy))
    assert np.allclose(m[abi.BG_CURLBY], -B.ddh(m[abi.BG_BZ], 0, sx))
    assert np.allclose(m[abi.BG_CURLBZ], B.ddh(m[abi.BG_BY], 0, sx) - B.ddh(m[abi.BG_BX], 1, sy))
    assert np.allclose(m[abi.BG_GRADP0_X], B.ddh(S.p0, 0, sx))
    # PML damping from the fast speed at the boundary planes
    fast = np.sqrt(S.c2 + va2)
    L = S.z[P.npmlbot - 1] - S.z[0]
    d0 = 2 * 3 * 3 * fast[:, :, 0] / L * 1.0  # plane 1: ((z(npml)-z(1))/L)^2 = 1
    assert np.allclose(-S.derived["az"][:, :, 0], d0)


def test_ddh_is_spectral():
    n = 32
    x = np.arange(n) / n
    f = np.sin(2 * np.pi * 3 * x)[:, None, None] * np.ones((1, 2, 3))
    d = B.ddh(f, 0, 2.0)
    assert np.allclose(d, 2.0 * 2 * np.pi * 3 * np.cos(2 * np.pi * 3 * x)[:, None, None], atol=1e-12)
    g = np.cos(np.pi * n * x)[:, None, None] * np.ones((1, 2, 3))  # Nyquist dropped
    assert np.abs(B.ddh(g, 0, 1.0)).max() < 1e-12
