"""Magnetic fields from vector potentials (sparc_amd/background.py):

* DBYD1, the compact x derivative of the 2-D set-up (dbyd1.f + tridag.f),
  bit for bit against the reference's own build (tests/golden/dbyd1_ref.npz,
  written from oracle/_ref by make_dbyd1_golden.py), every ibc 0..4 and the
  short-line branch;
* the 2-D field of physics2d.f90:124-140 (box = -dpsi/dz, boz = dpsi/dx by
  DBYD1) on psi = sin(2 pi x/L) zeta(z), against the closed form (with the
  (nx-1)/nx that DBYD1's index scaling puts on a periodic grid);
* the 3-D flux tube built as B = B_mean z + curl A: discrete divergence zero
  to rounding, the flux the same on every plane, B_z the Gaussian target,
  and A equal to the axisymmetric A_phi(r, z) = Phi (1 - e^(-r^2/R^2))/(2 pi r)
  less the uniform field's (B_mean/2) r near the tube.
"""
import os

import numpy as np
import pytest

from sparc_amd import background as B
from sparc_amd import driver as D

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("n1", [4, 12, 64, 256])
@pytest.mark.parametrize("ibc", range(5))
def test_dbyd1_matches_reference_build(n1, ibc):
    g = np.load(os.path.join(GOLD, "dbyd1_ref.npz"))
    out = B.dbyd1(g[f"B_{n1}_{ibc}"], ibc, g[f"A0_{n1}_{ibc}"].copy())
    assert np.array_equal(out, g[f"A_{n1}_{ibc}"])


def test_2d_field_from_vectorpsi_closed_form():
    nx, nz = 128, 100
    tab = B.polytrope_table(nz, 0.5, -12.0)
    z = tab[:, 0]
    st = 1.0 / B.dbyd2(z[None, :], 1)[0]
    L = 200e8
    sx = B.RSUN_R / L
    x = np.arange(nx) / nx
    zeta = np.exp(-((z - z.mean()) / (0.3 * np.ptp(z))) ** 2)
    psi = np.sin(2 * np.pi * x)[:, None, None] * zeta[None, None, :]
    box, boz, curl = B.magnetic_from_vectorpsi_2d(psi, st, sx)
    dzeta = B.ddz(np.broadcast_to(zeta, (1, 1, nz)).copy(), st, 1)[0, 0]
    want_bx = -np.sin(2 * np.pi * x)[:, None] * dzeta[None, :]
    # DBYD1 differentiates with respect to (i-1)/(nx-1) while the periodic grid
    # is x = (i-1)/nx (initialize.f90:195-197), so boz carries (nx-1)/nx
    want_bz = 2 * np.pi * sx * (nx - 1) / nx * np.cos(2 * np.pi * x)[:, None] * zeta[None, :]
    assert np.max(np.abs(box[:, 0] - want_bx)) < 1e-12 * np.max(np.abs(want_bx))
    # DBYD1 with ibc 0 zeroes the end points and its one-sided rows are not
    # periodic: the compact scheme meets the closed form away from them
    inner = slice(20, nx - 20)
    assert np.max(np.abs(boz[inner, 0] - want_bz[inner])) < 1e-6 * np.max(np.abs(want_bz))
    assert np.all(boz[0] == 0.0) and np.all(boz[-1] == 0.0)
    assert np.all(np.isfinite(curl))


@pytest.fixture(scope="module")
def tube():
    W = D.workload("tube")
    P = W.params
    f = B.flux_tube_vector_potential(W.tab, P.nx, P.ny, P.xlength, P.ylength, **W.extra["tube"])
    return W, f


def test_tube_divergence_free_by_construction(tube):
    W, f = tube
    P = W.params
    sx, sy = B.RSUN_R / P.xlength, B.RSUN_R / P.ylength
    st = 1.0 / B.dbyd2(np.asarray(W.tab[:, 0])[None, :], 1)[0]
    # d(B_mean)/dz = 0 exactly; the compact operator would return the residue
    # of its single-rounded closure constants on a constant
    div = B.ddh(f["bx"], 0, sx) + B.ddh(f["by"], 1, sy) + B.ddz(f["bz"] - f["b_mean"], st, 1)
    scale = np.max(np.abs(B.ddz(f["bz"], st, 1)))
    assert np.max(np.abs(div)) < 1e-11 * scale
    flux = f["bz"].sum(axis=(0, 1)) * (P.xlength / P.nx) * (P.ylength / P.ny)
    assert np.ptp(flux) < 1e-12 * np.abs(flux).mean()
    assert abs(flux.mean() - f["flux"].mean()) < 1e-6 * f["flux"].mean()


def test_tube_matches_the_axisymmetric_potential(tube):
    W, f = tube
    P = W.params
    kw = W.extra["tube"]
    zc = (W.tab[:, 0] - 1.0) * B.RSUN_R
    b0, R = B._tube_profile(zc, zc / B.M_CM, W.tab[:, 3], kw["b_ph"], kw["r_ph_mm"],
                            kw["decay_mm"], 1.5)
    x = (np.arange(P.nx) / P.nx - 0.5) * P.xlength
    y = (np.arange(P.ny) / P.ny - 0.5) * P.ylength
    X, Y = np.meshgrid(x, y, indexing="ij")
    r = np.hypot(X, Y)
    near = (r > 0) & (r < 2.5 * kw["r_ph_mm"] * B.M_CM)
    for k in (10, P.nz // 2, P.nz - 5):
        bz_t = b0[k] * np.exp(-(r / (R[k] * B.M_CM)) ** 2)
        assert np.max(np.abs(f["bz"][:, :, k] - bz_t)) < 1e-6 * b0[k]
        aphi = (-f["ax"][:, :, k] * Y + f["ay"][:, :, k] * X) / np.where(r > 0, r, 1.0)
        want = f["flux"][k] / (2 * np.pi * np.where(r > 0, r, 1.0)) * \
            (1 - np.exp(-(r / (R[k] * B.M_CM)) ** 2)) - 0.5 * f["b_mean"] * r
        # the periodic images of the tube 25 Mm away make the rest (<~0.5 %)
        assert np.max(np.abs(aphi - want)[near]) < 1e-2 * np.max(np.abs(want[near])), k
