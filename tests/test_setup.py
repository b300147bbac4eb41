"""Host set-up (product, numpy FP64) against the oracle's C restatement of
mp_Initialize_RHS and against the reference's shipped polytrope file."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from sparc_amd import background as B
from sparc_amd import driver as D
from tests.helpers import rel_err

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("name", ["tiny", "box1", "two_d"])
def test_init_rhs_matches_oracle(name):
    W = D.workload(name)
    P = W.params
    S = D.setup_for(W)
    sd = B.solar_data(W.tab)
    ref = O.init_rhs(P.nx, P.ny, P.nz, int(P.is2d), P.npmlbot, P.npmltop, B.DIML, B.DIMC,
                     sd["z"], sd["c2"], sd["rho0"], sd["p0"], sd["g"])
    for k in ("rhoinv", "c2rho0", "gradp0_z", "gradrho0_z", "az", "bzpml", "g", "unstretch",
              "stretch"):
        assert rel_err(S.derived[k], ref[k]) < 1e-14, k


def test_init_rhs_3d_background_matches_oracle():
    W = D.workload("tiny")
    P = W.params
    sd = B.solar_data(W.tab)
    rng = np.random.default_rng(2)
    c2 = sd["c2"][None, None, :] * (1 + 0.02 * rng.random((P.nx, P.ny, P.nz)))
    S = D.setup_for(W, c2_3d=c2)
    ref = O.init_rhs(P.nx, P.ny, P.nz, 0, P.npmlbot, P.npmltop, B.DIML, B.DIMC, sd["z"], c2,
                     sd["rho0"], sd["p0"], sd["g"])
    for k in ("c2rho0", "gradp0_z", "gradrho0_z", "az", "bzpml", "g"):
        assert rel_err(S.derived[k], ref[k]) < 1e-14, k


def test_polytrope_law_matches_shipped_file():
    """The shipped polytrope: rho ~ d^1.5, p ~ d^2.5, c^2 = Gamma p/rho, with
    d = c^2 (m+1)/(Gamma g); the generator reproduces its rows from d."""
    tab = np.loadtxt(os.path.join(GOLD, "polytrope.txt"))
    c, rho, p, g, gam = tab[:, 1], tab[:, 2], tab[:, 3], tab[:, 4], tab[:, 5]
    assert np.allclose(c ** 2, gam * p / rho, rtol=2e-5)
    d = c ** 2 * 2.5 / (gam * g)
    assert np.allclose(rho, 2.78e-7 * (d / 3.92e7) ** 1.5, rtol=2e-4)
    assert np.allclose(p, rho * g * d / 2.5, rtol=2e-5)
    # uniform acoustic travel time per row (constant c decrement in the file)
    dc = np.diff(c)
    assert np.allclose(dc, dc.mean(), rtol=1e-3)
    gen = B.polytrope_table(300, 0.5, -12.0)
    assert np.allclose(np.diff(gen[:, 1]), np.diff(gen[:, 1]).mean(), rtol=1e-9)
    assert np.allclose(gen[:, 1] ** 2, gen[:, 5] * gen[:, 3] / gen[:, 2], rtol=1e-12)


def test_solar_data_nondimensionalisation_of_shipped_file():
    tab = np.loadtxt(os.path.join(GOLD, "polytrope.txt"))
    sd = B.solar_data(tab)
    # dimrho = 10.0**(-3.0) is a default real (initialize.f90:755)
    assert B.DIMRHO == float(np.float32(1e-3)) and B.DIMRHO != 1e-3
    assert np.allclose(sd["rho0"] * B.DIMRHO, tab[:, 2], rtol=1e-15)
    assert np.allclose(sd["c2"], (tab[:, 1] / 1e6) ** 2, rtol=1e-15)
    assert np.allclose(sd["g"], tab[:, 4] * B.RSUN_R / 1e12, rtol=1e-15)
    assert B.RSUN_R == 69598945280.0 and B.PI_R == 3.1415927410125732


def test_decay_and_damping_tables():
    dx = B.decay_coeff(256)
    assert dx.shape == (129,) and abs(dx[0] * 256 - 1.0) < 1e-6 and dx[-1] < 1e-10
    r = B.damping_rates(64, 32, 50e8, 50e8, False)
    assert r.shape == (33, 32) and r[0, 0] == 0.0
    # |k| symmetry in y (damping.f90:27-35)
    assert np.array_equal(r[:, 1], r[:, 31])


def test_workload_timesteps_divide_the_cadence():
    for name in ("box1", "box2", "box2s", "box2l", "two_d"):
        P = D.workload(name).params
        assert abs(P.steps * P.timestep - 30.0) < 1e-9
        assert P.cadforcing_step == P.steps


def test_density_perturbation_set_up():
    """background.perturb_density: zero perturbation leaves every array
    unchanged; a bump moves c^2 rho0, 1/rho0 and rho0' with rho0 at fixed c
    and changes p0' by -g d(rho0) (hydrostatic), outside the PML layers."""
    W = D.workload("tiny")
    S = D.setup_for(W)
    P = W.params
    same = B.perturb_density(S, np.zeros(S.c2.shape))
    for k in ("rhoinv", "c2rho0", "gradp0_z", "gradrho0_z", "az", "bzpml", "g"):
        assert rel_err(same.derived[k], np.broadcast_to(S.derived[k], same.derived[k].shape)) \
            < 1e-15, k
    nz = P.nz
    h = np.zeros(S.c2.shape)
    h[:, :, P.npmlbot + 5:nz - P.npmltop - 5] = 0.01
    T = B.perturb_density(S, h)
    rho = np.broadcast_to(S.rho0, T.rho0.shape)
    assert np.allclose(T.rho0, rho * np.exp(h), rtol=1e-15)
    assert np.allclose(T.derived["c2rho0"], S.c2 * T.rho0, rtol=1e-15)
    assert np.allclose(T.derived["rhoinv"] * T.rho0, 1.0, rtol=1e-15)
    dp = T.derived["gradp0_z"] - S.derived["gradp0_z"]
    assert np.allclose(dp, -S.derived["g"][None, None, :] * (T.rho0 - rho), rtol=1e-12,
                       atol=1e-14 * np.abs(S.derived["gradp0_z"]).max())
    drho = B.ddz(np.asfortranarray(T.rho0 - rho), S.derived["stretch"], 1)
    inner = slice(P.npmlbot, nz - P.npmltop)
    assert np.allclose((T.derived["gradrho0_z"] - S.derived["gradrho0_z"])[:, :, inner],
                       drho[:, :, inner], rtol=1e-10, atol=1e-12 * np.abs(drho).max())
