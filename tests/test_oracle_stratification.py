"""Closed forms of a hydrostatic, adiabatic polytrope that pin the
stratification terms of the oracle's right-hand sides: gravity, the
background pressure and density gradients and the buoyancy rho' g of
options 4 and 10 (displacement.f90:532-535,589-597,663-849) and of options 2
and 8 (pml.f90:38-53,80-208).

The polytrope has constant g, p0' = -rho0 g (hydrostatic) and constant
Gamma_1, so for a vertical displacement xi = (0, 0, xi_z(z)) the momentum
right-hand side
    -(1/rho0) d/dz [-rho0 c^2 xi_z' - xi_z p0'] - g [-rho0 xi_z' - xi_z rho0'] / rho0
is, in the continuum:
    xi_z = 1 (a rigid lift)          ->  -g' = 0 (a translated hydrostatic
                                          atmosphere is in equilibrium);
    xi_z = z - z0 (a uniform stretch) -> (Gamma_1 p0)'/rho0 = -Gamma_1 g.
A dropped or sign-flipped gravity, p0' or rho0' term changes either value by
O(g (1 + |xi_z rho0'/rho0|)); the lift alone misses a consistent sign flip of
both Lagrangian terms, the stretch catches it.  For the (rho, v, p) form a
uniform upward velocity gives the Eulerian rates p' = -v_z p0' = rho0 g and
rho' = -v_z rho0' (the polytrope law rho0 ~ d^m, d the depth below its T=0
level).

The right-hand side is read from one LDDRK step of a state with v = 0 and a
tiny deltat: v(deltat)/deltat = RHS + O(deltat).  The 24 planes next to
each boundary are left out: the one-sided closures of the compact z
derivative (ibc 0 for p') are not the continuum there, and their effect
decays by a factor ~0.4 per plane into the interior (1e-8 at 24 planes).
The agreement on the interior is ~1e-11 of the term size; the bar is 1e-6.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sparc_amd import abi
from sparc_amd import background as B
from sparc_amd import driver as D

M = 1.5
GAMMA1 = 5.0 / 3.0
DT = 1e-7
KLO, KHI = 24, 24


def _case(option, nz=120):
    tab = B.polytrope_table(nz, 0.5, -12.0, m=M, gamma1=GAMMA1)
    P = D.Params(nx=32, ny=4, nz=nz, xlength=25e8, ylength=25e8 / 8, timestep=1.0,
                 option=option, npmlbot=0, npmltop=0, damp_waves=False, precision=abi.FP64)
    W = D.Workload("strat", P, tab, 10)
    S = D.setup_for(W)
    return W, S


def _rhs(option, S, W, state, g_scale=1.0, gp_scale=1.0, gr_scale=1.0):
    """One tiny step from ``state`` (field slot -> z-profile); returns the
    z-profile of every field's rate of change."""
    cfg = D.make_config(W.params, S, abi.FORWARD, 10)
    cfg.freq_filtering = 10 ** 6  # the spectral filter stays off at time 1
    cfg.deltat = DT
    o = O.Oracle(cfg)
    try:
        D.load_setup(o, S)
        nx, ny, nz = S.nx, S.ny, S.nz
        if option in (abi.OPT_MHD, abi.OPT_MHD8):
            for bid in (abi.BG_BX, abi.BG_BY, abi.BG_BZ, abi.BG_CURLBX, abi.BG_CURLBY,
                        abi.BG_CURLBZ, abi.BG_GRADP0_X, abi.BG_GRADP0_Y, abi.BG_GRADRHO0_X,
                        abi.BG_GRADRHO0_Y):
                o.set_background(bid, np.zeros(nz))
            o.set_profile(abi.PROF_SPONGE, np.zeros(nz))
        d = S.derived
        o.set_profile(abi.PROF_G, g_scale * d["g"])
        o.set_background(abi.BG_GRADP0_Z, gp_scale * d["gradp0_z"][0, 0])
        o.set_background(abi.BG_GRADRHO0_Z, gr_scale * d["gradrho0_z"][0, 0])
        nf = {abi.OPT_RHOVP: 5, abi.OPT_MHD8: 8}.get(option, 6)
        before = {}
        for f in range(nf):
            prof = state.get(f, np.zeros(nz))
            before[f] = prof
            o.put(abi.ARR_A, f, np.broadcast_to(prof, (nx, ny, nz)))
        o.step(1, 1)
        out = {}
        for f in range(nf):
            a = o.get(abi.ARR_A, f)
            assert np.ptp(a, axis=(0, 1)).max() <= 1e-12 * max(1.0, np.abs(a).max())
            out[f] = (a[0, 0] - before[f]) / DT
        return out
    finally:
        o.close()


def _profiles(S):
    rho = S.rho0[0, 0]
    p = S.p0[0, 0]
    g = S.g_raw
    # polytrope law: rho ~ d^m, p = rho g d / (m + 1)  =>  d = (m+1) p / (rho g)
    d = (M + 1.0) * p / (rho * g)
    drho = -M * rho / d
    return S.z, rho, p, g, drho


def _interior(x):
    return x[KLO:x.size - KHI]


def _worst(got, want, scale):
    return float(np.max(np.abs(_interior(got) - _interior(want)) / _interior(scale)))


DISPL = [abi.OPT_3D, abi.OPT_MHD]


@pytest.mark.parametrize("option", DISPL, ids=["option4", "option10"])
def test_rigid_lift_is_static(option):
    W, S = _case(option)
    z, rho, p, g, drho = _profiles(S)
    lift = {abi.XI_Z: np.ones(S.nz)}
    scale = g * (1.0 + np.abs(drho / rho))
    r = _rhs(option, S, W, lift)[abi.V_Z]
    assert _worst(r, 0.0 * r, scale) < 1e-6
    # planted mistakes: each stratification term flipped or dropped
    for kw in (dict(g_scale=-1.0), dict(gp_scale=0.0), dict(gr_scale=0.0), dict(gp_scale=-1.0)):
        bad = _rhs(option, S, W, lift, **kw)[abi.V_Z]
        assert _worst(bad, 0.0 * r, scale) > 0.1, kw


@pytest.mark.parametrize("option", DISPL, ids=["option4", "option10"])
def test_uniform_stretch_gives_gamma_g(option):
    W, S = _case(option)
    z, rho, p, g, drho = _profiles(S)
    zz = z - z[S.nz // 2]
    stretch = {abi.XI_Z: zz}
    scale = g * (1.0 + np.abs(zz * drho / rho))
    want = -GAMMA1 * g
    r = _rhs(option, S, W, stretch)[abi.V_Z]
    assert _worst(r, want, scale) < 1e-6
    for kw in (dict(g_scale=-1.0), dict(gp_scale=-1.0, gr_scale=-1.0), dict(gr_scale=0.0),
               dict(gp_scale=0.0)):
        bad = _rhs(option, S, W, stretch, **kw)[abi.V_Z]
        assert _worst(bad, want, scale) > 0.1, kw


RVP = [abi.OPT_RHOVP, abi.OPT_MHD8]


@pytest.mark.parametrize("option", RVP, ids=["option2", "option8"])
def test_eulerian_state_of_a_stretch_gives_gamma_g(option):
    """(rho, v, p) form: the Eulerian perturbations of xi_z = z - z0,
    rho' = -rho0 - xi_z rho0', p' = -Gamma_1 p0 + xi_z rho0 g (the analytic
    polytrope gradients), accelerate like the displacement form's stretch."""
    W, S = _case(option)
    z, rho, p, g, drho = _profiles(S)
    zz = z - z[S.nz // 2]
    state = {abi.RHO: -rho - zz * drho, abi.P: -GAMMA1 * p + zz * rho * g}
    scale = g * (1.0 + np.abs(zz * drho / rho))
    r = _rhs(option, S, W, state)[3]  # v_z is slot 3 of the (rho, v, p) layout
    assert _worst(r, -GAMMA1 * g, scale) < 1e-6
    bad = _rhs(option, S, W, state, g_scale=-1.0)[3]
    assert _worst(bad, -GAMMA1 * g, scale) > 0.1


@pytest.mark.parametrize("option", RVP, ids=["option2", "option8"])
def test_uniform_upflow_advects_the_background(option):
    """v = (0, 0, 1): the continuity and pressure equations give
    d rho/dt = -rho0' and d p/dt = -p0' = rho0 g (hydrostatic)."""
    W, S = _case(option)
    z, rho, p, g, drho = _profiles(S)
    up = {3: np.ones(S.nz)}
    r = _rhs(option, S, W, up)
    assert _worst(r[abi.RHO], -drho, np.abs(drho)) < 1e-6
    assert _worst(r[abi.P], rho * g, rho * g) < 1e-6
    bad = _rhs(option, S, W, up, gp_scale=-1.0, gr_scale=-1.0)
    assert _worst(bad[abi.RHO], -drho, np.abs(drho)) > 0.1
    assert _worst(bad[abi.P], rho * g, rho * g) > 0.1
