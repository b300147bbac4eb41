"""Pins of the oracle's option 17 (MP_QUIET_PML_2D_DISPL,
physics2d.f90:948-1105) that do not lean on its own restatement.

* Interior equivalence: on an x-z state that is the same on every y row, a
  step of option 17 equals a step of option 4 (displacement.f90:482-658,
  pinned by the plane-wave and stratification closed forms) on the planes
  where neither one's boundary handling reaches.  The two differ only in the
  z-derivative boundary conditions (ibc 3 with the top preset and ibc 2
  against ibc 1 and ibc 0), which planes they zero and where their z filter
  starts; the compact solve spreads a boundary row's change by a factor
  ~0.4 per plane, so the middle planes agree to rounding.
* The documented quirks, by their effect on a state the right-hand side
  does not change (v only, a tiny time step):
  - the 2-D z filter leaves planes 2..5 alone (physics2d.f90:530,553), the
    3-D one averages them;
  - the 2-D spectral filter fires when mod(time+1, ff) == 0 on all six
    fields, the 3-D one when mod(time, ff) == 0 on fields 1..5
    (physics2d.f90:564-586, physics.f90:795,813);
  - the top preset of option 17 (physics2d.f90:959-960) makes the top
    plane's divergence zero: a z-independent xi_x with xi_z = 0 has a
    uniform pressure in option 4 (no z force beyond the residue of the
    single-rounded closure constants) but in option 17 a pressure that
    drops to zero at the top, whose z force is composed here from the
    pinned primitives (ddx, and DBYD2 bit for bit against the reference
    build).
"""
import numpy as np
import pytest

from oracle import oracle as O
from sparc_amd import abi
from sparc_amd import background as B
from sparc_amd import driver as D

NZ = 120


def _oracle(option, ny, nx=32, deltat=None, ff=10 ** 6, nz=NZ, uniform=False):
    tab = B.polytrope_table(nz, 0.5, -12.0)
    P = D.Params(nx=nx, ny=ny, nz=nz, xlength=100e8, ylength=100e8 * ny / nx,
                 timestep=D.cfl_timestep(tab, 100e8 / nx, 100e8 / nx, 0.5), option=option,
                 npmlbot=0, npmltop=0, damp_waves=False, precision=abi.FP64)
    W = D.Workload("o17", P, tab, 10)
    S = D.setup_for(W)
    cfg = D.make_config(P, S, abi.FORWARD, 10)
    cfg.freq_filtering = ff
    if deltat is not None:
        cfg.deltat = deltat
    o = O.Oracle(cfg)
    D.load_setup(o, S)
    if uniform:  # no stratification: c = rho = 1, no gravity
        for bid in (abi.BG_RHO0, abi.BG_RHOINV, abi.BG_C2RHO0, abi.BG_C2):
            o.set_background(bid, np.ones(nz))
        for bid in (abi.BG_GRADP0_Z, abi.BG_GRADRHO0_Z):
            o.set_background(bid, np.zeros(nz))
        o.set_profile(abi.PROF_G, np.zeros(nz))
    return o, S, cfg


def _xz_state(nx, nz, seed):
    rng = np.random.default_rng(seed)
    x = np.arange(nx)[:, None] / nx
    z = np.linspace(0, 1, nz)[None, :]
    out = {}
    for f in (abi.XI_X, abi.XI_Z, abi.V_X, abi.V_Z):
        v = np.zeros((nx, nz))
        for _ in range(4):
            kx = rng.integers(0, 5)
            v += rng.standard_normal() * np.cos(2 * np.pi * kx * x + rng.uniform(0, 6)) * \
                np.cos(np.pi * rng.uniform(0.5, 5) * z + rng.uniform(0, 6))
        out[f] = 1e-3 * v
    return out


def test_option17_equals_option4_in_the_interior():
    o17, S17, c17 = _oracle(abi.OPT_2D, 1)
    o4, S4, c4 = _oracle(abi.OPT_3D, 4, deltat=c17.deltat)
    try:
        st = _xz_state(32, NZ, 5)
        for f, v in st.items():
            o17.put(abi.ARR_A, f, v[:, None, :])
            o4.put(abi.ARR_A, f, np.repeat(v[:, None, :], 4, axis=1))
        o17.step(1, 2)
        o4.step(1, 2)
        mid = slice(45, NZ - 45)
        for f in (abi.XI_X, abi.XI_Z, abi.V_X, abi.V_Z):
            a17 = o17.get(abi.ARR_A, f)[:, 0, :]
            a4 = o4.get(abi.ARR_A, f)
            assert np.max(np.abs(a4 - a4[:, :1, :])) <= 1e-15 * np.max(np.abs(a4))
            scale = np.max(np.abs(a17))
            assert np.max(np.abs(a17[:, mid] - a4[:, 0, mid])) < 1e-12 * scale, f
            # near the boundaries the two formulations do differ
            assert np.max(np.abs(a17[:, :4] - a4[:, 0, :4])) > 1e-6 * scale, f
        for f in (abi.XI_Y, abi.V_Y):
            assert not np.any(o17.get(abi.ARR_A, f))
    finally:
        o17.close()
        o4.close()


def _alternating(nx, nz):
    x = np.arange(nx) / nx
    return np.cos(2 * np.pi * x)[:, None] * ((-1.0) ** np.arange(1, nz + 1))[None, :]


@pytest.mark.parametrize("option,ny", [(abi.OPT_2D, 1), (abi.OPT_3D, 4)], ids=["2d", "3d"])
def test_z_filter_plane_range(option, ny):
    """(-1)^k in v_x: the 11-point filter removes it exactly (its symbol is
    zero at Nyquist), a 2-point average flips its sign, an untouched plane
    keeps it."""
    o, S, cfg = _oracle(option, ny, deltat=1e-12)
    try:
        f = _alternating(32, NZ)
        o.put(abi.ARR_A, abi.V_X, np.repeat(f[:, None, :], ny, axis=1))
        o.step(1, 1)
        got = o.get(abi.ARR_A, abi.V_X)[:, 0, :]
        k = np.arange(1, NZ + 1)
        if option == abi.OPT_2D:
            keep = (k == 1) | ((k >= 2) & (k <= 5)) | (k == NZ)
            flip = (k >= NZ - 4) & (k <= NZ - 1)
        else:
            keep = (k == 1) | (k == NZ)
            flip = ((k >= 2) & (k <= 5)) | ((k >= NZ - 4) & (k <= NZ - 1))
        kill = ~(keep | flip)
        assert np.max(np.abs(got[:, keep] - f[:, keep])) < 1e-9
        assert np.max(np.abs(got[:, flip] + f[:, flip])) < 1e-9
        assert np.max(np.abs(got[:, kill])) < 1e-9
    finally:
        o.close()


@pytest.mark.parametrize("option,ny", [(abi.OPT_2D, 1), (abi.OPT_3D, 4)], ids=["2d", "3d"])
def test_spectral_filter_cadence(option, ny):
    """A kx = 12 mode of 32 points (Fermi factor 1/(1 + e^(13 - 32/3)) =
    0.087) in v_x and v_z, smooth in z, stepped at a tiny deltat: the 2-D
    filter fires at time ff-1 on both, the 3-D one at time ff on v_x only,
    where the y pass also scales the y-uniform mode by 1/(1 + e^(1 - ny/3))."""
    ff = 6
    x = np.arange(32) / 32
    zprof = np.cos(np.pi * np.linspace(0, 1, NZ) * 0.5)
    f = np.cos(2 * np.pi * 12 * x)[:, None] * zprof[None, :]
    fac = 1.0 / (1.0 + np.exp(13 - 32 / 3.0))
    fires2d = {ff - 1: True, ff: False}
    fires3d = {ff - 1: False, ff: True}
    for t, fires in (fires2d if option == abi.OPT_2D else fires3d).items():
        o, S, cfg = _oracle(option, ny, deltat=1e-12, ff=ff)
        try:
            for fld in (abi.V_X, abi.V_Z):
                o.put(abi.ARR_A, fld, np.repeat(f[:, None, :], ny, axis=1))
            o.step(t, 1)
            mid = slice(10, NZ - 10)
            for fld in (abi.V_X, abi.V_Z):
                got = o.get(abi.ARR_A, fld)[:, 0, mid]
                damped = fires and (option == abi.OPT_2D or fld == abi.V_X)
                yfac = 1.0 / (1.0 + np.exp(1 - ny / 3.0)) if option == abi.OPT_3D else 1.0
                want = f[:, mid] * (fac * yfac if damped else 1.0)
                assert np.max(np.abs(got - want)) < 1e-6, (t, fld)
        finally:
            o.close()


def test_top_preset_zeroes_the_top_divergence():
    """Uniform medium, xi_x = sin(2 pi x / Lx) independent of z, xi_z = 0.
    Option 4 (ibc 1) sees a z-uniform pressure and no z force.  Option 17
    presets the top row of d xi_z/dz to -d xi_x/dx (in index units) and
    solves with ibc 3, so the divergence is zero on the top plane and the
    compact solve spreads the preset a few planes down; p = -div is then
    differentiated with ibc 2 (bottom row preset to 0), plane 1 of the RHS
    is zeroed and the step's z filter runs over planes 6..nz-1.  The
    expected rate is composed from ddx (pinned against numpy's FFT) and
    DBYD2 (pinned bit for bit against the reference build)."""
    nx = 32
    x = np.arange(nx) / nx
    xi = np.sin(2 * np.pi * x)[:, None] * np.ones(NZ)[None, :]
    res = {}
    for option, ny in ((abi.OPT_2D, 1), (abi.OPT_3D, 4)):
        o, S, cfg = _oracle(option, ny, deltat=1e-9, uniform=True)
        try:
            o.put(abi.ARR_A, abi.XI_X, np.repeat(xi[:, None, :], ny, axis=1))
            o.step(1, 1)
            res[option] = (o.get(abi.ARR_A, abi.V_Z)[:, 0, :] / cfg.deltat, S, cfg)
        finally:
            o.close()
    vz4, S, cfg = res[abi.OPT_3D]
    vz17, S, cfg = res[abi.OPT_2D]
    st, un = S.derived["stretch"], S.derived["unstretch"]
    f = xi[:, None, :]
    dxdx = O.ddx(f, cfg.stretchx)
    pre = np.zeros_like(f)
    pre[:, :, -1] = -dxdx[:, :, -1] * un[-1]
    div = dxdx + O.ddz(np.zeros_like(f), 3, st, preset=pre)
    assert np.max(np.abs(div[:, :, -1])) < 1e-12 * np.max(np.abs(dxdx))
    want = -O.ddz(-div, 2, st, preset=np.zeros_like(f))
    want[:, :, 0] = 0.0
    want = O.zfilter(want, 6)[:, 0, :]
    top = slice(NZ - 30, NZ)
    assert np.max(np.abs(want[:, top])) > 1.0
    assert np.max(np.abs(vz17 - want)) < 1e-6 * np.max(np.abs(want))
    # option 4's z force on its uniform pressure is only the residue of the
    # single-rounded closure constants (-43./96., 1./6., -1./96., dbyd2.f90:104-108)
    assert np.max(np.abs(vz4[:, top])) < 1e-3 * np.max(np.abs(vz17[:, top]))
