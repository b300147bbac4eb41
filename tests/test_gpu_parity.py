"""GPU engine (through the C ABI) against the FP64 oracle on the same seeded
inputs.  FP64 engine: agreement to FFT/tridiagonal rounding (<= 1e-10
relative); FP32 engine: wavefields within 1e-5 relative L2 (the stated
tolerance of the FP32 path)."""
import numpy as np
import pytest

from oracle import oracle as O
from sparc_amd import abi
from sparc_amd import driver as D
from tests.helpers import make_oracle, rel_err, rel_l2

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL32_OP = 2e-6     # one operator in FP32 (max-abs relative)
TOL32_L2 = 1e-5     # wavefield after a run, relative L2


@pytest.fixture(scope="module")
def tiny():
    W = D.workload("tiny")
    S = D.setup_for(W)
    return W, S


@pytest.fixture(scope="module")
def wide():
    W = D.workload("wide")
    S = D.setup_for(W)
    return W, S


@pytest.fixture(scope="module")
def tall():
    W = D.workload("tall")
    S = D.setup_for(W)
    return W, S


@pytest.mark.parametrize("prec,tol", [(abi.FP64, TOL64), (abi.FP32, TOL32_OP)])
def test_segmented_zsolve_matches_oracle(tall, prec, tol):
    """nz = 200 runs the z solve in column segments: coupled exactly with ud
    in LDS in FP32, by cold-start warm-up with ud in global memory in FP64
    (the column is too long for LDS); every ibc must still match DBYD2 to
    rounding."""
    W, S = tall
    eng, cfg = D.make_engine(W, S, precision=prec)
    try:
        st = S.derived["stretch"]
        for seed in (11, 12):
            f = _field(eng.shape, seed)
            for ibc in (0, 1, 2, 3):
                preset = _field(eng.shape, 20 + ibc)
                want = O.ddz(f, ibc, st, preset=preset if ibc in (2, 3) else None)
                got = eng.apply_op(abi.OP_DDZ, f, ibc=ibc, out=preset if ibc in (2, 3) else None)
                assert rel_err(got, want) < tol * 10, (seed, ibc)
    finally:
        eng.close()


@pytest.mark.parametrize("prec", [abi.FP64, abi.FP32])
def test_step_tall_matches_oracle(tall, prec):
    W, S = tall
    eng, o = _run_pair(W, S, 4, prec)
    try:
        for f in range(6):
            a, b = eng.get(abi.ARR_A, f), o.get(abi.ARR_A, f)
            assert np.linalg.norm(b) > 0, f
            if prec == abi.FP64:
                assert rel_err(a, b) < TOL64, f
            else:
                assert rel_l2(a, b) < TOL32_L2, f
    finally:
        eng.close()
        o.close()


@pytest.fixture(scope="module")
def two_d():
    W = D.workload("two_d")
    W.params.nx = 64
    W.params.nz = 60
    W.tab = __import__("sparc_amd.background", fromlist=["x"]).polytrope_table(60, 0.5, -12.0)
    W.params.xlength = 100e8
    W.params.ylength = 100e8
    S = D.setup_for(W)
    return W, S


def _field(shape, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    x = np.arange(nx) / nx
    y = np.arange(ny) / ny
    f = rng.standard_normal(shape)
    # smooth enough to have meaningful derivatives, random enough to exercise all modes
    f += 3 * np.cos(2 * np.pi * 2 * x)[:, None, None] * np.sin(2 * np.pi * y)[None, :, None]
    return np.asfortranarray(f)


@pytest.mark.parametrize("grid", ["tiny", "wide"])
@pytest.mark.parametrize("prec,tol", [(abi.FP64, TOL64), (abi.FP32, TOL32_OP)])
def test_operators_match_oracle(request, grid, prec, tol):
    """x, y, z derivatives, filters and the damping transform; the 256-wide
    grid runs the length-256 register four-step line kernels and the
    wavefront-per-line x derivative."""
    W, S = request.getfixturevalue(grid)
    eng, cfg = D.make_engine(W, S, precision=prec)
    try:
        f = _field(eng.shape, 1)
        st = S.derived["stretch"]
        assert rel_err(eng.apply_op(abi.OP_DDX, f), O.ddx(f, cfg.stretchx)) < tol
        assert rel_err(eng.apply_op(abi.OP_DDY, f), O.ddy(f, cfg.stretchy)) < tol
        for ibc in (0, 1, 2, 3):
            preset = _field(eng.shape, 5 + ibc)
            want = O.ddz(f, ibc, st, preset=preset if ibc in (2, 3) else None)
            got = eng.apply_op(abi.OP_DDZ, f, ibc=ibc, out=preset if ibc in (2, 3) else None)
            assert rel_err(got, want) < tol * 10, ibc
        assert rel_err(eng.apply_op(abi.OP_ZFILTER, f), O.zfilter(f, 2)) < tol
        assert rel_err(eng.apply_op(abi.OP_LOWPASS_X, f), O.lowpass(f, 0, S.decay_x)) < tol
        assert rel_err(eng.apply_op(abi.OP_LOWPASS_Y, f), O.lowpass(f, 1, S.decay_y)) < tol
        o, _ = make_oracle(W, S)
        want = o.damp_field(f, np.zeros_like(f))
        got = eng.apply_op(abi.OP_DAMP, f)
        assert rel_err(-got[:, :, 1:-1], want[:, :, 1:-1]) < tol
    finally:
        eng.close()


def _run_pair(W, S, nsteps, prec, t0=0, compensated=None):
    eng, cfg = D.make_engine(W, S, precision=prec, nsteps=nsteps, compensated=compensated)
    o, _ = make_oracle(W, S, nsteps=nsteps)
    mask, series = D.forward_source(W, nsteps + 8)
    for x in (eng, o):
        x.set_source(abi.SRC_SERIES, mask, series, 10 ** 9)
    eng.step(t0, nsteps)
    o.step(t0, nsteps)
    return eng, o


@pytest.mark.parametrize("prec,comp", [(abi.FP64, None), (abi.FP32, True), (abi.FP32, False)])
def test_step_3d_matches_oracle(tiny, prec, comp):
    W, S = tiny
    eng, o = _run_pair(W, S, 14, prec, compensated=comp)  # spectral filter at time 0 and 12
    try:
        for f in range(6):
            a, b = eng.get(abi.ARR_A, f), o.get(abi.ARR_A, f)
            assert np.linalg.norm(b) > 0, f
            if prec == abi.FP64:
                assert rel_err(a, b) < TOL64, f
            else:
                assert rel_l2(a, b) < TOL32_L2, f
        # acc = scr(4..6) of the last stage, the snapshot contract
        for f in range(3, 6):
            a, b = eng.get(abi.ARR_SCR, f), o.get(abi.ARR_SCR, f)
            assert (rel_err(a, b) < TOL64) if prec == abi.FP64 else (rel_l2(a, b) < TOL32_L2)
        for f in range(2):
            a, b = eng.get(abi.ARR_PMLVARS, f), o.get(abi.ARR_PMLVARS, f)
            assert (rel_err(a, b) < TOL64) if prec == abi.FP64 else (rel_l2(a, b) < 1e-4)
    finally:
        eng.close()
        o.close()


@pytest.mark.parametrize("prec", [abi.FP64, abi.FP32])
def test_step_3d_wide_matches_oracle(wide, prec):
    """Option 4 on 256 x 256 x 40: the length-256 x/y kernels in the stage
    (derivatives, RHS epilogue of the pressure gradient, damping) and the
    spectral filter at time 0."""
    W, S = wide
    eng, o = _run_pair(W, S, 3, prec)
    try:
        for arr, fields in ((abi.ARR_A, range(6)), (abi.ARR_SCR, range(6))):
            for f in fields:
                a, b = eng.get(arr, f), o.get(arr, f)
                assert np.linalg.norm(b) > 0, (arr, f)
                assert (rel_err(a, b) < TOL64) if prec == abi.FP64 else (rel_l2(a, b) < TOL32_L2)
    finally:
        eng.close()
        o.close()


def test_step_3d_wide_plane_damping_matches_oracle(wide):
    """Option 4 on 256 x 256 x 40 in FP32 with cfg.plane_damping: the damping
    of v_x, v_y, v_z by k_damp2d (one pass per plane) instead of the forward
    x, y and inverse x passes, and the x derivative of p written plainly."""
    from dataclasses import replace
    W0, S = wide
    W = D.Workload(W0.name, replace(W0.params, plane_damping=True), W0.tab, W0.nsteps)
    eng, o = _run_pair(W, S, 3, abi.FP32)
    try:
        for arr, fields in ((abi.ARR_A, range(6)), (abi.ARR_SCR, range(6))):
            for f in fields:
                a, b = eng.get(arr, f), o.get(arr, f)
                assert np.linalg.norm(b) > 0, (arr, f)
                assert rel_l2(a, b) < TOL32_L2, (arr, f)
    finally:
        eng.close()
        o.close()


@pytest.mark.parametrize("prec", [abi.FP64, abi.FP32])
def test_step_3d_wide_odd_planes_matches_oracle(prec):
    """Option 4 on 256 x 256 x 41: the damping passes pair planes 2q, 2q+1,
    so the last pair holds one plane (in the damping transforms and in the
    pass that writes RHS_vx with its damping), and the z solves' segments
    are of unequal length."""
    W = D.workload("wide_odd")
    S = D.setup_for(W)
    assert S.rho0.shape[2] == 41
    eng, o = _run_pair(W, S, 3, prec)
    try:
        for arr, fields in ((abi.ARR_A, range(6)), (abi.ARR_SCR, range(6))):
            for f in fields:
                a, b = eng.get(arr, f), o.get(arr, f)
                assert np.linalg.norm(b) > 0, (arr, f)
                assert (rel_err(a, b) < TOL64) if prec == abi.FP64 else (rel_l2(a, b) < TOL32_L2)
    finally:
        eng.close()
        o.close()


@pytest.mark.parametrize("prec", [abi.FP64, abi.FP32])
def test_step_3d_wide_3d_density_matches_oracle(wide, prec):
    """Option 4 on 256 x 256 x 40 with a 3-D density model (a 2 % bump): the
    density backgrounds, rhoinv among them, are 3-D arrays, so the 256-wide
    x kernels read rhoinv per point in the momentum-RHS derivative and in
    the pass that writes RHS_vx with its damping."""
    W, S0 = wide
    nx, ny, nz = S0.rho0.shape
    x = np.linspace(-1, 1, nx)[:, None, None]
    y = np.linspace(-1, 1, ny)[None, :, None]
    z = np.linspace(-1, 1, nz)[None, None, :]
    rho3 = S0.rho0 * (1.0 + 0.02 * np.exp(-(x ** 2 + y ** 2 + z ** 2) / 0.1))
    S = D.setup_for(W, rho0_3d=rho3)
    assert S.rho0.ndim == 3 and np.ptp(S.rho0[:, :, nz // 2]) > 0
    eng, o = _run_pair(W, S, 3, prec)
    try:
        for arr, fields in ((abi.ARR_A, range(6)), (abi.ARR_SCR, range(6))):
            for f in fields:
                a, b = eng.get(arr, f), o.get(arr, f)
                assert np.linalg.norm(b) > 0, (arr, f)
                assert (rel_err(a, b) < TOL64) if prec == abi.FP64 else (rel_l2(a, b) < TOL32_L2)
    finally:
        eng.close()
        o.close()


@pytest.mark.parametrize("prec", [abi.FP64, abi.FP32])
def test_step_2d_option17_matches_oracle(two_d, prec):
    W, S = two_d
    eng, o = _run_pair(W, S, 90, prec)  # (time+1) % 86 == 0 fires the 2-D spectral filter
    try:
        for f in (0, 2, 3, 5):
            a, b = eng.get(abi.ARR_A, f), o.get(abi.ARR_A, f)
            assert np.linalg.norm(b) > 0
            if prec == abi.FP64:
                assert rel_err(a, b) < TOL64, f
            else:
                assert rel_l2(a, b) < TOL32_L2, f
        for f in (1, 4):  # y components stay identically zero in 2-D
            assert not np.any(eng.get(abi.ARR_A, f))
    finally:
        eng.close()
        o.close()


@pytest.mark.parametrize("flow", [False, True])
def test_adjoint_lagrange_source_and_kernels_match_oracle(tiny, flow):
    """Forward run storing snapshots at the output cadence, adjoint run driven
    by a 7-point-Lagrange source at o_rad, kernel accumulation in FP64 (with
    the flow kernels, snapshots also keep v)."""
    from dataclasses import replace
    W0, S = tiny
    W = D.Workload(W0.name, replace(W0.params, flow_kernels=flow), W0.tab, W0.nsteps)
    nsteps = 18
    steps = W.params.steps
    nslots = nsteps // steps
    fe, _ = D.make_engine(W, S, nsteps=nsteps)
    fo, _ = make_oracle(W, S, nsteps=nsteps)
    mask, series = D.forward_source(W, nsteps + 8)
    for x in (fe, fo):
        x.set_source(abi.SRC_SERIES, mask, series, 10 ** 9)
        x.snapshots_alloc(nslots)
    for s in range(nslots):
        for x in (fe, fo):
            x.step(s * steps, steps)
            x.snapshot_store(s)
    ae, acfg = D.make_engine(W, S, mode=abi.ADJOINT, nsteps=nsteps)
    ao, _ = make_oracle(W, S, mode=abi.ADJOINT, nsteps=nsteps)
    rng = np.random.default_rng(9)
    nt = acfg.forcing_length + 8
    vr = np.zeros((W.params.nx, W.params.ny, nt))
    vr[:, :, :acfg.forcing_length] = rng.standard_normal((W.params.nx, W.params.ny,
                                                          acfg.forcing_length)) * 1e-6
    ae.snapshots_copy_from(fe)
    ao.snapshots_alloc(nslots)
    for s in range(nslots):
        for f in range(9 if flow else 6):
            ao.snapshot_put(s, f, fo.snapshot_get(s, f))
    grav = S.g_raw
    for x in (ae, ao):
        x.set_source(abi.SRC_LAGRANGE, vr, None, 10 ** 9)
        x.kernel_begin(grav)
    for s in range(nslots):
        for x in (ae, ao):
            x.step(s * steps, steps)
            x.kernel_accumulate(nslots - 1 - s)
    try:
        for f in range(3, 6):
            assert rel_err(ae.get(abi.ARR_A, f), ao.get(abi.ARR_A, f)) < TOL64
        kerns = (abi.KERN_C2, abi.KERN_RHO, abi.KERN_HESSIAN)
        if flow:
            kerns += (abi.KERN_VX, abi.KERN_VY, abi.KERN_VZ)
        else:
            with pytest.raises(abi.SparcError):
                ae.kernel_get(abi.KERN_VX)
        for k in kerns:
            a, b = ae.kernel_get(k), ao.kernel_get(k)
            assert np.linalg.norm(b) > 0
            assert rel_err(a, b) < TOL64, k
    finally:
        for x in (fe, fo, ae, ao):
            x.close()


def test_sample_plane_norm(tiny):
    W, S = tiny
    eng, o = _run_pair(W, S, 6, abi.FP64)
    try:
        a = eng.get(abi.ARR_A, abi.V_Z)
        k = eng.cfg.o_rad
        assert np.array_equal(eng.plane(abi.V_Z, k), a[:, :, k - 1])
        ijk = np.array([[0, 0, 1], [3, 5, k], [31, 31, 40]])
        assert np.array_equal(eng.sample(abi.V_Z, ijk), a[ijk[:, 0], ijk[:, 1], ijk[:, 2] - 1])
        assert abs(eng.norm(abi.V_Z) - o.norm(abi.V_Z)) <= 1e-10 * o.norm(abi.V_Z)
        with pytest.raises(abi.SparcError):
            eng.sample(abi.V_Z, np.array([[32, 0, 1]]))
    finally:
        eng.close()
        o.close()


@pytest.mark.parametrize("comp", [True, False])
def test_put_get_roundtrip_and_reset(tiny, comp):
    W, S = tiny
    eng, _ = D.make_engine(W, S, precision=abi.FP32, compensated=comp)
    try:
        f = _field(eng.shape, 3)
        eng.put(abi.ARR_A, abi.XI_Z, f)
        # the compensated state keeps hi + a 16-bit low word in steps of
        # ulp(hi)/2^16: at most one step (2^-39 relative) from each FP64 value,
        # half a step except where |v - hi| rounds to ulp/2 and the word clamps
        assert rel_err(eng.get(abi.ARR_A, abi.XI_Z), f) < (2e-12 if comp else 1e-7)
        if comp:
            lo = eng.get(abi.ARR_ALO, abi.XI_Z)
            hi = eng.get(abi.ARR_A, abi.XI_Z) - lo
            assert np.array_equal(hi, f.astype(np.float32).astype(np.float64))
            # low words put back under the same high words are unchanged
            eng.put(abi.ARR_ALO, abi.XI_Z, lo)
            assert np.array_equal(eng.get(abi.ARR_ALO, abi.XI_Z), lo)
        else:
            with pytest.raises(abi.SparcError):
                eng.get(abi.ARR_ALO, abi.XI_Z)
        eng.reset()
        assert not np.any(eng.get(abi.ARR_A, abi.XI_Z))
    finally:
        eng.close()


@pytest.fixture(scope="module")
def tiny_rhovp():
    W = D.workload("tiny")
    W.params.option = abi.OPT_RHOVP
    S = D.setup_for(W)
    return W, S


@pytest.mark.parametrize("prec,comp", [(abi.FP64, None), (abi.FP32, True)])
def test_step_option2_matches_oracle(tiny_rhovp, prec, comp):
    """Option 2 ((rho, v, p) form, pml.f90:12-76): source over 9 planes with
    source_dep, PML on dvz/dz and dp/dz, damping after the boundary zeroing."""
    W, S = tiny_rhovp
    eng, o = _run_pair(W, S, 14, prec, compensated=comp)
    try:
        for f in range(5):
            a, b = eng.get(abi.ARR_A, f), o.get(abi.ARR_A, f)
            assert np.linalg.norm(b) > 0, f
            if prec == abi.FP64:
                assert rel_err(a, b) < TOL64, f
            else:
                assert rel_l2(a, b) < TOL32_L2, f
        with pytest.raises(abi.SparcError):  # option 2 has five fields (dimfive = 5)
            eng.get(abi.ARR_A, 5)
        for f in range(2):
            a, b = eng.get(abi.ARR_PMLVARS, f), o.get(abi.ARR_PMLVARS, f)
            assert (rel_err(a, b) < TOL64) if prec == abi.FP64 else (rel_l2(a, b) < 1e-4)
    finally:
        eng.close()
        o.close()


def test_option2_random_state_matches_oracle(tiny_rhovp):
    """A random state (boundary planes included) stepped 3 times: exercises
    the damping of planes 1 and nz that option 2 applies after zeroing."""
    W, S = tiny_rhovp
    eng, _ = D.make_engine(W, S, precision=abi.FP64, nsteps=3)
    o, _ = make_oracle(W, S, nsteps=3)
    try:
        for f in range(5):
            x = _field(eng.shape, 40 + f) * 1e-3
            eng.put(abi.ARR_A, f, x)
            o.put(abi.ARR_A, f, x)
        eng.step(1, 3)
        o.step(1, 3)
        for f in range(5):
            assert rel_err(eng.get(abi.ARR_A, f), o.get(abi.ARR_A, f)) < TOL64, f
        with pytest.raises(abi.SparcError):
            eng.snapshots_alloc(1)
            eng.snapshot_store(0)
    finally:
        eng.close()
        o.close()


def _mhd_case(prec=abi.FP64, nsteps=6, option=abi.OPT_MHD, mode=abi.FORWARD):
    """The tiny polytrope box with option 10: bottom PML only, a smooth 3-D
    field B0 with v_A ~ 0.3 c, its curl, horizontal background gradients and a
    top sponge ramp (the reference's sponge is zero on boxes of <= 90 planes,
    so a ramp exercises the term here)."""
    from dataclasses import replace
    W0 = D.workload("tiny")
    P = replace(W0.params, option=option, npmltop=0 if option == abi.OPT_MHD else 10,
                precision=prec)
    W = D.Workload("tiny", P, W0.tab, W0.nsteps)
    S = D.setup_for(W)
    eng, cfg = D.make_engine(W, S, precision=prec, nsteps=nsteps, mode=mode)
    o, _ = make_oracle(W, S, nsteps=nsteps, mode=mode)
    nx, ny, nz = eng.shape
    x = np.arange(nx)[:, None, None] / nx
    y = np.arange(ny)[None, :, None] / ny
    zz = np.arange(nz)[None, None, :] / nz
    scale = 0.3 * np.sqrt(S.c2 * S.rho0)
    bgs = {
        abi.BG_BX: scale * (0.5 + 0.3 * np.cos(2 * np.pi * y)),
        abi.BG_BY: scale * 0.4 * np.sin(2 * np.pi * x) * np.cos(np.pi * zz),
        abi.BG_BZ: scale * (0.8 + 0.2 * np.cos(2 * np.pi * (x + y))),
        abi.BG_CURLBX: scale * 0.1 * np.sin(2 * np.pi * (x - y)),
        abi.BG_CURLBY: scale * 0.1 * np.cos(2 * np.pi * x),
        abi.BG_CURLBZ: scale * 0.1 * np.sin(2 * np.pi * y),
    }
    gp = np.broadcast_to(S.derived["gradp0_z"], (nx, ny, nz))
    gr = np.broadcast_to(S.derived["gradrho0_z"], (nx, ny, nz))
    bgs[abi.BG_GRADP0_X] = 0.05 * gp * np.sin(2 * np.pi * x)
    bgs[abi.BG_GRADP0_Y] = 0.05 * gp * np.cos(2 * np.pi * y)
    bgs[abi.BG_GRADRHO0_X] = 0.05 * gr * np.cos(2 * np.pi * x)
    bgs[abi.BG_GRADRHO0_Y] = 0.05 * gr * np.sin(2 * np.pi * y)
    sponge = np.zeros(nz)
    sponge[nz - 12:] = np.linspace(0.0, 50.0, 12)
    for e in (eng, o):
        for bid, arr in bgs.items():
            e.set_background(bid, np.broadcast_to(arr, (nx, ny, nz)))
        e.set_profile(abi.PROF_SPONGE, sponge)
    return W, S, eng, o


@pytest.mark.parametrize("prec", [abi.FP64, abi.FP32])
def test_step_option10_matches_oracle(prec):
    """Option 10 (linear MHD, displacement.f90:663-849): flux, curl, Lorentz
    force, bottom-only PML on six psi fields, top sponge, damping, source."""
    W, S, eng, o = _mhd_case(prec, nsteps=14)
    try:
        mask, series = D.forward_source(W, 22)
        for x in (eng, o):
            x.set_source(abi.SRC_SERIES, mask, series, 10 ** 9)
            x.step(0, 14)
        for f in range(6):
            a, b = eng.get(abi.ARR_A, f), o.get(abi.ARR_A, f)
            assert np.linalg.norm(b) > 0, f
            if prec == abi.FP64:
                assert rel_err(a, b) < TOL64, f
            else:
                assert rel_l2(a, b) < TOL32_L2, f
        for f in range(6):
            a, b = eng.get(abi.ARR_PMLVARS, f), o.get(abi.ARR_PMLVARS, f)
            assert np.linalg.norm(b) > 0, f
            assert (rel_err(a, b) < TOL64) if prec == abi.FP64 else (rel_l2(a, b) < 1e-4)
    finally:
        eng.close()
        o.close()


def test_option10_random_state_matches_oracle():
    """A random state everywhere (sponge and PML planes included), 3 steps."""
    W, S, eng, o = _mhd_case(abi.FP64, nsteps=3)
    try:
        for f in range(6):
            xr = _field(eng.shape, 60 + f) * 1e-4
            eng.put(abi.ARR_A, f, xr)
            o.put(abi.ARR_A, f, xr)
        eng.step(1, 3)
        o.step(1, 3)
        for f in range(6):
            assert rel_err(eng.get(abi.ARR_A, f), o.get(abi.ARR_A, f)) < TOL64, f
    finally:
        eng.close()
        o.close()


def _mhd2d_case(prec=abi.FP64, nsteps=6):
    """Option 18 on the 64 x 1 x 60 box of the option-17 tests: B0 = (box, boz)
    with v_A ~ 0.3 c, curl boy, horizontal background gradients, 2-D damping."""
    from dataclasses import replace
    tab = __import__("sparc_amd.background", fromlist=["x"]).polytrope_table(60, 0.5, -12.0)
    P = replace(D.workload("two_d").params, nx=64, nz=60, xlength=100e8, ylength=100e8,
                option=abi.OPT_MHD2D, damp_waves=True, precision=prec)
    W = D.Workload("two_d", P, tab, 400)
    S = D.setup_for(W)
    eng, cfg = D.make_engine(W, S, precision=prec, nsteps=nsteps)
    o, _ = make_oracle(W, S, nsteps=nsteps)
    nx, ny, nz = eng.shape
    x = np.arange(nx)[:, None, None] / nx
    zz = np.arange(nz)[None, None, :] / nz
    scale = 0.3 * np.sqrt(S.c2 * S.rho0)
    gp = np.broadcast_to(S.derived["gradp0_z"], (nx, ny, nz))
    gr = np.broadcast_to(S.derived["gradrho0_z"], (nx, ny, nz))
    bgs = {abi.BG_BX: scale * (0.5 + 0.3 * np.cos(2 * np.pi * x)),
           abi.BG_BZ: scale * (0.8 + 0.2 * np.sin(2 * np.pi * x) * np.cos(np.pi * zz)),
           abi.BG_CURLBY: scale * 0.1 * np.sin(2 * np.pi * x),
           abi.BG_GRADP0_X: 0.05 * gp * np.sin(2 * np.pi * x),
           abi.BG_GRADRHO0_X: 0.05 * gr * np.cos(2 * np.pi * x)}
    for e in (eng, o):
        for bid, arr in bgs.items():
            e.set_background(bid, np.broadcast_to(arr, (nx, ny, nz)))
    return W, S, eng, o


@pytest.mark.parametrize("prec", [abi.FP64, abi.FP32])
def test_step_option18_matches_oracle(prec):
    """Option 18 (2-D linear MHD, physics2d.f90:724-943) through a (time+1)%ff
    spectral-filter firing: ibc 3/2 presets, PML top and bottom on four psi
    fields, 2-D damping of v_x and v_z, plane 1 zeroed."""
    W, S, eng, o = _mhd2d_case(prec, nsteps=90)
    try:
        mask, series = D.forward_source(W, 98)
        for x in (eng, o):
            x.set_source(abi.SRC_SERIES, mask, series, 10 ** 9)
            x.step(0, 90)
        for f in (0, 2, 3, 5):
            a, b = eng.get(abi.ARR_A, f), o.get(abi.ARR_A, f)
            assert np.linalg.norm(b) > 0, f
            if prec == abi.FP64:
                assert rel_err(a, b) < TOL64, f
            else:
                assert rel_l2(a, b) < TOL32_L2, f
        for f in (1, 4):
            assert not np.any(eng.get(abi.ARR_A, f))
        for f in (0, 1, 2, 4):  # psivz, psip, psidzbx, psiinductionbx
            a, b = eng.get(abi.ARR_PMLVARS, f), o.get(abi.ARR_PMLVARS, f)
            assert np.linalg.norm(b) > 0, f
            assert (rel_err(a, b) < TOL64) if prec == abi.FP64 else (rel_l2(a, b) < 1e-4)
    finally:
        eng.close()
        o.close()


@pytest.mark.parametrize("prec", [abi.FP64, abi.FP32])
def test_step_option8_matches_oracle(prec):
    """Option 8 ((rho, v, p, b) linear MHD, pml.f90:80-208): induction, Lorentz
    force, PML top and bottom on six psi fields, source over 9 planes,
    damping after the boundary zeroing."""
    W, S, eng, o = _mhd_case(prec, nsteps=14, option=abi.OPT_MHD8)
    try:
        mask, series = D.forward_source(W, 22)
        for x in (eng, o):
            x.set_source(abi.SRC_SERIES, mask, series, 10 ** 9)
            x.step(0, 14)
        for f in range(8):
            a, b = eng.get(abi.ARR_A, f), o.get(abi.ARR_A, f)
            assert np.linalg.norm(b) > 0, f
            if prec == abi.FP64:
                assert rel_err(a, b) < TOL64, f
            else:
                assert rel_l2(a, b) < TOL32_L2, f
        for f in range(6):
            a, b = eng.get(abi.ARR_PMLVARS, f), o.get(abi.ARR_PMLVARS, f)
            assert np.linalg.norm(b) > 0, f
            assert (rel_err(a, b) < TOL64) if prec == abi.FP64 else (rel_l2(a, b) < 1e-4)
    finally:
        eng.close()
        o.close()


def test_option8_random_state_matches_oracle():
    W, S, eng, o = _mhd_case(abi.FP64, nsteps=3, option=abi.OPT_MHD8)
    try:
        for f in range(8):
            xr = _field(eng.shape, 70 + f) * 1e-4
            eng.put(abi.ARR_A, f, xr)
            o.put(abi.ARR_A, f, xr)
        eng.step(1, 3)
        o.step(1, 3)
        for f in range(8):
            assert rel_err(eng.get(abi.ARR_A, f), o.get(abi.ARR_A, f)) < TOL64, f
    finally:
        eng.close()
        o.close()


def test_mhd_adjoint_and_kernels_match_oracle():
    """Option 10 forward with snapshots, adjoint with the Lagrange source, and
    the kernel sums including the magnetic kernels K_B (kernels.f90:336-430)."""
    nsteps = 18
    fe, fo = _mhd_case(abi.FP64, nsteps=nsteps)[2:]
    W, S, ae, ao = _mhd_case(abi.FP64, nsteps=nsteps, mode=abi.ADJOINT)
    steps = W.params.steps
    nslots = nsteps // steps
    try:
        mask, series = D.forward_source(W, nsteps + 8)
        for x in (fe, fo):
            x.set_source(abi.SRC_SERIES, mask, series, 10 ** 9)
            x.snapshots_alloc(nslots)
        for sl in range(nslots):
            for x in (fe, fo):
                x.step(sl * steps, steps)
                x.snapshot_store(sl)
        rng = np.random.default_rng(19)
        nt = ae.cfg.forcing_length + 8
        vr = np.zeros((W.params.nx, W.params.ny, nt))
        vr[:, :, :ae.cfg.forcing_length] = rng.standard_normal(
            (W.params.nx, W.params.ny, ae.cfg.forcing_length)) * 1e-6
        ao.snapshots_alloc(nslots)
        for sl in range(nslots):
            for f in range(6):
                ao.snapshot_put(sl, f, fo.snapshot_get(sl, f))
        for x in (ae, ao):
            x.set_source(abi.SRC_LAGRANGE, vr, None, 10 ** 9)
            x.kernel_begin(S.g_raw)
        for sl in range(nslots):
            for x in (ae, ao):
                x.step(sl * steps, steps)
            ae.kernel_accumulate_from(fe, nslots - 1 - sl)
            ao.kernel_accumulate(nslots - 1 - sl)
        for k in (abi.KERN_C2, abi.KERN_RHO, abi.KERN_HESSIAN, abi.KERN_BX, abi.KERN_BY,
                  abi.KERN_BZ, abi.KERN_PSI):
            a, b = ae.kernel_get(k), ao.kernel_get(k)
            assert np.linalg.norm(b) > 0, k
            assert rel_err(a, b) < TOL64, k
    finally:
        for x in (fe, fo, ae, ao):
            x.close()


def test_option10_wilson_depression_planes_match_oracle():
    """Option 10 with per-pixel source planes (erad_2d from the Wilson-
    depression search, physics.f90:66-90): a density model whose layers sit
    one to two planes lower under part of the box moves the source there."""
    from sparc_amd import background as Bm
    W, S, eng, o = _mhd_case(abi.FP64, nsteps=10)
    try:
        nx, ny, nz = eng.shape
        e_rad = eng.cfg.e_rad
        rho_q = S.rho0[0, 0, :] if S.rho0.ndim == 3 else S.rho0
        x = np.arange(nx)[:, None, None] / nx
        shift = np.where(np.cos(2 * np.pi * x) > 0.3, 2, 0) + np.zeros((nx, ny, 1), dtype=int)
        k = np.arange(nz)[None, None, :]
        rho_dep = np.take(rho_q, np.clip(k + shift, 0, nz - 1))  # depressed layers
        planes = Bm.wilson_planes(rho_dep, rho_q, e_rad)
        assert set(np.unique(planes)) == {e_rad, e_rad - 2}
        mask, series = D.forward_source(W, 18)
        for e in (eng, o):
            e.set_source(abi.SRC_SERIES, mask, series, 10 ** 9)
            e.set_source_planes(planes)
            e.step(0, 10)
        for f in range(6):
            a, b = eng.get(abi.ARR_A, f), o.get(abi.ARR_A, f)
            assert np.linalg.norm(b) > 0, f
            assert rel_err(a, b) < TOL64, f
    finally:
        eng.close()
        o.close()


@pytest.fixture(scope="module")
def tube_wide_oracle():
    """Option 10 on the 256-wide flux-tube grid, 2 steps of the oracle (run
    once for both precisions)."""
    W = D.workload("tube_wide")
    S = D.setup_for(W)
    o, _ = make_oracle(W, S, nsteps=2)
    mask, series = D.forward_source(W, 10)
    o.set_source(abi.SRC_SERIES, mask, series, 10 ** 9)
    o.step(0, 2)
    want = [o.get(abi.ARR_A, f) for f in range(6)] + [o.get(abi.ARR_SCR, f) for f in range(6)]
    o.close()
    return W, S, (mask, series), want


@pytest.mark.parametrize("prec", [abi.FP64, abi.FP32])
def test_flux_tube_wide_option10_matches_oracle(tube_wide_oracle, prec):
    """Option 10 on 256 x 256 x 40 with the flux tube: the length-256 x and
    y kernels store the curl pieces combined onto the z pieces (b = curl
    of the flux assembled in place)."""
    W, S, (mask, series), want = tube_wide_oracle
    eng, cfg = D.make_engine(W, S, precision=prec, nsteps=2)
    try:
        eng.set_source(abi.SRC_SERIES, mask, series, 10 ** 9)
        eng.step(0, 2)
        got = [eng.get(abi.ARR_A, f) for f in range(6)] + [eng.get(abi.ARR_SCR, f) for f in range(6)]
        for i, (a, b) in enumerate(zip(got, want)):
            assert np.linalg.norm(b) > 0, i
            assert (rel_err(a, b) < TOL64) if prec == abi.FP64 else (rel_l2(a, b) < TOL32_L2), i
    finally:
        eng.close()


@pytest.mark.parametrize("prec", [abi.FP64, abi.FP32])
def test_flux_tube_option10_matches_oracle(prec):
    """Option 10 on the flux-tube background of the magnetic set-up (Alfven
    cap, curl terms, horizontal gradients, fast-speed PML, top sponge,
    Wilson-depression source planes) through load_setup, 20 steps."""
    W = D.workload("tube")
    S = D.setup_for(W)
    eng, cfg = D.make_engine(W, S, precision=prec, nsteps=20)
    o, _ = make_oracle(W, S, nsteps=20)
    try:
        mask, series = D.forward_source(W, 28)
        for x in (eng, o):
            x.set_source(abi.SRC_SERIES, mask, series, 10 ** 9)
            x.step(0, 20)
        for f in range(6):
            a, b = eng.get(abi.ARR_A, f), o.get(abi.ARR_A, f)
            assert np.linalg.norm(b) > 0, f
            assert (rel_err(a, b) < TOL64) if prec == abi.FP64 else (rel_l2(a, b) < TOL32_L2), f
    finally:
        eng.close()
        o.close()


@pytest.mark.parametrize("scale", [1e-25, 1e25], ids=["scale_1e-25", "scale_1e25"])
def test_fp32_state_range_is_not_a_constraint(tiny, scale):
    """FP32 relative precision does not depend on magnitude, so storing the
    state at 1e-25 or 1e25 times its nondimensional size (cfg.field_scale,
    the range lever) leaves the FP32 engine's agreement with the FP64 oracle
    where it is at scale 1: the nondimensional fields of the stratified box
    need no local rescaling to fit FP32."""
    from dataclasses import replace
    W0, S = tiny
    errs = {}
    for sc in (1.0, scale):
        W = D.Workload(W0.name, replace(W0.params, field_scale=sc), W0.tab, W0.nsteps)
        eng, o = _run_pair(W, S, 14, abi.FP32)
        try:
            errs[sc] = max(rel_l2(eng.get(abi.ARR_A, f), o.get(abi.ARR_A, f)) for f in range(6))
        finally:
            eng.close()
            o.close()
    print(f"  FP32 vs oracle, rel L2: scale 1 {errs[1.0]:.2e}, scale {scale:g} {errs[scale]:.2e}")
    assert errs[scale] < TOL32_L2 and errs[scale] < 3 * errs[1.0]
