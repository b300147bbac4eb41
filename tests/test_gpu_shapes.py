"""GPU engine (through the C ABI) against the FP64 oracle at the shapes the
benchmark configurations run, so the kernel instantiations that are timed
are the ones checked:

* 256x256x300 (the single-source forward + adjoint configuration): a forward
  step pair with snapshots, an adjoint step pair driven by the 7-point
  Lagrange source with kernel accumulation after each step.  The z solves run
  k_zsolve_b with its 153.6 KB column tile and k_zsolve with the 76.8 KB ud
  tile here and nowhere smaller.
* 64x64x100 for the full 400-step FP64 forward solve: v_z at the observation
  plane every output cadence and the final state.
* 128x128x150 and 384x384x40: the small grid of the serving mix and the
  384-wide line kernels (radix 3) of its large grid.
* 288-wide x lines and 288-long y lines: the generic (LDS) line kernels, with
  option 2's damping of the boundary planes.
* 64x64x41 and 96x64x40: the fused damping route on the register-free line
  kernels (odd nz, radix 3), with the next stage's damping forward pass done
  by the inverse pass in FP32.

Tolerances: FP64 engine <= 1e-10 max-abs relative; FP32 engine <= 1e-5
relative L2 on the wavefields and <= 1e-4 relative L2 on the kernels (the
stated bars of the FP32 path).
"""
import numpy as np
import pytest

from sparc_amd import abi
from sparc_amd import background as bgm
from sparc_amd import driver as D
from tests.helpers import make_oracle, rel_err, rel_l2

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL32_L2 = 1e-5
TOL32_KERN = 1e-4


def _smooth_state(shape, seed, amp=1e-3):
    """A random state of low-order Fourier modes in x and y and a smooth
    random profile in z, plus a small white part: every plane (boundary and
    PML planes included) carries signal."""
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    x = np.arange(nx)[:, None, None] / nx
    y = np.arange(ny)[None, :, None] / ny
    z = np.linspace(0.0, 1.0, nz)[None, None, :]
    f = np.zeros(shape)
    for _ in range(6):
        kx, ky = rng.integers(0, 6, size=2)
        kz = rng.uniform(0.5, 6.0)
        ph = rng.uniform(0, 2 * np.pi, size=3)
        f += rng.standard_normal() * (np.cos(2 * np.pi * kx * x + ph[0])
                                      * np.cos(2 * np.pi * ky * y + ph[1])
                                      * np.cos(np.pi * kz * z + ph[2]))
    f += 0.05 * rng.standard_normal(shape)
    return np.asfortranarray(amp * f)


def _grid_workload(nx, ny, nz, option=abi.OPT_3D, dx_mm=0.78125, precision=abi.FP64):
    """A polytrope box of the given grid with the horizontal spacing of the
    small parity boxes (the CFL time step stays that of a 25 Mm, 32-point
    box)."""
    tab = bgm.polytrope_table(nz, 0.5, -12.0)
    Lx, Ly = nx * dx_mm * bgm.M_CM, ny * dx_mm * bgm.M_CM
    dt = D.cfl_timestep(tab, dx_mm * bgm.M_CM, dx_mm * bgm.M_CM, 0.5)
    P = D.Params(nx=nx, ny=ny, nz=nz, xlength=Lx, ylength=Ly, timestep=dt, option=option,
                 precision=precision)
    return D.Workload(f"grid{nx}x{ny}x{nz}", P, tab, 10)


def _forward_pair_from_state(W, S, prec, nsteps, seed):
    nf = 5 if W.params.option == abi.OPT_RHOVP else 6
    eng, _ = D.make_engine(W, S, precision=prec, nsteps=nsteps + 2)
    o, _ = make_oracle(W, S, nsteps=nsteps + 2)
    mask, series = D.forward_source(W, nsteps + 10)
    for f in range(nf):
        x = _smooth_state(eng.shape, seed + f)
        eng.put(abi.ARR_A, f, x)
        o.put(abi.ARR_A, f, x)
    for x in (eng, o):
        x.set_source(abi.SRC_SERIES, mask, series, 10 ** 9)
        x.step(1, nsteps)
    return eng, o, nf


def _check(a, b, prec, tol32=TOL32_L2, what=""):
    assert np.linalg.norm(b) > 0, what
    if prec == abi.FP64:
        e = rel_err(a, b)
        print(f"  {what}: max rel err {e:.2e}")
        assert e < TOL64, (what, e)
    else:
        e = rel_l2(a, b)
        print(f"  {what}: rel L2 err {e:.2e}")
        assert e < tol32, (what, e)


# ---------------------------------------------------------------------------
# 256 x 256 x 300: forward, adjoint with the Lagrange source, kernels
# ---------------------------------------------------------------------------

class Box2Case:
    """The oracle's forward + adjoint pair on box2, computed once per module."""

    def __init__(self):
        self.W = D.workload("box2", precision=abi.FP64)
        W, self.S = self.W, D.setup_for(self.W)
        S = self.S
        self.nsteps = 2
        shape = (W.params.nx, W.params.ny, W.params.nz)
        self.fwd_init = {f: _smooth_state(shape, 100 + f) for f in range(6)}
        self.adj_init = {f: _smooth_state(shape, 200 + f) for f in range(3, 6)}
        fo, _ = make_oracle(W, S, nsteps=self.nsteps)
        mask, series = D.forward_source(W, self.nsteps + 10)
        self.mask, self.series = mask, series
        for f, x in self.fwd_init.items():
            fo.put(abi.ARR_A, f, x)
        fo.set_source(abi.SRC_SERIES, mask, series, 10 ** 9)
        fo.snapshots_alloc(self.nsteps)
        for s in range(self.nsteps):
            fo.step(s, 1)
            fo.snapshot_store(s)
        self.fwd_a = {f: fo.get(abi.ARR_A, f) for f in range(6)}
        self.fwd_acc = {f: fo.get(abi.ARR_SCR, f) for f in range(3, 6)}
        self.fwd_pml = {f: fo.get(abi.ARR_PMLVARS, f) for f in range(2)}
        self.snaps = {(s, f): fo.snapshot_get(s, f) for s in range(self.nsteps) for f in range(6)}
        fo.close()
        ao, acfg = make_oracle(W, S, mode=abi.ADJOINT, nsteps=self.nsteps)
        rng = np.random.default_rng(7)
        nt = acfg.forcing_length + 8
        vr = np.zeros((W.params.nx, W.params.ny, nt))
        vr[:, :, :acfg.forcing_length] = rng.standard_normal(
            (W.params.nx, W.params.ny, acfg.forcing_length)) * 1e-4
        self.vr = vr
        ao.snapshots_alloc(self.nsteps)
        for (s, f), x in self.snaps.items():
            ao.snapshot_put(s, f, x)
        for f, x in self.adj_init.items():
            ao.put(abi.ARR_A, f, x)
        ao.set_source(abi.SRC_LAGRANGE, vr, None, 10 ** 9)
        ao.kernel_begin(S.g_raw)
        for s in range(self.nsteps):
            ao.step(s, 1)
            ao.kernel_accumulate(self.nsteps - 1 - s)
        self.adj_a = {f: ao.get(abi.ARR_A, f) for f in range(3, 6)}
        self.kern = {k: ao.kernel_get(k) for k in (abi.KERN_C2, abi.KERN_RHO, abi.KERN_HESSIAN)}
        ao.close()


@pytest.fixture(scope="module")
def box2_case():
    return Box2Case()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("prec", [abi.FP64, abi.FP32], ids=["box2_256x256x300_fp64",
                                                          "box2_256x256x300_fp32"])
def test_box2_forward_adjoint_kernels_match_oracle(box2_case, prec):
    """Two forward steps storing snapshots, two adjoint steps with the
    Lagrange source and a kernel accumulation after each, at 256x256x300."""
    C = box2_case
    W, S = C.W, C.S
    fe, _ = D.make_engine(W, S, precision=prec, nsteps=C.nsteps)
    ae = None
    try:
        for f, x in C.fwd_init.items():
            fe.put(abi.ARR_A, f, x)
        fe.set_source(abi.SRC_SERIES, C.mask, C.series, 10 ** 9)
        fe.snapshots_alloc(C.nsteps)
        for s in range(C.nsteps):
            fe.step(s, 1)
            fe.snapshot_store(s)
        for f in range(6):
            _check(fe.get(abi.ARR_A, f), C.fwd_a[f], prec, what=("a", f))
        # acc = scr(4..6) enters only the kernels (the hessian, kernels.f90:276-277),
        # so the FP32 engine meets the kernel bar on it: a single RHS of this
        # random state carries grid-scale content that a stepped wavefield damps
        for f in range(3, 6):
            _check(fe.get(abi.ARR_SCR, f), C.fwd_acc[f], prec, tol32=TOL32_KERN, what=("acc", f))
        for f in range(2):
            _check(fe.get(abi.ARR_PMLVARS, f), C.fwd_pml[f], prec, tol32=1e-4, what=("psi", f))
        ae, _ = D.make_engine(W, S, mode=abi.ADJOINT, precision=prec, nsteps=C.nsteps)
        ae.snapshots_copy_from(fe)
        for f, x in C.adj_init.items():
            ae.put(abi.ARR_A, f, x)
        ae.set_source(abi.SRC_LAGRANGE, C.vr, None, 10 ** 9)
        ae.kernel_begin(S.g_raw)
        for s in range(C.nsteps):
            ae.step(s, 1)
            ae.kernel_accumulate(C.nsteps - 1 - s)
        for f in range(3, 6):
            _check(ae.get(abi.ARR_A, f), C.adj_a[f], prec, what=("adj a", f))
        for k, want in C.kern.items():
            _check(ae.kernel_get(k), want, prec, tol32=TOL32_KERN, what=("kernel", k))
    finally:
        fe.close()
        if ae is not None:
            ae.close()


# ---------------------------------------------------------------------------
# 64 x 64 x 100, the full 400-step FP64 forward solve
# ---------------------------------------------------------------------------

@pytest.mark.timeout(600)
def test_box1_64x64x100_fp64_400_steps_matches_oracle():
    """The plain FP64 forward solve in full: the Gaussian x Ricker source,
    400 steps, v_z at the observation plane every output cadence and the
    final 3-D state against the oracle."""
    W = D.workload("box1")
    S = D.setup_for(W)
    assert W.nsteps == 400 and W.params.precision == abi.FP64
    eng, cfg = D.make_engine(W, S)
    o, _ = make_oracle(W, S)
    try:
        mask, series = D.forward_source(W, W.nsteps + 8)
        for x in (eng, o):
            x.set_source(abi.SRC_SERIES, mask, series, 10 ** 9)
        steps = W.params.steps
        k = cfg.o_rad
        got, want = [], []
        t = 0
        while t < W.nsteps:
            n = min(steps, W.nsteps - t)
            eng.step(t, n)
            o.step(t, n)
            t += n
            got.append(eng.plane(abi.V_Z, k))
            want.append(o.get(abi.ARR_A, abi.V_Z)[:, :, k - 1])
        got, want = np.array(got), np.array(want)
        assert np.max(np.abs(want)) > 0
        assert rel_err(got, want) < TOL64
        for f in range(6):
            _check(eng.get(abi.ARR_A, f), o.get(abi.ARR_A, f), abi.FP64, what=("final a", f))
    finally:
        eng.close()
        o.close()


# ---------------------------------------------------------------------------
# the serving mix's grids
# ---------------------------------------------------------------------------

@pytest.mark.timeout(300)
@pytest.mark.parametrize("prec", [abi.FP64, abi.FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("grid", ["box2s_128x128x150", "box2l_lines_384x384x40"])
def test_serving_grids_match_oracle(grid, prec):
    if grid.startswith("box2s"):
        W = D.workload("box2s", precision=abi.FP64)
    else:
        W = _grid_workload(384, 384, 40)
    S = D.setup_for(W)
    eng, o, nf = _forward_pair_from_state(W, S, prec, 2, 300)
    try:
        for f in range(nf):
            _check(eng.get(abi.ARR_A, f), o.get(abi.ARR_A, f), prec, what=("a", f))
            _check(eng.get(abi.ARR_SCR, f), o.get(abi.ARR_SCR, f), prec, tol32=TOL32_KERN,
                   what=("scr", f))
    finally:
        eng.close()
        o.close()


# ---------------------------------------------------------------------------
# generic line lengths (LDS line kernels)
# ---------------------------------------------------------------------------

@pytest.mark.timeout(300)
@pytest.mark.parametrize("prec", [abi.FP64, abi.FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("case", ["opt2_288x32x40", "opt4_32x288x40", "opt4_288x32x40"])
def test_generic_line_lengths_match_oracle(case, prec):
    """x lines of 288 (option 2: the damping of planes 1 and nz that follows
    its boundary zeroing; option 4) and y lines of 288 (option 4) run the
    generic LDS line kernels."""
    opt, dims = case.split("_")
    nx, ny, nz = (int(v) for v in dims.split("x"))
    option = {"opt2": abi.OPT_RHOVP, "opt4": abi.OPT_3D}[opt]
    W = _grid_workload(nx, ny, nz, option=option)
    S = D.setup_for(W)
    eng, o, nf = _forward_pair_from_state(W, S, prec, 3, 400)
    try:
        assert W.params.damp_waves
        for f in range(nf):
            _check(eng.get(abi.ARR_A, f), o.get(abi.ARR_A, f), prec, what=("a", f))
    finally:
        eng.close()
        o.close()


# ---------------------------------------------------------------------------
# the fused damping route on line lengths other than 256
# ---------------------------------------------------------------------------

@pytest.mark.timeout(300)
@pytest.mark.parametrize("prec", [abi.FP64, abi.FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("dims", ["64x64x41", "96x64x40"])
def test_fused_damping_route_generic_lengths_match_oracle(dims, prec):
    """Option 4 with damping on fast line lengths other than 256: the damping
    forward pass stores RHS_xi, and in FP32 the segmented column z solve
    (too few columns for the streaming one) stages RHS_vz and the generic
    x-line damping inverse does the next stage's forward pass, RHS_xi
    alternating between two buffers over the five stages; an odd nz (the
    last plane pair has one plane) and 96-point x lines (radix 3)."""
    nx, ny, nz = (int(v) for v in dims.split("x"))
    W = _grid_workload(nx, ny, nz)
    S = D.setup_for(W)
    eng, o, nf = _forward_pair_from_state(W, S, prec, 3, 500)
    try:
        assert W.params.damp_waves
        for f in range(nf):
            _check(eng.get(abi.ARR_A, f), o.get(abi.ARR_A, f), prec, what=("a", f))
            _check(eng.get(abi.ARR_SCR, f), o.get(abi.ARR_SCR, f), prec, tol32=TOL32_KERN,
                   what=("scr", f))
    finally:
        eng.close()
        o.close()


# ---------------------------------------------------------------------------
# the streaming z solve's column splits (256-wide FP32 boxes)
# ---------------------------------------------------------------------------

@pytest.mark.timeout(300)
@pytest.mark.parametrize("nz", [96, 101, 128], ids=["split96", "split_odd101", "split128"])
def test_streaming_zsolve_column_splits_match_oracle(nz):
    """Option 4 on 256 x 256 x nz in FP32: the pressure and dp/dz z solves run
    as two walks per column, one from each end, each owning half of the
    column (equal halves at 96 and 128 planes, 50 + 51 at 101) and carrying
    its truncated backward sweeps 2G + 4 planes into the other half; the PML
    planes at both ends are each handled by one walk."""
    W = _grid_workload(256, 256, nz)
    S = D.setup_for(W)
    eng, o, nf = _forward_pair_from_state(W, S, abi.FP32, 2, 500)
    try:
        for f in range(nf):
            _check(eng.get(abi.ARR_A, f), o.get(abi.ARR_A, f), abi.FP32, what=("a", f))
            _check(eng.get(abi.ARR_SCR, f), o.get(abi.ARR_SCR, f), abi.FP32, tol32=TOL32_KERN,
                   what=("scr", f))
        for f in range(2):
            _check(eng.get(abi.ARR_PMLVARS, f), o.get(abi.ARR_PMLVARS, f), abi.FP32,
                   tol32=TOL32_KERN, what=("pmlvars", f))
    finally:
        eng.close()
        o.close()
