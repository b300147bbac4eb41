"""The reference's wired configuration driven from its own inputs, GPU
engine (through the C ABI) against the FP64 oracle under the same host
chain:

* params.i-format text with the reference run's grid and box (256 x 1 x 300
  over 400 Mm, 2 s steps, bottom PML of 10, no top PML) without the flux
  sheet (its background files are not shipped), so option 17;
* the reference's shipped background table (`polytrope`), Instruction file
  (`Spectral`: spectral, 2D), source pixel (master.pixels: -5 Mm), ridge
  windows (params.0, params.1) and frequency bands (filter.params.0, .1);
* forward recording the observation line every 30 s, ADJOINT_SOURCE_FILT
  against data from a +5 % sound-speed bump, the adjoint run on that
  source, kernel sums in the loop.  45 minutes of solar time (1351 steps)
  keep the oracle's share near a minute.

FP64 engine: traces, travel times, adjoint source and kernels to rounding;
FP32 engine (windows of the FP64 measurement reused, as the line search
does): traces within 5e-5 relative L2 (1.9e-5 measured on this observation
line one plane below the top boundary), travel times within 0.02 s, the
sound-speed kernel, with the adjoint on the oracle's adjoint source, within
1e-4 relative L2 on the model layers below the top 15 planes and 5e-4 over
all of them (the source and receiver layer at the top, where |K| peaks,
carries the FP32 error), and within 1e-3 end to end (the FP32 travel times'
error enters the source).
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from sparc_amd import abi
from sparc_amd import driver as D
from sparc_amd import pipeline as PL
from tests.helpers import rel_err, rel_l2

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
DATA = os.path.join(GOLD, "reference_data")

PARAMS = """
integer nx, ny, nz
parameter (nx = 256, ny = 1, nz = 300)
real*8 xlength, ylength, timestep
parameter (xlength = 400.0 * 10**(8), ylength = xlength, timestep = 2.0)
real*8 cadforcing
parameter( cadforcing = 30.0)
logical, parameter :: DAMP_WAVES = .TRUE.
logical, parameter :: magnetic = .false.
logical, parameter :: FLOWS = .FALSE.
real*8 nupeak, nuwidth
parameter(nupeak=0.0017, nuwidth = 0.0055)
logical, parameter :: DISPL = .truE.
logical USE_PML
parameter(USE_PML = .true.)
integer npmltop, npmlbot
parameter(npmlbot = 10, npmltop = 0)
real*8 solartime
parameter(solartime = 2.0)
real*8 obsheight
parameter (obsheight = 20000000.0)
real *8 outputcad
parameter (outputcad = 30.0)
real*8 excitdep
parameter( excitdep = - 50000000.0)
"""
NSTEPS = 15 * 90 + 1


class RefRun:
    def __init__(self):
        tab = np.loadtxt(os.path.join(GOLD, "polytrope.txt"))
        self.W, self.mode, self.data = D.reference_workload(
            PARAMS, DATA, 1, tab, instruction=os.path.join(DATA, "Spectral"), nsteps=NSTEPS)
        assert self.mode.kind == "forward" and self.W.params.option == abi.OPT_2D
        self.S0 = D.setup_for(self.W)
        self.src = self.W.src_xy_mm
        self.spec64 = PL.SolveSpec(W=self.W, receivers=np.zeros((0, 2), dtype=np.int64),
                                   maxtime=NSTEPS - 1, precision=abi.FP64)
        true = PL.bump_c2(self.S0, self.W, amp=0.05, radius_mm=5.0, depth_mm=3.0, x0_mm=10.0)
        St = D.setup_for(self.W, c2_3d=true)
        self.dat = PL.observe_planes(self.spec64, St, eng_factory=O.Oracle)
        self.ref = PL.filtered_solve(self.spec64, self.S0, self.dat, self.data.ridge,
                                     self.data.bands, self.src, eng_factory=O.Oracle)


@pytest.fixture(scope="module")
def run():
    return RefRun()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("prec", [abi.FP64, abi.FP32], ids=["fp64", "fp32"])
def test_reference_inputs_filtered_chain_matches_oracle(run, prec):
    R = run
    assert R.ref.meas.nmeas > 20  # f and p1 ridges on both sides of the source
    spec = PL.SolveSpec(W=R.W, receivers=np.zeros((0, 2), dtype=np.int64),
                        maxtime=NSTEPS - 1, precision=prec)
    got = PL.filtered_solve(spec, R.S0, R.dat, R.data.ridge, R.data.bands, R.src,
                            windows=None if prec == abi.FP64 else R.ref.meas.windows)
    tau_ref = np.array([m.tau_min for m in R.ref.meas.measurements])
    tau = np.array([m.tau_min for m in got.meas.measurements])
    assert got.meas.windows == R.ref.meas.windows
    print(f"  {tau.size} measurements, |tau| up to {np.abs(tau_ref).max() * 60:.2f} s")
    if prec == abi.FP64:
        assert rel_err(got.planes, R.ref.planes) < 1e-10
        assert np.max(np.abs(tau - tau_ref)) < 1e-8
        assert rel_err(got.meas.adj, R.ref.meas.adj) < 1e-8
        for k in (abi.KERN_C2, abi.KERN_RHO, abi.KERN_HESSIAN):
            assert np.linalg.norm(R.ref.kernel_raw[k]) > 0
            assert rel_err(got.kernel_raw[k], R.ref.kernel_raw[k]) < 1e-8, k
    else:
        e = rel_l2(got.planes, R.ref.planes)
        dtau = np.max(np.abs(tau - tau_ref)) * 60
        P = R.W.params
        inner = slice(P.npmlbot, P.nz - P.npmltop)
        deep = slice(P.npmlbot, P.nz - 15)

        def kerr(res, sl):
            return rel_l2(res.kernel_raw[abi.KERN_C2][:, :, sl], R.ref.kernel_raw[abi.KERN_C2][:, :, sl])
        same = PL.filtered_solve(spec, R.S0, R.dat, R.data.ridge, R.data.bands, R.src,
                                 windows=R.ref.meas.windows, adj=R.ref.meas.adj)
        ek, ek_same, ek_deep = kerr(got, inner), kerr(same, inner), kerr(same, deep)
        print(f"  FP32 traces rel L2 {e:.2e}, tau max diff {dtau:.2e} s; sound-speed kernel "
              f"rel L2 on the oracle's adjoint source {ek_same:.2e} (model layers), "
              f"{ek_deep:.2e} (below the top 15 planes); {ek:.2e} end to end")
        # the observation line is the plane below the top boundary of a box
        # without a top PML (the reference's setting), 1351 steps of 2 s: the
        # FP32 traces measure 1.9e-5 there, against 5.9e-6 for the 3-D
        # configuration's photosphere plane under its top PML
        assert e < 5e-5
        assert dtau < 0.02
        # on a common adjoint source the FP32 kernel meets the 1e-4 bar below
        # the source and receiver layer (5.7e-5 measured); in the top 15
        # planes, where the adjoint source enters and |K| peaks, the per-plane
        # errors are 1e-4 to 4.5e-4 (3.9e-4 over all model layers).  End to
        # end the source itself carries the travel times' FP32 error (up to
        # 2.8e-3 s on |tau| <= 20 s; 3.9e-4 measured)
        assert ek_deep < 1e-4
        assert ek_same < 5e-4
        assert ek < 1e-3
