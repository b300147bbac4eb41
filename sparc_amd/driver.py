"""Host driver of one solve: the kernel-mode sequence of driver.f90
(DETERMINE_STATUS, TIMESTEPPING, FORWARD_SOURCE) over the GPU engine.

``Params`` mirrors the compile-time switches of params.i as runtime values;
``workload()`` builds the synthetic configurations the benchmark and tests
use (adiabatic polytrope boxes with a Gaussian x Ricker source).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np

from . import abi
from . import background as bgm
from .background import M_CM, RSUN_R, sp


@dataclass
class Params:
    """Runtime mirror of params.i."""
    nx: int
    ny: int
    nz: int
    xlength: float                 # cm (params.i:27)
    ylength: float
    timestep: float                # s
    option: int = abi.OPT_3D
    npmlbot: int = 10              # params.i:120
    npmltop: int = 10
    damp_waves: bool = True        # params.i:52
    outputcad: float = 30.0        # params.i:140
    cadforcing: float = 30.0       # params.i:36
    excitdep: float = -1.5e7       # cm below the photosphere
    obsheight: float = 0.0         # cm above the photosphere
    precision: int = abi.FP32
    field_scale: float = 1.0
    compensated: bool = True       # FP32: hi + 16-bit lo state updated in FP64 arithmetic
    flow_kernels: bool = False     # snapshots keep v; kernel sums add K_v (kernels.f90:317-332)
    use_graphs: bool = True        # replay each step's stages + update as a captured HIP graph
    plane_damping: bool = False    # option 4, FP32, 256 x 256 planes: one-pass damping (k_damp2d)

    @property
    def is2d(self) -> bool:
        return self.option in (abi.OPT_2D, abi.OPT_MHD2D)

    @property
    def steps(self) -> int:
        """Output cadence in steps, floor(outputcad/timestep) (driver.f90:269);
        a 1e-9 guard keeps exact multiples exact."""
        return int(math.floor(self.outputcad / self.timestep + 1e-9))

    @property
    def cadforcing_step(self) -> int:
        return int(math.floor(self.cadforcing / self.timestep + 1e-9))

    @property
    def freq_filtering(self) -> int:
        return int(math.floor(60.0 / self.timestep + 1e-9))  # driver.f90:270


# Alfven-speed cap of the flux-tube cases: reduction cap2/(cap2 + v_A^2/c^2)
# bounds v_A by sqrt(cap2) c; 16 gives the 4 c of the 256^2 x 300 MHD case
# (the reference's own constant is 8, physics.f90:121)
TUBE_CAP2 = 16.0


def cfl_timestep(tab: np.ndarray, dx_cm: float, dy_cm: float, cfl: float,
                 cadence: float = 30.0, speed_factor: float = 1.0) -> float:
    """Largest timestep <= cfl * min(dz/c, dx/c_max, dy/c_max) that divides
    the output cadence an integer number of times; speed_factor scales c
    (the fast speed of a capped magnetic field)."""
    z_cm = (tab[:, 0] - 1.0) * RSUN_R
    c = tab[:, 1] * speed_factor
    dz = np.diff(z_cm)
    cmid = np.maximum(c[1:], c[:-1])
    lim = min(float(np.min(dz / cmid)), dx_cm / float(c.max()), dy_cm / float(c.max()))
    dt = cfl * lim
    m = math.ceil(cadence / dt - 1e-12)
    return cadence / m


@dataclass
class Workload:
    name: str
    params: Params
    tab: np.ndarray
    nsteps: int
    src_xy_mm: tuple = (0.0, 0.0)
    src_sigma_mm: float = 0.4
    ricker_nu: float = 3.0e-3
    extra: dict = field(default_factory=dict)


def workload(name: str, precision: int | None = None) -> Workload:
    """The synthetic cases of the benchmark configurations.

    box1   64 x 64 x 100, 50 Mm, z in [-12, +0.5] Mm, PML 10+10, CFL 0.5, 400 steps
    box2   256 x 256 x 300, 200 Mm, same polytrope, 2000 steps
    box2s  128 x 128 x 150, 100 Mm (the small grid of the serving mix)
    box2l  384 x 384 x 300, 300 Mm (the large grid of the serving mix)
    tiny   32 x 32 x 40, 25 Mm (parity tests)
    tall   32 x 32 x 200, 25 Mm (parity tests of the segmented z solve)
    wide   256 x 256 x 40, 200 Mm (parity tests of the length-256 line kernels)
    wide_odd  256 x 256 x 41: an odd plane count (the damping's plane pairs end
           with a single plane; z segments of unequal length)
    box3   box2 with option 10 and a flux tube (setup_for builds it): radius
           4 Mm, 3 kG at the photosphere, v_A capped at 4 c
    tube   32 x 32 x 100, 25 Mm, the flux tube of box3 with a 2 Mm radius
           (parity tests of the magnetic set-up)
    tube_wide  256 x 256 x 40, 200 Mm, the flux tube of box3 (parity tests of
           option 10 on the length-256 line kernels)
    column 32 x 32 x 100, 25 Mm, the vertical grid of box1 (adjoint/gradient tests:
           receivers and source lie below the top PML and filter layers)
    two_d  256 x 1 x 300, 400 Mm, option 17 (the reference's wired 2-D case)
    """
    if name == "box1":
        nx = ny = 64
        nz, L, nsteps, prec = 100, 50.0, 400, abi.FP64
    elif name == "box2":
        nx = ny = 256
        nz, L, nsteps, prec = 300, 200.0, 2000, abi.FP32
    elif name == "box2s":
        nx = ny = 128
        nz, L, nsteps, prec = 150, 100.0, 1000, abi.FP32
    elif name == "box2l":
        nx = ny = 384
        nz, L, nsteps, prec = 300, 300.0, 1000, abi.FP32
    elif name == "tiny":
        nx = ny = 32
        nz, L, nsteps, prec = 40, 25.0, 20, abi.FP64
    elif name == "tall":
        nx = ny = 32
        nz, L, nsteps, prec = 200, 25.0, 10, abi.FP64
    elif name in ("wide", "wide_odd"):
        nx = ny = 256
        nz, L, nsteps, prec = (40 if name == "wide" else 41), 200.0, 10, abi.FP64
    elif name in ("box3", "tube", "tube_wide"):
        big = name == "box3"
        nx = ny = 32 if name == "tube" else 256
        nz, L, nsteps = {"box3": (300, 200.0, 2000), "tube": (100, 25.0, 400),
                         "tube_wide": (40, 200.0, 20)}[name]
        tab = bgm.polytrope_table(nz, 0.5, -12.0)
        dx = L * M_CM / nx
        # the time step is set by the fast speed of the capped field
        dt = cfl_timestep(tab, dx, dx, 0.5, speed_factor=math.sqrt(1.0 + TUBE_CAP2))
        P = Params(nx=nx, ny=ny, nz=nz, xlength=L * M_CM, ylength=L * M_CM, timestep=dt,
                   option=abi.OPT_MHD, npmltop=0,
                   precision=(abi.FP32 if big else abi.FP64) if precision is None else precision)
        return Workload(name, P, tab, nsteps,
                        extra={"tube": dict(b_ph=3000.0, r_ph_mm=2.0 if name == "tube" else 4.0,
                                            decay_mm=10.0), "cap2": TUBE_CAP2})
    elif name == "column":
        nx = ny = 32
        nz, L, nsteps, prec = 100, 25.0, 400, abi.FP64
    elif name == "two_d":
        tab = bgm.polytrope_table(300, 0.5, -12.0)
        P = Params(nx=256, ny=1, nz=300, xlength=400.0 * M_CM, ylength=400.0 * M_CM,
                   timestep=cfl_timestep(tab, 400.0 * M_CM / 256, 1e30, 0.5),
                   option=abi.OPT_2D, npmlbot=10, npmltop=10, damp_waves=False,
                   precision=abi.FP64 if precision is None else precision)
        return Workload("two_d", P, tab, 400)
    else:
        raise ValueError(f"unknown workload {name}")
    tab = bgm.polytrope_table(nz, 0.5, -12.0)
    dx = L * M_CM / nx
    dt = cfl_timestep(tab, dx, dx, 0.5)
    P = Params(nx=nx, ny=ny, nz=nz, xlength=L * M_CM, ylength=L * M_CM, timestep=dt,
               precision=prec if precision is None else precision)
    return Workload(name, P, tab, nsteps)


def make_config(P: Params, S: bgm.Setup, mode: int, nsteps_total: int) -> abi.SparcConfig:
    c = abi.SparcConfig()
    c.abi_version = abi.ABI_VERSION
    c.nx, c.ny, c.nz = P.nx, P.ny, P.nz
    c.option = P.option
    c.precision = P.precision
    c.npmlbot, c.npmltop = P.npmlbot, P.npmltop
    # option 17 never damps (physics2d.f90:1071); option 18 does (:900)
    c.damp_waves = int(P.damp_waves and P.option != abi.OPT_2D)
    c.freq_filtering = P.freq_filtering
    if P.is2d:  # physics2d.f90:42,54 override the nearest-point search
        e_rad, o_rad = P.nz - 5, P.nz - 1
    else:
        e_rad = bgm.nearest_index(S.z, 1.0 + P.excitdep / RSUN_R)
        o_rad = bgm.nearest_index(S.z, 1.0 + P.obsheight / RSUN_R)
    c.e_rad, c.o_rad = e_rad, o_rad
    c.mode = mode
    c.cadforcing_step = P.cadforcing_step
    # initialize.f90:978-981 with maxtime = the solve length
    maxtime = P.cadforcing_step * (nsteps_total // P.cadforcing_step)
    c.forcing_length = maxtime // P.cadforcing_step + 1
    # driver.f90:280-281
    ns = int(math.floor(maxtime * P.timestep / P.cadforcing)) + 2
    c.num_steps = 2 * (ns // 2) + 2
    c.xlength, c.ylength = P.xlength, P.ylength
    c.timestep = P.timestep
    c.deltat = S.deltat
    c.stretchx, c.stretchy = S.stretchx, S.stretchy
    k = (o_rad if mode == abi.ADJOINT else e_rad) - 1
    c.delta_width = 2.0 / (S.z[k + 1] - S.z[k - 1])  # DETERMINE_STATUS (driver.f90:644,690)
    c.cadforcing = P.cadforcing
    c.field_scale = P.field_scale if P.precision == abi.FP32 else 1.0
    c.compensated = int(P.compensated and P.precision == abi.FP32)
    c.flow_kernels = int(P.flow_kernels)
    c.use_graphs = int(P.use_graphs)
    c.plane_damping = int(P.plane_damping)
    return c


def load_setup(eng, S: bgm.Setup) -> None:
    """Hands the set-up arrays to an engine (or an oracle with the same calls)."""
    for bid, arr in bgm.background_arrays(S).items():
        eng.set_background(bid, arr)
    d = S.derived
    eng.set_profile(abi.PROF_STRETCH, d["stretch"])
    eng.set_profile(abi.PROF_UNSTRETCH, d["unstretch"])
    eng.set_profile(abi.PROF_G, d["g"])
    eng.set_profile(abi.PROF_Z, S.z)
    if eng.cfg.option in (abi.OPT_RHOVP, abi.OPT_MHD8):
        eng.set_profile(abi.PROF_SOURCE_DEP, bgm.source_dep(S.z, eng.cfg.e_rad))
    if S.mhd is not None:
        # MHD backgrounds, the top sponge and the Wilson-depression source
        # planes (physics.f90:66-107) from a magnetic set-up
        for bid, arr in S.mhd.items():
            eng.set_background(bid, arr)
        if eng.cfg.option == abi.OPT_MHD:
            eng.set_profile(abi.PROF_SPONGE, bgm.sponge_z(S.z))
        if S.rho_q is not None and eng.cfg.option in (abi.OPT_MHD, abi.OPT_MHD8):
            k_ref = eng.cfg.e_rad if eng.cfg.mode == abi.FORWARD else eng.cfg.o_rad
            eng.set_source_planes(bgm.wilson_planes(S.rho0, S.rho_q, k_ref))
    eng.set_damping(S.rates)
    eng.set_filter(S.decay_x, S.decay_y)


def setup_for(W: Workload, c2_3d=None, rho0_3d=None) -> bgm.Setup:
    P = W.params
    if "tube" in W.extra:
        fields = bgm.flux_tube_vector_potential(W.tab, P.nx, P.ny, P.xlength, P.ylength,
                                                **W.extra["tube"])
        if c2_3d is not None:
            fields["c"] = np.sqrt(c2_3d) * bgm.DIMC
        return bgm.magnetic_setup(W.tab, fields, P.nx, P.ny, P.option, P.xlength, P.ylength,
                                  P.timestep, P.npmlbot, P.npmltop, cap2=W.extra["cap2"])
    return bgm.build_setup(W.tab, P.nx, P.ny, P.option, P.xlength, P.ylength, P.timestep,
                           P.npmlbot, P.npmltop, c2_3d=c2_3d, rho0_3d=rho0_3d)


def forward_source(W: Workload, nt: int):
    """mask(x,y) and time series of the forward source.  3-D: Gaussian
    (sigma 0.4 Mm) x Ricker (3 mHz); 2-D: the reference's sinc mask at the
    master pixel x = -5 Mm (data/master.pixels) x FORWARD_SOURCE."""
    P = W.params
    if P.is2d:  # the source pixel of master.pixels (data/master.pixels holds -5 Mm)
        mask = bgm.sinc_mask_2d(P.nx, P.xlength, W.extra.get("master_pixel_mm", -5.0))
        series = bgm.fwdsource_reference(nt, P.timestep)
    else:
        mask = bgm.gaussian_mask(P.nx, P.ny, P.xlength, P.ylength, W.src_xy_mm[0],
                                 W.src_xy_mm[1], W.src_sigma_mm)
        series = bgm.ricker_source(nt, P.timestep, nu=W.ricker_nu)
    return mask, series


def write_bundle(path: str, cfg: abi.SparcConfig, S: bgm.Setup, kind: int, mask: np.ndarray,
                 series: np.ndarray | None, active_until: int, maxtime: int, steps: int) -> None:
    """Writes the set-up bundle the Fortran host (sparc_amd/fortran/
    timestepping.f90) reads: the configuration, every array load_setup hands
    to an engine, the source, and the loop bounds."""
    import ctypes as C
    import struct

    def arr(a):
        return np.asfortranarray(a, dtype="<f8").ravel(order="F").tobytes()

    with open(path, "wb") as f:
        f.write(b"SPARCBN1")
        f.write(C.string_at(C.addressof(cfg), C.sizeof(cfg)))
        f.write(struct.pack("<ii", maxtime, steps))
        bgs = bgm.background_arrays(S)
        f.write(struct.pack("<i", len(bgs)))
        for bid, a in bgs.items():
            a = np.asarray(a)
            f.write(struct.pack("<iq", bid, a.size))
            f.write(arr(a))
        d = S.derived
        profs = [(abi.PROF_STRETCH, d["stretch"]), (abi.PROF_UNSTRETCH, d["unstretch"]),
                 (abi.PROF_G, d["g"]), (abi.PROF_Z, S.z)]
        f.write(struct.pack("<i", len(profs)))
        for pid, a in profs:
            f.write(struct.pack("<ii", pid, len(a)))
            f.write(arr(a))
        rates = np.asarray(S.rates) if S.rates is not None else np.zeros(0)
        f.write(struct.pack("<q", rates.size))
        f.write(arr(rates))
        for dtab in (S.decay_x, S.decay_y):
            f.write(struct.pack("<q", len(dtab)))
            f.write(arr(dtab))
        f.write(struct.pack("<iq", kind, np.asarray(mask).size))
        f.write(arr(mask))
        ser = np.zeros(0) if series is None else np.asarray(series)
        f.write(struct.pack("<i", ser.size))
        f.write(arr(ser))
        f.write(struct.pack("<i", active_until))


def make_engine(W: Workload, S: bgm.Setup, mode: int = abi.FORWARD, nsteps: int | None = None,
                precision: int | None = None, compensated: bool | None = None,
                use_graphs: bool | None = None):
    P = W.params if precision is None else replace(W.params, precision=precision)
    if compensated is not None:
        P = replace(P, compensated=compensated)
    if use_graphs is not None:
        P = replace(P, use_graphs=use_graphs)
    cfg = make_config(P, S, mode, nsteps if nsteps is not None else W.nsteps)
    eng = abi.Engine(cfg)
    load_setup(eng, S)
    return eng, cfg


def reference_workload(params_text: str, data_dir: str, src: int, tab: np.ndarray,
                       instruction: str | None = None, nsteps: int | None = None,
                       precision: int | None = None):
    """The reference's start-up from its own inputs (READ_PARAMETERS,
    initialize.f90:964-1087; DETERMINE_STATUS, driver.f90:552-702):
    params.i (as text) gives the grid, box, time step, PML, cadences and the
    option (its switches and the Instruction file's 2D line); the Instruction
    file gives the run mode; master.pixels in data_dir the source x of
    source `src` (1-based, argv 1); params.N / filter.params.N in data_dir
    the measurement windows and bands; `tab` is the background table named
    by file_data.  Returns (workload, run mode, data directory).  nsteps
    defaults to the reference's maxtime + 1 (solartime + 2/nupeak)."""
    import os

    from . import inputs as I
    mode = I.read_instruction(instruction) if instruction else I.RunMode("forward", False)
    v = I.parse_params_i(params_text)
    P = I.params_from_reference(params_text, test_in_2d=mode.test_in_2d,
                                precision=abi.FP64 if precision is None else precision)
    if tab.shape[0] != P.nz:
        raise ValueError(f"background table has {tab.shape[0]} rows, params.i nz = {P.nz}")
    if nsteps is None:
        sc = I.step_counts(float(v.get("solartime", 2.0)), float(v.get("nupeak", sp(0.0017))),
                           P.timestep, P.cadforcing)
        nsteps = sc.maxtime + 1
    data = I.read_data_dir(data_dir)
    x0 = I.master_pixel(os.path.join(data_dir, "master.pixels"), src)
    W = Workload(f"reference_src{src:02d}", P, tab, nsteps,
                 src_xy_mm=(x0, 0.0), extra={"master_pixel_mm": x0})
    return W, mode, data
