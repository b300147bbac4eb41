"""One source's forward + adjoint + kernel solve on the engine, following the
kernel-mode sequence of the reference:

  forward  TIMESTEPPING with the forward source (driver.f90:326 -> 398-547):
           maxtime+1 steps; after step i with mod(i, steps) == 0 the v_z
           samples at the receivers are recorded (vz_cc.bin, driver.f90:460-476)
           and the partial state is kept (WRITE_OUT_PARTIAL_STATE: xi, acc) --
           here in device memory instead of FITS files;
  measure  travel times and the adjoint source (ADJOINT_SOURCE_COMP,
           driver.f90:729-917) on the host in FP64;
  adjoint  TIMESTEPPING with the time-reversed adjoint source interpolated by
           LAGRANGE_INTERP at the receivers (driver.f90:350, physics.f90:678);
           the kernel sums of PRODUCE_KERNELS (kernels.f90:197-313) are
           accumulated in the loop, pairing forward stamp (k-1)*steps with
           adjoint stamp maxtime-(k-1)*steps (kernels.f90:203,241).

Scaling (nondimensional time t_n = t_s dimc/diml, volume in Rsun^3):
  misfit chi = 0.5 sum tau^2 [s^2];  a_i(t) = d chi / d u_i(t);
  the adjoint is driven at the receivers by a_i(T - t) (diml/dimc) delta^3 / rho0;
  K_lnc(x) = -2 rho0 c^2 sum_k div(xi_f) div(v_a) (steps * deltat)
  so that d chi = sum_cells K_lnc dlnc dV  (kernels.f90:537 with dimen = 1).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import abi
from . import background as bgm
from . import driver as D
from .measure import RingMeasurement


@dataclass
class SolveSpec:
    W: D.Workload
    receivers: np.ndarray  # (nrec, 2) 0-based (i, j)
    maxtime: int           # last time index (maxtime + 1 steps per solve)
    precision: int = abi.FP32
    window_width: float = 120.0  # s, Gaussian time window

    @property
    def steps(self) -> int:
        return self.W.params.steps

    @property
    def nt(self) -> int:
        return self.maxtime // self.steps + 1

    @property
    def cad(self) -> float:
        return self.steps * self.W.params.timestep


def _ijk(spec: SolveSpec, k: int) -> np.ndarray:
    r = spec.receivers
    return np.stack([r[:, 0], r[:, 1], np.full(len(r), k)], axis=1).astype(np.int32)


def sample(eng, field: int, ijk: np.ndarray) -> np.ndarray:
    if hasattr(eng, "sample"):
        return eng.sample(field, ijk)
    a = eng.get(abi.ARR_A, field)
    return a[ijk[:, 0], ijk[:, 1], ijk[:, 2] - 1]


def _run_to(eng, t: int, i_end: int) -> int:
    """Steps time indices t..i_end; returns the next time index."""
    if i_end >= t:
        eng.step(t, i_end - t + 1)
    return i_end + 1


def forward_solve(eng, spec: SolveSpec, store: bool = True, kstride: int | None = None) -> np.ndarray:
    """Runs the forward loop on an engine whose source is set; returns the
    receiver traces (nrec, nt) sampled after steps m*steps.  With store, the
    partial state after step j*kstride is kept in slot j (kstride defaults to
    the output cadence, the reference's snapshot stamps)."""
    steps = spec.steps
    ks = kstride or steps
    ijk = _ijk(spec, eng.cfg.o_rad)
    nslots = spec.maxtime // ks + 1
    if store:
        eng.snapshots_alloc(nslots)
    traces = np.zeros((len(spec.receivers), spec.nt))
    events = sorted(set(range(0, spec.maxtime + 1, steps)) | (set(range(0, spec.maxtime + 1, ks))
                                                            if store else set()))
    t = 0
    for i in events:
        t = _run_to(eng, t, i)
        if i % steps == 0:
            traces[:, i // steps] = sample(eng, abi.V_Z, ijk)
        if store and i % ks == 0:
            eng.snapshot_store(i // ks)
    return traces


def forward_planes(eng, spec: SolveSpec, store: bool = True) -> np.ndarray:
    """The forward loop recording the whole observation plane: v_z(:,:,o_rad)
    after every step with mod(time, steps) == 0, the vz_cc record of
    TIMESTEPPING (driver.f90:455-476), as (nx, ny, nt); with store, the
    partial state is kept at the same cadence."""
    P = spec.W.params
    steps = spec.steps
    if store:
        eng.snapshots_alloc(spec.maxtime // steps + 1)
    out = np.zeros((P.nx, P.ny, spec.nt))
    t = 0
    for i in range(0, spec.maxtime + 1, steps):
        t = _run_to(eng, t, i)
        k = i // steps
        if hasattr(eng, "plane"):
            out[:, :, k] = eng.plane(abi.V_Z, eng.cfg.o_rad)
        else:
            out[:, :, k] = eng.get(abi.ARR_A, abi.V_Z)[:, :, eng.cfg.o_rad - 1]
        if store:
            eng.snapshot_store(k)
    return out


def adjoint_vr(spec: SolveSpec, adj: np.ndarray, S: bgm.Setup) -> np.ndarray:
    """vr(nx, ny, nt+8) of the adjoint run: the time-reversed adjoint source
    (adj(:,1,nt-j+1) = ..., driver.f90:876) at the receiver pixels, as a
    force density (delta^3 through 1/(dx dy) here and delta_width in the
    engine) in nondimensional time units."""
    P = spec.W.params
    nt = spec.nt
    vr = np.zeros((P.nx, P.ny, nt + 8))
    dxn = 1.0 / (P.nx * S.stretchx)
    dyn = 1.0 / (P.ny * S.stretchy)
    tconv = bgm.DIML / bgm.DIMC
    for r, (i, j) in enumerate(spec.receivers):
        vr[i, j, :nt] += adj[r, ::-1] * tconv / (dxn * dyn)
    return vr


def adjoint_solve(eng, spec: SolveSpec, vr: np.ndarray, grav: np.ndarray,
                  kstride: int | None = None, aligned: bool = False, src=None) -> None:
    """Adjoint loop with the Lagrange source, accumulating the kernel sums
    against the forward snapshots held by ``eng``.

    Reference pairing (aligned=False): forward stamp i_f with adjoint stamp
    maxtime - i_f (kernels.f90:203,241).  The time-reversed source maps
    adjoint time t to forward time T' - t with T' = (maxtime + 1) dt
    (samples stamped after their step, driver.f90:455); the states after
    steps i_f and i_a are paired exactly when i_a = maxtime - 1 - i_f
    (aligned=True), one step earlier than the reference's pairing."""
    ks = kstride or spec.steps
    eng.set_source(abi.SRC_LAGRANGE, vr, None, 10 ** 9)
    eng.kernel_begin(grav)
    nslots = spec.maxtime // ks + 1
    pairs = []
    for j in range(nslots):
        ia = spec.maxtime - j * ks - (1 if aligned else 0)
        if ia >= 0:
            pairs.append((ia, j))
    pairs.sort()
    t = 0
    for ia, j in pairs:
        t = _run_to(eng, t, ia)
        if src is None:
            eng.kernel_accumulate(j)
        else:
            eng.kernel_accumulate_from(src, j)


def cell_volume(S: bgm.Setup) -> np.ndarray:
    """Nondimensional cell volume dx dy dz(k) (dz = unstretch/(nz-1))."""
    dxn = 1.0 / (S.nx * S.stretchx)
    dyn = 1.0 / (S.ny * S.stretchy)
    dz = S.derived["unstretch"] / (S.nz - 1)
    return dxn * dyn * dz


def kernel_lnc(raw_c2: np.ndarray, S: bgm.Setup, spec: SolveSpec,
               kstride: int | None = None) -> np.ndarray:
    """Sound-speed kernel K_lnc (per unit nondimensional volume) from the
    accumulated sum (kernels.f90:537: -2 rho0 c^2 sum * stepskern*timestep)."""
    dtn = (kstride or spec.steps) * S.deltat
    return -2.0 * S.rho0 * S.c2 * raw_c2 * dtn


def kernel_lnrho(raw_rho: np.ndarray, S: bgm.Setup, spec: SolveSpec,
                 kstride: int | None = None) -> np.ndarray:
    """Density kernel K_lnrho at fixed sound speed (per unit nondimensional
    volume) from the accumulated sum (kernels.f90:530: -rho0 sum *
    stepskern*timestep), so that d chi = sum_cells K_lnrho dlnrho dV for the
    hydrostatic perturbation of background.perturb_density."""
    dtn = (kstride or spec.steps) * S.deltat
    return -S.rho0 * raw_rho * dtn


def kernel_flow(raw_v: np.ndarray, S: bgm.Setup, spec: SolveSpec,
                kstride: int | None = None) -> np.ndarray:
    """Flow kernel component (per unit nondimensional volume and velocity)
    from the accumulated sum (kernels.f90:485: -2 rho0 sum * stepskern*timestep)."""
    dtn = (kstride or spec.steps) * S.deltat
    return -2.0 * S.rho0 * raw_v * dtn


def kernel_magnetic(raw_b: np.ndarray, S: bgm.Setup, spec: SolveSpec,
                    kstride: int | None = None) -> np.ndarray:
    """Magnetic kernel component from the accumulated sum (kernels.f90:504:
    sum * stepskern*timestep, nondimensional units)."""
    return raw_b * (kstride or spec.steps) * S.deltat


def make_engines(spec: SolveSpec, S: bgm.Setup, eng_factory=None):
    """Forward and adjoint engines (or oracles via eng_factory(cfg, S))."""
    W = spec.W
    P = W.params
    nsteps = spec.maxtime + 1
    fcfg = D.make_config(P, S, abi.FORWARD, nsteps)
    acfg = D.make_config(P, S, abi.ADJOINT, nsteps)
    for c in (fcfg, acfg):
        c.precision = spec.precision
        if spec.precision == abi.FP64:
            c.field_scale = 1.0
            c.compensated = 0
        else:
            c.compensated = int(P.compensated)
    if eng_factory is None:
        fe, ae = abi.Engine(fcfg), abi.Engine(acfg)
    else:
        fe, ae = eng_factory(fcfg), eng_factory(acfg)
    for e in (fe, ae):
        D.load_setup(e, S)
    mask, series = D.forward_source(W, nsteps + 8)
    fe.set_source(abi.SRC_SERIES, mask, series, 10 ** 9)
    return fe, ae


def copy_snapshots(dst, src, nslots: int) -> None:
    if hasattr(dst, "snapshots_copy_from"):
        dst.snapshots_copy_from(src)
        return
    dst.snapshots_alloc(nslots)
    nf = 9 if getattr(src.cfg, "flow_kernels", 0) else 6
    for s in range(nslots):
        for f in range(nf):
            dst.snapshot_put(s, f, src.snapshot_get(s, f))


def kernel_ids(cfg) -> list:
    """Kernel sums a solve with this configuration produces: c2, rho and the
    hessian always, the flow kernels when switched on, the magnetic kernels
    and the stream-function kernel for the MHD displacement options
    (kernels.f90:465-517)."""
    ids = [abi.KERN_C2, abi.KERN_RHO, abi.KERN_HESSIAN]
    if getattr(cfg, "flow_kernels", 0):
        ids += [abi.KERN_VX, abi.KERN_VY, abi.KERN_VZ]
    if cfg.option in (abi.OPT_MHD, abi.OPT_MHD2D):
        ids += [abi.KERN_BX, abi.KERN_BY, abi.KERN_BZ, abi.KERN_PSI]
    return ids


@dataclass
class SolveResult:
    traces: np.ndarray
    tau: np.ndarray
    misfit: float
    adj: np.ndarray
    kernel_raw: dict
    K_lnc: np.ndarray


def full_solve(spec: SolveSpec, S: bgm.Setup, obs: np.ndarray, meas: RingMeasurement | None = None,
               eng_factory=None, kstride: int | None = None, aligned: bool = False,
               exact_adjoint: bool = False) -> tuple[SolveResult, RingMeasurement]:
    """forward -> travel times vs obs -> adjoint source -> adjoint + kernels.
    Defaults follow the reference (kernel sums at the output cadence, the
    reference's snapshot pairing); kstride/aligned refine the quadrature."""
    fe, ae = make_engines(spec, S, eng_factory)
    try:
        syn = forward_solve(fe, spec, store=True, kstride=kstride)
        if meas is None:
            meas = RingMeasurement(syn, spec.cad, spec.window_width)
        tau = meas.travel_times(obs, syn)
        adj = meas.adjoint_source(syn, tau, obs=obs, exact=exact_adjoint)
        vr = adjoint_vr(spec, adj, S)
        if hasattr(ae, "kernel_accumulate_from"):  # read the forward bank in place
            adjoint_solve(ae, spec, vr, S.g_raw, kstride=kstride, aligned=aligned, src=fe)
        else:
            copy_snapshots(ae, fe, spec.maxtime // (kstride or spec.steps) + 1)
            adjoint_solve(ae, spec, vr, S.g_raw, kstride=kstride, aligned=aligned)
        raw = {k: ae.kernel_get(k) for k in kernel_ids(ae.cfg)}
        K = kernel_lnc(raw[abi.KERN_C2], S, spec, kstride=kstride)
        res = SolveResult(traces=syn, tau=tau, misfit=meas.misfit(tau), adj=adj, kernel_raw=raw,
                          K_lnc=K)
        return res, meas
    finally:
        fe.close()
        ae.close()


@dataclass
class FilteredResult:
    planes: np.ndarray      # (nx, ny, nt) synthetic v_z at the observation plane
    meas: object            # measure_filt.FiltResult: tau per pixel, windows, adjoint source
    kernel_raw: dict


def filtered_solve(spec: SolveSpec, S: bgm.Setup, dat: np.ndarray, ridge: dict, bands: dict,
                   src_mm: tuple, eng_factory=None, windows: dict | None = None,
                   receivers: np.ndarray | None = None, t1_min: float | None = None,
                   adj: np.ndarray | None = None) -> FilteredResult:
    """The reference's kernel-mode chain with the filtered measurement:
    forward recording the observation plane, ADJOINT_SOURCE_FILT against the
    data planes `dat` (ridge windows params.N, frequency bands
    filter.params.N), and the adjoint run driven by that source as the
    reference reads it back (vr(:,:,1:forcing_length) = source.fits,
    driver.f90:653-656), with the kernel sums accumulated in the loop.
    `adj` (nx, ny, nt) replaces the measured source in the adjoint run (the
    measurement is still made and returned)."""
    from . import measure_filt as MF
    P = spec.W.params
    fe, ae = make_engines(spec, S, eng_factory)
    try:
        planes = forward_planes(fe, spec, store=True)
        res = MF.adjoint_source_filt(planes, dat, spec.cad, P.xlength, P.ylength, src_mm,
                                     ridge, bands, t1_min=t1_min, receivers=receivers,
                                     windows=windows)
        vr = np.zeros((P.nx, P.ny, spec.nt + 8))
        vr[:, :, :spec.nt] = res.adj if adj is None else adj
        if hasattr(ae, "kernel_accumulate_from"):
            adjoint_solve(ae, spec, vr, S.g_raw, src=fe)
        else:
            copy_snapshots(ae, fe, spec.maxtime // spec.steps + 1)
            adjoint_solve(ae, spec, vr, S.g_raw)
        raw = {k: ae.kernel_get(k) for k in kernel_ids(ae.cfg)}
        return FilteredResult(planes=planes, meas=res, kernel_raw=raw)
    finally:
        fe.close()
        ae.close()


def observe_planes(spec: SolveSpec, S: bgm.Setup, eng_factory=None) -> np.ndarray:
    """Observation-plane v_z of a (true) model, every output cadence."""
    fe, ae = make_engines(spec, S, eng_factory)
    try:
        return forward_planes(fe, spec, store=False)
    finally:
        fe.close()
        ae.close()


def observe(spec: SolveSpec, S: bgm.Setup, eng_factory=None) -> np.ndarray:
    """Forward traces of a (true) model: the synthetic 'observed' data."""
    fe, ae = make_engines(spec, S, eng_factory)
    try:
        return forward_solve(fe, spec, store=False)
    finally:
        fe.close()
        ae.close()


def bump_c2(S0: bgm.Setup, W: D.Workload, amp: float = 0.02, radius_mm: float = 5.0,
            depth_mm: float = 3.0, x0_mm: float = 0.0, y0_mm: float = 0.0) -> np.ndarray:
    """c^2 of a +amp Gaussian sound-speed bump: c -> c (1 + amp exp(-r^2/R^2))
    centred at (x0, y0, -depth) Mm."""
    P = W.params
    x = ((np.arange(P.nx) / P.nx) - 0.5) * P.xlength / 1e8
    y = ((np.arange(P.ny) / P.ny) - 0.5) * P.ylength / 1e8 if P.ny > 1 else np.zeros(1)
    zmm = (S0.z - 1.0) * bgm.RSUN_R / 1e8
    r2 = ((x[:, None, None] - x0_mm) ** 2 + (y[None, :, None] - y0_mm) ** 2
          + (zmm[None, None, :] + depth_mm) ** 2)
    return S0.c2 * (1.0 + amp * np.exp(-r2 / radius_mm ** 2)) ** 2


# ---------------------------------------------------------------------------
# checkpointed forward + recomputing adjoint
# ---------------------------------------------------------------------------
def forward_checkpointed(eng, spec: SolveSpec, every: int) -> np.ndarray:
    """Forward loop keeping only restart states: checkpoint g = the state
    before step g*every (device memory).  Returns the receiver traces."""
    steps = spec.steps
    nck = spec.maxtime // every + 1
    eng.checkpoints_alloc(nck)
    ijk = _ijk(spec, eng.cfg.o_rad)
    traces = np.zeros((len(spec.receivers), spec.nt))
    events = sorted(set(range(0, spec.maxtime + 1, steps)) |
                    set(g * every - 1 for g in range(1, nck)) | {-1})
    t = 0
    for i in events:
        t = _run_to(eng, t, i)
        if i >= 0 and i % steps == 0:
            traces[:, i // steps] = sample(eng, abi.V_Z, ijk)
        if (i + 1) % every == 0 and (i + 1) // every < nck:
            eng.checkpoint_save((i + 1) // every)
    return traces


def adjoint_recompute(fe, ae, spec: SolveSpec, vr: np.ndarray, grav: np.ndarray, every: int,
                      kstride: int | None = None, aligned: bool = False) -> None:
    """Adjoint loop that regenerates the forward snapshots segment by segment
    from the checkpoints of ``fe`` (last segment first, as the adjoint
    consumes forward time backwards) and accumulates against them."""
    ks = kstride or spec.steps
    nslots = spec.maxtime // ks + 1
    pairs = []
    for j in range(nslots):
        ia = spec.maxtime - j * ks - (1 if aligned else 0)
        if ia >= 0:
            pairs.append((ia, j))
    pairs.sort()  # adjoint time ascending = forward slot descending
    seg_of = lambda j: (j * ks) // every
    per_seg = every // ks + 2
    fe.snapshots_alloc(per_seg)
    ae.set_source(abi.SRC_LAGRANGE, vr, None, 10 ** 9)
    ae.kernel_begin(grav)
    t_adj = 0
    q = 0
    while q < len(pairs):
        g = seg_of(pairs[q][1])
        group = []
        while q < len(pairs) and seg_of(pairs[q][1]) == g:
            group.append(pairs[q])
            q += 1
        # regenerate the forward snapshots of segment g
        fe.checkpoint_load(g)
        slots = sorted(j for _, j in group)
        local = {j: n for n, j in enumerate(slots)}
        t = g * every
        for j in slots:
            t = _run_to(fe, t, j * ks)
            fe.snapshot_store(local[j])
        for ia, j in group:
            t_adj = _run_to(ae, t_adj, ia)
            ae.kernel_accumulate_from(fe, local[j])


def full_solve_checkpointed(spec: SolveSpec, S: bgm.Setup, obs: np.ndarray,
                            meas: RingMeasurement | None, every: int,
                            kstride: int | None = None, aligned: bool = False,
                            exact_adjoint: bool = False):
    """full_solve with checkpoints every ``every`` steps instead of a
    snapshot bank for the whole run; returns (SolveResult, meas, peak device
    bytes of the two contexts)."""
    fe, ae = make_engines(spec, S)
    try:
        syn = forward_checkpointed(fe, spec, every)
        if meas is None:
            meas = RingMeasurement(syn, spec.cad, spec.window_width)
        tau = meas.travel_times(obs, syn)
        adj = meas.adjoint_source(syn, tau, obs=obs, exact=exact_adjoint)
        vr = adjoint_vr(spec, adj, S)
        adjoint_recompute(fe, ae, spec, vr, S.g_raw, every, kstride=kstride, aligned=aligned)
        raw = {k: ae.kernel_get(k) for k in kernel_ids(ae.cfg)}
        K = kernel_lnc(raw[abi.KERN_C2], S, spec, kstride=kstride)
        peak = fe.device_bytes + ae.device_bytes
        return (SolveResult(traces=syn, tau=tau, misfit=meas.misfit(tau), adj=adj,
                            kernel_raw=raw, K_lnc=K), meas, peak)
    finally:
        fe.close()
        ae.close()
