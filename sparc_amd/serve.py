"""Continuous-batching request loop for single-source forward + adjoint +
kernel solves on one GPU.

An inversion issues many independent solves: one per source pixel for the
gradient and one per source and line-search step size for the step
(full.py:34-99, linesearch.py:30-52, each a separate `sparc` process in the
reference).  Here they are requests to one server process that keeps them all
on the device:

* admission by HBM footprint: a request starts as soon as its contexts and
  its snapshot bank (or, when memory is short, its forward checkpoints) fit in
  the free budget, computed on the host by sparc_footprint before any
  allocation;
* requests that arrive while others run are admitted as slots and memory
  free up; every admitted request runs in its own worker thread on its own
  two HIP streams (forward, adjoint), so their launches interleave on the
  device (ctypes releases the GIL during each C call);
* the result of a request does not depend on what else is running: contexts
  share no device memory and every kernel is deterministic
  (tests/test_gpu_serve.py checks bit-equality alone vs in a mix).

Workload of the benchmark: Poisson arrivals (numpy default_rng(5)); 64
source pixels drawn from a 16 x 16 grid of positions, 4 line-search step
sizes (0.5, 1, 1.5, 2 times a model update); grid mix 40 % 128x128x150,
50 % 256x256x300, 10 % 384x384x300; durations uniform in 1000..3000 steps.
Observed data are the request's own synthetics delayed by 2 s (a stand-in
that needs no extra solve and gives travel times of about 2 s).
"""
from __future__ import annotations

import threading
import time
from collections import deque
from dataclasses import dataclass, replace

import numpy as np

from . import abi
from . import background as bgm
from . import driver as D
from . import pipeline as PL
from .measure import RingMeasurement, ring_pixels

STEP_SIZES = (0.5, 1.0, 1.5, 2.0)
GRID_MIX = (("box2s", 0.4), ("box2", 0.5), ("box2l", 0.1))


@dataclass
class Request:
    rid: int
    workload: str
    src_xy_mm: tuple
    alpha: float            # line-search step: model = base + alpha * update
    nsteps: int
    arrival: float = 0.0    # seconds after the loop starts
    precision: int = abi.FP32


@dataclass
class Result:
    rid: int
    misfit: float
    tau: np.ndarray
    kernel_norm: float
    K_lnc: np.ndarray | None
    mode: str               # "snapshots" or "checkpoints"
    reserved_bytes: int
    arrival: float
    start: float
    end: float

    @property
    def latency(self) -> float:
        return self.end - self.arrival


def c5_requests(n: int, rate_per_s: float, seed: int = 5, mix=GRID_MIX,
                steps_range=(1000, 3000), precision: int = abi.FP32,
                extent_mm: float = 15.0) -> list[Request]:
    """The first n requests of the benchmark's arrival stream; source pixels
    on a 16 x 16 grid over +-extent_mm around the box centre (inside every
    grid of the mix)."""
    rng = np.random.default_rng(seed)
    ax = np.linspace(-extent_mm, extent_mm, 16)
    grid = [(x, y) for x in ax for y in ax]
    pick = rng.choice(len(grid), size=64, replace=False)
    combos = [(grid[p], a) for p in pick for a in STEP_SIZES]
    order = rng.permutation(len(combos))
    names = [m[0] for m in mix]
    probs = np.array([m[1] for m in mix])
    t = 0.0
    out = []
    for r in range(n):
        (xy, alpha) = combos[order[r % len(combos)]]
        wl = names[rng.choice(len(names), p=probs / probs.sum())]
        nsteps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        t += float(rng.exponential(1.0 / rate_per_s))
        out.append(Request(r, wl, (float(xy[0]), float(xy[1])), float(alpha), nsteps, t,
                           precision))
    return out


def device_free_bytes() -> tuple[int, int]:
    """(free, total) device memory of the current GPU from the HIP runtime
    the engine itself links (hipMemGetInfo), so the query does not start a
    second runtime in the process."""
    import ctypes as C
    abi.lib()  # loads libamdhip64 as a dependency of the engine
    hip = C.CDLL("libamdhip64.so.7")
    free, total = C.c_size_t(0), C.c_size_t(0)
    rc = hip.hipMemGetInfo(C.byref(free), C.byref(total))
    if rc != 0:
        raise RuntimeError(f"hipMemGetInfo failed ({rc})")
    return int(free.value), int(total.value)


def mix_sample(nsteps: int = 1000, precision: int = abi.FP32) -> list[Request]:
    """A fixed request set with one request per grid of the mix (128^2 x 150,
    256^2 x 300, 384^2 x 300), all arriving at once, each nsteps long (the
    short end of the stream's 1000-3000): the serving figure the step
    benchmark reports beside its step rate."""
    picks = (("box2s", (-6.0, 4.0), 1.0), ("box2", (5.0, -3.0), 0.5), ("box2l", (2.0, 7.0), 1.5))
    return [Request(i, wl, xy, a, nsteps, 0.0, precision) for i, (wl, xy, a) in enumerate(picks)]


class ModelCache:
    """FP64 set-ups per (grid, step size): base polytrope plus alpha times a
    fixed update direction (+2 % sound speed, Gaussian of radius 5 Mm at 3 Mm
    depth).  Built once per model like the reference's per-iteration model
    files; thread-safe."""

    def __init__(self):
        self._lock = threading.Lock()
        self._setups: dict = {}
        self._base: dict = {}

    def workload(self, name: str, precision: int, src_xy_mm=(0.0, 0.0)) -> D.Workload:
        W = D.workload(name, precision=precision)
        W.src_xy_mm = tuple(src_xy_mm)
        return W

    def setup(self, name: str, alpha: float) -> bgm.Setup:
        key = (name, alpha)
        with self._lock:
            if key in self._setups:
                return self._setups[key]
        W = D.workload(name)
        base = self._base.get(name)
        if base is None:
            base = D.setup_for(W)
            self._base[name] = base
        if alpha == 0.0:
            S = base
        else:
            c2 = PL.bump_c2(base, W, amp=0.02 * alpha)
            S = D.setup_for(W, c2_3d=c2)
        with self._lock:
            self._setups[key] = S
        return S


def _delayed(u: np.ndarray, dt: float, shift_s: float) -> np.ndarray:
    n = u.shape[-1]
    f = np.fft.rfftfreq(n, d=dt)
    return np.fft.irfft(np.fft.rfft(u, axis=-1) * np.exp(-2j * np.pi * f * shift_s), n=n, axis=-1)


def request_spec(req: Request, cache: ModelCache, ring_mm: float = 10.0) -> PL.SolveSpec:
    W = cache.workload(req.workload, req.precision, req.src_xy_mm)
    P = W.params
    rec = ring_pixels(P.nx, P.ny, P.xlength, P.ylength, req.src_xy_mm[0], req.src_xy_mm[1],
                      ring_mm, 8)
    return PL.SolveSpec(W=W, receivers=rec, maxtime=req.nsteps - 1, precision=req.precision)


def footprint(spec: PL.SolveSpec, mode: str, checkpoint_every: int = 100) -> int:
    """Device bytes one request holds: two contexts plus the snapshot bank
    (mode "snapshots") or the forward checkpoints and one segment's snapshots
    (mode "checkpoints")."""
    P = spec.W.params
    c = abi.SparcConfig()
    c.abi_version = abi.ABI_VERSION
    c.nx, c.ny, c.nz = P.nx, P.ny, P.nz
    c.option = P.option
    c.precision = spec.precision
    c.npmlbot, c.npmltop = P.npmlbot, P.npmltop
    c.damp_waves = int(P.damp_waves and not P.is2d)
    c.freq_filtering = P.freq_filtering
    c.e_rad = c.o_rad = P.nz // 2
    c.compensated = int(P.compensated and spec.precision == abi.FP32)
    c.deltat = 1.0  # any valid step: the footprint does not depend on it
    c.field_scale = 1.0
    ctx = abi.footprint(c)
    ts = 4 if spec.precision == abi.FP32 else 8
    N = P.nx * P.ny * P.nz
    NP = P.nx * P.ny * max(P.npmlbot + P.npmltop, 1)
    if mode == "snapshots":
        extra = (spec.maxtime // spec.steps + 1) * 6 * N * ts
    else:
        nck = spec.maxtime // checkpoint_every + 1
        # a checkpoint slot: the six fields, their 2-byte low words when the
        # state is compensated, the two PML fields
        slot = (6 * N + 2 * NP) * ts + (6 * N * 2 if c.compensated else 0)
        extra = nck * slot + (checkpoint_every // spec.steps + 2) * 6 * N * ts
    return 2 * ctx + extra


def solve_request(req: Request, cache: ModelCache, mode: str = "snapshots",
                  checkpoint_every: int = 100, delay_s: float = 2.0) -> tuple:
    """One forward + measure + adjoint + kernel solve; returns (misfit, tau,
    K_lnc)."""
    spec = request_spec(req, cache)
    S = cache.setup(req.workload, req.alpha)
    fe, ae = PL.make_engines(spec, S)
    try:
        if mode == "snapshots":
            syn = PL.forward_solve(fe, spec, store=True)
        else:
            syn = PL.forward_checkpointed(fe, spec, checkpoint_every)
        obs = _delayed(syn, spec.cad, delay_s)
        meas = RingMeasurement(syn, spec.cad, 120.0)
        tau = meas.travel_times(obs, syn)
        adj = meas.adjoint_source(syn, tau)
        vr = PL.adjoint_vr(spec, adj, S)
        if mode == "snapshots":
            PL.adjoint_solve(ae, spec, vr, S.g_raw, src=fe)
        else:
            PL.adjoint_recompute(fe, ae, spec, vr, S.g_raw, checkpoint_every)
        K = PL.kernel_lnc(ae.kernel_get(abi.KERN_C2), S, spec)
        return meas.misfit(tau), tau, K
    finally:
        fe.close()
        ae.close()


class Server:
    """Admits requests in arrival order by HBM footprint and runs each in its
    own worker thread; ``run`` returns the results in completion order."""

    def __init__(self, budget_bytes: int, max_inflight: int = 8, checkpoint_every: int = 100,
                 keep_kernels: bool = False, cache: ModelCache | None = None):
        self.budget = int(budget_bytes)
        self.max_inflight = max_inflight
        self.checkpoint_every = checkpoint_every
        self.keep = keep_kernels
        self.cache = cache or ModelCache()
        self.used = 0
        self.peak_used = 0
        self.peak_inflight = 0
        self.errors: list = []

    def run(self, requests: list[Request], log=None) -> list[Result]:
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        cond = threading.Condition()
        inflight: dict = {}
        results: list[Result] = []
        t0 = time.perf_counter()

        def worker(req: Request, mode: str, need: int, started: float):
            try:
                chi, tau, K = solve_request(req, self.cache, mode, self.checkpoint_every)
                res = Result(req.rid, chi, tau, float(np.linalg.norm(K)), K if self.keep else None,
                             mode, need, req.arrival, started, time.perf_counter() - t0)
            except Exception as e:  # reported, the loop keeps serving
                res = None
                self.errors.append((req.rid, repr(e)))
            with cond:
                self.used -= need
                inflight.pop(req.rid, None)
                if res is not None:
                    results.append(res)
                    if log:
                        log(f"[serve] done rid={req.rid} {req.workload} {req.nsteps} steps "
                            f"mode={mode} latency={res.latency:.1f}s")
                cond.notify_all()

        with cond:
            while pending or inflight:
                now = time.perf_counter() - t0
                while pending and pending[0].arrival <= now and len(inflight) < self.max_inflight:
                    req = pending[0]
                    spec = request_spec(req, self.cache)
                    mode, need = None, 0
                    for m in ("snapshots", "checkpoints"):
                        b = footprint(spec, m, self.checkpoint_every)
                        if self.used + b <= self.budget:
                            mode, need = m, b
                            break
                    if mode is None:
                        if not inflight:  # cannot fit even alone
                            pending.popleft()
                            self.errors.append((req.rid, "request exceeds the HBM budget"))
                            continue
                        break
                    pending.popleft()
                    self.used += need
                    self.peak_used = max(self.peak_used, self.used)
                    th = threading.Thread(target=worker, args=(req, mode, need, now), daemon=True)
                    inflight[req.rid] = th
                    self.peak_inflight = max(self.peak_inflight, len(inflight))
                    th.start()
                timeout = None
                if pending and pending[0].arrival > now:
                    timeout = pending[0].arrival - now
                if pending or inflight:
                    cond.wait(timeout=timeout if timeout is not None else 1.0)
        return results


def summarize(results: list[Result]) -> dict:
    if not results:
        return {"solves": 0}
    lat = np.array([r.latency for r in results])
    first = min(r.arrival for r in results)
    last = max(r.end for r in results)
    return {"solves": len(results), "wall_s": last - first,
            "solves_per_hour": 3600.0 * len(results) / (last - first),
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "modes": {m: sum(r.mode == m for r in results) for m in ("snapshots", "checkpoints")}}


def plain_fp64(requests: list[Request], cache: ModelCache | None = None) -> dict:
    """The plain path: the same requests one after another in FP64 with the
    full snapshot bank, nothing overlapped."""
    return plain_sequential(requests, cache, abi.FP64)


def plain_sequential(requests: list[Request], cache: ModelCache | None = None,
                     precision: int = abi.FP64) -> dict:
    """The same requests one after another in the given precision with the
    full snapshot bank, nothing overlapped (no serving loop)."""
    cache = cache or ModelCache()
    t0 = time.perf_counter()
    per = []
    for r in requests:
        t = time.perf_counter()
        solve_request(replace(r, precision=precision), cache, "snapshots")
        per.append(time.perf_counter() - t)
    wall = time.perf_counter() - t0
    return {"solves": len(requests), "wall_s": wall, "solves_per_hour": 3600.0 * len(requests) / wall,
            "per_solve_s": per}
