#!/usr/bin/env python3
"""Benchmark of the forward time-stepping hot path on one MI355X.

Metric: Mpt-updates/s per RK step = nx*ny*nz*steps / wall-seconds, one step
being the reference's full 5-stage LDDRK STEP (5 RHS stages, a += dt*scr,
z filter, periodic spectral filter; step.f90:43-128), on the 256x256x300
polytrope box (200 Mm, z in [-12, +0.5] Mm, PML 10+10, Gaussian x Ricker
source, option 4 = MP_QUIET_PML_3D_DISPL with damping), FP32 state.
Inputs are resident in HBM before the timed region.

Prints ONE JSON line.  Besides the metric it carries:
  roofline      the dominant kernel class, its algorithmic bytes per launch /
                its mean launch time (HIP events on the engine's stream,
                recorded over a second pass of the same steps), vs 8 TB/s;
  fp64_step     the same workload stepped by the FP64 engine;
  serving       solves/hour and p50/p99 latency of the FP32 serving loop on
                one forward+adjoint+kernel request per grid of the mix
                (128^2 x 150, 256^2 x 300, 384^2 x 300; 1000 steps each),
                against the same requests one at a time in FP64;
  cpu_baseline  the FP64 CPU oracle (a C restatement of the reference, OpenMP)
                on a bounded sample of the same workload (25 s of RK steps).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# the kernel behind each timed class in the forward step of the 3-D box
KERNEL_SYMBOLS = {
    "xlines": "sparc::k_xy256_div<float, 2> -- x derivative of xi_x (one wavefront per line pair) and "
              "y derivative of xi_y (register four-step) in one launch",
    "ylines": "sparc::k_y256<float, 0 (YM_DERIV)> -- y derivative of xi_y, register four-step",
    "damp": "sparc::k_xw256<float, 2 (XM_DAMP_FWD), 4> -- damping, forward x of the stage temp "
            "a + beta scr of the velocities, also stored as RHS_xi; first stage only, the "
            "last pass of the previous stage does it for the other four "
            "(with --damp2d: sparc::k_damp2d<float>, the whole damping, one 256x256 plane per block)",
    "damp_y": "sparc::k_y256_damp_deriv<float> -- damping y pass of v_x, v_y, v_z and the y derivative "
              "of p (RHS_vy) in one launch",
    "damp_xi": "sparc::k_xw256_rhs_dampinv<float, 2, true|false> -- x derivative of p (RHS_vx) and "
               "the damping inverse x of v_x, v_y, v_z subtracted from RHS_v in one launch; in "
               "stages 1-4 (true) also the next stage's damping forward x pass and RHS_xi",
    "zsolve": "sparc::k_zsolve<float, 7 (EPI_RHS3N), 8, true> -- z solve of p (ud in LDS) + RHS_vz",
    "zsolve_p": "sparc::k_zstream_pv -- streaming z solves, two walks per 64-column group (one "
                "from each end): d xi_z/dz with the pressure, then dp/dz with RHS_vz "
                "(without damping or on other line sizes: sparc::k_zsolve_b<float, 1 (EPI_P3D), 16, true>)",
    "update": "sparc::k_update_filter_c -- a += dt*scr with the z filter (FP32 high words, 16-bit low words), all fields per launch",
}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="box2")
    ap.add_argument("--precision", default="fp32", choices=("fp32", "fp64"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true",
                    help="skip the FP64 step rate of the same workload")
    ap.add_argument("--serving-mix", type=int, default=1000, metavar="STEPS",
                    help="steps per request of the in-line serving figure (one request per "
                         "grid of the mix; 0 = skip)")
    ap.add_argument("--traffic-json", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="HBM bytes per launch measured by rocprofv3 --pmc (optional)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (calibrates the "
                         "FETCH_SIZE/WRITE_SIZE counters against known bytes) and exit")
    ap.add_argument("--dft-probe", type=int, default=0, metavar="REPS",
                    help="only time the 128-wide x derivative of the 128x128x150 box as the "
                         "engine's FFT line kernel and as a dense FP32 (and bf16) GEMM on the "
                         "matrix cores, REPS launches each, and exit")
    ap.add_argument("--serving", type=int, default=0, metavar="N",
                    help="serving benchmark instead: the first N requests of the arrival "
                         "stream through the FP32 serving loop, then through the plain FP64 path")
    ap.add_argument("--serving-rate", type=float, default=1.0, help="arrivals per second")
    ap.add_argument("--solve", type=int, default=0, metavar="STEPS",
                    help="one forward + adjoint + kernel solve of STEPS steps on the workload "
                         "(observed data from a +2%% sound-speed bump, 8 receivers on a 10 Mm "
                         "ring) instead of the step benchmark")
    ap.add_argument("--traffic-from", default="", metavar="DIR",
                    help="write the counter traffic per launch (--traffic-out) from rocprofv3 "
                         "--pmc csv output: DIR/fetch and DIR/write (FETCH_SIZE, WRITE_SIZE over "
                         "a bench run), DIR/copy_fetch and DIR/copy_write (the same counters "
                         "over --copy-probe 256), with this run's algorithmic bytes per class")
    ap.add_argument("--traffic-out", default=os.path.join(ROOT, "profiles", "traffic.json"))
    ap.add_argument("--damp2d", action="store_true",
                    help="the damping of 256x256 FP32 planes by k_damp2d, one pass per plane "
                         "(cfg.plane_damping)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.dft_probe:
        return dft_probe(args.dft_probe)
    if args.serving:
        if args.gpus != 1:
            print("bench.py measures one GPU only", file=sys.stderr)
            return 2
        return serving_bench(args.serving, args.serving_rate)
    if args.solve:
        if args.gpus != 1:
            print("bench.py measures one GPU only", file=sys.stderr)
            return 2
        return solve_bench(args.workload, args.precision, args.solve)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        return 2

    import numpy as np

    from sparc_amd import abi
    from sparc_amd import driver as D

    prec = abi.FP32 if args.precision == "fp32" else abi.FP64
    W = D.workload(args.workload, precision=prec)
    if args.damp2d:
        W.params.plane_damping = True
    P = W.params
    S = D.setup_for(W)
    total = args.warmup + args.steps
    eng, cfg = D.make_engine(W, S, nsteps=max(total, W.nsteps))
    mask, series = D.forward_source(W, max(total, W.nsteps) + 8)
    eng.set_source(abi.SRC_SERIES, mask, series, 10 ** 9)
    N = P.nx * P.ny * P.nz
    print(f"[bench] workload={W.name} grid={P.nx}x{P.ny}x{P.nz} dt={P.timestep:.4f}s "
          f"precision={args.precision} warmup={args.warmup} steps={args.steps}",
          file=sys.stderr, flush=True)

    eng.step(0, args.warmup)
    eng.sync()
    t0 = time.perf_counter()
    eng.step(args.warmup, args.steps)
    eng.sync()
    t1 = time.perf_counter()
    wall = t1 - t0
    mpts = N * args.steps / wall / 1e6
    finite = bool(np.isfinite(eng.norm(abi.V_Z)))
    print(f"[bench] timed {args.steps} steps in {wall:.3f}s -> {mpts:.1f} Mpt/s", file=sys.stderr,
          flush=True)

    # second pass of the same steps with HIP-event brackets per launch
    eng.profile(True)
    eng.step(args.warmup + args.steps, args.steps)
    prof = eng.profile_read()
    eng.profile(False)
    kernels = {k: {"ms_total": v[0], "launches": v[2],
                   "us_per_launch": 1e3 * v[0] / v[2] if v[2] else None,
                   "gbs": v[1] / (v[0] * 1e-3) / 1e9 if v[0] > 0 else None}
               for k, v in prof.items() if v[2]}
    dom = max(prof, key=lambda k: prof[k][0])
    ms, by, nl = prof[dom]
    achieved = by / (ms * 1e-3) / 1e9
    traffic = None
    try:
        with open(args.traffic_json) as f:
            tj = json.load(f)
        ent = tj.get(f"{W.name}:{args.precision}:{dom}")
        # counter bytes of the same kernel only: an entry naming other
        # kernels, or whose algorithmic bytes differ from this run's, was
        # measured on other code
        same = ent and ent.get("kernel") == " + ".join(CLASS_KERNELS.get(dom, ()))
        if same and abs(ent["algorithmic_bytes_per_launch"] * nl / by - 1.0) < 0.01:
            traffic = ent["hbm_bytes_per_launch"]
        elif ent:
            print(f"[bench] {args.traffic_json} entry for {dom} is stale ({ent.get('kernel')}, "
                  f"{ent['algorithmic_bytes_per_launch']:.4g} vs {by / nl:.4g} algorithmic bytes): "
                  "traffic not reported", file=sys.stderr, flush=True)
    except (OSError, ValueError):
        pass
    roof = {"bound": "hbm", "kernel": dom, "kernel_symbol": KERNEL_SYMBOLS.get(dom, dom),
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic, "algorithmic_bytes_per_launch": by / nl,
            "us_per_launch": 1e3 * ms / nl, "kernels": kernels}
    # bytes per point and step: 18 field arrays moved per field and step, 5
    # 3-D background arrays per stage (15 with a magnetic field), and the
    # compensated state's 2-byte low words read and written once per field
    mhd = P.option in (abi.OPT_MHD, abi.OPT_MHD8)
    nbg = 15 if mhd else 5
    step_algo = ((18 * 6 + 5 * nbg) * (4 if prec == abi.FP32 else 8)
                 + (6 * 2 * 2 if cfg.compensated else 0))
    opt = "option 10" if P.option == abi.OPT_MHD else ("option 8" if P.option == abi.OPT_MHD8
                                                        else "option 4")
    eng.close()

    fp64 = None
    if prec == abi.FP32 and not args.no_fp64:
        fp64 = fp64_step_rate(W, S, args.warmup, min(args.steps, 20))
    serving = None
    if args.serving_mix:
        serving = serving_mix(args.serving_mix)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(W)

    line = {
        "metric": f"Mpt-updates/s per RK step (forward, {opt}, {P.nx}x{P.ny}x{P.nz})",
        "value": round(mpts, 2),
        "unit": "Mpt/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == abi.FP32 else "f64",
        "data": "synthetic (polytrope background" + (" with a 3 kG flux tube" if mhd else "")
                + ", Gaussian x Ricker source)",
        "config": {"workload": f"{W.name}: {P.nx}x{P.ny}x{P.nz} polytrope box, "
                               f"{P.xlength / 1e8:.0f} Mm, {opt} with PML {P.npmlbot}+{P.npmltop}, "
                               f"damping, filters, dt={P.timestep:.4f}s",
                   "nx": P.nx, "ny": P.ny, "nz": P.nz, "parallelism": "single GPU",
                   "finite": finite,
                   "compensated_fp32_state": bool(cfg.compensated),
                   "step_bytes_per_point": step_algo,
                   "step_roofline_frac": round(mpts * 1e6 * step_algo / 1e9 / HBM_PEAK_GBS, 4)},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if args.traffic_from:
        write_traffic(args.traffic_from, args.traffic_out, f"{W.name}:{args.precision}", prof)
    if fp64 is not None:
        line["fp64_step"] = fp64
    if serving is not None:
        line["serving"] = serving
    print(json.dumps(line))
    return 0


def fp64_step_rate(W, S, warmup: int, steps: int) -> dict:
    """The same workload stepped by the FP64 engine (the plain path's
    arithmetic): Mpt/s over `steps` steps after `warmup`."""
    from sparc_amd import abi
    from sparc_amd import driver as D
    P = W.params
    total = warmup + steps
    eng, cfg = D.make_engine(W, S, precision=abi.FP64, nsteps=max(total, W.nsteps))
    try:
        mask, series = D.forward_source(W, max(total, W.nsteps) + 8)
        eng.set_source(abi.SRC_SERIES, mask, series, 10 ** 9)
        eng.step(0, warmup)
        eng.sync()
        t0 = time.perf_counter()
        eng.step(warmup, steps)
        eng.sync()
        wall = time.perf_counter() - t0
    finally:
        eng.close()
    N = P.nx * P.ny * P.nz
    print(f"[bench] FP64: {steps} steps in {wall:.3f}s -> {N * steps / wall / 1e6:.1f} Mpt/s",
          file=sys.stderr, flush=True)
    return {"value": round(N * steps / wall / 1e6, 2), "unit": "Mpt/s", "steps": steps,
            "ms_per_step": round(1e3 * wall / steps, 4), "dtype": "f64"}


def serving_mix(nsteps: int) -> dict:
    """Solves/hour of the FP32 serving loop and of the plain FP64 and FP32
    paths (the same requests one at a time) on serve.mix_sample: one request
    per grid of the mix, 384x384x300 included."""
    from sparc_amd import abi
    from sparc_amd import serve as SV

    reqs = SV.mix_sample(nsteps)
    cache = SV.ModelCache()
    for r in reqs:
        cache.setup(r.workload, r.alpha)
    free, _ = SV.device_free_bytes()
    srv = SV.Server(int(0.9 * free), max_inflight=8, cache=cache)
    res = srv.run(reqs, log=lambda m: print(m, file=sys.stderr, flush=True))
    fast = SV.summarize(res)
    print(f"[bench] serving loop: {fast}", file=sys.stderr, flush=True)
    plain = SV.plain_fp64(reqs, cache)
    print(f"[bench] plain FP64: {plain}", file=sys.stderr, flush=True)
    plain32 = SV.plain_sequential(reqs, cache, abi.FP32)
    print(f"[bench] plain FP32 (sequential): {plain32}", file=sys.stderr, flush=True)
    sph = fast.get("solves_per_hour", 0.0)
    return {"metric": "forward+adjoint+kernel solves/hour", "unit": "solves/h",
            "requests": [f"{r.workload} x {r.nsteps} steps" for r in reqs],
            "errors": len(srv.errors),
            "fp32_loop": {"solves_per_hour": round(sph, 1),
                          "wall_s": round(fast.get("wall_s", 0.0), 2),
                          "p50_latency_s": round(fast.get("p50_latency_s", 0.0), 2),
                          "p99_latency_s": round(fast.get("p99_latency_s", 0.0), 2)},
            "plain_fp64": {"solves_per_hour": round(plain["solves_per_hour"], 1),
                           "wall_s": round(plain["wall_s"], 2),
                           "per_solve_s": [round(x, 2) for x in plain["per_solve_s"]]},
            "plain_fp32": {"solves_per_hour": round(plain32["solves_per_hour"], 1),
                           "wall_s": round(plain32["wall_s"], 2),
                           "per_solve_s": [round(x, 2) for x in plain32["per_solve_s"]]},
            "speedup_vs_plain_fp64": round(sph / plain["solves_per_hour"], 3),
            "speedup_vs_plain_fp32": round(sph / plain32["solves_per_hour"], 3)}


def serving_bench(n: int, rate: float) -> int:
    """Solves/hour of the FP32 serving loop against the plain FP64 path on the
    same n requests (forward + adjoint + kernel each)."""
    from sparc_amd import serve as SV

    reqs = SV.c5_requests(n, rate_per_s=rate)
    cache = SV.ModelCache()
    t0 = time.perf_counter()
    for r in reqs:  # the models of an inversion step exist before its solves
        cache.setup(r.workload, r.alpha)
    setup_s = time.perf_counter() - t0
    free, total = SV.device_free_bytes()
    budget = int(0.9 * free)
    print(f"[serving] {n} requests, set-up {setup_s:.1f}s, budget {budget / 2**30:.0f} GiB",
          file=sys.stderr, flush=True)
    srv = SV.Server(budget, max_inflight=8, cache=cache)
    res = srv.run(reqs, log=lambda m: print(m, file=sys.stderr, flush=True))
    if srv.errors:
        print(f"[serving] errors: {srv.errors}", file=sys.stderr, flush=True)
    fast = SV.summarize(res)
    print(f"[serving] FP32 loop: {fast}", file=sys.stderr, flush=True)
    plain = SV.plain_fp64(reqs, cache)
    print(f"[serving] plain FP64: {plain}", file=sys.stderr, flush=True)
    line = {
        "metric": "forward+adjoint+kernel solves/hour (FP32 serving loop)",
        "value": round(fast.get("solves_per_hour", 0.0), 2),
        "unit": "solves/h",
        "n_gpus": 1,
        "higher_is_better": True,
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (polytrope models, observations = synthetics delayed by 2 s)",
        "config": {"workload": f"first {n} requests of the arrival stream (Poisson {rate}/s, "
                               "default_rng(5); grid mix 40/50/10 % of 128x128x150 / "
                               "256x256x300 / 384x384x300; 1000-3000 steps; 8 ring receivers)",
                   "requests": [(r.workload, r.nsteps) for r in reqs],
                   "max_inflight": 8, "hbm_budget_gib": round(budget / 2**30, 1)},
        "serving": {k: v for k, v in fast.items()},
        "peak_inflight": srv.peak_inflight,
        "peak_reserved_gib": round(srv.peak_used / 2**30, 2),
        "plain_fp64": {"solves_per_hour": round(plain["solves_per_hour"], 2),
                       "wall_s": round(plain["wall_s"], 2),
                       "per_solve_s": [round(x, 2) for x in plain["per_solve_s"]]},
        "speedup_vs_plain_fp64": round(fast.get("solves_per_hour", 0.0)
                                       / plain["solves_per_hour"], 3),
    }
    print(json.dumps(line))
    return 0


# the kernels of each timed class on the default path of the 256x256x300 FP32
# box, as rocprofv3 names them
CLASS_KERNELS = {
    "zsolve_p": ("k_zstream_pv",),
    "damp": ("k_xw256<float, 2, 4>",),
    "damp_y": ("k_y256_damp_deriv<float>",),
    "damp_xi": ("k_xw256_rhs_dampinv<float, 2, true>", "k_xw256_rhs_dampinv<float, 2, false>"),
    "update": ("k_update_filter_c",),
    "xlines": ("k_xy256_div<float, 2>",),
    "ylines": ("k_y256<float, 0>",),
}


def _pmc_per_kernel(directory: str) -> dict:
    """{kernel name: (dispatches, mean counter value per dispatch)} from the
    *counter_collection.csv files rocprofv3 writes under directory."""
    import csv
    import glob
    per = {}
    for f in glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True):
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                key = (row["Kernel_Name"], row.get("Dispatch_Id", row.get("Correlation_Id")))
                per[key] = per.get(key, 0.0) + float(row["Counter_Value"])
    out = {}
    for (name, _), v in per.items():
        n, t = out.get(name, (0, 0.0))
        out[name] = (n + 1, t + v)
    return {k: (n, t / n) for k, (n, t) in out.items()}


def write_traffic(directory: str, out_path: str, prefix: str, prof: dict) -> None:
    """HBM bytes per launch of each class: FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch) scaled by what they read for a device-to-device copy of known
    size, averaged over the class's dispatches."""
    fe, wr = _pmc_per_kernel(os.path.join(directory, "fetch")), _pmc_per_kernel(
        os.path.join(directory, "write"))
    cf, cw = _pmc_per_kernel(os.path.join(directory, "copy_fetch")), _pmc_per_kernel(
        os.path.join(directory, "copy_write"))
    copy_bytes = 256 * (1 << 20)
    ck = [k for k in cf if "copyBuffer" in k]
    if not ck:
        raise RuntimeError(f"no copy kernel under {directory}/copy_fetch")
    fscale = copy_bytes / (1024.0 * cf[ck[0]][1])
    wscale = copy_bytes / (1024.0 * cw[ck[0]][1])
    res = {"_calibration": {"copy_bytes": copy_bytes, "fetch_scale": fscale, "write_scale": wscale,
                            "note": "FETCH_SIZE and WRITE_SIZE (KiB per dispatch) scaled by what "
                                    "they read for a 256 MiB device-to-device copy"}}
    for cls, names in CLASS_KERNELS.items():
        if cls not in prof or not prof[cls][2]:
            continue
        rows = [(k, fe[k], wr.get(k)) for k in fe if any(nm in k for nm in names)]
        if not rows or any(w is None for _, _, w in rows):
            continue
        nd = sum(f[0] for _, f, _ in rows)
        fb = sum(f[0] * f[1] for _, f, _ in rows) / nd * 1024.0 * fscale
        wb = sum(w[0] * w[1] for _, _, w in rows) / nd * 1024.0 * wscale
        ms, by, nl = prof[cls]
        res[f"{prefix}:{cls}"] = {"kernel": " + ".join(names), "hbm_bytes_per_launch": fb + wb,
                                  "fetch_bytes": fb, "write_bytes": wb,
                                  "algorithmic_bytes_per_launch": by / nl, "dispatches": nd}
    with open(out_path, "w") as f:
        json.dump(res, f, indent=1)
    print(f"[bench] counter traffic of {len(res) - 1} classes -> {out_path}", file=sys.stderr,
          flush=True)


def dft_probe(reps: int) -> int:
    """The x derivative of the 128 x 128 x 150 box (19200 lines of 128) two
    ways: the engine's FFT line kernel (HIP events around each launch), and
    the same operator as a dense product Y = X D^T on the matrix cores
    (torch.matmul, FP32 and bf16), D being the engine's own derivative matrix
    (the operator applied to unit impulses).  Prints one JSON line with the
    times per launch and the FP32 product's difference from the FFT result."""
    import numpy as np
    import torch

    from sparc_amd import abi
    from sparc_amd import driver as D

    W = D.workload("box2s", precision=abi.FP32)
    S = D.setup_for(W)
    eng, _ = D.make_engine(W, S, nsteps=2)
    nx, ny, nz = eng.shape
    # D[:, j] = d/dx of the unit impulse at x = j (line y = j of plane 0)
    imp = np.zeros((nx, ny, nz))
    for j in range(min(nx, ny)):
        imp[j, j, 0] = 1.0
    d_op = eng.apply_op(abi.OP_DDX, imp)[:, :nx, 0]
    rng = np.random.default_rng(3)
    f = rng.standard_normal((nx, ny, nz))
    g = eng.apply_op(abi.OP_DDX, f)
    eng.profile(True)
    for _ in range(reps):
        eng.apply_op(abi.OP_DDX, f)
    prof = eng.profile_read()
    eng.profile(False)
    eng.close()
    ms, by, nl = prof["xlines"]
    us_fft = 1e3 * ms / nl
    # lines as rows: X[l, i] = f[i, y, z], l = y + ny z
    X = torch.tensor(f.transpose(1, 2, 0).reshape(-1, nx), dtype=torch.float32, device="cuda")
    Dt = torch.tensor(d_op.T.copy(), dtype=torch.float32, device="cuda")
    res = {}
    for name, dt in (("fp32", torch.float32), ("bf16", torch.bfloat16)):
        Xd, Dd = X.to(dt), Dt.to(dt).contiguous()
        Y = torch.matmul(Xd, Dd)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            Y = torch.matmul(Xd, Dd)
        e1.record()
        torch.cuda.synchronize()
        us = 1e3 * e0.elapsed_time(e1) / reps
        y = Y.float().cpu().numpy().reshape(ny, nz, nx).transpose(2, 0, 1)
        err = float(np.linalg.norm(y - g) / np.linalg.norm(g))
        flops = 2.0 * X.shape[0] * nx * nx
        res[name] = {"us_per_launch": round(us, 2), "rel_l2_vs_fft_kernel": err,
                     "tflops": round(flops / (us * 1e-6) / 1e12, 1)}
    line = {"probe": "x derivative of 128x128x150 (19200 lines of 128)",
            "fft_line_kernel": {"us_per_launch": round(us_fft, 2),
                                "gbs": round(by / (ms * 1e-3) / 1e9, 1), "launches": nl},
            "dense_gemm": res, "reps": reps}
    print(json.dumps(line))
    return 0


def copy_probe(mib: int) -> int:
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n,
                      "copies": 5}))
    return 0


def solve_bench(workload: str, precision: str, nsteps: int) -> int:
    """Wall time of one forward + adjoint + kernel solve (the kernel-mode
    sequence of driver.f90) on the workload, with the kernels it returns."""
    import numpy as np

    from sparc_amd import abi
    from sparc_amd import driver as D
    from sparc_amd import pipeline as PL
    from sparc_amd.measure import ring_pixels

    prec = abi.FP32 if precision == "fp32" else abi.FP64
    W = D.workload(workload, precision=prec)
    P = W.params
    S0 = D.setup_for(W)
    rec = ring_pixels(P.nx, P.ny, P.xlength, P.ylength, W.src_xy_mm[0], W.src_xy_mm[1], 10.0, 8)
    maxtime = max(P.steps, nsteps // P.steps * P.steps)
    spec = PL.SolveSpec(W=W, receivers=rec, maxtime=maxtime, precision=prec)
    print(f"[solve] workload={W.name} grid={P.nx}x{P.ny}x{P.nz} option={P.option} "
          f"steps={maxtime + 1} precision={precision}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    obs = PL.observe(spec, D.setup_for(W, c2_3d=PL.bump_c2(S0, W)))
    t1 = time.perf_counter()
    print(f"[solve] observed data in {t1 - t0:.1f}s", file=sys.stderr, flush=True)
    res, _ = PL.full_solve(spec, S0, obs)
    t2 = time.perf_counter()
    names = {abi.KERN_C2: "c2", abi.KERN_RHO: "rho", abi.KERN_HESSIAN: "hessian",
             abi.KERN_VX: "vx", abi.KERN_VY: "vy", abi.KERN_VZ: "vz", abi.KERN_BX: "bx",
             abi.KERN_BY: "by", abi.KERN_BZ: "bz", abi.KERN_PSI: "psi"}
    kern = {names[k]: {"l2": float(np.linalg.norm(v)), "finite": bool(np.all(np.isfinite(v)))}
            for k, v in res.kernel_raw.items()}
    line = {"metric": "forward + adjoint + kernel solve wall time", "value": round(t2 - t1, 3),
            "unit": "s", "higher_is_better": False, "n_gpus": 1,
            "dtype": "f32" if prec == abi.FP32 else "f64", "data": "synthetic",
            "config": {"workload": W.name, "grid": [P.nx, P.ny, P.nz], "option": P.option,
                       "steps_per_solve": maxtime + 1, "receivers": len(rec)},
            "fwd_plus_adj_mpt_per_s": round(P.nx * P.ny * P.nz * (maxtime + 1) * 2 / (t2 - t1) / 1e6, 1),
            "misfit": res.misfit, "tau_s": [round(float(t), 4) for t in res.tau],
            "kernels": kern}
    print(json.dumps(line))
    return 0


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(W, min_seconds: float = 25.0, max_steps: int = 10):
    """FP64 CPU oracle (C restatement of the reference step, OpenMP) timed on
    RK steps of the same workload: steps until min_seconds have passed (at
    most max_steps), after one untimed step.  The reference itself cannot be
    built here (it needs FFTW3), so this port stands in for it."""
    try:
        from oracle import oracle as O
        from sparc_amd import abi
        from sparc_amd import driver as D
        S = D.setup_for(W)
        cfg = D.make_config(W.params, S, abi.FORWARD, W.nsteps)
        cfg.precision = abi.FP64
        o = O.Oracle(cfg)
        D.load_setup(o, S)
        mask, series = D.forward_source(W, W.nsteps + 8)
        o.set_source(abi.SRC_SERIES, mask, series, 10 ** 9)
        o.step(0, 1)  # warm (time 0 also runs the spectral filter)
        t0 = time.perf_counter()
        n = 0
        while n < max_steps:
            o.step(1 + n, 1)
            n += 1
            if time.perf_counter() - t0 >= min_seconds:
                break
        dt = time.perf_counter() - t0
        o.close()
        P = W.params
        threads = O.omp_threads()
        return {"value": round(P.nx * P.ny * P.nz * n / dt / 1e6, 4), "unit": "Mpt/s",
                "cores": threads, "kind": "port", "cpu_model": cpu_model(),
                "sample": f"{n} RK steps (time 1..{n}) of {P.nx}x{P.ny}x{P.nz}, FP64, "
                          f"{dt:.1f} s on {threads} OpenMP threads of {cpu_model()}; "
                          f"the oracle (C restatement of the reference's step), since the "
                          f"reference needs FFTW3, which is absent"}
    except Exception as e:  # the baseline is reported, never required
        return {"value": None, "unit": "Mpt/s", "cores": None, "kind": "port",
                "sample": f"unavailable: {e}"}


if __name__ == "__main__":
    sys.exit(main())
