"""Host-side parts of the serving loop (no device): the arrival stream and
the footprint model used for admission."""
import numpy as np

from sparc_amd import abi
from sparc_amd import serve as SV


def test_arrival_stream_is_deterministic_and_follows_the_mix():
    a = SV.c5_requests(400, rate_per_s=0.5)
    b = SV.c5_requests(400, rate_per_s=0.5)
    assert [(r.workload, r.nsteps, r.arrival, r.src_xy_mm, r.alpha) for r in a] == \
           [(r.workload, r.nsteps, r.arrival, r.src_xy_mm, r.alpha) for r in b]
    frac = {w: sum(r.workload == w for r in a) / len(a) for w, _ in SV.GRID_MIX}
    assert abs(frac["box2s"] - 0.4) < 0.08 and abs(frac["box2"] - 0.5) < 0.08
    assert abs(frac["box2l"] - 0.1) < 0.05
    assert all(1000 <= r.nsteps <= 3000 for r in a)
    assert {r.alpha for r in a} == set(SV.STEP_SIZES)
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert abs(gaps.mean() - 2.0) < 0.4  # Poisson at 0.5 per second


def test_footprints_order_and_scale():
    cache = SV.ModelCache()
    reqs = {w: SV.Request(0, w, (0.0, 0.0), 1.0, 2000) for w in ("box2s", "box2", "box2l")}
    snap = {w: SV.footprint(SV.request_spec(r, cache), "snapshots") for w, r in reqs.items()}
    ck = {w: SV.footprint(SV.request_spec(r, cache), "checkpoints") for w, r in reqs.items()}
    assert snap["box2s"] < snap["box2"] < snap["box2l"]
    for w in reqs:
        assert ck[w] < snap[w]
    # 2000 steps at the 30 s cadence of the 256^2 x 300 box: 47 FP32 slots of 6 fields
    N = 256 * 256 * 300
    assert snap["box2"] - 2 * abi.footprint(_cfg(256, 256, 300)) == 47 * 6 * N * 4


def _cfg(nx, ny, nz):
    c = abi.SparcConfig()
    c.abi_version = abi.ABI_VERSION
    c.nx, c.ny, c.nz = nx, ny, nz
    c.option = abi.OPT_3D
    c.precision = abi.FP32
    c.npmlbot = c.npmltop = 10
    c.damp_waves = 1
    c.freq_filtering = 86
    c.e_rad = c.o_rad = nz // 2
    c.compensated = 1
    c.deltat = 1.0
    c.field_scale = 1.0
    return c
