"""The serving loop: each request's result is the same alone and in a mix of
concurrent requests (bit for bit: the contexts share no device memory and
every kernel is deterministic), and admission respects the HBM budget."""
import numpy as np
import pytest

from sparc_amd import abi
from sparc_amd import serve as SV


def _requests():
    # sources inside the 25 Mm box of the tiny grid
    return SV.c5_requests(5, rate_per_s=1e6, mix=(("tiny", 1.0),), steps_range=(90, 150),
                          extent_mm=6.0)


@pytest.mark.gpu
def test_results_same_alone_and_in_mix():
    reqs = _requests()
    cache = SV.ModelCache()
    alone = {r.rid: SV.solve_request(r, cache, "snapshots") for r in reqs}
    srv = SV.Server(budget_bytes=64 << 30, max_inflight=5, keep_kernels=True, cache=cache)
    res = srv.run(reqs)
    assert not srv.errors, srv.errors
    assert len(res) == len(reqs)
    assert srv.peak_inflight > 1
    for r in res:
        chi, tau, K = alone[r.rid]
        assert r.mode == "snapshots"
        assert r.misfit == chi
        assert np.array_equal(r.tau, tau)
        assert np.array_equal(r.K_lnc, K)
        assert np.linalg.norm(K) > 0


@pytest.mark.gpu
def test_checkpoint_mode_when_memory_is_short():
    """A budget below one request's snapshot bank but above its checkpoint
    footprint admits it in the checkpointed mode, with the same kernel up to
    the identical recomputation (bit-equal)."""
    r = _requests()[0]
    cache = SV.ModelCache()
    spec = SV.request_spec(r, cache)
    snap = SV.footprint(spec, "snapshots", 30)
    ck = SV.footprint(spec, "checkpoints", 30)
    assert ck < snap
    srv = SV.Server(budget_bytes=(ck + snap) // 2, max_inflight=2, checkpoint_every=30,
                    keep_kernels=True, cache=cache)
    res = srv.run([r])
    assert not srv.errors, srv.errors
    assert res[0].mode == "checkpoints"
    chi, tau, K = SV.solve_request(r, cache, "snapshots")
    assert res[0].misfit == chi
    assert np.array_equal(res[0].K_lnc, K)


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_full_grid_mix_same_alone_and_in_mix():
    """One request on each grid of the serving mix (128^2 x 150, 256^2 x 300,
    384^2 x 300), 200 steps each, all arriving at once: every result in the
    concurrent loop equals the same request run alone, bit for bit, and the
    reserved HBM never exceeds the admission budget (half of what the
    device reports free, so the 384^2 x 300 request has to wait for room
    or run beside the others within it)."""
    reqs = SV.mix_sample(200)
    assert sorted(r.workload for r in reqs) == ["box2", "box2l", "box2s"]
    cache = SV.ModelCache()
    alone = {r.rid: SV.solve_request(r, cache, "snapshots") for r in reqs}
    free, _ = SV.device_free_bytes()
    srv = SV.Server(budget_bytes=free // 2, max_inflight=3, keep_kernels=True, cache=cache)
    res = srv.run(reqs)
    assert not srv.errors, srv.errors
    assert len(res) == len(reqs)
    assert srv.peak_used <= srv.budget
    print(f"\n  peak in flight {srv.peak_inflight}, peak reserved "
          f"{srv.peak_used / 2**30:.2f} of {srv.budget / 2**30:.2f} GiB; modes "
          f"{sorted(r.mode for r in res)}")
    for r in res:
        chi, tau, K = alone[r.rid]
        assert r.misfit == chi
        assert np.array_equal(r.tau, tau)
        assert np.array_equal(r.K_lnc, K)
        assert np.linalg.norm(K) > 0
