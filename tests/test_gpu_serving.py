"""The serving loop on the GPU: every request's result is the same alone and
in the mix (bitwise: each request's kernels read only its own state), the
admission keeps the resident footprint within the budget, and the latency
summary is well formed."""

import numpy as np
import pytest

from snapy_amd import serving

pytestmark = pytest.mark.gpu


def small_stream():
    reqs = serving.sample_requests(8, seed=5, rate=50.0, kmin=6, kmax=14, sizes="small")
    kinds = {r.kind for r in reqs}
    assert kinds == {"sod", "density_current", "swe"}, kinds
    return reqs


def test_same_alone_and_in_the_mix():
    reqs = small_stream()
    print("serving mix of", [(r.kind, r.params) for r in reqs], flush=True)
    mixed = {r.rid: r for r in serving.Server(realtime=False, max_slots=8).run(reqs)}
    assert len(mixed) == len(reqs)
    for req in reqs:
        alone = serving.Server(realtime=False, max_slots=1).run([req])[0]
        print("request", req.rid, req.kind, flush=True)
        assert alone.finite
        assert np.array_equal(alone.state, mixed[req.rid].state), req


def test_admission_respects_the_budget():
    reqs = small_stream()
    blocks = [serving.make_block(r) for r in reqs]
    need = [b.hbm_footprint() for b in blocks]
    del blocks
    budget = max(need) + min(need)  # room for the largest and one small one
    srv = serving.Server(budget_bytes=budget, realtime=False, max_slots=8)
    res = srv.run(reqs)
    assert len(res) == len(reqs)
    assert all(used <= budget for _, _, _, used in srv.log)
    summary = serving.summarize(res, max(r.done for r in res))
    print(summary, flush=True)
    assert summary["all_finite"]
    assert 0.0 <= summary["p50_latency_s"] <= summary["p99_latency_s"]


def test_tubes_batched_equal_their_own_launches():
    """The shock tubes of the stream as one work list give the same states,
    bit for bit, as each tube on its own captured stages."""
    reqs = [r for r in serving.sample_requests(24, seed=9, rate=50.0, kmin=5, kmax=12, sizes="small")
            if r.kind == "sod"][:5]
    assert len(reqs) >= 3
    print("tubes", [r.params["n"] for r in reqs], flush=True)
    a = {r.rid: r for r in serving.Server(realtime=False, max_slots=8, batch_tubes=True).run(reqs)}
    b = {r.rid: r for r in serving.Server(realtime=False, max_slots=8, batch_tubes=False).run(reqs)}
    for r in reqs:
        assert a[r.rid].finite
        assert np.array_equal(a[r.rid].state, b[r.rid].state), r


def test_preempted_requests_resume_bit_for_bit():
    """A long tube is parked (written to a host restart part, its memory
    freed) when a short request arrives that does not fit beside it; it
    resumes from the part later and ends with the same bits as without the
    preemption."""
    long_req = serving.Request(0, "sod", {"n": 2000}, 3, 5000, arrival=0.0)
    short_req = serving.Request(1, "sod", {"n": 600}, 4, 40, arrival=0.1)
    other = serving.Request(2, "density_current", {"nx": 64, "nz": 16, "dT": -10.0}, 5, 30, arrival=0.12)
    reqs = [long_req, short_req, other]
    need = [serving.make_block(r).hbm_footprint() for r in reqs]
    print("preemption footprints", need, flush=True)
    srv = serving.Server(budget_bytes=need[0] + max(need[1:]) - 1, realtime=True, max_slots=4, preempt=True)
    got = {r.rid: r for r in srv.run(reqs)}
    assert any(e[0] == "preempt" and e[1] == 0 for e in srv.log), srv.log
    ref = {r.rid: r for r in serving.Server(realtime=False, max_slots=4).run(reqs)}
    for r in reqs:
        print("request", r.rid, r.kind, flush=True)
        assert got[r.rid].finite
        assert np.array_equal(got[r.rid].state, ref[r.rid].state), r


def test_density_currents_grouped_equal_alone():
    """Density-current requests of one state form stepped as one group (one
    launch per kernel over requests of different widths and heights),
    joining and leaving at different rounds (four slots for six requests,
    different step counts), and each ends with the same bits as alone."""
    reqs = []
    for i, (prec, nx, nz, dT, steps) in enumerate(
        [("float64", 64, 16, -15.0, 9), ("float64", 96, 24, -8.0, 5), ("float32", 48, 16, -12.0, 7),
         ("float64", 32, 32, -19.0, 6), ("float32", 64, 24, -6.0, 8), ("float32", 96, 16, -10.0, 4)]
    ):
        reqs.append(serving.Request(i, "density_current", {"nx": nx, "nz": nz, "dT": dT, "precision": prec},
                                    100 + i, steps, 0.0))
    print("groups of", [(r.params["precision"], r.params["nx"], r.params["nz"], r.steps) for r in reqs], flush=True)
    srv = serving.Server(realtime=False, max_slots=4, batch_groups=True)
    a = {r.rid: r for r in srv.run(reqs)}
    built = [e for e in srv.log if e[0] == "group"]
    print("groups", built, flush=True)
    assert len(built) >= 2
    b = {r.rid: r for r in serving.Server(realtime=False, max_slots=4, batch_groups=False).run(reqs)}
    for r in reqs:
        assert a[r.rid].finite
        assert np.array_equal(a[r.rid].state, b[r.rid].state), r
