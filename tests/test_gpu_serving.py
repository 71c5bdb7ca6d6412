"""Serving form: continuous batching admits requests as memory frees up, every
request's result is bit-identical to the same request run alone, and large
requests run in lanes of their own without holding short ones back."""
import math

import numpy as np
import pytest

import lpmkokkos_amd as L
from lpmkokkos_amd import ensemble as E
from lpmkokkos_amd import serve as S

pytestmark = pytest.mark.gpu


def test_requests_match_alone_under_admission_pressure():
    reqs = S.make_requests(10, rate=200.0, seed=11, level_probs={2: 0.4, 3: 0.4, 4: 0.2}, steps=3)
    biggest = max(S.request_bytes(r.spec) for r in reqs)
    # room for about three requests at a time: later arrivals wait for retirements
    # level-4 requests take the two large lanes, the rest share the batched lane
    srv = S.Server(budget_bytes=3.2 * biggest, keep_results=True, large_level=4, large_lanes=2)
    res = srv.run(reqs)
    srv.close()
    assert res["requests"] == 10
    assert 0 < res["large_lane_requests"] < 10
    assert res["large_lanes_max_busy"] <= res["large_lanes"] == 2
    assert res["latency_p99_s"] >= res["latency_p50_s"] > 0
    assert max(r.admitted - r.arrival for r in reqs) > 0  # some request queued
    for r in reqs:
        c = E.make_member(r.spec)
        c.advance(0.005, 2 * math.pi, L.SCHEME_LAGR, L.PREC_FP32C, nsteps=r.steps)
        s = c.get_state()
        for k in ("vx", "vzeta", "vu", "fx", "fzeta", "fu"):
            assert np.array_equal(s[k], r.result[k]), (r.rid, k)
        c.close()


def test_budget_rejects_oversized_request():
    reqs = S.make_requests(2, rate=10.0, seed=1, level_probs={3: 1.0}, steps=1)
    srv = S.Server(budget_bytes=1e4)
    with pytest.raises(ValueError):
        srv.run(reqs)
    srv.close()


def test_large_requests_wait_for_a_lane():
    """Five large requests (level 4 with large_level 4) and one large lane:
    they run one after another on the lane's stream, short requests arriving
    meanwhile are admitted past the waiting ones, and every result equals the
    request run alone."""
    big = [S.Request(rid=k, arrival=0.0, spec=E.MemberSpec(level=4, vortex_seed=30 + k), steps=4) for k in range(5)]
    small = [S.Request(rid=5 + k, arrival=0.0, spec=E.MemberSpec(level=2, vortex_seed=40 + k), steps=4)
             for k in range(3)]
    srv = S.Server(keep_results=True, large_level=4, large_lanes=1)
    res = srv.run(big + small)
    srv.close()
    assert res["requests"] == 8 and res["large_lanes_max_busy"] == 1
    assert max(r.finished for r in small) < max(r.finished for r in big)
    for r in big + small:
        c = E.make_member(r.spec)
        c.advance(0.005, 2 * math.pi, L.SCHEME_LAGR, L.PREC_FP32C, nsteps=r.steps)
        s = c.get_state()
        for k in ("vx", "vzeta", "vu", "fx", "fzeta", "fu"):
            assert np.array_equal(s[k], r.result[k]), (r.rid, k)
        c.close()


def test_large_lane_does_not_hold_short_requests():
    """A level-7 request (about 3 s of device time) arrives first; level-2
    requests arriving while it runs finish in a small fraction of its time,
    because they step in the batched lane while the large one runs in its own."""
    big = S.Request(rid=0, arrival=0.0, spec=E.MemberSpec(level=7, vortex_seed=5), steps=20)
    small = [S.Request(rid=1 + k, arrival=0.2 + 0.05 * k, spec=E.MemberSpec(level=2, vortex_seed=10 + k), steps=20)
             for k in range(8)]
    srv = S.Server(large_level=7, batch_levels=(), large_lanes=1)  # two streams: distinct hardware queues
    res = srv.run([big] + small)
    srv.close()
    assert res["requests"] == 9 and res["large_lane_requests"] == 1
    worst_small = max(r.latency for r in small)
    print(f"large request {big.latency:.2f} s, slowest short request {worst_small:.3f} s")
    assert big.finished > max(r.finished for r in small)  # the short ones overtook it
    assert worst_small < 0.25 * big.latency


def test_suspend_and_resume_match_alone(tmp_path):
    """A server stopped after a few batched steps checkpoints its batched
    requests and queue; a second server resumes them.  Every request's result,
    whichever server finished it, equals the request run alone, bit for bit."""
    reqs = S.make_requests(8, rate=400.0, seed=21, level_probs={2: 0.4, 3: 0.4, 4: 0.2}, steps=6)
    a = S.Server(keep_results=True, large_level=4)
    res_a = a.run(reqs, stop_after_steps=3, checkpoint_dir=str(tmp_path))
    a.close()
    finished = {r.rid: r for r in reqs if r.result is not None}
    assert res_a["requests"] == len(finished) and res_a["suspended"] == 8 - len(finished) > 0
    resumed = S.load_manifest(str(tmp_path))
    assert any(r.checkpoint and 0 < r.steps_done < r.steps for r in resumed)
    b = S.Server(keep_results=True, large_level=4)
    res_b = b.run(resumed)
    b.close()
    assert res_b["requests"] == len(resumed) and res_b["suspended"] == 0
    finished.update({r.rid: r for r in resumed})
    assert sorted(finished) == list(range(8))
    for rid, r in finished.items():
        c = E.make_member(r.spec)
        c.advance(0.005, 2 * math.pi, L.SCHEME_LAGR, L.PREC_FP32C, nsteps=r.steps)
        s = c.get_state()
        for k in ("vx", "vzeta", "vu", "fx", "fzeta", "fu"):
            assert np.array_equal(s[k], r.result[k]), (rid, k)
        c.close()


def test_size_class_lanes_step_independently():
    """Batched lanes by size class: level-2 requests in the (0..5) lane finish
    while a level-6 request in the (6..6) lane is still stepping, and every
    result equals the request run alone."""
    mid = S.Request(rid=0, arrival=0.0, spec=E.MemberSpec(level=6, vortex_seed=50), steps=6)
    small = [S.Request(rid=1 + k, arrival=0.0, spec=E.MemberSpec(level=2, vortex_seed=60 + k), steps=6)
             for k in range(4)]
    srv = S.Server(keep_results=True, large_level=7, batch_levels=(5,), large_lanes=0)
    res = srv.run([mid] + small)
    srv.close()
    assert res["batch_lanes"] == [[0, 5], [6, 6]] and res["requests"] == 5
    assert max(r.finished for r in small) < mid.finished
    for r in [mid] + small:
        c = E.make_member(r.spec)
        c.advance(0.005, 2 * math.pi, L.SCHEME_LAGR, L.PREC_FP32C, nsteps=r.steps)
        s = c.get_state()
        for k in ("vx", "vzeta", "vu", "fx", "fzeta", "fu"):
            assert np.array_equal(s[k], r.result[k]), (r.rid, k)
        c.close()


@pytest.mark.timeout(300)
def test_level8_request_in_a_large_lane_beside_batched_requests():
    """The serving configuration's largest size: a level-8 request (1.31M
    sources) runs in a large lane while requests of levels 4-7 are batched in
    their size-class lanes; every result, the level-8 one included, equals
    the request run alone, bit for bit."""
    reqs = [S.Request(rid=0, arrival=0.0, spec=E.MemberSpec(level=8, rh_n=3, nvort=6, vortex_seed=80), steps=2)]
    reqs += [S.Request(rid=1 + k, arrival=0.05 * k, spec=E.MemberSpec(level=lev, rh_n=2 + k % 5, nvort=4 + k,
                                                                     vortex_seed=90 + k), steps=3)
             for k, lev in enumerate([4, 5, 6, 7, 4, 5])]
    srv = S.Server(keep_results=True)  # default lanes: levels <= 5, 6, 7 batched; level 8 large
    res = srv.run(reqs)
    srv.close()
    assert res["requests"] == 7 and res["large_lane_requests"] == 1
    assert res["levels"] == {4: 2, 5: 2, 6: 1, 7: 1, 8: 1}
    for r in reqs:
        c = E.make_member(r.spec)
        c.advance(0.005, 2 * math.pi, L.SCHEME_LAGR, L.PREC_FP32C, nsteps=r.steps)
        s = c.get_state()
        for k in ("vx", "vzeta", "vu", "fx", "fzeta", "fu"):
            assert np.array_equal(s[k], r.result[k]), (r.rid, k)
        c.close()


def test_large_request_without_large_lanes_is_refused():
    reqs = [S.Request(rid=0, arrival=0.0, spec=E.MemberSpec(level=4, vortex_seed=1), steps=1)]
    srv = S.Server(large_level=4, large_lanes=0)
    with pytest.raises(ValueError):
        srv.run(reqs)
    srv.close()


def test_mixed_seed_requests_match_alone():
    """Cubed-sphere and icosahedral requests share the size-class lanes: each
    is sized from its own seed (memory charge, pair count) and its result equals
    the request run alone, bit for bit."""
    cs = L.SEED_CUBED_SPHERE
    specs = [E.MemberSpec(level=3, seed=cs, rh_n=3, vortex_seed=70), E.MemberSpec(level=3, vortex_seed=71),
             E.MemberSpec(level=4, seed=cs, nvort=5, vortex_seed=72), E.MemberSpec(level=2, vortex_seed=73),
             E.MemberSpec(level=4, vortex_seed=74)]
    reqs = [S.Request(rid=k, arrival=0.01 * k, spec=s, steps=4) for k, s in enumerate(specs)]
    srv = S.Server(keep_results=True, large_level=5, batch_levels=(3,))
    res = srv.run(reqs)
    srv.close()
    assert res["requests"] == 5 and res["batch_lanes"] == [[0, 3], [4, 4]] and res["large_lane_requests"] == 0
    assert res["pairs"] == sum(4 * S.pairs_per_step(s.level, s.seed) for s in specs)
    for r in reqs:
        c = E.make_member(r.spec)
        c.advance(0.005, 2 * math.pi, L.SCHEME_LAGR, L.PREC_FP32C, nsteps=r.steps)
        s = c.get_state()
        assert s["vx"].shape[0] == S.tree_sizes(r.spec.level, r.spec.seed)[0], r.rid
        for k in ("vx", "vzeta", "vu", "fx", "fzeta", "fu"):
            assert np.array_equal(s[k], r.result[k]), (r.rid, k)
        c.close()
