"""Serving form: continuous batching admits requests as memory frees up, and
every request's result is bit-identical to the same request run alone."""
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
    srv = S.Server(budget_bytes=3.2 * biggest, keep_results=True)
    res = srv.run(reqs)
    srv.close()
    assert res["requests"] == 10
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
