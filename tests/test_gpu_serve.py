"""Served form on the device: continuous batching of a Poisson request stream.

Bar: every request's result is bitwise the one it has when served alone (the mix changes
neither its tile nor its sums), results match a plain single-solve run, and the latency
statistics are consistent.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device visible")


def _stream():
    from cubedsphere.serve import random_requests
    # high arrival rate: requests overlap, join mid-run and leave at their own step counts
    return random_requests(11, 8, rate=400.0, n_range=(20, 72), steps=(10, 70), f64_fraction=0.25)


def test_served_results_equal_alone():
    from cubedsphere.serve import Server, serve_alone
    reqs = _stream()
    srv = Server(chunk=8, max_slots=5)
    done = srv.run(reqs)
    assert sorted(done) == [r.rid for r in reqs]
    assert srv.stats["peak_slots"] > 1                     # requests did share rounds
    for r in reqs:
        alone = serve_alone(r, chunk=8)
        assert np.array_equal(done[r.rid].result, alone), r.rid


def test_served_result_matches_plain_solve():
    from cubedsphere.batch import make_member
    from cubedsphere.serve import Server, default_tile
    reqs = _stream()
    done = Server(chunk=8).run(reqs)
    for r in reqs[:3]:
        sol = make_member(r.N, r.case, r.nsteps, dtype=r.dtype, tile=default_tile(r.N), device_grid=True)
        sol.init_vars_adv()
        sol.step(r.nsteps)
        assert np.array_equal(done[r.rid].result, sol.state()), r.rid


def test_served_stats_are_consistent():
    from cubedsphere.serve import Server
    reqs = _stream()
    srv = Server(chunk=8, budget_bytes=64 << 20)           # a small budget forces queueing
    done = srv.run(reqs)
    st = srv.stats
    assert st["requests"] == len(reqs)
    assert 0 < st["p50_latency_s"] <= st["p99_latency_s"]
    assert all(c.admitted >= c.arrival - 1e-9 and c.done > c.admitted for c in done.values())
    assert st["peak_bytes"] <= (64 << 20) or st["peak_slots"] == 1
    assert st["cell_updates"] == sum(6.0 * r.N ** 2 * r.case.ntr * r.nsteps for r in reqs)
