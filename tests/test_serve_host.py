"""Host logic of the served form on the CPU: the seeded request stream and the memory
estimate used for admission (no device calls)."""
import math

import numpy as np

from cubedsphere.serve import Request, Server, default_tile, random_requests


def test_request_stream_is_seeded_and_in_range():
    a = random_requests(3, 200, rate=50.0)
    b = random_requests(3, 200, rate=50.0)
    assert [(r.arrival, r.N, r.nsteps, r.dtype, r.case.ntr) for r in a] == \
           [(r.arrival, r.N, r.nsteps, r.dtype, r.case.ntr) for r in b]
    arr = np.array([r.arrival for r in a])
    assert np.all(np.diff(arr) > 0)
    gaps = np.diff(np.concatenate([[0.0], arr]))
    assert 0.5 / 50.0 < gaps.mean() < 2.0 / 50.0           # Poisson: mean gap 1/rate
    Ns = np.array([r.N for r in a])
    assert Ns.min() >= 48 and Ns.max() <= 1024
    # log-uniform: about half of the draws below the geometric mean of the range
    assert 0.3 < np.mean(Ns < math.sqrt(48 * 1024)) < 0.7
    assert all(1 <= r.case.ntr <= 4 and 50 <= r.nsteps <= 2000 for r in a)
    assert {r.dtype for r in a} == {"f32", "f64"}


def test_memory_estimate_scales_with_size_and_precision():
    r = random_requests(1, 1, rate=1.0)[0]
    small = Request(0, 0.0, 64, r.case, 100, "f32")
    big = Request(1, 0.0, 512, r.case, 100, "f32")
    dbl = Request(2, 0.0, 512, r.case, 100, "f64")
    eb = Server.estimate_bytes
    assert eb(small) < eb(big) < eb(dbl)
    M = 520
    state = 2 * r.case.ntr * 6 * M * M * 4
    assert eb(big) > state


def test_tile_rule_is_fixed_per_n():
    assert [default_tile(n) for n in (48, 95, 96, 255, 256, 1024)] == [1, 1, 0, 0, 2, 2]
