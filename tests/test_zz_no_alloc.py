"""csadv_diagnostics_noalloc through the C ABI: the allocation-free variant of csadv_diagnostics,
whose block partials live in a fixed device array of the library.  Bar: its output is bitwise
csadv_diagnostics' (the same kernels on the same inputs) on two calls in a row, the second with
a smaller input; more tracers than the array holds are rejected with CSADV_E_RANGE.
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device visible")


def _run(N, ntr, nsteps):
    from cubedsphere.cases import Case
    from cubedsphere.solver import Advection
    rng = np.random.default_rng(N)
    tracers = [{"field": "gaussians", "centres": [(rng.uniform(-3, 3), rng.uniform(-1, 1))], "b": 8.0}
               for _ in range(ntr)]
    case = Case(tracers, ("solid_body", -0.7, 5.0), "noalloc")
    sol = Advection(N, case, 0.3 * (np.pi / 2 / N), device="cuda:0").init_vars_adv()
    sol.step(nsteps)
    return sol


def _call(sol, name, tracer0, ntracers, qexact, limit):
    from cubedsphere import _lib
    out = torch.full((ntracers, 4), -1.0, dtype=torch.float64, device="cuda:0")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    qe = ctypes.c_void_p(qexact.data_ptr()) if qexact is not None else None
    rc = getattr(sol.lib, name)(sol.handle, 0, tracer0, ntracers, qe, ctypes.c_double(limit),
                                ctypes.c_void_p(out.data_ptr()), stream)
    _lib.check(rc, sol.handle, name)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_diagnostics_equal_the_scratch_variant_on_two_calls():
    big = _run(64, 5, 6)
    qe = torch.from_numpy(np.stack([big.exact(n) for n in range(5)])).to("cuda:0")
    a = _call(big, "csadv_diagnostics_noalloc", 0, 5, qe, 0.0)
    b = _call(big, "csadv_diagnostics", 0, 5, qe, 0.0)
    assert np.array_equal(a, b) and np.all(np.isfinite(a)) and np.all(a[:, 0] > 0)
    small = _run(24, 2, 4)                     # the second call: a smaller input
    a2 = _call(small, "csadv_diagnostics_noalloc", 1, 1, None, 0.0)
    b2 = _call(small, "csadv_diagnostics", 1, 1, None, 0.0)
    assert np.array_equal(a2, b2)
    assert abs(a2[0, 1] - np.abs(small.state()[1]).max()) <= 1e-15 * a2[0, 1]


def test_noalloc_diagnostics_reject_more_tracers_than_the_array_holds():
    from cubedsphere import _lib
    sol = _run(16, 17, 1)
    out = torch.zeros((17, 4), dtype=torch.float64, device="cuda:0")
    rc = sol.lib.csadv_diagnostics_noalloc(sol.handle, 0, 0, 17, None, ctypes.c_double(0.0),
                                           ctypes.c_void_p(out.data_ptr()),
                                           ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == -4
    assert b"at most 16 tracers" in sol.lib.csadv_last_error(sol.handle)
    ok = _call(sol, "csadv_diagnostics_noalloc", 1, 16, None, 0.0)   # 16 is within the limit
    ref = _call(sol, "csadv_diagnostics", 1, 16, None, 0.0)
    assert np.array_equal(ok, ref)
    assert _lib.load() is sol.lib
