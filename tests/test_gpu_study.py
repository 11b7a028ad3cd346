"""The reference's acceptance procedure on the device: the convergence study of
src/advection_error.py:21-188 (its four schemes, N = 16, 32, 64, Tf = 5, ic = 2, vf = 1).

Bar: at N = 16 and 32 each scheme's (L_inf, L_1, L_2) equals the oracle's within
2e-12 max|q| (a norm of the error moves by at most the state's max difference, which is
within 1e-12 relative); through N = 64 the errors fall at the schemes' orders: the RK2
schemes at 2 or more, the RK1 schemes at 0.9 or more (RK1 departure points are first
order in time at a fixed Courant number, and the oracle shows rates of 1.0-1.4 there).
"""
import numpy as np
import pytest
import torch

from helpers import run_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device visible")


class _Par:
    ic, vf = 2, 1


def test_convergence_study_matches_oracle_and_converges():
    from cubedsphere import advection
    from cubedsphere.cases import reference_case
    from oracle import timestep as ots
    Ns, err, rates = advection.error_analysis_adv(_Par, ntest=3)
    assert Ns == [16, 32, 64] and err.shape == (3, 4, 3) and rates.shape == (2, 4, 3)
    for d, (recon, opsplit, et, mt, dp, mf) in enumerate(advection.STUDY_SCHEMES):
        for i in range(2):
            dt = advection.STUDY_DT0[1] * 0.5 ** i
            n = int(5.0 / dt)
            grid, sim = run_oracle(Ns[i], reference_case(2, 1), dt, n, recon=recon, opsplit=opsplit, et=et,
                                   mt=mt, dp=dp, mf=mf)
            I = slice(grid.i0, grid.iend)
            qe = ots.exact_interior(grid, sim, n * dt)
            ref = np.array(ots.compute_errors(sim.Q[I, I, :], qe))
            assert np.all(np.abs(err[i, d] - ref) <= 2e-12 * np.abs(qe).max()), (advection.STUDY_NAMES[d], Ns[i])
        floor = 2.0 if dp == 2 else 0.9
        assert rates[:, d, :].min() >= floor, (advection.STUDY_NAMES[d], rates[:, d, :])
