"""Output side on the device: the lat-lon nearest-cell remap (src/interpolation.py:38-149,
src/cs_datastruct.py:569-582) and the reference's output files (src/output.py:16-150).

Bar: on the 720 x 1440 plotting grid the device's (i, j, panel) tables equal the oracle's at
every point whose equiangular coordinates lie more than 1e-9 cells from a cell edge and whose
largest Cartesian component leads the next by more than 1e-12 (closer than that, the device's
and numpy's atan / cos may round to either neighbour).  That excludes the two pole rows and the
lon = +-pi columns, whose points sit on a cell edge at even N (0.42% of the grid), and leaves
over 99%; the remapped values are the cell values at the device's indices, bitwise.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device visible")


def _solve(N, nsteps, dtype="f64"):
    from cubedsphere.cases import reference_case
    from cubedsphere.solver import Advection
    sol = Advection(N, reference_case(2, 1), 0.01, dtype=dtype, device="cuda:0", max_steps=nsteps).init_vars_adv()
    sol.step(nsteps)
    return sol


def test_latlon_remap_matches_oracle_tables_and_gathers_cell_values():
    from oracle import remap
    N = 24
    sol = _solve(N, 5)
    pre, _ = sol.remap_latlon(tables=True)       # with the mass-fixer total still pending
    st = sol.state()[0]                          # folds it: [6][N][N]
    q_ll, (i, j, p) = sol.remap_latlon(tables=True)
    ll = remap.LatLonGrid(720, 1440)
    oi, oj, op = remap.ll2cs(N, ll)
    c = np.sort(np.stack([abs(ll.X), abs(ll.Y), abs(ll.Z)], axis=-1), axis=-1)
    safe = (remap.edge_distance(N, ll) > 1e-9) & (c[..., 2] - c[..., 1] > 1e-12)
    assert safe.mean() > 0.99
    assert np.array_equal(i[safe], oi[safe].astype(np.int32))
    assert np.array_equal(j[safe], oj[safe].astype(np.int32))
    assert np.array_equal(p[safe], op[safe].astype(np.int32))
    assert np.array_equal(q_ll, st[p, j, i])
    assert np.allclose(pre, q_ll, rtol=0, atol=1e-14 * np.abs(st).max())
    qe = sol.exact(0)
    assert np.array_equal(sol.remap_latlon(field=qe), np.asarray(qe)[p, j, i])


def test_output_adv_writes_the_reference_files(tmp_path):
    from cubedsphere import advection
    from cubedsphere.grid import Grid
    g = Grid(16)
    sim = advection.adv_simulation_par(g, 0.025, 0.25, 2, 1, 1, 3, 1, 1, 3, 1, 3)
    linf, l1, l2 = advection.adv_sphere(g, sim, datadir=str(tmp_path))
    stem = tmp_path / ("gnomonic_equiangular_cs_16_" + advection.output_name(sim))
    err_ll = np.load(str(stem) + ".npy", allow_pickle=False)
    txt = np.loadtxt(str(stem))
    assert err_ll.shape == (advection.NLON, advection.NLAT)
    assert np.allclose(txt[:3], [linf, l1, l2], rtol=1e-15, atol=0)
    assert txt[3] == pytest.approx(sim.CFL) and 0 < sim.CFL < 1
    assert txt[5] == 0.025 and txt[6] == 0.25
    assert np.abs(err_ll).max() <= linf * (1 + 1e-12)
