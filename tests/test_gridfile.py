"""Grid files in the reference's layout (cubedsphere.gridfile), on the CPU.

The layout is the one src/plot.py:520-669 writes and src/cs_datastruct.py:44-190 reads: the
variables, their dimensions and types are checked name by name, and the values against the
oracle's grid (src/cs_transform.py maps and tangents restated in oracle/grid.py): positions
bitwise, lon on the circle, tangent vectors to rounding.
"""
import numpy as np
import pytest
from scipy.io import netcdf_file

from cubedsphere import gridfile
from cubedsphere.grid import Grid
from oracle import grid as og

# src/plot.py:568-585: variable -> dimensions
FAMILY_DIMS = {"po": ("ix", "jy"), "pc": ("ix2", "jy2"), "pu": ("ix", "jy2"), "pv": ("ix2", "jy")}
SCALARS_F8 = ("dx", "dy", "a")
SCALARS_I4 = ("ng", "ngl", "ngr", "N", "i0", "iend", "j0", "jend")


def _oracle_family(N, fam):
    g = og.CubedSphere(N)
    x = {"po": (g.x_pu, g.x_pu), "pc": (g.x_pc, g.y_pc), "pu": (g.x_pu, g.y_pc), "pv": (g.x_pc, g.y_pv)}[fam]
    pos = og.equiangular_gnomonic_map(*x)
    return pos, og.equiangular_tg_xdir(*x), og.equiangular_tg_ydir(*x)


@pytest.mark.parametrize("N", [8, 17])
def test_grid_file_has_the_reference_layout_and_values(N, tmp_path):
    path = str(tmp_path / (gridfile.grid_name(N) + ".nc"))
    gridfile.save(Grid(N), path)
    n = N + 8
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.title.decode() == "gnomonic_equiangular_cs_%d" % N
        assert {k: v for k, v in f.dimensions.items()} == {
            "panel": 6, "ix": n + 1, "jy": n + 1, "ix2": n, "jy2": n, "coorddim": 5, "r3dim": 3, "ed": 4}
        v = f.variables
        for name in SCALARS_F8:
            assert v[name].dimensions == () and v[name].typecode() == "d"
        for name in SCALARS_I4:
            assert v[name].dimensions == () and v[name].typecode() == "i"
        assert int(v["N"].getValue()) == N and int(v["ng"].getValue()) == 8
        assert (int(v["i0"].getValue()), int(v["iend"].getValue())) == (4, N + 4)
        assert float(v["dx"].getValue()) == float(v["dy"].getValue()) == (np.pi / 4 - (-np.pi / 4)) / N
        assert float(v["a"].getValue()) == np.pi / 4
        for fam, dims in FAMILY_DIMS.items():
            pos, tx, ty = _oracle_family(N, fam)
            for name, ref in ((fam, pos), ("ex_" + fam, tx), ("ey_" + fam, ty)):
                var = v[name]
                assert var.dimensions == dims + ("panel", "coorddim") and var.typecode() == "d"
                got = np.asarray(var[:])
                if name == fam:
                    # X, Y, Z and lat bitwise; lon bitwise up to the branch cut at +-pi
                    for k in (0, 1, 2, 4):
                        assert np.array_equal(got[..., k], ref[k]), (name, k)
                    dlon = np.abs(got[..., 3] - ref[3])
                    assert np.all((dlon == 0) | (np.abs(dlon - 2 * np.pi) < 1e-15)), name
                else:
                    scale = max(np.abs(r).max() for r in ref)
                    for k in range(3):
                        assert np.abs(got[..., k] - ref[k]).max() <= 1e-14 * scale, (name, k)
                    # the reference's tangent ``point`` objects keep lon = lat = 0
                    assert not got[..., 3:].any(), name


def test_load_round_trip_and_rejects_other_grids(tmp_path):
    N = 12
    path = str(tmp_path / "g.nc")
    gridfile.save(Grid(N), path)
    g = gridfile.load(path)
    assert g.N == N and g.i0 == 4 and g.iend == N + 4
    # a file whose centres are not the equiangular grid's (an equidistant grid's, N=12)
    bad = str(tmp_path / "bad.nc")
    gridfile.save(Grid(N), bad)
    with netcdf_file(bad, "a", mmap=False) as f:
        ge = og.CubedSphere(N, projection="gnomonic_equidistant")
        pc = np.asarray(f.variables["pc"][:])
        pc[..., 0], pc[..., 1], pc[..., 2] = ge.pc.X, ge.pc.Y, ge.pc.Z
        f.variables["pc"][:] = pc
    with pytest.raises(ValueError, match="pc is not the equiangular grid"):
        gridfile.load(bad)
    # another N in the header than in the arrays
    other = str(tmp_path / "other.nc")
    gridfile.save(Grid(N), other)
    with netcdf_file(other, "a", mmap=False) as f:
        f.variables["N"][()] = N + 1
    with pytest.raises(ValueError):
        gridfile.load(other)


def test_gridload_reads_the_file_when_present_and_generates_otherwise(tmp_path, capsys):
    """cubed_sphere(N, gridload) as src/cs_datastruct.py:45-60: a present file is read (and checked),
    a missing one is reported and the grid generated."""
    d = str(tmp_path)
    g = gridfile.cubed_sphere(16, gridload=True, griddir=d)
    assert "Couldnt find the grid file" in capsys.readouterr().out and g.N == 16
    gridfile.save(g, str(tmp_path / (gridfile.grid_name(16) + ".nc")))
    g2 = gridfile.cubed_sphere(16, gridload=True, griddir=d, verbose=True)
    assert "Loading grid file" in capsys.readouterr().out and g2.N == 16
    assert np.array_equal(g2.pc.X, g.pc.X)
    assert gridfile.cubed_sphere(16, gridload=False, griddir=d).N == 16
