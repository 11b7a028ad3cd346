"""Write golden vectors for the two reference functions that no closed form pins bit-for-bit.

Run in the development container only (the reference tree is not present anywhere
else):  ``python tests/golden/make_golden.py``.

Imported from the reference (read-only, never copied):
  src/lagrange.py:28-163     lagrange_poly_ghostcell_pc  -- stencil clamping + weights
  src/halo_data.py:15-400    get_halo_data_interpolation{,_NS,_WE} -- panel orientation
Both modules import without stand-ins (numpy, matplotlib and the reference's own
constants / cs_transform / sphgeo only).  Inputs (grid centre coordinates, random
fields) come from the oracle and are stored in the fixture next to the outputs.
"""
import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_SRC = "/root/reference/src"


def main():
    os.environ.setdefault("MPLBACKEND", "Agg")
    sys.path.insert(0, REPO)
    from oracle.grid import CubedSphere

    sys.path.insert(0, REF_SRC)
    import halo_data as ref_halo  # noqa: E402
    import lagrange as ref_lagrange  # noqa: E402

    out = {}
    rng = np.random.default_rng(20240501)
    for N in (8, 16, 48):
        grid = CubedSphere(N)
        cs = types.SimpleNamespace(
            N=N, ng=grid.ng, ngl=grid.ngl, ngr=grid.ngr, i0=grid.i0, iend=grid.iend, j0=grid.j0,
            jend=grid.jend, projection=grid.projection, dy=grid.dy,
            pc=types.SimpleNamespace(X=grid.pc.X, Y=grid.pc.Y, Z=grid.pc.Z))
        sim = types.SimpleNamespace(degree=3)
        ref_lagrange.lagrange_poly_ghostcell_pc(cs, sim)
        w = sim.lagrange_poly_ghost_pc
        kmin, kmax = sim.stencil_ghost_pc
        if N <= 16:
            out["N%d_pcX" % N], out["N%d_pcY" % N], out["N%d_pcZ" % N] = grid.pc.X, grid.pc.Y, grid.pc.Z
        for s, side in enumerate(("east", "west", "north", "south")):
            out["N%d_w_%s" % (N, side)] = np.ascontiguousarray(w[s])
            out["N%d_kmin_%s" % (N, side)] = np.ascontiguousarray(kmin[s])
            out["N%d_kmax_%s" % (N, side)] = np.ascontiguousarray(kmax[s])
        if N <= 16:
            M = N + grid.ng
            Q = rng.standard_normal((M, M, 6))
            Qx = rng.standard_normal((M, M, 6))
            Qy = rng.standard_normal((M, M, 6))
            out["N%d_Q" % N], out["N%d_Qx" % N], out["N%d_Qy" % N] = Q, Qx, Qy
            he, hw, hn, hs = ref_halo.get_halo_data_interpolation(Q, cs)
            out["N%d_halo_east" % N], out["N%d_halo_west" % N] = he, hw
            out["N%d_halo_north" % N], out["N%d_halo_south" % N] = hn, hs
            hn, hs = ref_halo.get_halo_data_interpolation_NS(Qx, Qy, cs)
            out["N%d_haloNS_north" % N], out["N%d_haloNS_south" % N] = hn, hs
            he, hw = ref_halo.get_halo_data_interpolation_WE(Qx, Qy, cs)
            out["N%d_haloWE_east" % N], out["N%d_haloWE_west" % N] = he, hw
    path = os.path.join(HERE, "ref_lagrange_halo.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
