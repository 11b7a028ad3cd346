"""The general step's ghost reads, on the CPU: a model of the index arithmetic of the two passes.

The passes of the general step (ET-S72, ET-PL07) store no ghost cell: a lane in a ghost column
(interior rows) or a ghost row (interior columns) reads the neighbour panel's cell directly, at
element offset a + y * b with one (a, b) per row regime, built from the halo orientation table
(``lane_lines`` in csrc/csadv_general.hpp, restated here line by line).  This checks, for every
position a pass reads, every work item of every tile shape and N = 8, 9, 17, 24, 61 (the
smallest N, odd N, two column strips), that:

* every load and store is in bounds of its array;
* the value read is the one the oracle's copy halo (oracle/interp.py:ghost_cells_adjacent_panels,
  src/interpolation.py:320-340) leaves at that position -- Q in pass 1; in pass 2 the copy between
  the passes, Qy's E/W ghosts and Qx's N/S ghosts (src/discrete_operators.py:76-78);
* the last column strip keeps the four columns the east-side stencils need.
"""
import itertools

import numpy as np
import pytest

from cubedsphere.topology import orientation
from oracle import timestep as ots
from oracle.interp import ghost_cells_adjacent_panels

NG, WAVE, STRIP_MAX = 4, 64, 58
SEG_ROWS = [8, 4, 16, 32, 6, 12, 11, 5, 10, 14, 17, 23]      # tile id -> rows (csrc/csadv_plan.hpp)


def tiling(N, tile):
    """csadv_plan.cpp:tiling with strip_count."""
    strips = (N + STRIP_MAX - 1) // STRIP_MAX
    while N - (strips - 1) * ((N + strips - 1) // strips) < 4 and (N + strips - 1) // strips > 1:
        strips += 1
    sw = (N + strips - 1) // strips
    sl = min(SEG_ROWS[tile], N)
    return sw, strips, sl, (N + sl - 1) // sl


def lane_lines(O, ld, ps, N, p, x, own, cols, ew, rows, ns):
    """csadv_general.hpp:lane_lines: (a, b) per row regime (south ghost, interior, north ghost)."""
    i0, iend = NG, N + NG
    lines = [None, None, None]
    if i0 <= x < iend:
        lines[1] = (own + p * ps + x, ld)
        if rows:
            os_, on = O[p, 3], O[p, 2]
            bs = ns + os_[0] * ps + (os_[4] + os_[6] * x) * ld + os_[1] + os_[3] * x
            ss = os_[5] * ld + os_[2]
            bn = ns + on[0] * ps + (on[4] + on[6] * x) * ld + on[1] + on[3] * x
            sn = on[5] * ld + on[2]
            lines[0] = (bs + (i0 - 1) * ss, -ss)
            lines[2] = (bn - iend * sn, sn)
    elif cols and i0 - 3 <= x < iend + 3:
        west = x < i0
        d = i0 - 1 - x if west else x - iend
        o = O[p, 1 if west else 0]
        lines[1] = (ew + o[0] * ps + (o[4] + o[5] * d) * ld + o[1] + o[2] * d, o[6] * ld + o[3])
    return lines


@pytest.mark.parametrize("N", [8, 9, 17, 24, 61])
def test_ghost_reads_are_the_copy_halo_and_in_bounds(N):
    M = N + 2 * NG
    ld, ldu = (N + 2 * NG + 3) // 4 * 4, (N + 2 * NG + 4) // 4 * 4
    ps = M * ld
    cell = 6 * ps
    O = orientation(N)
    grid, _ = ots.make(N, {"ic": 2, "vf": 1}, 0.01)
    # fields of element offsets in the oracle's layout [i, j, p]: Qx at 0, Qy at + cell
    ii, jj, pp = np.meshgrid(np.arange(M), np.arange(M), np.arange(6), indexing="ij")
    offs = pp * ps + jj * ld + ii
    Qx, Qy = offs.astype(float), (offs + cell).astype(float)
    ghost_cells_adjacent_panels(Qx, Qy, grid)                # between the passes
    Q = offs.astype(float)
    ghost_cells_adjacent_panels(Q, Q, grid)                  # the step's halo of Q
    i0, iend = NG, N + NG
    reads = 0
    for tile in range(len(SEG_ROWS)):
        sw, strips, sl, segs = tiling(N, tile)
        assert N - sw * (strips - 1) >= min(N, 4)
        for p, sx, sg in itertools.product(range(6), range(strips), range(segs)):
            ti, tj = NG + sx * sw, NG + sg * sl
            nx, ny = min(sw, iend - ti), min(sl, iend - tj)
            ylast = tj + ny + 3
            for lane in range(WAVE):
                x = ti - 3 + lane
                pass1 = lane_lines(O, ld, ps, N, p, x, 0, True, 0, True, 0)
                qy = lane_lines(O, ld, ps, N, p, x, cell, True, 0 if p in (4, 5) else cell, False, 0)
                qx = lane_lines(O, ld, ps, N, p, x, 0, False, 0, True, cell if p in (1, 3) else 0)
                xc, xe = min(x, M - 1), min(x, M)
                for y in range(tj - 3, ylast + 1):
                    reg = 0 if y < i0 else (1 if y < iend else 2)
                    # step-invariant tables and winds ([M][ld], [M][ldu], y-edges [M+1][ld])
                    assert 0 <= y * ld + xc < M * ld and 0 <= y * ldu + xe < M * ldu
                    assert 0 <= (max(y - 3, 0) + 1) * ld + xc < (M + 1) * ld
                    for lines, limit, ref in ((pass1, cell, Q), (qy, 2 * cell, Qy), (qx, 2 * cell, Qx)):
                        ab = lines[reg]
                        if ab is None:
                            continue
                        o = ab[0] + y * ab[1]
                        assert 0 <= o < limit, (N, p, x, y)
                        if x < M:
                            assert o == ref[x, y, p], (N, p, x, y, o, ref[x, y, p])
                            reads += 1
                    if 3 <= lane < 3 + nx and tj <= y < tj + ny:
                        assert 0 <= cell + p * ps + x + y * ld < 2 * cell        # Qx, Qy stores
    assert reads > 0
