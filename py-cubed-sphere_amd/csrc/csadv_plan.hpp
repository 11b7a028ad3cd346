// Host-side planning of libcsadv: descriptor validation, tile shapes, the XCD-banded work list
// of the step kernel and the general step's work-buffer layout.  Plain C++17 with no HIP
// dependency, so that it builds with g++ under the address and undefined-behaviour sanitizers
// on its own; the kernels share its layout structs and constants.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "csadv.h"

#if defined(__HIPCC__) || defined(__HIP__)
#define CSADV_HD __host__ __device__
#else
#define CSADV_HD
#endif

namespace csadv {

constexpr int NG = 4;                // ghost layers per side (src/cs_datastruct.py:225-231)
constexpr int BLOCK = 256;           // threads per block of the whole-array kernels (4 waves)
constexpr int WAVE = 64;             // CDNA4 wavefront
constexpr int kXcds = 8;             // MI355X: 8 XCDs; workgroup b is dispatched to XCD b % 8
constexpr int STRIP_MAX = WAVE - 6;  // interior columns of one wavefront (3-cell halo each side)
constexpr int kDiagBlocks = 256;     // blocks of a diagnostics reduction (partials per tracer)
constexpr int kDiagWords = 8;        // doubles per diagnostics partial
constexpr int kMaxTracers = 64;      // tracers per solve

// tile id -> rows per wavefront segment (the strip width is always the widest that fits)
constexpr int kSegRows[] = {8, 4, 16, 32, 6, 12, 11, 5, 10, 14, 17, 23};
constexpr int kNumTiles = sizeof(kSegRows) / sizeof(kSegRows[0]);

// Device copy of one solve: the public descriptor plus derived strides.
struct Solve {
  csadv_solve_desc d;
  int32_t M;
  int32_t strip_w, seg_l;            // one wavefront: strip_w columns x seg_l rows of a panel
  int32_t strips, segs, items;       // strips and segments per panel, items = strips * segs
  int64_t pstride, tstride;          // cell arrays: panel / tracer strides
  int64_t pstride_u, pstride_v;      // x-edge / y-edge arrays: panel strides
  int32_t part_per_tracer;           // partial sums per tracer (= 6 * items)
  int32_t pad;
};

// One work item of the step kernel: (solve, panel + 8 * tracer, strip, segment); solve -1 pads.
// Same layout as HIP's int4.
struct Item {
  int32_t x, y, z, w;
};

// The general step's work buffer of one solve (dtype T), used by one tracer at a time:
// element offsets of the first pass's Qx and Qy and, with MF-PR, the second pass's div (cell
// arrays [6][M][ld]), the second pass's cube-edge fluxes with MF-AF ([6][4][N]), then a tail
// of n_red doubles: the MF-PR partial sums of the wavefronts, and their total in the last one.
struct WorkLayout {
  int64_t cell;                          // elements of one cell array
  int64_t o_qx, o_qy, o_div, o_side;     // element offsets
  int64_t o_red, n_red;                  // offset (in T units, 8-aligned) and doubles of the tail
  int64_t total;                         // elements of T before the tail (= o_red)
};

// Column strips of a panel of N columns: the fewest balanced strips (ceil(N / strips) columns,
// the last one the rest) of at most STRIP_MAX columns whose last strip keeps at least
// kMinLastStrip columns (a stencil beside the panel's east side reaches five columns back).
constexpr int kMinLastStrip = 4;
CSADV_HD inline int strip_count(int N) {
  int strips = (N + STRIP_MAX - 1) / STRIP_MAX;
  while (N - (strips - 1) * ((N + strips - 1) / strips) < kMinLastStrip && (N + strips - 1) / strips > 1) ++strips;
  return strips;
}

// Most wavefronts of one tracer's pass at size N over every tile shape (the shortest segment).
CSADV_HD inline int64_t general_max_items(int N) {
  int rows = kSegRows[0];
  for (int k = 1; k < kNumTiles; ++k) rows = kSegRows[k] < rows ? kSegRows[k] : rows;
  const int64_t strips = strip_count(N), segs = (N + rows - 1) / rows;
  return 6 * strips * segs;
}

CSADV_HD inline WorkLayout work_layout(int N, int ld, int ldu) {
  (void)ldu;
  WorkLayout w{};
  const int M = N + 2 * NG;
  w.cell = (int64_t)6 * M * ld;
  int64_t o = 0;
  w.o_qx = o; o += w.cell;
  w.o_qy = o; o += w.cell;
  w.o_div = o; o += w.cell;
  w.o_side = o; o += (int64_t)6 * 4 * N;
  o = (o + 7) / 8 * 8;
  w.o_red = o;
  w.n_red = general_max_items(N) + 1;
  w.total = o;
  return w;
}

// Row strides the general step expects (and csadv_work_bytes sizes its buffer for).
inline int general_ld(int N) { return (N + 2 * NG + 3) / 4 * 4; }
inline int general_ldu(int N) { return (N + 2 * NG + 4) / 4 * 4; }

// The options of one solve that need the general step: ET-S72 and ET-PL07 (ET-DG runs the
// fused step, with MF-AF through its cube-edge fix-up).
inline bool is_general(const csadv_solve_desc &d) { return d.et != 3; }
// Bytes of one tracer's Qx and Qy in the general step's work buffer (two cell arrays [6][M][ld]):
// the passes address them from one base with 32-bit byte offsets.
inline int64_t general_pair_bytes(int N, int dtype) {
  return int64_t(2) * 6 * (N + 2 * NG) * general_ld(N) * (dtype == CSADV_F64 ? 8 : 4);
}

// Elements (of the state's dtype) of the work buffer the fused step needs: MF-AF keeps the
// second pass's cube-edge fluxes, [ntr][6][4][N]; none otherwise.
inline int64_t fused_work_elems(const csadv_solve_desc &d) {
  return d.mf == 2 ? (int64_t)d.ntr * 6 * 4 * d.N : 0;
}

// Descriptor checks of csadv_create / csadv_update: CSADV_OK, or an error code and message.
int validate(const csadv_solve_desc &d, std::string &msg);

// Solves of one handle share dtype, recon, split, dp and the step kind.
int check_compatible(const csadv_solve_desc *solves, int nsolve, std::string &msg);

// Strips (columns) and segments (rows) per panel of tile `tile` at size N.
void tiling(int N, int tile, int &strip_w, int &strips, int &seg_l, int &segs);

// The descriptor with its derived strides; tile = the handle's tile when d.tile < 0.
Solve derive(const csadv_solve_desc &d, int tile);

// The step kernel's work list: every (solve, panel, tracer group, strip, segment) once, the
// row segments of each solve cut into kXcds contiguous bands, one per XCD, and the bands
// interleaved so that workgroup b (on XCD b % kXcds) takes its item from band b % kXcds
// (padded with solve -1 to equal length).  group: tracers per item (the FP32 multi-solve
// step kernel marches two).
std::vector<Item> work_list(const std::vector<Solve> &solves, int group = 1);

// The per-step ghost fill's work list of a multi-solve handle: (solve, tracer, cell block) for
// every block of BLOCK ghost cells of every tracer, each tracer's blocks followed by its
// correction-total block (cell block -1, MF-PR only), tracer layer by tracer layer (all solves'
// blocks of tracer 0 first) as in a dense (blocks, solves, tracers) grid, without that grid's
// blocks past a solve's cells or tracers.
std::vector<Item> ghost_list(const std::vector<Solve> &solves);

// Sizes of the public entry points (see include/csadv.h).
int partials_per_tracer(int N, int tile);
int work_bytes(int N, int dtype, int64_t *bytes);
int tile_rows(int tile);

}  // namespace csadv
