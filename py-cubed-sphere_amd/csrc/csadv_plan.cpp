// Host-side planning of libcsadv (no HIP): descriptor validation, tile shapes, the XCD-banded
// work list and the general step's buffer sizes.  See csadv_plan.hpp.
#include "csadv_plan.hpp"

#include <algorithm>

namespace csadv {

int validate(const csadv_solve_desc &d, std::string &msg) {
  if (d.N < 8 || d.N > 8192) { msg = "N must be in [8, 8192]"; return CSADV_E_RANGE; }
  if (d.ntr < 1 || d.ntr > kMaxTracers) { msg = "ntr must be in [1, 64]"; return CSADV_E_RANGE; }
  if (d.dtype != CSADV_F64 && d.dtype != CSADV_F32) { msg = "dtype"; return CSADV_E_ARG; }
  const int M = d.N + 2 * NG;
  if (d.ld < M || d.ldu < M + 1) { msg = "row strides too small"; return CSADV_E_ARG; }
  if (d.recon < 1 || d.recon > 4) { msg = "recon must be 1..4"; return CSADV_E_ARG; }
  if (d.dp != 1 && d.dp != 2) { msg = "dp must be 1 or 2"; return CSADV_E_ARG; }
  if (!((d.split == 1 || d.split == 2) && d.mt == 1) && !(d.split == 3 && d.mt == 2)) {
    msg = "invalid (opsplit, mt) pair: valid pairs are (1|2, 1) and (3, 2)";
    return CSADV_E_ARG;
  }
  if (d.et < 1 || d.et > 3) { msg = "et must be 1 (ET-S72), 2 (ET-PL07) or 3 (ET-DG)"; return CSADV_E_ARG; }
  if (d.mf < 1 || d.mf > 3) { msg = "mf must be 1 (none), 2 (MF-AF) or 3 (MF-PR)"; return CSADV_E_ARG; }
  if (is_general(d) && !d.work) {
    msg = "ET-S72 and ET-PL07 run the general step, which needs desc.work (csadv_work_bytes)";
    return CSADV_E_ARG;
  }
  if (!is_general(d) && d.mf == 2 && !d.work) {
    msg = "ET-DG with MF-AF needs desc.work for the cube-edge fluxes (ntr * 24 * N elements of the dtype)";
    return CSADV_E_ARG;
  }
  if (is_general(d) && (d.ld != general_ld(d.N) || d.ldu != general_ldu(d.N))) {
    msg = "the general step expects ld = round_up(N+8, 4) and ldu = round_up(N+9, 4) (the csadv_work_bytes layout)";
    return CSADV_E_ARG;
  }
  if (is_general(d) && general_pair_bytes(d.N, d.dtype) >= (int64_t(1) << 32)) {
    msg = "the general step (ET-S72, ET-PL07) addresses a tracer's Qx and Qy with 32-bit byte offsets: "
          "N up to 6680 in FP64";
    return CSADV_E_RANGE;
  }
  if (d.nbasis < 1 || d.nbasis > 8) { msg = "nbasis must be 1..8"; return CSADV_E_RANGE; }
  if (d.unsteady && d.max_steps < 1) { msg = "unsteady winds need a coefficient table"; return CSADV_E_ARG; }
  if (!d.q0 || !d.q1 || !d.uavg || !d.vavg || !d.gc || !d.gu || !d.gv || !d.rgc || !d.gc64 || !d.lagw ||
      !d.lagk || !d.orient || !d.wbu || !d.wbv || !d.coef || !d.ctl) {
    msg = "null device pointer in descriptor";
    return CSADV_E_ARG;
  }
  if (d.dp == 2 && (!d.usgn || !d.vsgn)) { msg = "dp=2 needs the usgn and vsgn wind arrays"; return CSADV_E_ARG; }
  if (d.mf == 3 && (!d.corr || !d.partials)) { msg = "MF-PR needs corr and partials"; return CSADV_E_ARG; }
  if (!is_general(d) && !d.gfix) { msg = "the fused step (ET-DG) needs gfix"; return CSADV_E_ARG; }
  if (is_general(d) && d.offset) { msg = "state offsets apply to the fused step only (desc.offset must be NULL)"; return CSADV_E_ARG; }
  if (d.n_strip != 6 * 4 * 3 * d.N || d.n_corner != 6 * 2 * 3 * 6 || !d.strip_tgt || !d.strip_src || !d.strip_w ||
      !d.corner_tgt || !d.corner_src || !d.corner_w) {
    msg = "ghost recipes missing or of the wrong size (expected 72 N strip and 216 corner cells)";
    return CSADV_E_ARG;
  }
  if (((uintptr_t)d.strip_src | (uintptr_t)d.corner_src | (uintptr_t)d.strip_w | (uintptr_t)d.corner_w) % 16 != 0) {
    msg = "strip_src, corner_src, strip_w and corner_w must be 16-byte aligned";
    return CSADV_E_ARG;
  }
  if (d.stop < 0) { msg = "stop must be >= 0"; return CSADV_E_RANGE; }
  if (d.tile < -1 || d.tile >= kNumTiles) { msg = "tile must be -1 or a valid tile id"; return CSADV_E_RANGE; }
  if (!(d.dt > 0.0) || !(d.dx > 0.0) || !(d.dy > 0.0)) { msg = "dt, dx, dy must be positive"; return CSADV_E_ARG; }
  return CSADV_OK;
}

int check_compatible(const csadv_solve_desc *solves, int nsolve, std::string &msg) {
  for (int s = 1; s < nsolve; ++s) {
    const csadv_solve_desc &a = solves[0], &b = solves[s];
    if (b.dtype != a.dtype || b.recon != a.recon || b.split != a.split || b.dp != a.dp ||
        is_general(b) != is_general(a)) {
      msg = "solve " + std::to_string(s) +
            ": all solves of one handle share dtype, recon, split, dp and the step kind (fused or general)";
      return CSADV_E_ARG;
    }
  }
  return CSADV_OK;
}

void tiling(int N, int tile, int &strip_w, int &strips, int &seg_l, int &segs) {
  strips = strip_count(N);
  strip_w = (N + strips - 1) / strips;
  seg_l = std::min(kSegRows[tile], N);
  segs = (N + seg_l - 1) / seg_l;
}

Solve derive(const csadv_solve_desc &d, int tile) {
  Solve s{};
  s.d = d;
  s.M = d.N + 2 * NG;
  tiling(d.N, d.tile >= 0 ? d.tile : tile, s.strip_w, s.strips, s.seg_l, s.segs);
  s.items = s.strips * s.segs;
  s.pstride = (int64_t)s.M * d.ld;
  s.tstride = 6 * s.pstride;
  s.pstride_u = (int64_t)s.M * d.ldu;
  s.pstride_v = (int64_t)(s.M + 1) * d.ld;
  s.part_per_tracer = 6 * s.items;
  return s;
}

std::vector<Item> work_list(const std::vector<Solve> &solves, int group) {
  // segment rows are cut into kXcds contiguous bands, one per XCD: the band's halo rows and
  // its share of the panel-invariant metric then stay in that XCD's L2 for all six panels
  // (the tracers of one segment are neighbours in a band, so they share its wind and metric
  // rows in L2; the work item packs panel and tracer as p + 8 * tracer, the first tracer of a
  // group of `group` consecutive tracers, the last group of a solve holding the rest)
  std::vector<std::vector<Item>> band(kXcds);
  for (int s = 0; s < (int)solves.size(); ++s) {
    const Solve &S = solves[s];
    for (int sg = 0; sg < S.segs; ++sg)
      for (int p = 0; p < 6; ++p)
        for (int sx = 0; sx < S.strips; ++sx)
          for (int tr = 0; tr < S.d.ntr; tr += group) band[sg * kXcds / S.segs].push_back(Item{s, p + 8 * tr, sx, sg});
  }
  size_t blen = 0;
  for (auto &bd : band) blen = std::max(blen, bd.size());
  std::vector<Item> work;
  work.reserve(blen * kXcds);
  for (size_t i = 0; i < blen; ++i)
    for (int x = 0; x < kXcds; ++x) work.push_back(i < band[x].size() ? band[x][i] : Item{-1, 0, 0, 0});
  return work;
}

std::vector<Item> ghost_list(const std::vector<Solve> &solves) {
  int maxtr = 0;
  for (auto &S : solves) maxtr = std::max(maxtr, S.d.ntr);
  std::vector<Item> g;
  for (int tr = 0; tr < maxtr; ++tr)
    for (int s = 0; s < (int)solves.size(); ++s) {
      const Solve &S = solves[s];
      if (tr >= S.d.ntr) continue;
      const int blocks = (S.d.n_strip + S.d.n_corner + BLOCK - 1) / BLOCK;
      for (int b = 0; b < blocks; ++b) g.push_back(Item{s, tr, b, 0});
      if (S.d.mf == 3) g.push_back(Item{s, tr, -1, 0});
    }
  return g;
}

int partials_per_tracer(int N, int tile) {
  if (tile < 0 || tile >= kNumTiles || N < 1) return CSADV_E_ARG;
  int sw, ns, sl, nsg;
  tiling(N, tile, sw, ns, sl, nsg);
  return 6 * ns * nsg;
}

int work_bytes(int N, int dtype, int64_t *bytes) {
  if (!bytes || N < 8 || N > 8192 || (dtype != CSADV_F64 && dtype != CSADV_F32)) return CSADV_E_ARG;
  const WorkLayout w = work_layout(N, general_ld(N), general_ldu(N));
  const int64_t s = dtype == CSADV_F64 ? 8 : 4;
  *bytes = w.total * s + w.n_red * 8;
  return CSADV_OK;
}

int tile_rows(int tile) {
  if (tile < 0 || tile >= kNumTiles) return CSADV_E_ARG;
  return kSegRows[tile];
}

}  // namespace csadv
