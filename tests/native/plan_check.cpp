// Checks of libcsadv's host-side planning (py-cubed-sphere_amd/csrc/csadv_plan.cpp), built with
// g++ under the address and undefined-behaviour sanitizers by tests/test_native_plan.py:
// descriptor validation on valid and broken descriptors, tile shapes over the whole N range,
// the XCD-banded work list of random batches (every item exactly once, bands contiguous and
// per XCD), and the general step's buffer layout.  Prints "plan_check ok <cases>" on success.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "csadv_plan.hpp"

using namespace csadv;

static int failures = 0;
#define CHECK(cond, ...)                                   \
  do {                                                     \
    if (!(cond)) {                                         \
      ++failures;                                          \
      if (failures < 20) {                                 \
        std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
        std::fprintf(stderr, __VA_ARGS__);                 \
        std::fprintf(stderr, "\n");                        \
      }                                                    \
    }                                                      \
  } while (0)

// A valid descriptor of size N (pointers are distinct non-null dummies, 16-byte aligned).
static csadv_solve_desc valid(int N, int ntr, std::vector<char> &arena) {
  csadv_solve_desc d{};
  d.N = N;
  d.ntr = ntr;
  d.dtype = CSADV_F64;
  d.ld = general_ld(N);
  d.ldu = general_ldu(N);
  d.recon = 3; d.dp = 1; d.split = 1; d.et = 3; d.mt = 1; d.mf = 3;
  d.nbasis = 1; d.unsteady = 0; d.max_steps = 1;
  d.dt = 0.01; d.dx = d.dy = 3.14159 / 2 / N; d.a2 = 1.0;
  char *p = arena.data();
  auto next = [&]() { p += 16; return (void *)p; };
  d.q0 = next(); d.q1 = next(); d.uavg = next(); d.vavg = next();
  d.gc = next(); d.gu = next(); d.gv = next(); d.rgc = next(); d.gfix = next();
  d.gc64 = (const double *)next(); d.lagw = (const double *)next(); d.lagk = (const int32_t *)next();
  d.orient = (const int32_t *)next(); d.wbu = (double *)next(); d.wbv = (double *)next();
  d.n_strip = 72 * N; d.n_corner = 216;
  d.strip_tgt = (const int32_t *)next(); d.strip_src = (const int32_t *)next(); d.strip_w = (const double *)next();
  d.corner_tgt = (const int32_t *)next(); d.corner_src = (const int32_t *)next(); d.corner_w = (const double *)next();
  d.coef = (const double *)next(); d.corr = (double *)next(); d.partials = (double *)next(); d.ctl = (int32_t *)next();
  d.tile = -1; d.stop = 0;
  return d;
}

int main() {
  std::vector<char> arena(4096);
  int cases = 0;
  std::string msg;
  // ---- validation: a valid descriptor passes; each broken field is rejected with a message
  for (int N : {8, 9, 48, 384, 1024, 6680, 6681, 8192}) {
    csadv_solve_desc d = valid(N, 4, arena);
    CHECK(validate(d, msg) == CSADV_OK, "valid N=%d rejected: %s", N, msg.c_str());
    ++cases;
    std::vector<csadv_solve_desc> bad;
    auto mut = [&](auto f) { csadv_solve_desc b = d; f(b); bad.push_back(b); };
    mut([](csadv_solve_desc &b) { b.N = 7; });
    mut([](csadv_solve_desc &b) { b.N = 8193; });
    mut([](csadv_solve_desc &b) { b.ntr = 0; });
    mut([](csadv_solve_desc &b) { b.ntr = 65; });
    mut([](csadv_solve_desc &b) { b.dtype = 2; });
    mut([](csadv_solve_desc &b) { b.ld = b.N + 7; });
    mut([](csadv_solve_desc &b) { b.ldu = b.N + 8; });
    mut([](csadv_solve_desc &b) { b.recon = 5; });
    mut([](csadv_solve_desc &b) { b.dp = 3; });
    mut([](csadv_solve_desc &b) { b.split = 3; b.mt = 1; });
    mut([](csadv_solve_desc &b) { b.split = 1; b.mt = 2; });
    mut([](csadv_solve_desc &b) { b.et = 0; });
    mut([](csadv_solve_desc &b) { b.mf = 4; });
    mut([](csadv_solve_desc &b) { b.et = 2; b.work = nullptr; });
    mut([&](csadv_solve_desc &b) { b.et = 2; b.work = arena.data(); b.ld += 4; });
    mut([](csadv_solve_desc &b) { b.nbasis = 9; });
    mut([](csadv_solve_desc &b) { b.unsteady = 1; b.max_steps = 0; });
    mut([](csadv_solve_desc &b) { b.q1 = nullptr; });
    mut([](csadv_solve_desc &b) { b.dp = 2; b.usgn = nullptr; });
    mut([](csadv_solve_desc &b) { b.partials = nullptr; });
    mut([](csadv_solve_desc &b) { b.gfix = nullptr; });
    mut([](csadv_solve_desc &b) { b.n_strip -= 1; });
    mut([](csadv_solve_desc &b) { b.strip_w = (const double *)((const char *)b.strip_w + 8); });
    mut([](csadv_solve_desc &b) { b.stop = -1; });
    mut([](csadv_solve_desc &b) { b.tile = kNumTiles; });
    mut([](csadv_solve_desc &b) { b.dt = 0.0; });
    mut([](csadv_solve_desc &b) { b.dx = -1.0; });
    for (size_t k = 0; k < bad.size(); ++k) {
      msg.clear();
      const int rc = validate(bad[k], msg);
      CHECK(rc < 0 && !msg.empty(), "broken descriptor %zu of N=%d accepted", k, N);
      ++cases;
    }
    // the general step with its work buffer and strides passes up to its addressing limit
    // (a tracer's Qx and Qy within 4 GiB of their base) and is rejected past it
    csadv_solve_desc g = d;
    g.et = 2; g.split = 3; g.mt = 2; g.mf = 1; g.work = arena.data();
    const bool fits = general_pair_bytes(N, g.dtype) < (int64_t(1) << 32);
    const int grc = validate(g, msg);
    CHECK(fits ? grc == CSADV_OK : grc == CSADV_E_RANGE, "general N=%d dtype %d: %d %s", N, g.dtype, grc, msg.c_str());
    // the passes' stores go through buffers of one cell array each, whose byte size must stay
    // below the dropped-store offset (2^31) wherever the general step is accepted
    for (int dt : {CSADV_F64, CSADV_F32}) {
      g.dtype = dt;
      const int64_t cell_bytes = int64_t(6) * (N + 2 * NG) * general_ld(N) * (dt == CSADV_F64 ? 8 : 4);
      if (validate(g, msg) == CSADV_OK) CHECK(cell_bytes < (int64_t(1) << 31), "general N=%d dtype %d: cell array of %lld bytes", N, dt, (long long)cell_bytes);
      ++cases;
    }
    g.dtype = d.dtype;
  }
  // ---- handle compatibility
  {
    csadv_solve_desc a = valid(48, 1, arena), b = valid(96, 3, arena);
    csadv_solve_desc two[2] = {a, b};
    CHECK(check_compatible(two, 2, msg) == CSADV_OK, "compatible pair rejected");
    two[1].dtype = CSADV_F32;
    CHECK(check_compatible(two, 2, msg) < 0, "mixed dtypes accepted");
    two[1] = b; two[1].et = 1;
    CHECK(check_compatible(two, 2, msg) < 0, "mixed step kinds accepted");
    cases += 3;
  }
  // ---- tile shapes cover every panel for all N and tiles
  for (int N = 8; N <= 8192; ++N) {
    for (int t = 0; t < kNumTiles; ++t) {
      int sw, ns, sl, nsg;
      tiling(N, t, sw, ns, sl, nsg);
      CHECK(sw >= 1 && sw <= STRIP_MAX && (int64_t)sw * ns >= N && (int64_t)sw * (ns - 1) < N,
            "strips N=%d tile=%d: %d x %d", N, t, ns, sw);
      CHECK(N - sw * (ns - 1) >= kMinLastStrip, "last strip N=%d: %d columns", N, N - sw * (ns - 1));
      CHECK(sl >= 1 && (int64_t)sl * nsg >= N && (int64_t)sl * (nsg - 1) < N, "segments N=%d tile=%d", N, t);
      CHECK(partials_per_tracer(N, t) == 6 * ns * nsg, "partials N=%d", N);
      ++cases;
    }
  }
  CHECK(partials_per_tracer(48, -1) < 0 && partials_per_tracer(48, kNumTiles) < 0, "bad tile accepted");
  CHECK(tile_rows(-1) < 0 && tile_rows(kNumTiles) < 0 && tile_rows(0) == kSegRows[0], "tile_rows");
  // ---- work lists of random batches
  std::mt19937 rng(2024);
  for (int trial = 0; trial < 120; ++trial) {
    const int group = 1 + trial % 2;             // tracers per item (FP64 1, FP32 fused step 2)
    const int nsolve = 1 + (int)(rng() % 70);
    std::vector<Solve> solves;
    for (int s = 0; s < nsolve; ++s) {
      const int N = 8 + (int)(rng() % 600);
      csadv_solve_desc d = valid(N, 1 + (int)(rng() % 8), arena);
      d.tile = (int)(rng() % (kNumTiles + 1)) - 1;
      solves.push_back(derive(d, (int)(rng() % kNumTiles)));
    }
    const std::vector<Item> work = work_list(solves, group);
    CHECK(work.size() % kXcds == 0, "work list length %zu", work.size());
    std::set<std::tuple<int, int, int, int>> seen;
    std::map<std::pair<int, int>, int> band_of;   // (solve, segment) -> XCD
    int64_t expect = 0;
    for (auto &S : solves) expect += (int64_t)6 * S.strips * S.segs * ((S.d.ntr + group - 1) / group);
    int64_t real = 0;
    for (size_t b = 0; b < work.size(); ++b) {
      const Item &it = work[b];
      if (it.x < 0) continue;
      ++real;
      CHECK(it.x < nsolve, "solve index");
      const Solve &S = solves[it.x];
      const int p = it.y & 7, tr = it.y >> 3;
      CHECK(p < 6 && tr < S.d.ntr && tr % group == 0 && it.z < S.strips && it.w < S.segs, "item out of range");
      CHECK(seen.insert({it.x, it.y, it.z, it.w}).second, "item listed twice");
      const int xcd = (int)(b % kXcds);
      auto key = std::make_pair(it.x, it.w);
      auto f = band_of.find(key);
      if (f == band_of.end()) band_of[key] = xcd;
      else CHECK(f->second == xcd, "segment split across XCDs");
      CHECK(xcd == it.w * kXcds / S.segs, "segment %d of %d on XCD %d", it.w, S.segs, xcd);
    }
    CHECK(real == expect, "work list holds %lld of %lld items", (long long)real, (long long)expect);
    ++cases;
    // the ghost fill's list: every cell block of every tracer once, the correction-total block
    // (-1) of a tracer exactly with MF-PR, tracer layers in order, nothing past a solve
    if (trial % 3 == 0)
      for (auto &S : solves) S.d.mf = 1 + (int)(rng() % 3);
    const std::vector<Item> gw = ghost_list(solves);
    std::set<std::tuple<int, int, int>> gseen;
    int64_t gexpect = 0;
    for (auto &S : solves)
      gexpect += (int64_t)S.d.ntr * ((S.d.n_strip + S.d.n_corner + BLOCK - 1) / BLOCK + (S.d.mf == 3 ? 1 : 0));
    CHECK((int64_t)gw.size() == gexpect, "ghost list holds %zu of %lld items", gw.size(), (long long)gexpect);
    int last_tr = 0;
    for (const Item &it : gw) {
      CHECK(it.x >= 0 && it.x < nsolve, "ghost item solve");
      const Solve &S = solves[it.x];
      const int blocks = (S.d.n_strip + S.d.n_corner + BLOCK - 1) / BLOCK;
      CHECK(it.y >= 0 && it.y < S.d.ntr, "ghost item tracer");
      CHECK(it.z >= -1 && it.z < blocks && (it.z >= 0 || S.d.mf == 3), "ghost item block %d", it.z);
      CHECK(it.y >= last_tr, "ghost items not tracer layer by layer");
      last_tr = it.y;
      CHECK(gseen.insert({it.x, it.y, it.z}).second, "ghost item listed twice");
    }
    ++cases;
  }
  // ---- general step layout: arrays disjoint, in bounds, FP64 tail aligned
  for (int N : {8, 9, 17, 48, 384, 768, 8192}) {
    const WorkLayout w = work_layout(N, general_ld(N), general_ldu(N));
    const int64_t offs[] = {w.o_qx, w.o_qy, w.o_div, w.o_side, w.o_red};
    const int64_t lens[] = {w.cell, w.cell, w.cell, (int64_t)24 * N};
    for (int k = 0; k + 1 < 5; ++k) CHECK(offs[k] + lens[k] <= offs[k + 1], "layout overlap N=%d array %d", N, k);
    CHECK(w.o_red % 8 == 0 && w.total == w.o_red, "reduction tail N=%d", N);
    // the tail holds every wavefront's partial of any tile shape, and their total
    for (int t = 0; t < kNumTiles; ++t) CHECK(partials_per_tracer(N, t) < w.n_red, "partials N=%d tile %d", N, t);
    int64_t f64 = 0, f32 = 0;
    CHECK(work_bytes(N, CSADV_F64, &f64) == CSADV_OK && work_bytes(N, CSADV_F32, &f32) == CSADV_OK, "work_bytes");
    CHECK(f64 == w.total * 8 + w.n_red * 8 && f32 == w.total * 4 + w.n_red * 8, "work_bytes N=%d", N);
    ++cases;
  }
  int64_t dummy;
  CHECK(work_bytes(7, CSADV_F64, &dummy) < 0 && work_bytes(48, 5, &dummy) < 0 && work_bytes(48, 0, nullptr) < 0,
        "work_bytes accepted bad arguments");
  if (failures) {
    std::fprintf(stderr, "plan_check: %d failure(s)\n", failures);
    return 1;
  }
  std::printf("plan_check ok %d\n", cases);
  return 0;
}
