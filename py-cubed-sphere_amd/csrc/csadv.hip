// libcsadv: MI355X-native dimension-split finite-volume advection on the cubed sphere.
//
// Kernels (all launched on the caller's stream; none allocates):
//   k_ghost_edges / k_ghost_corners   full 4-layer duo-grid ghost fill (src/interpolation.py:154-315),
//                                     used by csadv_ghost_fill and for the wind basis
//   k_ghost                           per-step ghost fill of the 3 layers a step reads, from
//                                     flattened recipes, plus the MF-PR sum of the last step
//   k_step<T,RECON,RK2,SPLIT,WIND>    the step: both directional passes of the split, PPM edges,
//                                     upwind fluxes, update and the mass-fixer partial sums of one
//                                     (solve, panel, tracer, column strip, row segment) per
//                                     wavefront (src/discrete_operators.py:18-101, src/flux.py,
//                                     src/reconstruction_1d.py); k_step1 is the same for a handle
//                                     of one solve, whose descriptor is a kernel argument
//   gp::k_gp_* (csadv_general.hip)    the general step (ET-S72, ET-PL07, MF-AF): one kernel per
//                                     stage of src/discrete_operators.py:18-101
//   k_wind_e2c / k_wind_c2e           duo-grid vector ghost fill (src/interpolation.py:347-532)
//   k_wind_avg                        time-averaged winds (src/averaged_velocity.py:14-62)
//   k_settle / k_finalize / k_clear_corr, k_mass / k_norms   state read-out and diagnostics
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "csadv_device.hpp"

using namespace csadv;

// ------------------------------------------------------------------ ghost fill
// One block per (panel, side, depth) strip of one field set; threads run along the edge.
template <typename T>
__global__ void k_ghost_edges(const FieldSet *sets) {
  const FieldSet &F = sets[blockIdx.y];
  const int p = blockIdx.x >> 4, side = (blockIdx.x >> 2) & 3, d = blockIdx.x & 3;
  const int N = F.N, M = F.M, ld = F.ld;
  const int c = F.cur ? (*F.cur & 1) : 0;
  T *base = reinterpret_cast<T *>(F.base[c]);
  const int32_t *o = F.orient + (p * 4 + side) * 7;
  for (int kk = threadIdx.x; kk < N; kk += blockDim.x) {
    const int k = kk + NG;
    const int K = F.lagk[d * M + k];
    const double *w = F.lagw + ((int64_t)d * M + k) * 4;
    int64_t src[4];
    double gs[4];
    int q = 0, si, sj;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      halo_src(o, d, K + l, q, si, sj);
      src[l] = (int64_t)q * F.pstride + (int64_t)sj * ld + si;
      gs[l] = F.corr ? F.gc64[(int64_t)sj * ld + si] : 0.0;
    }
    int ti, tj;
    ghost_target(side, d, k, N, ti, tj);
    const int64_t tgt = (int64_t)p * F.pstride + (int64_t)tj * ld + ti;
    for (int f = 0; f < F.nfield; ++f) {
      T *Q = base + (int64_t)f * F.fstride;
      const double cf = F.corr ? F.corr[(F.cur ? (*F.cur & 1) * F.nfield : 0) + f] : 0.0;
      double acc = 0.0;
#pragma unroll
      for (int l = 0; l < 4; ++l) acc += ((double)Q[src[l]] + cf * gs[l]) * w[l];
      Q[tgt] = (T)acc;
    }
  }
}

// One block per panel: 2 sides (E, W) x 4 depths x 8 corner rows; sources may be the
// neighbours' edge strips written by k_ghost_edges (src/interpolation.py:250-314).
template <typename T>
__global__ void k_ghost_corners(const FieldSet *sets) {
  const FieldSet &F = sets[blockIdx.y];
  const int p = blockIdx.x;
  const int t = threadIdx.x;
  if (t >= 64) return;
  const int side = t >> 5, d = (t >> 3) & 3, kk = t & 7;
  const int N = F.N, M = F.M, ld = F.ld;
  const int k = kk < 4 ? kk : N + NG + (kk - 4);
  const int c = F.cur ? (*F.cur & 1) : 0;
  T *base = reinterpret_cast<T *>(F.base[c]);
  const int32_t *o = F.orient + (p * 4 + side) * 7;
  const int K = F.lagk[d * M + k];
  const double *w = F.lagw + ((int64_t)d * M + k) * 4;
  int64_t src[4];
  double gs[4];
  int q = 0, si, sj;
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    halo_src(o, d, K + l, q, si, sj);
    src[l] = (int64_t)q * F.pstride + (int64_t)sj * ld + si;
    const bool interior = si >= NG && si < N + NG && sj >= NG && sj < N + NG;
    gs[l] = (F.corr && interior) ? F.gc64[(int64_t)sj * ld + si] : 0.0;
  }
  int ti, tj;
  ghost_target(side, d, k, N, ti, tj);
  const int64_t tgt = (int64_t)p * F.pstride + (int64_t)tj * ld + ti;
  for (int f = 0; f < F.nfield; ++f) {
    T *Q = base + (int64_t)f * F.fstride;
    const double cf = F.corr ? F.corr[(F.cur ? (*F.cur & 1) * F.nfield : 0) + f] : 0.0;
    double acc = 0.0;
#pragma unroll
    for (int l = 0; l < 4; ++l) acc += ((double)Q[src[l]] + cf * gs[l]) * w[l];
    Q[tgt] = (T)acc;
  }
}

// ------------------------------------------------------------------ step ghost fill
// Ghost cells a step reads (3 layers) from flattened duo-grid recipes: every value is a
// weighted sum of interior cells of other panels (edge strips: 4 sources; corners: 16, the
// two interpolation levels of src/interpolation.py:250-314 multiplied out).  The sum is taken
// over the stored values S; the step adds the correction total to every cell it reads as
// C gfix, where gfix is g inside the panel and sum_l w_l g_l at a ghost cell (the same table
// for the six panels), so that sum_l w_l (S_l + C g_l) = sum_l w_l S_l + C gfix.  The ghost
// values therefore never wait for the correction total, which one more block per
// (solve, tracer) reduces beside them.
// grid: x = recipe blocks + 1 (the last: correction total), y = solve, z = tracer.
// Control words of a solve (device int32 ctl[4], zero at init), each written by one kernel
// only and read by the other, so no kernel reads a word it writes:
//   ctl[1]  steps completed           written by k_step (one lane), read by k_ghost, k_wind_avg
//   ctl[2]  index of the running step written by k_ghost (one lane), read by k_step
//   ctl[3]  1 = the partials of the last step are not yet added to the correction total
// The state of step n lives in q[n & 1].  With MF-PR the stored interior values S are not
// corrected during stepping: Q = S + C g, where C (FP64, per tracer) is the running total of
// the mass-fixer corrections, kept in corr[(n & 1) * ntr + tracer] for step n.  Folding a
// correction of ~1e-9 into an FP32 state at every step would round it away where |S| is
// large and keep it where S is small, a bias that grows with the step count.
// Sum over a block of BLOCK threads in a fixed order (within each wavefront by halving
// distances, then the wavefronts in index order); every thread gets the total.
__device__ __forceinline__ double block_sum_fixed(double v, double *red) {
  v = warp_sum(v);
  if ((threadIdx.x & (WAVE - 1)) == 0) red[threadIdx.x / WAVE] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < BLOCK / WAVE; ++w) s += red[w];
  __syncthreads();
  return s;
}

// Sum of a tracer's mass-fixer partials (the last step's div g per wavefront), in a fixed
// order so that every run gets the same bits: every thread issues its eight loads of a batch
// before adding them in the order i, i+BLOCK, ... (masked past the end), then block_sum_fixed.
__device__ double partial_sum(const double *P, int n, double *red) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 8 * BLOCK) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = i + u * BLOCK < n ? P[i + u * BLOCK] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (i + u * BLOCK < n) acc += v[u];
  }
  return block_sum_fixed(acc, red);
}

// Mass-fixer correction total C after `steps` steps, from the stored totals c0 / c1 (slots of
// even / odd steps) and the partials' sum: with the partials of the last step pending,
// C = C(steps - 1) + dt m0 / a2, m0 = sum of div g over the interior
// (src/discrete_operators.py:98-101); otherwise C is the stored total of this step.
__device__ __forceinline__ double corr_total(double dt, double a2, double c0, double c1, double sum, int pending,
                                             int steps) {
  const double held = (steps & 1) ? c1 : c0, prev = (steps & 1) ? c0 : c1;
  return pending ? prev + dt * sum / a2 : held;
}

template <typename T>
__device__ __forceinline__ void ghost_cells(const Solve &S, double *red) {
  const csadv_solve_desc &D = S.d;
  // Every descriptor field of this kernel is requested up front: the empty asm below takes them
  // all, so their scalar loads are issued together before the first branch (one round trip,
  // where loading each field at its first use took four dependent ones).  The asm is neither
  // volatile nor a memory clobber, so the loads after it may still be scalar loads.
  const int ntr = D.ntr, mf = D.mf, stop = D.stop, n_strip = D.n_strip, n_corner = D.n_corner;
  const int part_n = S.part_per_tracer;
  const int64_t tstride = S.tstride;
  const double dt = D.dt, a2 = D.a2;
  int32_t *const ctl = D.ctl;
  double *const corr = D.corr;
  const double *const partials = D.partials;
  void *const q0 = D.q0, *const q1 = D.q1;
  const int32_t *const strip_src = D.strip_src, *const strip_tgt = D.strip_tgt;
  const int32_t *const corner_src = D.corner_src, *const corner_tgt = D.corner_tgt;
  const double *const strip_w = D.strip_w, *const corner_w = D.corner_w;
  const unsigned last_block = gridDim.x - 1;
  // one tracer per grid layer (tracers run side by side); the tracer index passes through the
  // asm (unchanged), so it exists only once every field is in registers
  int tr = blockIdx.z;
  asm("" : "=s"(tr) : "0"(tr), "s"(ntr), "s"(mf), "s"(stop), "s"(n_strip), "s"(n_corner), "s"(part_n),
      "s"(tstride), "s"(dt), "s"(a2), "s"(ctl), "s"(corr), "s"(partials), "s"(q0), "s"(q1), "s"(strip_src),
      "s"(strip_tgt), "s"(corner_src), "s"(corner_tgt), "s"(strip_w), "s"(corner_w), "s"(last_block));
  if (tr >= ntr) return;
  const int steps = ctl[1];            // requested before the recipes, whose wait it would join
  if (blockIdx.x == last_block) {      // the correction total after the last step
    if (mf != 3) return;
    // the partials and both stored totals are requested with the control words, not after
    // them: no load of this block waits for another
    const double c0 = corr[tr], c1 = corr[ntr + tr];
    const int pending = ctl[3];
    const double sum = partial_sum(partials + (int64_t)tr * part_n, part_n, red);
    if (stop > 0 && steps >= stop) return;
    if (threadIdx.x == 0 && pending) corr[(steps & 1) * ntr + tr] = corr_total(dt, a2, c0, c1, sum, 1, steps);
    return;
  }
  const int ncell = n_strip + n_corner;
  if ((int)(blockIdx.x * BLOCK) >= ncell && blockIdx.x != 0) return;
  const int idx = blockIdx.x * BLOCK + threadIdx.x;
  const bool active = idx < ncell;
  const bool strip = idx < n_strip;
  const int c = strip ? idx : idx - n_strip;
  // the recipe of this cell does not depend on the step: request all of it first, so that
  // its latency overlaps the read of the step counter
  int s16[16];
  double w16[16];
  int tgt = 0;
  if (active) {
    if (strip) {
      const int4 s4 = reinterpret_cast<const int4 *>(strip_src)[c];
      s16[0] = s4.x; s16[1] = s4.y; s16[2] = s4.z; s16[3] = s4.w;
      const double2 *w2 = reinterpret_cast<const double2 *>(strip_w + 4 * c);
      const double2 a = w2[0], b = w2[1];
      w16[0] = a.x; w16[1] = a.y; w16[2] = b.x; w16[3] = b.y;
      tgt = strip_tgt[c];
    } else {
#pragma unroll
      for (int l = 0; l < 4; ++l) {
        const int4 s4 = reinterpret_cast<const int4 *>(corner_src + 16 * c)[l];
        s16[4 * l] = s4.x; s16[4 * l + 1] = s4.y; s16[4 * l + 2] = s4.z; s16[4 * l + 3] = s4.w;
      }
      const double2 *w2 = reinterpret_cast<const double2 *>(corner_w + 16 * c);
#pragma unroll
      for (int l = 0; l < 8; ++l) {
        const double2 a = w2[l];
        w16[2 * l] = a.x; w16[2 * l + 1] = a.y;
      }
      tgt = corner_tgt[c];
    }
  }
  const bool lead = blockIdx.x == 0 && tr == 0 && threadIdx.x == 0;
  if (stop > 0 && steps >= stop) {          // solve finished: k_step reads ctl[2] and skips it too
    if (lead) ctl[2] = steps;
    return;
  }
  T *Qt = reinterpret_cast<T *>((steps & 1) ? q1 : q0) + tr * tstride;
  double acc = 0.0;                         // in recipe order
  if (active) {
    if (strip) {
      const T v0 = Qt[s16[0]], v1 = Qt[s16[1]], v2 = Qt[s16[2]], v3 = Qt[s16[3]];
      acc += (double)v0 * w16[0];
      acc += (double)v1 * w16[1];
      acc += (double)v2 * w16[2];
      acc += (double)v3 * w16[3];
    } else {
      T v[16];
#pragma unroll
      for (int l = 0; l < 16; ++l) v[l] = Qt[s16[l]];
#pragma unroll
      for (int l = 0; l < 16; ++l) acc += (double)v[l] * w16[l];
    }
    Qt[tgt] = (T)acc;
  }
  if (lead) ctl[2] = steps;
}

// Add the last step's partials to the correction total (once; used by finalize and the
// diagnostics).
__global__ void k_settle(Solve *solves, int first) {
  __shared__ double red[BLOCK];
  Solve &S = solves[first + blockIdx.x];
  csadv_solve_desc &D = S.d;
  if (D.mf != 3 || !D.ctl[3]) return;
  const int steps = D.ctl[1];
  for (int tr = 0; tr < D.ntr; ++tr) {
    const double sum = partial_sum(D.partials + (int64_t)tr * S.part_per_tracer, S.part_per_tracer, red);
    const double cf = corr_total(D.dt, D.a2, D.corr[tr], D.corr[D.ntr + tr], sum, 1, steps);
    if (threadIdx.x == 0) D.corr[(steps & 1) * D.ntr + tr] = cf;
  }
  __syncthreads();
  if (threadIdx.x == 0) D.ctl[3] = 0;
}

template <typename T>
__global__ __launch_bounds__(BLOCK) void k_ghost(const Solve *solves) {
  __shared__ double red[BLOCK];
  ghost_cells<T>(solves[blockIdx.y], red);
}

// A handle of one solve: the descriptor travels in the kernel arguments.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_ghost1(const Solve one) {
  __shared__ double red[BLOCK];
  ghost_cells<T>(one, red);
}

// Work item = (solve, panel, tracer, strip, segment): one wavefront marches down strip_w interior
// columns (lane l <-> column ti-3+l, so the 3-cell halo of each side rides in the outer
// lanes) over seg_l interior rows.  Iteration y:
//   * column stage at y-edge e = y-3 from the 6-row windows (rows y-6 .. y-1): y-fluxes of
//     Q and of Qx, Qy of row y-4, x-flux of Qy (neighbours by DPP lane shifts), update of
//     row y-4 and its mass-fixer partial sum;
//   * row stage for row y: x-flux of Q, split half-step Qx.
// The windows are rings indexed by the iteration's phase modulo 6 (the loop is unrolled by
// six, so every slot is a fixed register and nothing is shifted); row quantities that the
// column stage needs four iterations later (metric, 1/g, wind) ride in the same rings.
// The inputs of row y+PF are requested at the end of iteration y, PF = 2 rows ahead, so their
// memory latency overlaps a whole iteration of arithmetic: the row's state, correction metric
// and y-edge wind into one of two rotating buffers, the step-invariant tables and winds
// straight into their ring slots.
// (src/discrete_operators.py:18-101, src/flux.py, src/reconstruction_1d.py,
// src/advection_timestep.py:43)
template <typename T>
struct StepIn {     // inputs of iteration y at the lane's column: row y and y-edge y-3
  T q;                       // row y: Q
  T gf;                      // row y: metric the correction total multiplies (gfix)
  T raw;                     // row y: the stored value before the correction total is added
  T qa;                      // first rows only: Q of row y in the other state buffer
  T v, vs;                   // y-edge y-3: wind, upwind-side wind
};

// SPLIT is the splitting (1 AVLT, 2 L04, 3 PL07); the metric treatment follows from it, since
// the valid pairs are (AVLT or L04, MT-0) and (PL07, MT-PL07) (src/discrete_operators.py:48-82).
template <typename T, int RECON, bool RK2, int SPLIT, bool WIND = false, bool XPULL = false>
struct March {
  static constexpr int split = SPLIT;
  // rows whose inputs are in flight ahead of the one being processed (a third row of state
  // measured slower: its extra row offsets cost more than the longer lead saves)
  static constexpr int PF = 2;
  static constexpr bool mt0 = SPLIT != 3;
  // ---- inputs (per wavefront and tracer)
  const T *Q, *Qa, *U, *US, *V, *VS, *gc, *gu, *gv, *rgc, *gfix;
  // WIND: the winds of the step are formed here from the one-field FP64 basis (WU, WV) and its
  // coefficient ws at the step's time, as k_wind_avg forms them for RK1, instead of read
  const double *WU, *WV;
  double ws;
  T *Qo;
  int ld, ldu, M, iend, x, xc, xe, tj, ylast;
  int am2, am1, ap1, ap2;    // byte addresses of the lanes -2, -1, +1, +2 (mod 64) for lane pulls
  bool out_lane, need_col;
  T dtdx, dtdy;
  double cf;
  // ---- state: rings of rows y-6 .. y-1 (slot of row y-6+i in iteration phase K: (K+i) % 6).
  // The step-invariant row inputs (metric, 1/g, winds) are loaded straight into the slot their
  // row will hold: with PF = 2 the loads of row y+2 are issued at the end of iteration y, when
  // that slot's previous row, y-4, has had its last use.  gvr holds gv at y-edges by the same
  // phase (edge y-2 in slot K % 6), whose previous edge, y-6, was last read at iteration y-2.
  T qw[6], xw[6], gcw[6], rgw[6], uw[6], usw[6], guw[6], gvr[6];
  T sw[6];                   // FP32: stored values S of the rows (the update is S + flux)
  T g1prev, g2prev, vprev;
  T pa_A, pa_B, pa_q6, pb_A, pb_B, pb_q6;   // row y-4 as the left (upwind) cell of edge y-3: Q, Qx
  double part;

  // Inputs of row y, whose ring slot is SLOT.  SPEC: also read Q from the other state buffer
  // (Qa), for the rows requested before the step counter has arrived.
  template <bool SPEC, int SLOT>
  __device__ __forceinline__ void load(int y, StepIn<T> &r) {
    static_assert(PF == 2, "the ring slots are reused two rows ahead");
    const int yc = min(y, M - 1);
    const uint32_t o = (uint32_t)(yc * ld + xc), ou = (uint32_t)(yc * ldu + xe);
    if constexpr (WIND) uw[SLOT] = T(wind_of_basis(ws, ld_at(WU, ou)));
    else uw[SLOT] = ld_at(U, ou);
    if constexpr (RK2) usw[SLOT] = ld_at(US, ou);
    guw[SLOT] = ld_at(gu, ou);
    gcw[SLOT] = ld_at(gc, o);
    rgw[SLOT] = ld_at(rgc, o);
    r.gf = ld_at(gfix, o);
    const bool reach = need_col && y >= 1 && y < iend + 3;
    if constexpr (std::is_same<T, float>::value) {
      // FP32: Q of every lane is read (clamped row and column: rows past ylast are requested
      // too) and zeroed outside the stencils' reach by a select, with no branch around the load
      // (measured faster for FP32, slower for FP64, whose registers it raises)
      const T qv = ld_at(Q, o);
      r.q = reach ? qv : T(0);
      if constexpr (SPEC) {
        const T qa = ld_at(Qa, o);
        r.qa = reach ? qa : T(0);
      }
    } else {
      r.q = T(0);    // rows outside the stencils' reach
      r.qa = T(0);
      if (reach) {
        r.q = ld_at(Q, (uint32_t)(y * ld + x));
        if constexpr (SPEC) r.qa = ld_at(Qa, (uint32_t)(y * ld + x));
      }
    }
    const int ec = min(max(y - 3, 0), M - 1);
    if constexpr (WIND) r.v = T(wind_of_basis(ws, ld_at(WV, (uint32_t)(ec * ld + xc))));
    else r.v = ld_at(V, (uint32_t)(ec * ld + xc));
    r.vs = RK2 ? ld_at(VS, (uint32_t)(ec * ld + xc)) : r.v;
    gvr[SLOT] = ld_at(gv, (uint32_t)((ec + 1) * ld + xc));
  }

  // the upwind-side x-wind of the row in a slot (RK1: the wind itself)
  __device__ __forceinline__ T wind_sign(int slot) const { return RK2 ? usw[slot] : uw[slot]; }

  // every cell the stencils read carries the mass-fixer correction total C: Q = S + C gfix,
  // gfix being g inside the panel and, at a ghost cell, the weighted metric of the cells its
  // value was interpolated from (in the solve's precision: C gfix is ~1e-6 of q at most, so
  // its FP32 rounding is far below that of the sum, which is rounded once).  Lanes and rows
  // outside the stencils' reach hold 0 + C gfix, which no stored cell reads.
  // (with no correction total, cf = 0 and q + 0 gfix = q: no branch on the mass fixer)
  __device__ __forceinline__ void correct(StepIn<T> &r) const {
    r.raw = r.q;
    r.q = r.q + T(cf) * r.gf;
  }

  // Iteration y in phase K.  STAGE: how much of the column stage has begun, fixed by the row's
  // offset from the item's first row y0 = tj - 3: parabolas of row y-3 from y = tj+2 on
  // (STAGE >= 1), fluxes through y-edge y-3 from tj+3 on (>= 2), the update of row y-4 from
  // tj+4 on (3).  The first seven rows are peeled with their stage; every later row is 3, so
  // the loop carries no per-row test.
  template <int K, int STAGE>
  __device__ __forceinline__ void body(const int y, StepIn<T> &in) {
    constexpr int s0 = K % 6, s1 = (K + 1) % 6, s2 = (K + 2) % 6, s3 = (K + 3) % 6, s4 = (K + 4) % 6,
                  s5 = (K + 5) % 6;
    correct(in);
    const T gvm1 = gvr[s4], gv0 = gvr[s5], gvp1 = gvr[s0];   // gv at y-edges y-4, y-3, y-2
    // ---- column stage: parabolas of row y-3, y-edge e = y-3, output row r = y-4
    if constexpr (STAGE >= 1) {
      const T gce = gcw[s3];
      const Parab<T> a = parabola<T, RECON>(qw[s1], qw[s2], qw[s3], qw[s4], qw[s5], gv0, gvp1, gce, mt0);
      const Parab<T> b = parabola<T, RECON>(xw[s1], xw[s2], xw[s3], xw[s4], xw[s5], gv0, gvp1, gce, mt0);
      if constexpr (STAGE >= 2) {
        const T v = in.v;
        const bool vup = in.vs >= T(0);
        const T cy = v * dtdy;
        const T g1 = upwind<T>(vup, pa_A, pa_B, pa_q6, a, cy, gv0, mt0);     // times dt/dy
        const T g2 = upwind<T>(vup, pb_A, pb_B, pb_q6, b, cy, gv0, mt0);
        if constexpr (STAGE >= 3) {
          const T gcm = gcw[s2];               // row y-4
          const T dG1 = g1prev - g1;
          T d1 = T(0), d2 = T(0);
          if (split != 1) {
            d1 = gv0 * cy;
            d2 = gvm1 * (vprev * dtdy);
          }
          const T rg = rgw[s2];
          const T qy = split_update<T>(split, qw[s2], dG1, gcm, rg, d1, d2);
          const T ur = uw[s2];
          const bool urp = wind_sign(s2) >= T(0);
          const T gur = guw[s2];
          T qy_m2, qy_m1, qy_p1, qy_p2;
          xwindow<T, XPULL>(qy, am2, am1, ap1, ap2, qy_m2, qy_m1, qy_p1, qy_p2);
          const Parab<T> c = parabola<T, RECON>(qy_m2, qy_m1, qy, qy_p1, qy_p2, gur, from_right(gur), gcm, mt0);
          const T f2 = upwind_x<T>(urp, c, ur * dtdx, gur, from_right(gur), mt0);
          const T dF2 = f2 - from_right(f2);
          const T dG2 = g2prev - g2;
          // div = -(dF2 + dG2) / (dt g);  Q - dt div = Q + (dF2 + dG2) / g
          const T flx = dF2 + dG2;
          if (out_lane) {
            // the correction total stays out of the stored state: S(n+1) = S(n) + flux
            T out;
            if constexpr (std::is_same<T, float>::value) out = sw[s2] + flx * rg;
            else out = (qw[s2] + flx * rg) - T(cf) * gcm;
            st_at(Qo, (uint32_t)((y - 4) * ld + x), out);
            part -= (double)flx;               // used with MF-PR only
          }
        }
        g1prev = g1;
        g2prev = g2;
        vprev = v;
      }
      pa_A = a.rA; pa_B = a.rB; pa_q6 = a.q6;
      pb_A = b.rA; pb_B = b.rB; pb_q6 = b.q6;
    }
    // ---- row stage: row y
    const T q = in.q, u = uw[s0], gu0 = guw[s0], gc0 = gcw[s0];
    const bool up = wind_sign(s0) >= T(0);
    const T gup1 = from_right(gu0);
    T q_m2, q_m1, q_p1, q_p2;
    xwindow<T, XPULL>(q, am2, am1, ap1, ap2, q_m2, q_m1, q_p1, q_p2);
    const Parab<T> c = parabola<T, RECON>(q_m2, q_m1, q, q_p1, q_p2, gu0, gup1, gc0, mt0);
    const T f1 = upwind_x<T>(up, c, u * dtdx, gu0, gup1, mt0);
    const T dF1 = f1 - from_right(f1);
    T c1 = T(0), c2 = T(0);
    if (split != 1) {
      c1 = gup1 * (from_right(u) * dtdx);
      c2 = gu0 * (u * dtdx);
    }
    const T qx = split_update<T>(split, q, dF1, gc0, rgw[s0], c1, c2);
    // row y replaces row y-6 in the rings of computed values (the loaded ones are there already)
    qw[s0] = q; xw[s0] = qx;
    if constexpr (std::is_same<T, float>::value) sw[s0] = in.raw;
    // rows past ylast are read in bounds (clamped rows, masked Q) and never used
    load<false, (K + PF) % 6>(y + PF, in);
  }

  // The first PF rows' inputs, requested before the step counter is known: Q from both state
  // buffers (Q = q0, Qa = q1), the rest from the step-invariant tables and winds.
  __device__ __forceinline__ void start(StepIn<T> *b) {
#pragma unroll
    for (int k = 0; k < 6; ++k) qw[k] = xw[k] = gcw[k] = rgw[k] = uw[k] = usw[k] = guw[k] = gvr[k] = sw[k] = T(0);
    g1prev = g2prev = vprev = T(0);   // gv is complete from y = tj-1 on; the column stage starts at tj+2
    pa_A = pa_B = pa_q6 = pb_A = pb_B = pb_q6 = T(0);
    part = 0.0;
    // y0 = tj - 3 has phase 0; ylast - y0 = ny + 6 >= PF: every initial row exists
    load<true, 0>(tj - 3, b[0]);
    load<true, 1>(tj - 2, b[1]);
  }

  // The step's input buffer is known: keep the first rows' Q from it and read the rest there.
  __device__ __forceinline__ void resolve(StepIn<T> *b, bool odd) {
#pragma unroll
    for (int k = 0; k < PF; ++k) b[k].q = odd ? b[k].qa : b[k].q;
    Q = odd ? Qa : Q;
  }

  __device__ __forceinline__ void run(StepIn<T> *b) {
    // rows y0 .. y0+6 (ylast - y0 = ny + 6 >= 7: they all exist), the column stage starting
    int y = tj - 3;
    body<0, 0>(y, b[0]); body<1, 0>(y + 1, b[1]); body<2, 0>(y + 2, b[0]); body<3, 0>(y + 3, b[1]);
    body<4, 0>(y + 4, b[0]); body<5, 1>(y + 5, b[1]); body<0, 2>(y + 6, b[0]);
    // rows y0+7 on: the full step, six rows (phases 1 .. 5, 0) per iteration
    for (y += 7; y + 5 <= ylast; y += 6) {
      body<1, 3>(y, b[1]); body<2, 3>(y + 1, b[0]); body<3, 3>(y + 2, b[1]);
      body<4, 3>(y + 3, b[0]); body<5, 3>(y + 4, b[1]); body<0, 3>(y + 5, b[0]);
    }
    if (y <= ylast) body<1, 3>(y, b[1]);
    if (y + 1 <= ylast) body<2, 3>(y + 1, b[0]);
    if (y + 2 <= ylast) body<3, 3>(y + 2, b[1]);
    if (y + 3 <= ylast) body<4, 3>(y + 3, b[0]);
    if (y + 4 <= ylast) body<5, 3>(y + 4, b[1]);
  }
};

template <typename T, int RECON, bool RK2, int SPLIT, bool WIND, bool XPULL>
__device__ __forceinline__ void step_item(const Solve &S, const int4 wk) {
  const int lane = threadIdx.x;
  const csadv_solve_desc &D = S.d;
  const int p = wk.y & 7, tr = wk.y >> 3;    // panel, tracer
  const int N = D.N;
  const int ti = NG + wk.z * S.strip_w, tj = NG + wk.w * S.seg_l;
  const int iend = N + NG;
  const int nx = min(S.strip_w, iend - ti), ny = min(S.seg_l, iend - tj);
  const int item = (p * S.segs + wk.w) * S.strips + wk.z;
  const bool mfpr = D.mf == 3;
  March<T, RECON, RK2, SPLIT, WIND, XPULL> m;
  m.ld = D.ld;
  m.ldu = D.ldu;
  m.M = S.M;
  m.iend = iend;
  m.x = ti - 3 + lane;
  m.xc = min(m.x, m.M - 1);          // clamped (lanes past the array)
  m.xe = min(m.x, m.M);
  m.tj = tj;
  m.ylast = tj + ny + 3;
  m.am2 = ((lane + 62) & 63) << 2;
  m.am1 = ((lane + 63) & 63) << 2;
  m.ap1 = ((lane + 1) & 63) << 2;
  m.ap2 = ((lane + 2) & 63) << 2;
  m.out_lane = lane >= 3 && lane < 3 + nx;
  m.need_col = m.x >= 1 && m.x < iend + 3;   // columns the stencils reach
  m.dtdx = T(D.dt / D.dx);
  m.dtdy = T(D.dt / D.dy);
  m.U = reinterpret_cast<const T *>(D.uavg) + p * S.pstride_u;
  m.V = reinterpret_cast<const T *>(D.vavg) + p * S.pstride_v;
  // RK2 picks the upwind side from separate (un-averaged) winds; RK1 from the winds themselves
  m.US = RK2 ? reinterpret_cast<const T *>(D.usgn) + p * S.pstride_u : m.U;
  m.VS = RK2 ? reinterpret_cast<const T *>(D.vsgn) + p * S.pstride_v : m.V;
  m.gc = reinterpret_cast<const T *>(D.gc);
  m.gu = reinterpret_cast<const T *>(D.gu);
  m.gv = reinterpret_cast<const T *>(D.gv);
  m.rgc = reinterpret_cast<const T *>(D.rgc);
  m.gfix = reinterpret_cast<const T *>(D.gfix);
  const int64_t off = tr * S.tstride + p * S.pstride;
  m.Q = reinterpret_cast<const T *>(D.q0) + off;
  m.Qa = reinterpret_cast<const T *>(D.q1) + off;
  StepIn<T> b[March<T, RECON, RK2, SPLIT, WIND, XPULL>::PF];
  int step;
  if constexpr (WIND) {
    // the winds of the first rows need the step's coefficient: the step counter comes first
    step = D.ctl[2];
    const int row = min(step, D.max_steps - 1);     // as k_wind_avg (one basis field)
    m.ws = D.coef[row * 2];
    m.WU = D.wbu + p * S.pstride_u;
    m.WV = D.wbv + p * S.pstride_v;
    m.start(b);
  } else {
    m.start(b);
    // read after the first rows are requested: a wait for any scalar load waits for all of them
    // (they return out of order), so reading these first would hold back the rows' loads
    step = D.ctl[2];
  }
  const double c0 = mfpr ? D.corr[tr] : 0.0, c1 = mfpr ? D.corr[D.ntr + tr] : 0.0;
  if (D.stop > 0 && step >= D.stop) return;   // solve finished
  const bool odd = step & 1;                  // the state of step n lives in q[n & 1]
  m.resolve(b, odd);
  m.Qo = reinterpret_cast<T *>(odd ? D.q0 : D.q1) + off;
  m.cf = odd ? c1 : c0;
  m.run(b);
  if (mfpr) {
    const double part = warp_sum(m.part) / D.dt;   // sum of div g over the wavefront's cells
    if (lane == 0) D.partials[(int64_t)tr * S.part_per_tracer + item] = part;
  }
  if (item == 0 && tr == 0 && lane == 0) {   // hand the step over to the next kernels of the stream
    D.ctl[1] = step + 1;
    if (mfpr) D.ctl[3] = 1;
  }
}

// Bring every cache line of the kernel-argument descriptor into the scalar cache with loads
// issued together: an empty asm takes one word of each line and passes v through unchanged, so
// v exists only once all of them have arrived (one scalar round trip); the field loads at their
// first use then hit the cache instead of waiting for their lines one after another (the
// one-solve step kernel's start took four such dependent round trips).  The asm is neither
// volatile nor a memory clobber, so the loads after it may still be scalar loads.  (For a
// descriptor in global memory the compiler reads the words with vector loads, which warm no
// scalar cache: the multi-solve kernel does not use it.)
__device__ __forceinline__ int touch_solve(const Solve &S, int v) {
  static_assert(sizeof(Solve) == 424, "one word in each 64-byte line of the descriptor");
  const char *b = reinterpret_cast<const char *>(&S);
  int w[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) __builtin_memcpy(&w[k], b + (k < 7 ? 64 * k : 420), 4);
  asm("" : "=s"(v) : "0"(v), "s"(w[0]), "s"(w[1]), "s"(w[2]), "s"(w[3]), "s"(w[4]), "s"(w[5]), "s"(w[6]), "s"(w[7]));
  return v;
}

template <typename T, int RECON, bool RK2, int SPLIT, bool WIND>
__global__ __launch_bounds__(WAVE) void k_step(const Solve *__restrict__ solves, const int4 *__restrict__ work) {
  const int4 wk = work[blockIdx.x];
  if (wk.x < 0) return;              // padding item of the XCD-banded work list
  step_item<T, RECON, RK2, SPLIT, WIND, false>(solves[wk.x], wk);
}

// Work item b of a single-solve handle, computed instead of read: the list the host builds
// (csadv_update) gives workgroup b, which runs on XCD b % kXcds, item b / kXcds of that XCD's
// band; band x holds the row segments sg with sg * kXcds / segs == x, each segment listing
// its (panel, strip, tracer) items with the tracer fastest.  Past the band's end: padding.
__device__ __forceinline__ bool single_item(const Solve &S, int b, int4 &wk) {
  const int x = b % kXcds, i = b / kXcds;
  const int lo = (x * S.segs + kXcds - 1) / kXcds, hi = ((x + 1) * S.segs + kXcds - 1) / kXcds;
  const int ntr = S.d.ntr, row = S.strips * ntr, per = 6 * row;
  if (i >= (hi - lo) * per) return false;
  const int r = i % per, r2 = r % row;
  wk = make_int4(0, r / row + 8 * (r2 % ntr), r2 / ntr, lo + i / per);
  return true;
}

// A handle of one solve: the descriptor travels in the kernel arguments and the work item is
// computed, so the first loads of a wavefront depend on nothing but the arguments.
// FP32 takes its x windows through the crossbar (measured 3% faster at N=768 with 4 tracers);
// FP64 and the multi-solve kernel keep the DPP shifts (measured 0.5-4% slower with the crossbar).
template <typename T, int RECON, bool RK2, int SPLIT, bool WIND>
__global__ __launch_bounds__(WAVE) void k_step1(const Solve one) {
  int4 wk;
  if (!single_item(one, touch_solve(one, blockIdx.x), wk)) return;
  step_item<T, RECON, RK2, SPLIT, WIND, std::is_same<T, float>::value>(one, wk);
}

// Fold the correction totals into the current buffer (interior only); k_clear_corr then
// clears them.
template <typename T>
__global__ void k_finalize(Solve *solves, int first) {
  Solve &S = solves[first + blockIdx.y];
  csadv_solve_desc &D = S.d;
  if (D.mf != 3) return;
  const int N = D.N, ld = D.ld;
  T *Q = reinterpret_cast<T *>((D.ctl[1] & 1) ? D.q1 : D.q0);
  const int64_t tot = (int64_t)6 * N * N;
  for (int tr = 0; tr < D.ntr; ++tr) {
    const double cf = D.corr[(D.ctl[1] & 1) * D.ntr + tr];
    if (cf == 0.0) continue;
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * blockDim.x) {
      const int p = (int)(idx / ((int64_t)N * N));
      const int rem = (int)(idx - (int64_t)p * N * N);
      const int j = rem / N + NG, i = rem % N + NG;
      T &v = Q[tr * S.tstride + p * S.pstride + (int64_t)j * ld + i];
      v = (T)((double)v + cf * D.gc64[(int64_t)j * ld + i]);
    }
  }
}

__global__ void k_clear_corr(Solve *solves, int first) {
  csadv_solve_desc &D = solves[first + blockIdx.x].d;
  if (D.mf == 3)
    for (int k = threadIdx.x; k < 2 * D.ntr; k += blockDim.x) D.corr[k] = 0.0;
}

// ------------------------------------------------------------------ lat-lon output remap
// Nearest-cell remap onto the plotting grid (src/cs_datastruct.py:569-582,
// src/interpolation.py:38-149).  Point (a, b) has lon = -pi + a * (2pi / (nlon-1)) and
// lat = -pi/2 + b * (pi / (nlat-1)), formed as numpy.linspace forms them (product, then sum,
// each rounded; the last point is the end value).  Its panel is the face of its largest
// Cartesian component; the reference assigns the face masks in the order 0..5, so a later face
// wins a tie.  Its cell is floor((x + a)/dx), floor((y + a)/dy) of the inverse equiangular map
// (clamped to the panel, where the reference would index out of range).  out[a][b] is the
// tracer's value Q = S + C g (field == null) or field[p][j][i]; ix / jy / panel, when given,
// receive the cell indices (the reference's ll2cs tables).
template <typename T>
__global__ void k_remap_latlon(const Solve *solves, int solve, int tracer, const double *field, int nlat, int nlon,
                               double *out, int32_t *ix, int32_t *jy, int32_t *pn) {
  const Solve &S = solves[solve];
  const csadv_solve_desc &D = S.d;
  const int N = D.N, ld = D.ld;
  const double pi = 3.141592653589793, a = pi * 0.5 * 0.5, dx = 2.0 * a / N;
  const double lon_step = (pi - (-pi)) / (nlon - 1), lat_step = (pi * 0.5 - (-pi * 0.5)) / (nlat - 1);
  const T *Q = reinterpret_cast<const T *>((D.ctl[1] & 1) ? D.q1 : D.q0) + tracer * S.tstride;
  const double cf = (D.corr && D.mf == 3) ? D.corr[(D.ctl[1] & 1) * D.ntr + tracer] : 0.0;
  const int64_t total = (int64_t)nlon * nlat;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < total; k += (int64_t)gridDim.x * blockDim.x) {
    const int ia = (int)(k / nlat), ib = (int)(k - (int64_t)ia * nlat);
    const double lon = ia == nlon - 1 ? pi : __dadd_rn(__dmul_rn((double)ia, lon_step), -pi);
    const double lat = ib == nlat - 1 ? pi * 0.5 : __dadd_rn(__dmul_rn((double)ib, lat_step), -pi * 0.5);
    const double cl = cos(lat);
    const double X = __dmul_rn(cl, cos(lon)), Y = __dmul_rn(cl, sin(lon)), Z = sin(lat);
    const double ax = fabs(X), ay = fabs(Y), az = fabs(Z), pm = fmax(fmax(ax, ay), az);
    int p = 0;
    if (pm == ax && X > 0) p = 0;
    if (pm == ay && Y > 0) p = 1;
    if (pm == ax && X < 0) p = 2;
    if (pm == ay && Y < 0) p = 3;
    if (pm == az && Z > 0) p = 4;
    if (pm == az && Z <= 0) p = 5;
    // inverse equiangular map of face p (src/cs_transform.py:99-126): tan x = nx / d, tan y = ny / d
    const double d = (p == 0 || p == 2) ? X : ((p == 1 || p == 3) ? Y : Z);
    const double nx = (p == 0 || p == 2) ? Y : ((p == 1 || p == 3) ? -X : (p == 4 ? Y : -Y));
    const double ny = p == 0 ? Z : (p == 1 ? Z : (p == 2 ? -Z : (p == 3 ? -Z : -X)));
    const double x = atan(nx / d), y = atan(ny / d);
    const int i = min(max((int)floor((x - (-a)) / dx), 0), N - 1);
    const int j = min(max((int)floor((y - (-a)) / dx), 0), N - 1);
    if (out) {
      double v;
      if (field) {
        v = field[((int64_t)p * N + j) * N + i];
      } else {
        const int ci = i + NG, cj = j + NG;
        v = (double)Q[p * S.pstride + (int64_t)cj * ld + ci] + cf * D.gc64[(int64_t)cj * ld + ci];
      }
      out[k] = v;
    }
    if (ix) ix[k] = i;
    if (jy) jy[k] = j;
    if (pn) pn[k] = p;
  }
}

// ------------------------------------------------------------------ winds
// Conversion-factor tables (FP64), per panel:
//   conv_pc [4][6][M][ld]      ex_lon, ex_lat, ey_lon, ey_lat at centres
//   conv_pu [5][6][M][ldu]     ... and det at x-edges
//   conv_pv [5][6][M+1][ld]    ... and det at y-edges
struct WindArgs {
  const Solve *S;
  const double *conv_pc, *conv_pu, *conv_pv;
  double *wpc;            // [nbasis][2][6][M][ld]: u_lon, v_lat at centres
};

__device__ __forceinline__ WindArgs wind_args(const Solve *S) {
  return WindArgs{S, S->d.conv_pc, S->d.conv_pu, S->d.conv_pv, S->d.wpc};
}

// Edge winds -> centres in the 4-deep boundary frame, then contravariant -> lat-lon
// (src/interpolation.py:347-425).  grid: x = frame cell, y = panel, z = basis.
__global__ void k_wind_e2c(const Solve *solves, int s) {
  const WindArgs A = wind_args(solves + s);
  const Solve &S = *A.S;
  const csadv_solve_desc &D = S.d;
  const int N = D.N, M = S.M, ld = D.ld, ldu = D.ldu;
  const int p = blockIdx.y, b = blockIdx.z;
  const int i0 = NG, iend = N + NG;
  const double a1 = 5.0 / 16.0, a2 = 15.0 / 16.0, a3 = -5.0 / 16.0, a4 = 1.0 / 16.0;
  const double b1 = -1.0 / 16.0, b2 = 9.0 / 16.0, b3 = 9.0 / 16.0, b4 = -1.0 / 16.0;
  const double *u = D.wbu + (int64_t)b * 6 * S.pstride_u + p * S.pstride_u;
  const double *v = D.wbv + (int64_t)b * 6 * S.pstride_v + p * S.pstride_v;
  const int64_t cps = (int64_t)M * ld;            // conv_pc panel stride
  const int64_t cfs = 6 * cps;                    // conv_pc factor stride
  double *ul = A.wpc + ((int64_t)b * 2 + 0) * 6 * S.pstride + p * S.pstride;
  double *vl = A.wpc + ((int64_t)b * 2 + 1) * 6 * S.pstride + p * S.pstride;
  // every interior cell within 4 of a panel edge
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < N * N; idx += gridDim.x * blockDim.x) {
    const int i = idx % N + i0, j = idx / N + i0;
    const bool frame = (i < i0 + 4) || (i >= iend - 4) || (j < i0 + 4) || (j >= iend - 4);
    if (!frame) continue;
    const double *ur = u + (int64_t)j * ldu;
    double uc, vc;
    if (i == i0) uc = a1 * ur[i0] + a2 * ur[i0 + 1] + a3 * ur[i0 + 2] + a4 * ur[i0 + 3];
    else if (i == iend - 1) uc = a4 * ur[iend - 3] + a3 * ur[iend - 2] + a2 * ur[iend - 1] + a1 * ur[iend];
    else uc = b1 * ur[i - 1] + b2 * ur[i] + b3 * ur[i + 1] + b4 * ur[i + 2];
    auto vv = [&](int e) { return v[(int64_t)e * ld + i]; };
    if (j == i0) vc = a1 * vv(i0) + a2 * vv(i0 + 1) + a3 * vv(i0 + 2) + a4 * vv(i0 + 3);
    else if (j == iend - 1) vc = a4 * vv(iend - 3) + a3 * vv(iend - 2) + a2 * vv(iend - 1) + a1 * vv(iend);
    else vc = b1 * vv(j - 1) + b2 * vv(j) + b3 * vv(j + 1) + b4 * vv(j + 2);
    const int64_t o = p * cps + (int64_t)j * ld + i;
    const double exlon = A.conv_pc[o], exlat = A.conv_pc[cfs + o];
    const double eylon = A.conv_pc[2 * cfs + o], eylat = A.conv_pc[3 * cfs + o];
    ul[(int64_t)j * ld + i] = exlon * uc + eylon * vc;      // src/sphgeo.py:130-133
    vl[(int64_t)j * ld + i] = exlat * uc + eylat * vc;
  }
}

// Centre lat-lon winds -> ghost x-edges (rows outside the panel, plus edge columns i0-1 and
// iend+1) and ghost y-edges, then lat-lon -> contravariant (src/interpolation.py:436-532).
__global__ void k_wind_c2e(const Solve *solves, int s) {
  const WindArgs A = wind_args(solves + s);
  const Solve &S = *A.S;
  const csadv_solve_desc &D = S.d;
  const int N = D.N, M = S.M, ld = D.ld, ldu = D.ldu;
  const int p = blockIdx.y, b = blockIdx.z;
  const int i0 = NG, iend = N + NG;
  const double c1 = 9.0 / 16.0, c2 = -1.0 / 16.0;
  const double *ul = A.wpc + ((int64_t)b * 2 + 0) * 6 * S.pstride + p * S.pstride;
  const double *vl = A.wpc + ((int64_t)b * 2 + 1) * 6 * S.pstride + p * S.pstride;
  double *u = D.wbu + (int64_t)b * 6 * S.pstride_u + p * S.pstride_u;
  double *v = D.wbv + (int64_t)b * 6 * S.pstride_v + p * S.pstride_v;
  const int64_t ups = (int64_t)M * ldu, ufs = 6 * ups;
  const int64_t vps = (int64_t)(M + 1) * ld, vfs = 6 * vps;
  // x-edges: positions (e, j) with e in [i0-1, iend+1]; ghost rows for e in [i0, iend],
  // every row for e = i0-1 and e = iend+1
  const int nu = (N + 3) * M;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < nu + (N + 3) * M; idx += gridDim.x * blockDim.x) {
    if (idx < nu) {
      const int e = idx % (N + 3) + i0 - 1, j = idx / (N + 3);
      const bool ghostrow = j < i0 || j >= iend;
      if (!(e == i0 - 1 || e == iend + 1 || ghostrow)) continue;
      const double *ur = ul + (int64_t)j * ld, *vr = vl + (int64_t)j * ld;
      const double lon = c1 * (ur[e] + ur[e - 1]) + c2 * (ur[e + 1] + ur[e - 2]);
      const double lat = c1 * (vr[e] + vr[e - 1]) + c2 * (vr[e + 1] + vr[e - 2]);
      const int64_t o = p * ups + (int64_t)j * ldu + e;
      const double eylon = A.conv_pu[2 * ufs + o], eylat = A.conv_pu[3 * ufs + o];
      const double det = A.conv_pu[4 * ufs + o];
      u[(int64_t)j * ldu + e] = (eylat * lon - eylon * lat) / det;     // src/sphgeo.py:120-125
    } else {
      const int id2 = idx - nu;
      const int e = id2 % (N + 3) + i0 - 1, i = id2 / (N + 3);
      const bool ghostcol = i < i0 || i >= iend;
      if (!(e == i0 - 1 || e == iend + 1 || ghostcol)) continue;
      const double lon = c1 * (ul[(int64_t)e * ld + i] + ul[(int64_t)(e - 1) * ld + i]) +
                         c2 * (ul[(int64_t)(e + 1) * ld + i] + ul[(int64_t)(e - 2) * ld + i]);
      const double lat = c1 * (vl[(int64_t)e * ld + i] + vl[(int64_t)(e - 1) * ld + i]) +
                         c2 * (vl[(int64_t)(e + 1) * ld + i] + vl[(int64_t)(e - 2) * ld + i]);
      const int64_t o = p * vps + (int64_t)e * ld + i;
      const double exlon = A.conv_pv[o], exlat = A.conv_pv[vfs + o];
      const double det = A.conv_pv[4 * vfs + o];
      v[(int64_t)e * ld + i] = (-exlat * lon + exlon * lat) / det;
    }
  }
}

// Time-averaged winds of the current step from the FP64 basis (src/averaged_velocity.py:14-62).
// u(t) = sum_m s_m(t) W_m; RK1: u_avg = u(t_{n-1}); RK2: upwind-interpolated extrapolation.
// Writes dtype uavg / usgn on x-edges e in [i0-1, iend+1] (all rows) and likewise for v.
template <typename T>
__global__ void k_wind_avg(Solve *solves, int all) {
  Solve &S = solves[blockIdx.z];
  const csadv_solve_desc &D = S.d;
  const int N = D.N, M = S.M, ld = D.ld, ldu = D.ldu, K = D.nbasis;
  const int p = blockIdx.y;
  const int i0 = NG, iend = N + NG;
  if (!all && !D.unsteady) return;                                  // steady: set at init
  if (D.stop > 0 && D.ctl[1] >= D.stop) return;                    // solve finished
  const int row = D.unsteady ? min(D.ctl[1], D.max_steps - 1) : 0;   // steps completed
  const double *s1 = D.coef + (int64_t)row * 2 * K;   // coefficients at t_{n-1}
  const double *s0 = s1 + K;                          // at t_{n-2}
  const double dto2 = 0.5 * D.dt;
  const bool rk2 = D.dp == 2;
  const int64_t bs_u = 6 * S.pstride_u, bs_v = 6 * S.pstride_v;
  const double *Wu = D.wbu + p * S.pstride_u;
  const double *Wv = D.wbv + p * S.pstride_v;
  T *ua = reinterpret_cast<T *>(D.uavg) + p * S.pstride_u;
  T *va = reinterpret_cast<T *>(D.vavg) + p * S.pstride_v;
  T *us = D.usgn ? reinterpret_cast<T *>(D.usgn) + p * S.pstride_u : nullptr;
  T *vs = D.vsgn ? reinterpret_cast<T *>(D.vsgn) + p * S.pstride_v : nullptr;
  auto comb = [&](const double *W, int64_t bs, int64_t off, const double *s) {
    if (K == 1) return wind_of_basis(s[0], W[off]);   // the same bits as the step's own (WIND)
    double acc = 0.0;
    for (int m = 0; m < K; ++m) acc += s[m] * W[m * bs + off];
    return acc;
  };
  const int ne = N + 1;   // edges i0 .. iend
  // one block per row of edges: x-edge rows c = 0 .. M-1 ([row][edge], edges i0 .. iend), then
  // y-edge rows e = i0 .. iend ([edge][column], columns 0 .. M-1); threads run along the row,
  // so consecutive threads take consecutive addresses and no index is divided
  const int r = blockIdx.x;
  if (r >= M + ne) return;                    // the grid covers the largest solve
  const bool xdir = r < M;
  const int len = xdir ? ne : M;
  for (int t = threadIdx.x; t < len; t += blockDim.x) {
    const int e = xdir ? t + i0 : r - M + i0;   // edge index along the flux direction
    const int c = xdir ? r : t;                 // row (x-edges) / column (y-edges)
    const int64_t bs = xdir ? bs_u : bs_v;
    const double *W = xdir ? Wu : Wv;
    const int64_t off = xdir ? (int64_t)c * ldu + e : (int64_t)e * ld + c;
    const int64_t step = xdir ? 1 : ld;
    const double rd = xdir ? 1.0 / D.dx : 1.0 / D.dy;
    const double u = comb(W, bs, off, s1);
    double avg = u;
    if (rk2) {
      auto ui = [&](int64_t o) {
        const double un = comb(W, bs, o, s1);
        const double uo = comb(W, bs, o, s0);
        return 1.5 * un - 0.5 * uo;
      };
      const double a = (u * dto2) * rd;
      if (u >= 0.0) avg = (1.0 - a) * ui(off) + a * ui(off - step);
      else avg = -a * ui(off + step) + (1.0 + a) * ui(off);
    }
    if (xdir) { ua[off] = (T)avg; if (us) us[off] = (T)u; }
    else { va[off] = (T)avg; if (vs) vs[off] = (T)u; }
  }
}

// ------------------------------------------------------------------ diagnostics
// Compensated (Neumaier) FP64 sums.
struct KSum { double s, c; };
__host__ __device__ __forceinline__ void kadd(KSum &a, double x) {
  const double t = a.s + x;
  if (fabs(a.s) >= fabs(x)) a.c += (a.s - t) + x; else a.c += (x - t) + a.s;
  a.s = t;
}

template <typename T>
__global__ void k_mass(const Solve *solves, int solve, int tracer, double *partial) {
  const Solve &S = solves[solve];
  const csadv_solve_desc &D = S.d;
  const int N = D.N, ld = D.ld;
  const T *Q = reinterpret_cast<const T *>((D.ctl[1] & 1) ? D.q1 : D.q0) + tracer * S.tstride;
  const double cf = D.corr ? D.corr[(D.ctl[1] & 1) * D.ntr + tracer] : 0.0;
  const double dxdy = D.dx * D.dy;
  KSum acc{0.0, 0.0};
  const int64_t tot = (int64_t)6 * N * N;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * blockDim.x) {
    const int p = (int)(idx / ((int64_t)N * N));
    const int rem = (int)(idx - (int64_t)p * N * N);
    const int j = rem / N + NG, i = rem % N + NG;
    const double g = D.gc64[(int64_t)j * ld + i];
    const double q = (double)Q[p * S.pstride + (int64_t)j * ld + i] + cf * g;
    kadd(acc, q * g * dxdy);
  }
  __shared__ double ss[BLOCK], sc[BLOCK];
  ss[threadIdx.x] = acc.s;
  sc[threadIdx.x] = acc.c;
  __syncthreads();
  if (threadIdx.x == 0) {
    KSum b{0.0, 0.0};
    for (int t = 0; t < (int)blockDim.x; ++t) { kadd(b, ss[t]); kadd(b, sc[t]); }
    partial[2 * blockIdx.x] = b.s;
    partial[2 * blockIdx.x + 1] = b.c;
  }
}

template <typename T>
__global__ void k_norms(const Solve *solves, int solve, int tracer, const double *qex, double *partial) {
  const Solve &S = solves[solve];
  const csadv_solve_desc &D = S.d;
  const int N = D.N, ld = D.ld;
  const T *Q = reinterpret_cast<const T *>((D.ctl[1] & 1) ? D.q1 : D.q0) + tracer * S.tstride;
  const double cf = D.corr ? D.corr[(D.ctl[1] & 1) * D.ntr + tracer] : 0.0;
  KSum s1{0.0, 0.0}, s2{0.0, 0.0};
  double mx = 0.0;
  const int64_t tot = (int64_t)6 * N * N;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * blockDim.x) {
    const int p = (int)(idx / ((int64_t)N * N));
    const int rem = (int)(idx - (int64_t)p * N * N);
    const int j = rem / N + NG, i = rem % N + NG;
    const double q = (double)Q[p * S.pstride + (int64_t)j * ld + i] + cf * D.gc64[(int64_t)j * ld + i];
    const double e = fabs(qex[idx] - q);
    kadd(s1, e);
    kadd(s2, e * e);
    mx = fmax(mx, e);
  }
  __shared__ double a[BLOCK], b[BLOCK], c[BLOCK], d[BLOCK], m[BLOCK];
  a[threadIdx.x] = s1.s; b[threadIdx.x] = s1.c; c[threadIdx.x] = s2.s; d[threadIdx.x] = s2.c; m[threadIdx.x] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    KSum t1{0.0, 0.0}, t2{0.0, 0.0};
    double mm = 0.0;
    for (int t = 0; t < (int)blockDim.x; ++t) { kadd(t1, a[t]); kadd(t1, b[t]); kadd(t2, c[t]); kadd(t2, d[t]); mm = fmax(mm, m[t]); }
    partial[5 * blockIdx.x + 0] = t1.s; partial[5 * blockIdx.x + 1] = t1.c;
    partial[5 * blockIdx.x + 2] = t2.s; partial[5 * blockIdx.x + 3] = t2.c;
    partial[5 * blockIdx.x + 4] = mm;
  }
}

#include "csadv_general.hip"

// ------------------------------------------------------------------ host side
struct csadv_s {
  std::vector<csadv_solve_desc> desc;
  std::vector<Solve> host;
  Solve *dev = nullptr;           // device descriptors (caller-visible scratch is not used)
  int4 *work = nullptr;           // step work items
  FieldSet *qsets = nullptr;      // ghost-fill sets for the tracers
  FieldSet *wsets = nullptr;      // ghost-fill sets for the lat-lon centre winds
  int nwork = 0;
  int tile = 0;
  int dtype = 0;
  int recon = 3;
  bool rk2 = false;               // dp = 2 (separate upwind-side winds)
  int split = 1;                  // splitting of every solve (its metric treatment follows)
  bool general = false;           // ET-S72 / ET-PL07 / MF-AF: the general step
  bool wind1 = false;             // every solve RK1 with a one-field time-dependent wind: the
                                  // step kernel forms its winds (no per-step k_wind_avg)
  double *red = nullptr;          // small reduction scratch (mass / norms)
  std::vector<int> done;          // host copy of the steps completed per solve (ctl[1])
  std::vector<hipEvent_t> ev;     // profiling event pairs (start, end) around the step kernel
  int ev_used = 0;
  int graph_steps = 0;            // steps per captured graph (0 = plain launches)
  hipStream_t cap = nullptr;      // private stream used only for capture
  hipGraphExec_t gexec = nullptr; // instantiated graph of graph_steps steps
  std::string err;
};

namespace {

// tile id -> rows per wavefront segment (the strip width is always the widest that fits)
constexpr int kSegRows[] = {8, 4, 16, 32, 6, 12, 11, 5, 10, 14, 17, 23};
constexpr int kNumTiles = sizeof(kSegRows) / sizeof(kSegRows[0]);

void tiling(int N, int tile, int &strip_w, int &strips, int &seg_l, int &segs) {
  strips = (N + STRIP_MAX - 1) / STRIP_MAX;
  strip_w = (N + strips - 1) / strips;          // balanced strips of at most 58 columns
  seg_l = std::min(kSegRows[tile], N);
  segs = (N + seg_l - 1) / seg_l;
}
constexpr int kRedBlocks = 256;

int fail(csadv_t *h, int code, const std::string &msg) {
  if (h) h->err = msg;
  return code;
}

#define HIPCHK(h, call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail((h), CSADV_E_HIP, std::string(#call " failed: ") + hipGetErrorString(e_)); \
  } while (0)

int validate(const csadv_solve_desc &d, std::string &msg) {
  if (d.N < 8 || d.N > 8192) { msg = "N must be in [8, 8192]"; return CSADV_E_RANGE; }
  if (d.ntr < 1 || d.ntr > 64) { msg = "ntr must be in [1, 64]"; return CSADV_E_RANGE; }
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
  if ((d.et != 3 || d.mf == 2) && !d.work) {
    msg = "ET-S72, ET-PL07 and MF-AF run the general step, which needs desc.work (csadv_work_bytes)";
    return CSADV_E_ARG;
  }
  if ((d.et != 3 || d.mf == 2) && (d.ld != (M + 3) / 4 * 4 || d.ldu != (M + 4) / 4 * 4)) {
    msg = "the general step expects ld = round_up(N+8, 4) and ldu = round_up(N+9, 4) (the csadv_work_bytes layout)";
    return CSADV_E_ARG;
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
  if (d.et == 3 && d.mf != 2 && !d.gfix) { msg = "the fused step (ET-DG) needs gfix"; return CSADV_E_ARG; }
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

template <typename T, int R, int SP>
hipError_t launch_step_s(csadv_t *h, hipStream_t st) {
  const bool one = h->host.size() == 1;
  if (h->wind1) {                    // RK1 only
    if (one) hipLaunchKernelGGL((k_step1<T, R, false, SP, true>), dim3(h->nwork), dim3(WAVE), 0, st, h->host[0]);
    else hipLaunchKernelGGL((k_step<T, R, false, SP, true>), dim3(h->nwork), dim3(WAVE), 0, st, h->dev, h->work);
    return hipGetLastError();
  }
  if (h->rk2) {
    if (one) hipLaunchKernelGGL((k_step1<T, R, true, SP, false>), dim3(h->nwork), dim3(WAVE), 0, st, h->host[0]);
    else hipLaunchKernelGGL((k_step<T, R, true, SP, false>), dim3(h->nwork), dim3(WAVE), 0, st, h->dev, h->work);
  } else {
    if (one) hipLaunchKernelGGL((k_step1<T, R, false, SP, false>), dim3(h->nwork), dim3(WAVE), 0, st, h->host[0]);
    else hipLaunchKernelGGL((k_step<T, R, false, SP, false>), dim3(h->nwork), dim3(WAVE), 0, st, h->dev, h->work);
  }
  return hipGetLastError();
}

template <typename T, int R>
hipError_t launch_step_r(csadv_t *h, hipStream_t st) {
  if (h->split == 2) return launch_step_s<T, R, 2>(h, st);
  if (h->split == 3) return launch_step_s<T, R, 3>(h, st);
  return launch_step_s<T, R, 1>(h, st);
}

template <typename T>
hipError_t launch_step_dt(csadv_t *h, hipStream_t st) {
  switch (h->recon) {
    case 1: return launch_step_r<T, 1>(h, st);
    case 2: return launch_step_r<T, 2>(h, st);
    case 4: return launch_step_r<T, 4>(h, st);
    default: return launch_step_r<T, 3>(h, st);
  }
}

int gp_blocks(int64_t total) { return (int)std::min<int64_t>(4096, (total + BLOCK - 1) / BLOCK); }

void drop_graph(csadv_t *h) {
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  h->gexec = nullptr;
}

int step_ghosts(csadv_t *h, hipStream_t st) {
  const int ns = (int)h->desc.size();
  int maxc = 0;
  for (auto &d : h->desc) maxc = std::max(maxc, d.n_strip + d.n_corner);
  int maxtr = 1;
  for (auto &d : h->desc) maxtr = std::max(maxtr, d.ntr);
  // recipe blocks of the largest solve, then the block of the correction total
  const dim3 grid((maxc + BLOCK - 1) / BLOCK + 1, ns, maxtr);
  const bool one = ns == 1;
  if (h->dtype == CSADV_F64) {
    if (one) hipLaunchKernelGGL(k_ghost1<double>, grid, dim3(BLOCK), 0, st, h->host[0]);
    else hipLaunchKernelGGL(k_ghost<double>, grid, dim3(BLOCK), 0, st, h->dev);
  } else {
    if (one) hipLaunchKernelGGL(k_ghost1<float>, grid, dim3(BLOCK), 0, st, h->host[0]);
    else hipLaunchKernelGGL(k_ghost<float>, grid, dim3(BLOCK), 0, st, h->dev);
  }
  HIPCHK(h, hipGetLastError());
  return CSADV_OK;
}

int ghost_fill(csadv_t *h, hipStream_t st) {
  const int ns = (int)h->desc.size();
  if (h->dtype == CSADV_F64) {
    hipLaunchKernelGGL(k_ghost_edges<double>, dim3(96, ns), dim3(BLOCK), 0, st, h->qsets);
    hipLaunchKernelGGL(k_ghost_corners<double>, dim3(6, ns), dim3(64), 0, st, h->qsets);
  } else {
    hipLaunchKernelGGL(k_ghost_edges<float>, dim3(96, ns), dim3(BLOCK), 0, st, h->qsets);
    hipLaunchKernelGGL(k_ghost_corners<float>, dim3(6, ns), dim3(64), 0, st, h->qsets);
  }
  HIPCHK(h, hipGetLastError());
  return CSADV_OK;
}

int wind_avg(csadv_t *h, hipStream_t st, bool all) {
  const int ns = (int)h->desc.size();
  bool any = all;
  for (auto &d : h->desc) any = any || d.unsteady;
  if (!any) return CSADV_OK;
  int maxN = 0;
  for (auto &d : h->desc) maxN = std::max(maxN, d.N);
  const int blocks = (maxN + 2 * NG) + (maxN + 1);   // x-edge rows, then y-edge rows
  if (h->dtype == CSADV_F64)
    hipLaunchKernelGGL(k_wind_avg<double>, dim3(blocks, 6, ns), dim3(BLOCK), 0, st, h->dev, (int)all);
  else
    hipLaunchKernelGGL(k_wind_avg<float>, dim3(blocks, 6, ns), dim3(BLOCK), 0, st, h->dev, (int)all);
  HIPCHK(h, hipGetLastError());
  return CSADV_OK;
}

}  // namespace

extern "C" {

int csadv_abi_version(void) { return CSADV_ABI_VERSION; }

int csadv_partials_per_tracer(int32_t N, int32_t tile) {
  if (tile < 0 || tile >= kNumTiles || N < 1) return CSADV_E_ARG;
  int sw, ns, sl, nsg;
  tiling(N, tile, sw, ns, sl, nsg);
  return 6 * ns * nsg;
}

int csadv_work_bytes(int32_t N, int32_t dtype, int64_t *bytes) {
  if (!bytes || N < 8 || N > 8192 || (dtype != CSADV_F64 && dtype != CSADV_F32)) return CSADV_E_ARG;
  const int M = N + 2 * NG;
  const int ld = (M + 3) / 4 * 4, ldu = (M + 4) / 4 * 4;
  const gp::Work w = gp::layout(N, ld, ldu);
  const int64_t s = dtype == CSADV_F64 ? 8 : 4;
  *bytes = w.total * s + (int64_t)gp::kRedGen * 8;
  return CSADV_OK;
}

int csadv_tile_rows(int32_t tile) {
  if (tile < 0 || tile >= kNumTiles) return CSADV_E_ARG;
  return kSegRows[tile];
}

int csadv_create(const csadv_solve_desc *solves, int32_t nsolve, int32_t tile, csadv_t **out) {
  if (!out) return CSADV_E_ARG;
  *out = nullptr;
  if (!solves || nsolve < 1 || nsolve > 4096) return CSADV_E_ARG;
  if (tile < 0 || tile >= kNumTiles) return CSADV_E_ARG;
  csadv_t *h = new csadv_t();
  h->tile = tile;
  h->dtype = solves[0].dtype;
  h->recon = solves[0].recon;
  h->rk2 = solves[0].dp == 2;
  h->split = solves[0].split;
  h->general = solves[0].et != 3 || solves[0].mf == 2;
  for (int s = 0; s < nsolve; ++s) {
    std::string msg;
    int rc = validate(solves[s], msg);
    if (rc == CSADV_OK && (solves[s].dtype != h->dtype || solves[s].recon != h->recon || solves[s].split != h->split ||
                           solves[s].dp != solves[0].dp ||
                           (solves[s].et != 3 || solves[s].mf == 2) != h->general)) {
      rc = CSADV_E_ARG;
      msg = "all solves of one handle share dtype, recon, split, dp and the step kind (fused or general)";
    }
    if (rc != CSADV_OK) {
      // the handle is returned so that csadv_last_error() can explain the failure
      h->err = "solve " + std::to_string(s) + ": " + msg;
      *out = h;
      return rc;
    }
    h->desc.push_back(solves[s]);
  }
  h->done.assign(nsolve, 0);
  *out = h;
  hipError_t e;
  if ((e = hipMalloc(&h->dev, sizeof(Solve) * nsolve)) != hipSuccess ||
      (e = hipMalloc(&h->qsets, sizeof(FieldSet) * nsolve)) != hipSuccess ||
      (e = hipMalloc(&h->wsets, sizeof(FieldSet) * nsolve)) != hipSuccess ||
      (e = hipMalloc(&h->red, sizeof(double) * 5 * kRedBlocks + 64)) != hipSuccess)
    return fail(h, CSADV_E_HIP, std::string("descriptor allocation: ") + hipGetErrorString(e));
  return csadv_update(h, solves, nsolve, nullptr);
}

int csadv_update(csadv_t *h, const csadv_solve_desc *solves, int32_t nsolve, void *stream) {
  if (!h) return CSADV_E_ARG;
  if (!solves || nsolve != (int)h->desc.size()) return fail(h, CSADV_E_ARG, "solve count mismatch");
  hipStream_t st = (hipStream_t)stream;
  drop_graph(h);   // launch shapes may change with the descriptors
  h->host.clear();
  std::vector<std::vector<int4>> band(kXcds);
  std::vector<FieldSet> sets;
  for (int s = 0; s < nsolve; ++s) {
    std::string msg;
    int rc = validate(solves[s], msg);
    if (rc != CSADV_OK) return fail(h, rc, "solve " + std::to_string(s) + ": " + msg);
    h->desc[s] = solves[s];
    Solve S = derive(solves[s], h->tile);
    h->host.push_back(S);
    // segment rows are cut into kXcds contiguous bands, one per XCD: the band's halo rows and
    // its share of the panel-invariant metric then stay in that XCD's L2 for all six panels
    // (the tracers of one segment are neighbours in a band, so they share its wind and
    // metric rows in L2; the work item packs panel and tracer as p + 8 * tracer)
    for (int sg = 0; sg < S.segs; ++sg)
      for (int p = 0; p < 6; ++p)
        for (int sx = 0; sx < S.strips; ++sx)
          for (int tr = 0; tr < solves[s].ntr; ++tr)
            band[sg * kXcds / S.segs].push_back(make_int4(s, p + 8 * tr, sx, sg));
    FieldSet F{};
    F.base[0] = solves[s].q0;
    F.base[1] = solves[s].q1;
    F.cur = solves[s].ctl + 1;
    F.nfield = solves[s].ntr;
    F.N = solves[s].N;
    F.M = S.M;
    F.ld = solves[s].ld;
    F.fstride = S.tstride;
    F.pstride = S.pstride;
    F.corr = solves[s].mf == 3 ? solves[s].corr : nullptr;
    F.gc64 = solves[s].gc64;
    F.lagw = solves[s].lagw;
    F.lagk = solves[s].lagk;
    F.orient = solves[s].orient;
    sets.push_back(F);
  }
  // workgroup b runs on XCD b % kXcds: interleave the bands, padding them to equal length
  size_t blen = 0;
  for (auto &bd : band) blen = std::max(blen, bd.size());
  std::vector<int4> work;
  work.reserve(blen * kXcds);
  for (size_t i = 0; i < blen; ++i)
    for (int x = 0; x < kXcds; ++x) work.push_back(i < band[x].size() ? band[x][i] : make_int4(-1, 0, 0, 0));
  if ((int)work.size() != h->nwork) {
    if (h->work) (void)hipFree(h->work);
    h->work = nullptr;
    HIPCHK(h, hipMalloc(&h->work, sizeof(int4) * work.size()));
    h->nwork = (int)work.size();
  }
  h->wind1 = !h->general;
  for (int s = 0; s < nsolve; ++s)
    h->wind1 = h->wind1 && solves[s].unsteady && solves[s].dp == 1 && solves[s].nbasis == 1;
  HIPCHK(h, hipMemcpyAsync(h->work, work.data(), sizeof(int4) * work.size(), hipMemcpyHostToDevice, st));
  HIPCHK(h, hipMemcpyAsync(h->dev, h->host.data(), sizeof(Solve) * nsolve, hipMemcpyHostToDevice, st));
  HIPCHK(h, hipMemcpyAsync(h->qsets, sets.data(), sizeof(FieldSet) * nsolve, hipMemcpyHostToDevice, st));
  // steps already completed by each solve (a handle may be built over solves that have run)
  std::vector<int32_t> ctl(4 * nsolve);
  for (int s = 0; s < nsolve; ++s)
    HIPCHK(h, hipMemcpyAsync(ctl.data() + 4 * s, solves[s].ctl, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HIPCHK(h, hipStreamSynchronize(st));
  for (int s = 0; s < nsolve; ++s) h->done[s] = ctl[4 * s + 1];
  return CSADV_OK;
}

int csadv_wind_init(csadv_t *h, void *stream) {
  if (!h || !h->dev) return CSADV_E_STATE;
  hipStream_t st = (hipStream_t)stream;
  const int ns = (int)h->desc.size();
  std::vector<FieldSet> sets;
  for (int s = 0; s < ns; ++s) {
    const csadv_solve_desc &d = h->desc[s];
    if (!d.conv_pc || !d.conv_pu || !d.conv_pv || !d.wpc)
      return fail(h, CSADV_E_ARG, "wind init needs conv_pc, conv_pu, conv_pv and wpc");
    const Solve &S = h->host[s];
    FieldSet F{};
    F.base[0] = F.base[1] = d.wpc;
    F.cur = nullptr;
    F.nfield = 2 * d.nbasis;
    F.N = d.N;
    F.M = S.M;
    F.ld = d.ld;
    F.fstride = S.tstride;
    F.pstride = S.pstride;
    F.corr = nullptr;
    F.gc64 = d.gc64;
    F.lagw = d.lagw;
    F.lagk = d.lagk;
    F.orient = d.orient;
    sets.push_back(F);
  }
  HIPCHK(h, hipMemcpyAsync(h->wsets, sets.data(), sizeof(FieldSet) * ns, hipMemcpyHostToDevice, st));
  for (int s = 0; s < ns; ++s) {
    const csadv_solve_desc &d = h->desc[s];
    if (d.et != 3) {
      // ET-S72 / ET-PL07 have no duo-grid wind halo; RK2 copies the neighbours' winds next to
      // each side (src/edges_treatment.py:296-347)
      if (d.dp == 2)
        hipLaunchKernelGGL(gp::k_gp_wind_rk2_copy, dim3(gp_blocks((int64_t)d.nbasis * 24 * d.N)), dim3(BLOCK), 0, st,
                           h->dev, s);
      continue;
    }
    const int blocks = std::min(1024, (d.N * d.N + BLOCK - 1) / BLOCK);
    hipLaunchKernelGGL(k_wind_e2c, dim3(blocks, 6, d.nbasis), dim3(BLOCK), 0, st, h->dev, s);
    hipLaunchKernelGGL(k_ghost_edges<double>, dim3(96, 1), dim3(BLOCK), 0, st, h->wsets + s);
    hipLaunchKernelGGL(k_ghost_corners<double>, dim3(6, 1), dim3(64), 0, st, h->wsets + s);
    const int b2 = std::min(1024, (2 * (d.N + 3) * (d.N + 8) + BLOCK - 1) / BLOCK);
    hipLaunchKernelGGL(k_wind_c2e, dim3(b2, 6, d.nbasis), dim3(BLOCK), 0, st, h->dev, s);
  }
  HIPCHK(h, hipGetLastError());
  int rc = wind_avg(h, st, true);
  if (rc) return rc;
  HIPCHK(h, hipStreamSynchronize(st));
  return CSADV_OK;
}

int csadv_ghost_fill(csadv_t *h, void *stream) {
  if (!h || !h->dev) return CSADV_E_STATE;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_settle, dim3((int)h->desc.size()), dim3(BLOCK), 0, st, h->dev, 0);   // totals current
  return ghost_fill(h, st);
}

}  // extern "C"

namespace {

// ET-PL07 edge values: reconstruction of the cells next to the panel sides with the
// extrapolation, cube-edge averages, their write-back with the ghost parabolas
template <typename T, int R>
void gp_edges(csadv_t *h, hipStream_t st, int s, int tr, int fx_field, int fy_field) {
  const csadv_solve_desc &d = h->desc[s];
  const int64_t sides = (int64_t)6 * 4 * d.N;
  hipLaunchKernelGGL((gp::k_gp_extrap<T, R>), dim3(gp_blocks(sides)), dim3(BLOCK), 0, st, h->dev, s, tr, fx_field,
                     fy_field);
  hipLaunchKernelGGL((gp::k_gp_edge_avg<T>), dim3(gp_blocks(sides)), dim3(BLOCK), 0, st, h->dev, s, tr);
  hipLaunchKernelGGL((gp::k_gp_edge_ghosts<T>), dim3(gp_blocks(sides)), dim3(BLOCK), 0, st, h->dev, s, tr);
}

template <typename T>
void gp_copy_halo(csadv_t *h, hipStream_t st, int s, int tr, int mode) {
  const int M = h->desc[s].N + 2 * NG;
  hipLaunchKernelGGL((gp::k_gp_copy_gather<T>), dim3(gp_blocks((int64_t)96 * M)), dim3(BLOCK), 0, st, h->dev, s, tr,
                     mode);
  hipLaunchKernelGGL((gp::k_gp_copy_scatter<T>), dim3(gp_blocks((int64_t)96 * M)), dim3(BLOCK), 0, st, h->dev, s, tr,
                     mode);
}

// One step of every solve by the general step (src/discrete_operators.py:18-101 stage by stage).
template <typename T, int R>
int enqueue_general_r(csadv_t *h, hipStream_t st) {
  const int ns = (int)h->desc.size();
  bool any_dg = false;
  for (auto &d : h->desc) any_dg = any_dg || d.et == 3;
  if (any_dg) {                  // ET-DG solves: the full duo-grid halo of Q
    const int rc = ghost_fill(h, st);
    if (rc) return rc;
  }
  for (int s = 0; s < ns; ++s) {
    const csadv_solve_desc &d = h->desc[s];
    const int N = d.N, M = d.N + 2 * NG;
    const int64_t edges = (int64_t)6 * M * (N + 1), cells = (int64_t)6 * M * M;
    for (int tr = 0; tr < d.ntr; ++tr) {
      // ET-PL07 rewrites edge values next to the panel sides between the reconstruction and
      // the flux, which reads those from the arrays; every other edge value the flux kernel
      // reconstructs itself
      const bool fuse = d.et != 2;
      // x- and y-edge fluxes in one launch: gp_blocks(edges / 2) blocks per direction, two
      // edge items per lane
      auto flux = [&](int fx_field, int fy_field) {
        const int nbx = gp_blocks((edges + 1) / 2);
        if (fuse)
          hipLaunchKernelGGL((gp::k_gp_flux<T, R, true>), dim3(2 * nbx), dim3(BLOCK), 0, st, h->dev, s, tr, fx_field,
                             fy_field, nbx);
        else
          hipLaunchKernelGGL((gp::k_gp_flux<T, R, false>), dim3(2 * nbx), dim3(BLOCK), 0, st, h->dev, s, tr, fx_field,
                             fy_field, nbx);
      };
      if (d.et != 3) gp_copy_halo<T>(h, st, s, tr, 0);
      if (!fuse) {
        gp_edges<T, R>(h, st, s, tr, 0, 0);
      }
      flux(0, 0);
      hipLaunchKernelGGL((gp::k_gp_split<T>), dim3(gp_blocks((cells + 1) / 2)), dim3(BLOCK), 0, st, h->dev, s, tr);
      if (d.et != 3) gp_copy_halo<T>(h, st, s, tr, 1);
      // second pass with the arguments swapped: x-fluxes of Qy, y-fluxes of Qx
      if (!fuse) {
        gp_edges<T, R>(h, st, s, tr, 2, 1);
      }
      flux(2, 1);
      if (d.mf == 2) {           // MF-AF: gather the cube-edge averages; k_gp_div_update reads them
        const int64_t sides = (int64_t)6 * 4 * N;
        hipLaunchKernelGGL((gp::k_gp_flux_avg<T>), dim3(gp_blocks(sides)), dim3(BLOCK), 0, st, h->dev, s, tr);
      }
      if (d.mf == 3) {           // MF-PR: the global sum of div g sits between div and the update
        hipLaunchKernelGGL((gp::k_gp_div<T>), dim3(gp::kRedGen), dim3(BLOCK), 0, st, h->dev, s, tr);
        hipLaunchKernelGGL((gp::k_gp_update<T>), dim3(gp_blocks(cells)), dim3(BLOCK), 0, st, h->dev, s, tr);
      } else {
        if (d.mf == 2)
          hipLaunchKernelGGL((gp::k_gp_div_update<T, true>), dim3(gp_blocks((cells + 1) / 2)), dim3(BLOCK), 0, st, h->dev,
                             s, tr);
        else
          hipLaunchKernelGGL((gp::k_gp_div_update<T, false>), dim3(gp_blocks((cells + 1) / 2)), dim3(BLOCK), 0, st, h->dev,
                             s, tr);
      }
    }
    hipLaunchKernelGGL(gp::k_gp_advance, dim3(1), dim3(64), 0, st, h->dev, s);
  }
  HIPCHK(h, hipGetLastError());
  return CSADV_OK;
}

template <typename T>
int enqueue_general_t(csadv_t *h, hipStream_t st) {
  switch (h->recon) {
    case 1: return enqueue_general_r<T, 1>(h, st);
    case 2: return enqueue_general_r<T, 2>(h, st);
    case 4: return enqueue_general_r<T, 4>(h, st);
    default: return enqueue_general_r<T, 3>(h, st);
  }
}

// The kernels of one step, in order, on stream st (event pairs only when armed).
int enqueue_step(csadv_t *h, hipStream_t st, bool allow_prof) {
  if (h->general) {
    const int rc = wind_avg(h, st, false);
    if (rc) return rc;
    return h->dtype == CSADV_F64 ? enqueue_general_t<double>(h, st) : enqueue_general_t<float>(h, st);
  }
  int rc = h->wind1 ? CSADV_OK : wind_avg(h, st, false);
  if (rc) return rc;
  rc = step_ghosts(h, st);
  if (rc) return rc;
  const bool prof = allow_prof && 2 * (h->ev_used + 1) <= (int)h->ev.size();
  if (prof) HIPCHK(h, hipEventRecord(h->ev[2 * h->ev_used], st));
  hipError_t e = h->dtype == CSADV_F64 ? launch_step_dt<double>(h, st) : launch_step_dt<float>(h, st);
  if (e != hipSuccess) return fail(h, CSADV_E_HIP, std::string("step kernel: ") + hipGetErrorString(e));
  if (prof) {
    HIPCHK(h, hipEventRecord(h->ev[2 * h->ev_used + 1], st));
    ++h->ev_used;
  }
  return CSADV_OK;
}

// Capture graph_steps steps on the private stream and instantiate them.
int build_graph(csadv_t *h) {
  if (!h->cap) HIPCHK(h, hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking));
  HIPCHK(h, hipStreamBeginCapture(h->cap, hipStreamCaptureModeRelaxed));
  int rc = CSADV_OK;
  for (int k = 0; k < h->graph_steps && rc == CSADV_OK; ++k) rc = enqueue_step(h, h->cap, false);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(h->cap, &g);
  if (rc != CSADV_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(h, CSADV_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    h->gexec = nullptr;
    return fail(h, CSADV_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
  }
  return CSADV_OK;
}

}  // namespace

extern "C" {

int csadv_set_graph(csadv_t *h, int32_t steps_per_graph) {
  if (!h) return CSADV_E_ARG;
  if (steps_per_graph < 0 || steps_per_graph > 4096) return fail(h, CSADV_E_RANGE, "steps_per_graph must be in [0, 4096]");
  if (steps_per_graph != h->graph_steps) drop_graph(h);
  h->graph_steps = steps_per_graph;
  return CSADV_OK;
}

int csadv_step(csadv_t *h, int32_t nsteps, void *stream) {
  if (!h || !h->dev) return CSADV_E_STATE;
  if (nsteps < 0) return fail(h, CSADV_E_ARG, "nsteps < 0");
  hipStream_t st = (hipStream_t)stream;
  const int ns = (int)h->desc.size();
  int k = 0;
  if (h->graph_steps > 0 && h->ev.empty()) {
    for (; k + h->graph_steps <= nsteps; k += h->graph_steps) {
      if (!h->gexec) {
        const int rc = build_graph(h);
        if (rc) return rc;
      }
      HIPCHK(h, hipGraphLaunch(h->gexec, st));
    }
  }
  for (; k < nsteps; ++k) {
    const int rc = enqueue_step(h, st, true);
    if (rc) return rc;
  }
  for (int s = 0; s < ns; ++s) {
    const int stop = h->desc[s].stop;
    h->done[s] += nsteps;
    if (stop > 0 && h->done[s] > stop) h->done[s] = std::max(stop, h->done[s] - nsteps);
  }
  return CSADV_OK;
}

int csadv_finalize(csadv_t *h, int32_t solve, void *stream) {
  if (!h || !h->dev) return CSADV_E_STATE;
  hipStream_t st = (hipStream_t)stream;
  const int ns = (int)h->desc.size();
  if (solve < -1 || solve >= ns) return fail(h, CSADV_E_ARG, "finalize: bad solve index");
  const int first = solve < 0 ? 0 : solve, count = solve < 0 ? ns : 1;
  hipLaunchKernelGGL(k_settle, dim3(count), dim3(BLOCK), 0, st, h->dev, first);
  if (h->dtype == CSADV_F64)
    hipLaunchKernelGGL(k_finalize<double>, dim3(256, count), dim3(BLOCK), 0, st, h->dev, first);
  else
    hipLaunchKernelGGL(k_finalize<float>, dim3(256, count), dim3(BLOCK), 0, st, h->dev, first);
  hipLaunchKernelGGL(k_clear_corr, dim3(count), dim3(64), 0, st, h->dev, first);
  HIPCHK(h, hipGetLastError());
  return CSADV_OK;
}

int csadv_mass(csadv_t *h, int32_t solve, int32_t tracer, double *out, void *stream) {
  if (!h || !h->dev) return CSADV_E_STATE;
  if (solve < 0 || solve >= (int)h->desc.size() || tracer < 0 || tracer >= h->desc[solve].ntr || !out)
    return fail(h, CSADV_E_ARG, "mass: bad solve / tracer / output");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_settle, dim3(1), dim3(BLOCK), 0, st, h->dev, solve);
  if (h->dtype == CSADV_F64)
    hipLaunchKernelGGL(k_mass<double>, dim3(kRedBlocks), dim3(BLOCK), 0, st, h->dev, solve, tracer, h->red);
  else
    hipLaunchKernelGGL(k_mass<float>, dim3(kRedBlocks), dim3(BLOCK), 0, st, h->dev, solve, tracer, h->red);
  HIPCHK(h, hipGetLastError());
  std::vector<double> part(2 * kRedBlocks);
  HIPCHK(h, hipMemcpyAsync(part.data(), h->red, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st));
  HIPCHK(h, hipStreamSynchronize(st));
  KSum acc{0.0, 0.0};
  for (int b = 0; b < kRedBlocks; ++b) { kadd(acc, part[2 * b]); kadd(acc, part[2 * b + 1]); }
  out[0] = acc.s + acc.c;
  return CSADV_OK;
}

int csadv_remap_latlon(csadv_t *h, int32_t solve, int32_t tracer, const double *field, int32_t nlat, int32_t nlon,
                       double *out, int32_t *ix, int32_t *jy, int32_t *panel, void *stream) {
  if (!h || !h->dev) return CSADV_E_STATE;
  if (solve < 0 || solve >= (int)h->desc.size() || tracer < 0 || tracer >= h->desc[solve].ntr)
    return fail(h, CSADV_E_ARG, "remap: bad solve / tracer");
  if (nlat < 2 || nlon < 2 || (int64_t)nlat * nlon > ((int64_t)1 << 31))
    return fail(h, CSADV_E_RANGE, "remap: nlat, nlon must be >= 2 and nlat * nlon <= 2^31");
  hipStream_t st = (hipStream_t)stream;
  if (!field) hipLaunchKernelGGL(k_settle, dim3(1), dim3(BLOCK), 0, st, h->dev, solve);   // totals current
  const int blocks = (int)std::min<int64_t>(4096, ((int64_t)nlat * nlon + BLOCK - 1) / BLOCK);
  if (h->dtype == CSADV_F64)
    hipLaunchKernelGGL(k_remap_latlon<double>, dim3(blocks), dim3(BLOCK), 0, st, h->dev, solve, tracer, field, nlat,
                       nlon, out, ix, jy, panel);
  else
    hipLaunchKernelGGL(k_remap_latlon<float>, dim3(blocks), dim3(BLOCK), 0, st, h->dev, solve, tracer, field, nlat,
                       nlon, out, ix, jy, panel);
  HIPCHK(h, hipGetLastError());
  return CSADV_OK;
}

int csadv_norms(csadv_t *h, int32_t solve, int32_t tracer, const double *qexact, double *out, void *stream) {
  if (!h || !h->dev) return CSADV_E_STATE;
  if (solve < 0 || solve >= (int)h->desc.size() || tracer < 0 || tracer >= h->desc[solve].ntr || !out || !qexact)
    return fail(h, CSADV_E_ARG, "norms: bad solve / tracer / pointer");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_settle, dim3(1), dim3(BLOCK), 0, st, h->dev, solve);
  if (h->dtype == CSADV_F64)
    hipLaunchKernelGGL(k_norms<double>, dim3(kRedBlocks), dim3(BLOCK), 0, st, h->dev, solve, tracer, qexact, h->red);
  else
    hipLaunchKernelGGL(k_norms<float>, dim3(kRedBlocks), dim3(BLOCK), 0, st, h->dev, solve, tracer, qexact, h->red);
  HIPCHK(h, hipGetLastError());
  std::vector<double> part(5 * kRedBlocks);
  HIPCHK(h, hipMemcpyAsync(part.data(), h->red, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st));
  HIPCHK(h, hipStreamSynchronize(st));
  KSum s1{0.0, 0.0}, s2{0.0, 0.0};
  double mx = 0.0;
  for (int b = 0; b < kRedBlocks; ++b) {
    kadd(s1, part[5 * b]); kadd(s1, part[5 * b + 1]);
    kadd(s2, part[5 * b + 2]); kadd(s2, part[5 * b + 3]);
    mx = std::max(mx, part[5 * b + 4]);
  }
  const double n = 6.0 * h->desc[solve].N * h->desc[solve].N;
  out[0] = mx;
  out[1] = (s1.s + s1.c) / n;
  out[2] = std::sqrt((s2.s + s2.c) / n);
  return CSADV_OK;
}

int csadv_current(csadv_t *h, int32_t solve) {
  if (!h || solve < 0 || solve >= (int)h->done.size()) return CSADV_E_ARG;
  return h->done[solve] & 1;
}

int csadv_profile(csadv_t *h, int32_t capacity) {
  if (!h) return CSADV_E_ARG;
  if (capacity < 0 || capacity > (1 << 16)) return fail(h, CSADV_E_RANGE, "profile capacity");
  for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
  h->ev.clear();
  h->ev_used = 0;
  for (int i = 0; i < 2 * capacity; ++i) {
    hipEvent_t e;
    HIPCHK(h, hipEventCreate(&e));
    h->ev.push_back(e);
  }
  return CSADV_OK;
}

int csadv_profile_read(csadv_t *h, double *total_ms, int32_t *launches) {
  if (!h || !total_ms || !launches) return CSADV_E_ARG;
  double tot = 0.0;
  for (int k = 0; k < h->ev_used; ++k) {
    HIPCHK(h, hipEventSynchronize(h->ev[2 * k + 1]));
    float ms = 0.f;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev[2 * k], h->ev[2 * k + 1]));
    tot += ms;
  }
  *total_ms = tot;
  *launches = h->ev_used;
  h->ev_used = 0;
  return CSADV_OK;
}

// Mean duration of the step kernel over reps back-to-back launches captured in one graph
// and bracketed by one event pair (measurement only).  The step kernel takes its step index
// from ctl[2], which only the ghost fill advances, so every repeat reads the same input and
// writes the same output: the state afterwards is that of a single launch.
int csadv_time_step_kernel(csadv_t *h, int32_t reps, void *stream, double *us_per_launch) {
  if (!h || !us_per_launch || reps < 1 || reps > 4096) return CSADV_E_ARG;
  if (h->general) return fail(h, CSADV_E_STATE, "the general step has no single step kernel to time");
  hipStream_t st = (hipStream_t)stream;
  if (!h->cap) HIPCHK(h, hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking));
  HIPCHK(h, hipStreamBeginCapture(h->cap, hipStreamCaptureModeRelaxed));
  hipError_t e = hipSuccess;
  for (int k = 0; k < reps && e == hipSuccess; ++k)
    e = h->dtype == CSADV_F64 ? launch_step_dt<double>(h, h->cap) : launch_step_dt<float>(h, h->cap);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(h->cap, &g);
  if (e != hipSuccess || ec != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(h, CSADV_E_HIP, std::string("timing capture: ") + hipGetErrorString(e != hipSuccess ? e : ec));
  }
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(h, CSADV_E_HIP, std::string("timing graph: ") + hipGetErrorString(e));
  hipEvent_t a = nullptr, b = nullptr;
  float ms = 0.f;
  if ((e = hipEventCreate(&a)) == hipSuccess && (e = hipEventCreate(&b)) == hipSuccess &&
      (e = hipGraphLaunch(x, st)) == hipSuccess &&                          // warm-up replay
      (e = hipEventRecord(a, st)) == hipSuccess && (e = hipGraphLaunch(x, st)) == hipSuccess &&
      (e = hipEventRecord(b, st)) == hipSuccess && (e = hipEventSynchronize(b)) == hipSuccess)
    e = hipEventElapsedTime(&ms, a, b);
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  (void)hipGraphExecDestroy(x);
  if (e != hipSuccess) return fail(h, CSADV_E_HIP, std::string("timing run: ") + hipGetErrorString(e));
  *us_per_launch = 1e3 * (double)ms / reps;
  return CSADV_OK;
}

const char *csadv_last_error(csadv_t *h) { return h ? h->err.c_str() : "null handle"; }

void csadv_destroy(csadv_t *h) {
  if (!h) return;
  if (h->dev) (void)hipFree(h->dev);
  if (h->work) (void)hipFree(h->work);
  if (h->qsets) (void)hipFree(h->qsets);
  if (h->wsets) (void)hipFree(h->wsets);
  if (h->red) (void)hipFree(h->red);
  for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
  drop_graph(h);
  if (h->cap) (void)hipStreamDestroy(h->cap);
  delete h;
}

}  // extern "C"
