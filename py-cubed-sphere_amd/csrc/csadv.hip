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
//   gp::k_gp_* (csadv_general.hpp)    the general step (ET-S72, ET-PL07): one kernel per
//                                     stage of src/discrete_operators.py:18-101
//   k_wind_e2c / k_wind_c2e           duo-grid vector ghost fill (src/interpolation.py:347-532)
//   k_wind_avg                        time-averaged winds (src/averaged_velocity.py:14-62)
//   k_settle / k_finalize / k_clear_corr, k_diag / k_diag_final   state read-out and diagnostics
// The fused step kernel lives in csadv_step.hpp (instantiated by step_<dtype>_r<recon>.hip), the
// general step in csadv_general.hpp (general_<dtype>_s<split>.hip), the host-side planning in
// csadv_plan.cpp.
#include "csadv_internal.hpp"

#include <cstdio>
#include <cstring>


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

// blk: the block of BLOCK ghost cells, or -1 for the tracer's correction-total block; tr: tracer.
template <typename T>
__device__ __forceinline__ void ghost_cells(const Solve &S, double *red, int blk, int tr) {
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
  // the tracer index passes through the asm (unchanged), so it exists only once every field is
  // in registers
  asm("" : "=s"(tr) : "0"(tr), "s"(ntr), "s"(mf), "s"(stop), "s"(n_strip), "s"(n_corner), "s"(part_n),
      "s"(tstride), "s"(dt), "s"(a2), "s"(ctl), "s"(corr), "s"(partials), "s"(q0), "s"(q1), "s"(strip_src),
      "s"(strip_tgt), "s"(corner_src), "s"(corner_tgt), "s"(strip_w), "s"(corner_w), "s"(blk));
  if (tr >= ntr) return;
  const int steps = ctl[1];            // requested before the recipes, whose wait it would join
  const int halted = ctl[0];           // set by a diagnostics call whose blow-up limit was exceeded
  const bool done = (stop > 0 && steps >= stop) || halted != 0;
  if (blk < 0) {                       // the correction total after the last step
    if (mf != 3) return;
    // the partials and both stored totals are requested with the control words, not after
    // them: no load of this block waits for another
    const double c0 = corr[tr], c1 = corr[ntr + tr];
    const int pending = ctl[3];
    const double sum = partial_sum(partials + (int64_t)tr * part_n, part_n, red);
    if (done) return;
    if (threadIdx.x == 0 && pending) corr[(steps & 1) * ntr + tr] = corr_total(dt, a2, c0, c1, sum, 1, steps);
    return;
  }
  const int ncell = n_strip + n_corner;
  if (blk * BLOCK >= ncell && blk != 0) return;
  const int idx = blk * BLOCK + (int)threadIdx.x;
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
  const bool lead = blk == 0 && tr == 0 && threadIdx.x == 0;
  if (done) {       // solve finished (k_step reads ctl[2] and skips it too) or halted (ctl[2] < 0)
    if (lead) ctl[2] = halted ? -1 : steps;
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

// A handle of several solves: workgroup b takes item b of the ghost work list (ghost_list).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_ghost(const Solve *solves, const int4 *gwork) {
  __shared__ double red[BLOCK];
  const int4 it = gwork[blockIdx.x];
  ghost_cells<T>(solves[it.x], red, it.z, it.y);
}

// A handle of one solve: the descriptor travels in the kernel arguments.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_ghost1(const Solve one) {
  __shared__ double red[BLOCK];
  // grid (cell blocks + 1, 1, tracers): the last block of each layer is the correction total's
  ghost_cells<T>(one, red, blockIdx.x == gridDim.x - 1 ? -1 : (int)blockIdx.x, (int)blockIdx.z);
}

// Fold the correction totals (interior) and the FP32 state offsets (every cell: a ghost value is
// a weighted sum with weights summing to 1, so it takes the offset as it is) into the current
// buffer; k_clear_corr then clears them.
template <typename T>
__global__ void k_finalize(Solve *solves, int first) {
  Solve &S = solves[first + blockIdx.y];
  csadv_solve_desc &D = S.d;
  if (D.mf != 3 && !D.offset) return;
  const int N = D.N, M = S.M, ld = D.ld;
  T *Q = reinterpret_cast<T *>((D.ctl[1] & 1) ? D.q1 : D.q0);
  const int64_t tot = (int64_t)6 * M * M, mm = (int64_t)M * M;
  for (int tr = 0; tr < D.ntr; ++tr) {
    const double cf = D.mf == 3 ? D.corr[(D.ctl[1] & 1) * D.ntr + tr] : 0.0;
    const double off = D.offset ? D.offset[tr] : 0.0;
    if (cf == 0.0 && off == 0.0) continue;
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * blockDim.x) {
      const int p = (int)(idx / mm);
      const int rem = (int)(idx - p * mm);
      const int j = rem / M, i = rem % M;
      const bool interior = i >= NG && i < N + NG && j >= NG && j < N + NG;
      if (!interior && off == 0.0) continue;
      T &v = Q[tr * S.tstride + p * S.pstride + (int64_t)j * ld + i];
      const double c = interior ? cf * D.gc64[(int64_t)j * ld + i] : 0.0;
      v = (T)(((double)v + off) + c);
    }
  }
}

__global__ void k_clear_corr(Solve *solves, int first) {
  csadv_solve_desc &D = solves[first + blockIdx.x].d;
  if (D.mf == 3)
    for (int k = threadIdx.x; k < 2 * D.ntr; k += blockDim.x) D.corr[k] = 0.0;
  if (D.offset)
    for (int k = threadIdx.x; k < D.ntr; k += blockDim.x) D.offset[k] = 0.0;
}

// ------------------------------------------------------------------ MF-AF on the fused step
// The fused step updates every cell with its own second-pass fluxes and, with MF-AF, keeps the
// fluxes through the 24 panel sides in work ([tracer][panel][E, W, N, S][N]); this kernel
// replaces them by the cube-edge averages in the boundary cells of the new state (af_fix_cell).
// grid: x = boundary cells of the 6 panels, y = solve, z = tracer.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_af_fix(const Solve *solves) {
  const Solve &S = solves[blockIdx.y];
  const csadv_solve_desc &D = S.d;
  const int tr = blockIdx.z;
  if (D.mf != 2 || tr >= D.ntr) return;
  const int run = D.ctl[2];                 // the step k_ghost started (< 0: the solve is halted)
  if (run < 0 || (D.stop > 0 && run >= D.stop)) return;
  af_fix_cell<T>(S, tr, run, reinterpret_cast<const T *>(D.work) + (int64_t)tr * 24 * D.N, blockIdx.x * BLOCK + threadIdx.x);
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
  const double off = D.offset ? D.offset[tracer] : 0.0;
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
        v = ((double)Q[p * S.pstride + (int64_t)cj * ld + ci] + off) + cf * D.gc64[(int64_t)cj * ld + ci];
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
  if ((D.stop > 0 && D.ctl[1] >= D.stop) || D.ctl[0]) return;       // solve finished or halted
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
__device__ __forceinline__ void kmerge(KSum &a, const KSum &b) {
  kadd(a, b.s);
  a.c += b.c;
}

// Butterfly reductions over the 64 lanes of a wavefront (cross-lane shuffles, no LDS): lane 0
// ends with the wavefront's total, combined in a fixed order.
__device__ __forceinline__ void wave_ksum(KSum &a) {
  for (int off = WAVE / 2; off > 0; off >>= 1) {
    const KSum b{__shfl_xor(a.s, off, WAVE), __shfl_xor(a.c, off, WAVE)};
    kmerge(a, b);
  }
}
__device__ __forceinline__ double wave_max(double v) {
  for (int off = WAVE / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, WAVE));
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
  return v;
}

// The 8 words of a diagnostics partial: mass (sum, compensation), sum |e| (sum, comp.),
// sum e^2 (sum, comp.), max |e| and the count of NaN values of e (fmax passes over NaN).
// Wavefront totals by shuffles, then thread 0 merges the block's wavefronts in index order:
// the result does not depend on scheduling.
__device__ __forceinline__ void diag_block(KSum m, KSum a1, KSum a2, double mx, double nan, double *o) {
  __shared__ double red[BLOCK / WAVE][kDiagWords];
  wave_ksum(m);
  wave_ksum(a1);
  wave_ksum(a2);
  mx = wave_max(mx);
  nan = wave_sum(nan);
  const int w = threadIdx.x / WAVE;
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    red[w][0] = m.s; red[w][1] = m.c; red[w][2] = a1.s; red[w][3] = a1.c;
    red[w][4] = a2.s; red[w][5] = a2.c; red[w][6] = mx; red[w][7] = nan;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    KSum M{red[0][0], red[0][1]}, A1{red[0][2], red[0][3]}, A2{red[0][4], red[0][5]};
    double X = red[0][6], B = red[0][7];
    for (int k = 1; k < BLOCK / WAVE; ++k) {
      kmerge(M, KSum{red[k][0], red[k][1]});
      kmerge(A1, KSum{red[k][2], red[k][3]});
      kmerge(A2, KSum{red[k][4], red[k][5]});
      X = fmax(X, red[k][6]);
      B += red[k][7];
    }
    o[0] = M.s; o[1] = M.c; o[2] = A1.s; o[3] = A1.c; o[4] = A2.s; o[5] = A2.c; o[6] = X; o[7] = B;
  }
}

// Diagnostics of tracers tracer0 .. tracer0 + gridDim.y - 1 of one solve (grid: kDiagBlocks x
// tracers): per block the compensated FP64 sums of the mass Q g dx dy (src/diagnostics.py:14-26)
// and of |e| and e^2, and the max |e| (src/errors.py:99-113), with e = qexact - Q when the exact
// field is given ([tracers][6][N][N] FP64) and e = Q otherwise.  Q = S + C g (the pending
// mass-fixer total; + the FP32 state offset), read from the solve's current buffer.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_diag(const Solve *solves, int solve, int tracer0, const double *qex,
                                                double *part) {
  const Solve &S = solves[solve];
  const csadv_solve_desc &D = S.d;
  const int tr = tracer0 + blockIdx.y;
  const int N = D.N, ld = D.ld;
  const int cur = D.ctl[1] & 1;
  const T *Q = reinterpret_cast<const T *>(cur ? D.q1 : D.q0) + tr * S.tstride;
  const double cf = (D.corr && D.mf == 3) ? D.corr[cur * D.ntr + tr] : 0.0;
  const double off = D.offset ? D.offset[tr] : 0.0;
  const double dxdy = D.dx * D.dy;
  const int64_t cells = (int64_t)6 * N * N, nn = (int64_t)N * N;
  const double *qe = qex ? qex + (int64_t)blockIdx.y * cells : nullptr;
  KSum m{0.0, 0.0}, a1{0.0, 0.0}, a2{0.0, 0.0};
  double mx = 0.0, nan = 0.0;
  for (int64_t idx = blockIdx.x * (int64_t)BLOCK + threadIdx.x; idx < cells; idx += (int64_t)gridDim.x * BLOCK) {
    const int p = (int)(idx / nn);
    const int rem = (int)(idx - p * nn);
    const int j = rem / N + NG, i = rem % N + NG;
    const double g = D.gc64[(int64_t)j * ld + i];
    const double q = ((double)Q[p * S.pstride + (int64_t)j * ld + i] + off) + cf * g;
    kadd(m, q * g * dxdy);
    const double e = fabs(qe ? qe[idx] - q : q);
    kadd(a1, e);
    kadd(a2, e * e);
    mx = fmax(mx, e);
    nan += e != e ? 1.0 : 0.0;
  }
  diag_block(m, a1, a2, mx, nan, part + ((int64_t)blockIdx.y * kDiagBlocks + blockIdx.x) * kDiagWords);
}

// The tracers' totals from the block partials (one block per tracer, thread t takes partial t):
// out[4 k .. 4 k + 3] = mass, Linf, L1, L2 (L1, L2 as unweighted means over the 6 N^2 cells;
// Linf is NaN when any e is, as numpy's amax gives).  limit > 0 and Linf > limit -- or a NaN --
// halt the solve (ctl[0] = 1): its step kernels leave it untouched from then on, as the
// reference stops a run whose error exceeds 100 (src/output.py:44-47).
__global__ __launch_bounds__(BLOCK) void k_diag_final(Solve *solves, int solve, const double *part, double limit,
                                                      double *out) {
  static_assert(kDiagBlocks == BLOCK, "one partial per thread");
  csadv_solve_desc &D = solves[solve].d;
  const double *o = part + ((int64_t)blockIdx.x * kDiagBlocks + threadIdx.x) * kDiagWords;
  __shared__ double tot[kDiagWords];
  diag_block(KSum{o[0], o[1]}, KSum{o[2], o[3]}, KSum{o[4], o[5]}, o[6], o[7], tot);
  __syncthreads();
  if (threadIdx.x == 0) {
    const double n = 6.0 * D.N * D.N;
    const double linf = tot[7] > 0.0 ? __builtin_nan("") : tot[6];
    double *r = out + 4 * blockIdx.x;
    r[0] = tot[0] + tot[1];
    r[1] = linf;
    r[2] = (tot[2] + tot[3]) / n;
    r[3] = sqrt((tot[4] + tot[5]) / n);
    if (limit > 0.0 && (linf > limit || tot[7] > 0.0)) D.ctl[0] = 1;
  }
}

// general-step kernels that are not templates (the templates: csadv_general.hpp)
namespace gp {

// ET-S72 / ET-PL07 wind halo for RK2 (src/edges_treatment.py:305-347), on every wind basis
// field: the ghost edge one outside each panel side takes the neighbour's normal wind at its
// edge one inside the shared edge, times the axis sign across the edge.
__global__ void k_gp_wind_rk2_copy(const Solve *solves, int s) {
  const Solve &S = solves[s];
  const csadv_solve_desc &D = S.d;
  const int ld = D.ld, ldu = D.ldu, i0 = NG, iend = D.N + NG;
  const int64_t bu = 6 * S.pstride_u, bv = 6 * S.pstride_v;
  const unsigned N = D.N, total = (unsigned)D.nbasis * 24u * N;
  for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const unsigned u = t / N;                        // u = (b * 6 + p) * 4 + side
    const int k = (int)(t % N) + i0;
    const int side = u & 3, p = (u >> 2) % 6, b = (u >> 2) / 6;
    double *U = D.wbu + b * bu, *V = D.wbv + b * bv;
    const int32_t *o = D.orient + (p * 4 + side) * 7;
    int q, i, j, sq, sign;
    halo_src(o, 0, k, q, i, j);
    across(o, side, sq, sign);
    // the neighbour's normal wind one edge inside the shared edge, x-edges (E/W) or y-edges (N/S);
    // both addresses are formed and one is selected
    const int ne = (sq == 0 || sq == 2) ? iend - 1 : i0 + 1;
    const double *nu = U + q * S.pstride_u + (int64_t)j * ldu + ne;
    const double *nv = V + q * S.pstride_v + (int64_t)ne * ld + i;
    const double nb = *(sq < 2 ? nu : nv);
    const double val = sign > 0 ? nb : -nb;
    // our ghost edge one outside the side
    const int te = (side == 0 || side == 2) ? iend + 1 : i0 - 1;
    double *tu = U + p * S.pstride_u + (int64_t)k * ldu + te;
    double *tv = V + p * S.pstride_v + (int64_t)te * ld + k;
    *(side < 2 ? tu : tv) = val;
  }
}

}  // namespace gp



int fail(csadv_t *h, int code, const std::string &msg) {
  if (h) h->err = msg;
  return code;
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

namespace {

template <typename T>
hipError_t launch_step_dt(csadv_t *h, hipStream_t st) {
  switch (h->recon) {
    case 1: return launch_step_r<T, 1>(h, st);
    case 2: return launch_step_r<T, 2>(h, st);
    case 4: return launch_step_r<T, 4>(h, st);
    default: return launch_step_r<T, 3>(h, st);
  }
}

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
    else hipLaunchKernelGGL(k_ghost<double>, dim3(h->ngwork), dim3(BLOCK), 0, st, h->dev, h->gwork);
  } else {
    if (one) hipLaunchKernelGGL(k_ghost1<float>, grid, dim3(BLOCK), 0, st, h->host[0]);
    else hipLaunchKernelGGL(k_ghost<float>, dim3(h->ngwork), dim3(BLOCK), 0, st, h->dev, h->gwork);
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

int csadv_partials_per_tracer(int32_t N, int32_t tile) { return partials_per_tracer(N, tile); }

int csadv_work_bytes(int32_t N, int32_t dtype, int64_t *bytes) { return work_bytes(N, dtype, bytes); }

int csadv_tile_rows(int32_t tile) { return tile_rows(tile); }

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
  h->general = is_general(solves[0]);
  int maxtr = 1;
  for (int s = 0; s < nsolve; ++s) {
    std::string msg;
    const int rc = validate(solves[s], msg);
    if (rc != CSADV_OK) {
      // the handle is returned so that csadv_last_error() can explain the failure
      h->err = "solve " + std::to_string(s) + ": " + msg;
      *out = h;
      return rc;
    }
    h->desc.push_back(solves[s]);
    maxtr = std::max(maxtr, solves[s].ntr);
  }
  std::string msg;
  const int rc = check_compatible(solves, nsolve, msg);
  *out = h;
  if (rc != CSADV_OK) return fail(h, rc, msg);
  h->done.assign(nsolve, 0);
  h->red_tracers = maxtr;
  hipError_t e;
  if ((e = hipMalloc(&h->dev, sizeof(Solve) * nsolve)) != hipSuccess ||
      (e = hipMalloc(&h->qsets, sizeof(FieldSet) * nsolve)) != hipSuccess ||
      (e = hipMalloc(&h->wsets, sizeof(FieldSet) * nsolve)) != hipSuccess ||
      (e = hipMalloc(&h->red, sizeof(double) * diag_scratch_words(maxtr))) != hipSuccess)
    return fail(h, CSADV_E_HIP, std::string("descriptor allocation: ") + hipGetErrorString(e));
  return csadv_update(h, solves, nsolve, nullptr);
}

int csadv_update(csadv_t *h, const csadv_solve_desc *solves, int32_t nsolve, void *stream) {
  if (!h) return CSADV_E_ARG;
  if (!solves || nsolve != (int)h->desc.size()) return fail(h, CSADV_E_ARG, "solve count mismatch");
  hipStream_t st = (hipStream_t)stream;
  drop_graph(h);   // launch shapes may change with the descriptors
  h->host.clear();
  std::vector<FieldSet> sets;
  for (int s = 0; s < nsolve; ++s) {
    std::string msg;
    int rc = validate(solves[s], msg);
    if (rc == CSADV_OK && solves[s].ntr > h->red_tracers) {
      rc = CSADV_E_RANGE;
      msg = "ntr grew past the handle's diagnostics scratch (create a new handle)";
    }
    if (rc != CSADV_OK) return fail(h, rc, "solve " + std::to_string(s) + ": " + msg);
    h->desc[s] = solves[s];
    const Solve S = derive(solves[s], h->tile);
    h->host.push_back(S);
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
  {
    std::string msg;
    const int rc = check_compatible(solves, nsolve, msg);
    if (rc != CSADV_OK) return fail(h, rc, msg);
  }
  // workgroup b runs on XCD b % kXcds (the list interleaves one band of row segments per XCD)
  static_assert(sizeof(Item) == sizeof(int4), "work items are read as int4");
  // (the FP32 multi-solve step kernel marches two tracers per item: kStepGroup in csadv_step.hpp)
  const int group = (!h->general && h->dtype == CSADV_F32 && nsolve > 1) ? 2 : 1;
  const std::vector<Item> work = work_list(h->host, group);
  if ((int)work.size() != h->nwork) {
    if (h->work) (void)hipFree(h->work);
    h->work = nullptr;
    HIPCHK(h, hipMalloc(&h->work, sizeof(int4) * work.size()));
    h->nwork = (int)work.size();
  }
  h->mfaf = false;
  for (int s = 0; s < nsolve; ++s) h->mfaf = h->mfaf || (!h->general && solves[s].mf == 2);
  h->wind1 = !h->general;
  for (int s = 0; s < nsolve; ++s)
    h->wind1 = h->wind1 && solves[s].unsteady && solves[s].dp == 1 && solves[s].nbasis == 1;
  HIPCHK(h, hipMemcpyAsync(h->work, work.data(), sizeof(int4) * work.size(), hipMemcpyHostToDevice, st));
  if (nsolve > 1) {                  // the ghost fill of a multi-solve handle reads its own list
    const std::vector<Item> gw = ghost_list(h->host);
    if ((int)gw.size() != h->ngwork) {
      if (h->gwork) (void)hipFree(h->gwork);
      h->gwork = nullptr;
      HIPCHK(h, hipMalloc(&h->gwork, sizeof(int4) * gw.size()));
      h->ngwork = (int)gw.size();
    }
    HIPCHK(h, hipMemcpyAsync(h->gwork, gw.data(), sizeof(int4) * gw.size(), hipMemcpyHostToDevice, st));
  }
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

// Scratch of csadv_diagnostics_noalloc: the block partials of up to kScratchTracers tracers in a
// file-scope device array (256 KiB on every GPU the library runs on, held until the process
// ends).  In host code `scratch` is only a stand-in for the symbol: its device address comes
// from hipGetSymbolAddress, once per device, and only that address is handed to kernels.  Every
// kernel of a call writes all the partials it reads, so nothing has to start at zero; calls that
// overlap on one GPU (two streams at once) share the array and race.
constexpr int kScratchTracers = 16;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[(int64_t)kScratchTracers * kDiagBlocks * kDiagWords];
double *g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up on its first use there).
int scratch_address(csadv_t *h, double **ptr) {
  int dev = -1;
  HIPCHK(h, hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(h, CSADV_E_RANGE, "diagnostics scratch: device index out of range");
  if (!g_scratch[dev]) {
    void *p = nullptr;
    HIPCHK(h, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[dev] = static_cast<double *>(p);
  }
  *ptr = g_scratch[dev];
  return CSADV_OK;
}

// Device outputs of csadv_mass / csadv_norms: after the partials of the handle's largest solve.
double *diag_out(csadv_t *h) { return h->red + (int64_t)h->red_tracers * kDiagBlocks * kDiagWords; }

int diag_args(csadv_t *h, int solve, int tracer0, int ntracers, double limit, const double *out) {
  if (solve < 0 || solve >= (int)h->desc.size()) return fail(h, CSADV_E_ARG, "diagnostics: bad solve index");
  if (tracer0 < 0 || ntracers < 1 || tracer0 + ntracers > h->desc[solve].ntr)
    return fail(h, CSADV_E_ARG, "diagnostics: tracers out of range of the solve");
  if (!out) return fail(h, CSADV_E_ARG, "diagnostics: null output");
  if (!(limit >= 0.0)) return fail(h, CSADV_E_ARG, "diagnostics: blow-up limit must be >= 0 (0 = none)");
  return CSADV_OK;
}

// k_settle (the pending mass-fixer total), the block partials into scratch (diag_scratch_words
// doubles for ntracers) and the totals into out ([ntracers][4] device doubles); no host
// synchronisation and no allocation, so a graph may capture it.
int diagnostics(csadv_t *h, int solve, int tracer0, int ntracers, const double *qexact, double limit, double *out,
                double *scratch, hipStream_t st) {
  hipLaunchKernelGGL(k_settle, dim3(1), dim3(BLOCK), 0, st, h->dev, solve);
  if (h->dtype == CSADV_F64)
    hipLaunchKernelGGL(k_diag<double>, dim3(kDiagBlocks, ntracers), dim3(BLOCK), 0, st, h->dev, solve, tracer0, qexact,
                       scratch);
  else
    hipLaunchKernelGGL(k_diag<float>, dim3(kDiagBlocks, ntracers), dim3(BLOCK), 0, st, h->dev, solve, tracer0, qexact,
                       scratch);
  hipLaunchKernelGGL(k_diag_final, dim3(ntracers), dim3(BLOCK), 0, st, h->dev, solve, scratch, limit, out);
  HIPCHK(h, hipGetLastError());
  return CSADV_OK;
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
  if (h->mfaf) {   // MF-AF: the cube-edge flux averages into the boundary cells
    int maxN = 0, maxtr = 1;
    for (auto &d : h->desc) { maxN = std::max(maxN, d.N); maxtr = std::max(maxtr, d.ntr); }
    const dim3 grid((6 * 4 * (maxN - 1) + BLOCK - 1) / BLOCK, (int)h->desc.size(), maxtr);
    if (h->dtype == CSADV_F64) hipLaunchKernelGGL(k_af_fix<double>, grid, dim3(BLOCK), 0, st, h->dev);
    else hipLaunchKernelGGL(k_af_fix<float>, grid, dim3(BLOCK), 0, st, h->dev);
    HIPCHK(h, hipGetLastError());
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
  double *dout = diag_out(h);
  const int rc = diagnostics(h, solve, tracer, 1, nullptr, 0.0, dout, h->red, st);
  if (rc) return rc;
  HIPCHK(h, hipMemcpyAsync(out, dout, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(h, hipStreamSynchronize(st));
  return CSADV_OK;
}

int csadv_diagnostics(csadv_t *h, int32_t solve, int32_t tracer0, int32_t ntracers, const double *qexact,
                      double blowup_limit, double *out, void *stream) {
  if (!h || !h->dev) return CSADV_E_STATE;
  const int rc = diag_args(h, solve, tracer0, ntracers, blowup_limit, out);
  if (rc) return rc;
  return diagnostics(h, solve, tracer0, ntracers, qexact, blowup_limit, out, h->red, (hipStream_t)stream);
}

int csadv_diagnostics_noalloc(csadv_t *h, int32_t solve, int32_t tracer0, int32_t ntracers, const double *qexact,
                              double blowup_limit, double *out, void *stream) {
  if (!h || !h->dev) return CSADV_E_STATE;
  int rc = diag_args(h, solve, tracer0, ntracers, blowup_limit, out);
  if (rc) return rc;
  if (ntracers > kScratchTracers)
    return fail(h, CSADV_E_RANGE, "csadv_diagnostics_noalloc: at most " + std::to_string(kScratchTracers) +
                                      " tracers per call (its scratch is a fixed device array)");
  double *ptr = nullptr;
  rc = scratch_address(h, &ptr);
  if (rc) return rc;
  return diagnostics(h, solve, tracer0, ntracers, qexact, blowup_limit, out, ptr, (hipStream_t)stream);
}

int csadv_status(csadv_t *h, int32_t solve, int32_t *steps_done, int32_t *halted, void *stream) {
  if (!h || !h->dev) return CSADV_E_STATE;
  if (solve < 0 || solve >= (int)h->desc.size()) return fail(h, CSADV_E_ARG, "status: bad solve index");
  hipStream_t st = (hipStream_t)stream;
  int32_t ctl[2] = {0, 0};
  HIPCHK(h, hipMemcpyAsync(ctl, h->desc[solve].ctl, sizeof(ctl), hipMemcpyDeviceToHost, st));
  HIPCHK(h, hipStreamSynchronize(st));
  h->done[solve] = ctl[1];
  if (steps_done) *steps_done = ctl[1];
  if (halted) *halted = ctl[0];
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
  double *dout = diag_out(h);
  const int rc = diagnostics(h, solve, tracer, 1, qexact, 0.0, dout, h->red, st);
  if (rc) return rc;
  HIPCHK(h, hipMemcpyAsync(out, dout + 1, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(h, hipStreamSynchronize(st));
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

// Mean duration of the step kernel over reps rounds of one launch per handle, all captured in one
// graph between one event pair (measurement only): with handles over distinct state buffers
// whose total exceeds the last-level cache, every launch streams its state from HBM.  Each
// launch redoes its handle's current step, as in csadv_time_step_kernel.
int csadv_time_step_kernel_rotating(csadv_t *const *handles, int32_t nhandles, int32_t reps, void *stream,
                                    double *us_per_launch) {
  if (!handles || nhandles < 1 || !us_per_launch || reps < 1 || (int64_t)reps * nhandles > 8192) return CSADV_E_ARG;
  csadv_t *h0 = handles[0];
  if (!h0) return CSADV_E_ARG;
  for (int k = 0; k < nhandles; ++k) {
    if (!handles[k] || !handles[k]->dev) return fail(h0, CSADV_E_STATE, "rotating timing: a handle is not created");
    if (handles[k]->general) return fail(h0, CSADV_E_STATE, "the general step has no single step kernel to time");
  }
  hipStream_t st = (hipStream_t)stream;
  if (!h0->cap) HIPCHK(h0, hipStreamCreateWithFlags(&h0->cap, hipStreamNonBlocking));
  HIPCHK(h0, hipStreamBeginCapture(h0->cap, hipStreamCaptureModeRelaxed));
  hipError_t e = hipSuccess;
  for (int r = 0; r < reps && e == hipSuccess; ++r)
    for (int k = 0; k < nhandles && e == hipSuccess; ++k)
      e = handles[k]->dtype == CSADV_F64 ? launch_step_dt<double>(handles[k], h0->cap)
                                         : launch_step_dt<float>(handles[k], h0->cap);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(h0->cap, &g);
  if (e != hipSuccess || ec != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(h0, CSADV_E_HIP, std::string("timing capture: ") + hipGetErrorString(e != hipSuccess ? e : ec));
  }
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(h0, CSADV_E_HIP, std::string("timing graph: ") + hipGetErrorString(e));
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
  if (e != hipSuccess) return fail(h0, CSADV_E_HIP, std::string("timing run: ") + hipGetErrorString(e));
  *us_per_launch = 1e3 * (double)ms / ((double)reps * nhandles);
  return CSADV_OK;
}

const char *csadv_last_error(csadv_t *h) { return h ? h->err.c_str() : "null handle"; }

void csadv_destroy(csadv_t *h) {
  if (!h) return;
  if (h->dev) (void)hipFree(h->dev);
  if (h->work) (void)hipFree(h->work);
  if (h->gwork) (void)hipFree(h->gwork);
  if (h->qsets) (void)hipFree(h->qsets);
  if (h->wsets) (void)hipFree(h->wsets);
  if (h->red) (void)hipFree(h->red);
  for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
  drop_graph(h);
  if (h->cap) (void)hipStreamDestroy(h->cap);
  delete h;
}

}  // extern "C"
