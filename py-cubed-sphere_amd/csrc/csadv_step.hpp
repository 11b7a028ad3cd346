// The fused step kernel (k_step, k_step1) and its launch: both directional passes of the
// split, PPM edges, upwind fluxes, update and mass-fixer partials of one (solve, panel,
// tracer, column strip, row segment) per wavefront (src/discrete_operators.py:18-101,
// src/flux.py, src/reconstruction_1d.py).  Included by the step_<dtype>_r<recon>.hip
// translation units, one per precision and reconstruction, which instantiate launch_step_r.
#pragma once
#include <utility>

#include "csadv_internal.hpp"

// Work item = (solve, panel, tracer, strip, segment): one wavefront marches down strip_w interior
// columns (lane l <-> column ti-3+l, so the 3-cell halo of each side rides in the outer
// lanes) over seg_l interior rows.  Iteration y:
//   * column stage at y-edge e = y-3 from the windows of rows y-5 .. y-1: y-fluxes of
//     Q and of Qx, Qy of row y-4, x-flux of Qy (neighbours by DPP lane shifts), update of
//     row y-4 and its mass-fixer partial sum;
//   * row stage for row y: x-flux of Q, split half-step Qx.
// The per-row values, loaded or computed, live in rings indexed by the row's offset from the
// item's first row modulo the ring depth (the loop is unrolled by RD, a multiple of every
// depth, so every slot is a fixed register and nothing is shifted).  The inputs of row y+PF
// are requested at the end of iteration y, so their memory latency overlaps PF-1 whole
// iterations of arithmetic.
// (src/discrete_operators.py:18-101, src/flux.py, src/reconstruction_1d.py,
// src/advection_timestep.py:43)

// Rows whose inputs are in flight ahead of the row being processed.
template <typename T> constexpr int kStepPrefetch = 2;

// SPLIT is the splitting (1 AVLT, 2 L04, 3 PL07); the metric treatment follows from it, since
// the valid pairs are (AVLT or L04, MT-0) and (PL07, MT-PL07) (src/discrete_operators.py:48-82).
// NT: tracers per wavefront.  FP32 marches two tracers of one (panel, strip, segment) together:
// the loads, wind products, metric shifts and addresses of a row serve both (the FP32 launches
// are bound by issue), and every tracer's values are formed by the same expressions as alone,
// which the FP32 units fuse per expression, so the bits are the same.
template <typename T, int RECON, bool RK2, int SPLIT, bool WIND = false, bool XPULL = false, int NT = 1>
struct March {
  static constexpr int split = SPLIT;
  static constexpr int PF = kStepPrefetch<T>;
  // Two arrangements of the per-row values, by precision (measured):
  //  * EXACT (FP64, where a single solve leaves about two wavefronts per SIMD, so each
  //    wavefront's waits are exposed): every per-row value, loaded or computed, sits in a ring
  //    with one slot more than its lifetime needs, and the output row is stored after the next
  //    row's loads by a store every lane issues.  The spare slot keeps a new value's register
  //    apart from the one it replaces, whatever order the compiler gives the two stages of an
  //    iteration; without it the register allocator rotated values at the loop's back-edge, and
  //    a rotated value still in flight made the wavefront wait there for every outstanding
  //    load.  Loads and stores share one counter, so a store ahead of the next row's loads, or
  //    one behind a branch (whose issue the wait has to doubt), holds each row's wait back.
  //    Together: 6-8% off the N=384 step kernel.
  //  * otherwise (FP32, whose launches hold many wavefronts per SIMD and are bound by issue):
  //    the loaded state of a row rotates through PF slots, and the output is stored where it
  //    is formed.  The EXACT arrangement holds more registers here (126 against 96 in the
  //    batched launch's kernel), which costs a wavefront per SIMD and 5% of that launch.
  static constexpr bool EXACT = !std::is_same<T, float>::value;
  // Ring depth: a row's loaded inputs live from the end of iteration r-PF to iteration r+4, its
  // computed values from iteration r to r+5 (EXACT: one slot more).  RS: depth of the rings of
  // the loaded state (Q, gfix, y-winds), which live from the end of iteration r-PF to r.
  static constexpr int RD = PF + 4 + (EXACT ? 1 : 0);
  static constexpr int RS = EXACT ? RD : PF;
  static_assert(RD % RS == 0, "the state slot of a row follows from its ring phase");
  static_assert(NT == 1 || !EXACT, "two tracers per wavefront in FP32 only");
  static constexpr bool mt0 = SPLIT != 3;
  // ring slot of row y + j in phase K (the phase of row y is (y - y0) mod RD)
  static constexpr int sl(int K, int j) { return ((K + j) % RD + RD) % RD; }
  // ---- inputs (per wavefront; Q, Qa per tracer)
  const T *Q[NT], *Qa[NT], *U, *US, *V, *VS, *gc, *gu, *gv, *rgc, *gfix;
  // WIND: the winds of the step are formed here from the one-field FP64 basis (WU, WV) and its
  // coefficient ws at the step's time, as k_wind_avg forms them for RK1, instead of read
  const double *WU, *WV;
  double ws;
  T *Qo[NT];                       // the output panel of each tracer ([M][ld])
  __amdgpu_buffer_rsrc_t qo_buf;   // the same, for the stores of the EXACT arrangement
  int ld, ldu, M, iend, x, xc, xe, tj, ylast;
  int am2, am1, ap1, ap2;    // byte addresses of the lanes -2, -1, +1, +2 (mod 64) for lane pulls
  bool out_lane, need_col;
  T dtdx, dtdy;
  double cf[NT];
  T off[NT];                 // FP32: the tracer's state offset (Q = S + C gfix + off); FP64: 0
  // MF-AF: the second pass's fluxes through the cube edges go to fside ([4][N]: E, W, N, S of
  // this panel and tracer), for k_af_fix to replace them by the cube-edge averages; null otherwise.
  // w_lane / e_lane: the lanes on the panel's west / east cube edge (first / last strip)
  T *fside[NT];
  bool w_lane, e_lane;
  int N;
  // ---- state: rings of rows (slot of row y+j in phase K: sl(K, j); in the state rings,
  // sl(K, j) % RS).  The inputs of row y+PF are loaded at the end of iteration y straight into
  // the slot their row will hold, whose previous row has had its last use by then.
  //   loaded state, row r: qr Q as stored (S: the correction total is not in it), gfr the metric
  //                   the correction total multiplies (gfix), qar Q of the other state buffer
  //                   (first rows only); vr / vsr y-wind and its upwind-side wind at y-edge r-3
  //   loaded tables, row r: gcw metric, rgw 1/g, uw / usw x-wind and its upwind-side wind,
  //                   guw x-edge metric, gvr gv at y-edge r-2
  //   computed, row r: qw Q (corrected), xw the split half-step Qx, sw (FP32) Q as stored
  //   (per tracer: qr, qar, qw, xw, sw, the flux carries and the partial sum)
  T qr[NT][RS], gfr[RS], qar[NT][RS], vr[RS], vsr[RS];
  T qw[NT][RD], xw[NT][RD], gcw[RD], rgw[RD], uw[RD], usw[RD], guw[RD], gvr[RD];
  T sw[NT][RD];
  T g1prev[NT], g2prev[NT], vprev;
  // row y-4 as the left (upwind) cell of edge y-3: Q, Qx
  T pa_A[NT], pa_B[NT], pa_q6[NT], pb_A[NT], pb_B[NT], pb_q6[NT];
  double part[NT];

  // Inputs of row y into ring slot SLOT.  SPEC: also read Q from the other state buffer (Qa),
  // for the rows requested before the step counter has arrived.
  template <bool SPEC, int SLOT>
  __device__ __forceinline__ void load(int y) {
    const int yc = min(y, M - 1);
    const uint32_t o = (uint32_t)(yc * ld + xc), ou = (uint32_t)(yc * ldu + xe);
    if constexpr (WIND) uw[SLOT] = T(wind_of_basis(ws, ld_at(WU, ou)));
    else uw[SLOT] = ld_at(U, ou);
    if constexpr (RK2) usw[SLOT] = ld_at(US, ou);
    guw[SLOT] = ld_at(gu, ou);
    gcw[SLOT] = ld_at(gc, o);
    rgw[SLOT] = ld_at(rgc, o);
    constexpr int SS = SLOT % RS;          // state slot
    gfr[SS] = ld_at(gfix, o);
    // Q of every lane, at the clamped row and column (rows past ylast are requested too): no
    // branch around the load, and the value outside the stencils' reach is replaced by zero
    // where the row is used (see reach), not here, where the select would wait for the data
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if constexpr (EXACT) {
        qr[t][SS] = ld_at(Q[t], o);
        if constexpr (SPEC) qar[t][SS] = ld_at(Qa[t], o);
      } else {   // zeroed outside the stencils' reach here (fewer registers than where it is used)
        const bool rch = reach(y);
        const T qv = ld_at(Q[t], o);
        qr[t][SS] = rch ? qv : T(0);
        if constexpr (SPEC) {
          const T qa = ld_at(Qa[t], o);
          qar[t][SS] = rch ? qa : T(0);
        }
      }
    }
    const int ec = min(max(y - 3, 0), M - 1);
    if constexpr (WIND) vr[SS] = T(wind_of_basis(ws, ld_at(WV, (uint32_t)(ec * ld + xc))));
    else vr[SS] = ld_at(V, (uint32_t)(ec * ld + xc));
    if constexpr (RK2) vsr[SS] = ld_at(VS, (uint32_t)(ec * ld + xc));
    gvr[SLOT] = ld_at(gv, (uint32_t)((ec + 1) * ld + xc));
  }

  // whether row y at this lane's column is within the stencils' reach (the columns and rows of
  // the panel's 3-cell halo and interior; the array's outer ghost layer is never read)
  __device__ __forceinline__ bool reach(int y) const { return need_col && y >= 1 && y < iend + 3; }

  // the upwind-side winds of the row in a slot (RK1: the winds themselves)
  __device__ __forceinline__ T wind_sign(int slot) const { return RK2 ? usw[slot] : uw[slot]; }
  __device__ __forceinline__ T vwind_sign(int slot) const { return RK2 ? vsr[slot] : vr[slot]; }

  // every cell the stencils read carries the mass-fixer correction total C: Q = S + C gfix,
  // gfix being g inside the panel and, at a ghost cell, the weighted metric of the cells its
  // value was interpolated from (in the solve's precision: C gfix is ~1e-6 of q at most, so
  // its FP32 rounding is far below that of the sum, which is rounded once).  Lanes and rows
  // outside the stencils' reach hold 0 + C gfix, which no stored cell reads.
  // (with no correction total, cf = 0 and q + 0 gfix = q: no branch on the mass fixer)
  __device__ __forceinline__ T corrected(T raw, T gf, int t) const {
    if constexpr (std::is_same<T, float>::value) return (raw + off[t]) + T(cf[t]) * gf;
    else return raw + T(cf[t]) * gf;
  }

  // Iteration y in phase K.  STAGE: how much of the column stage has begun, fixed by the row's
  // offset from the item's first row y0 = tj - 3: parabolas of row y-3 from y = tj+2 on
  // (STAGE >= 1), fluxes through y-edge y-3 from tj+3 on (>= 2), the update of row y-4 from
  // tj+4 on (3).  The first seven rows are peeled with their stage; every later row is 3, so
  // the loop carries no per-row test.
  template <int K, int STAGE>
  __device__ __forceinline__ void body(const int y) {
    // slots of rows y, y-5 .. y-1, and of row y in the state rings
    constexpr int s0 = sl(K, 0), s1 = sl(K, -5), s2 = sl(K, -4), s3 = sl(K, -3), s4 = sl(K, -2), s5 = sl(K, -1);
    constexpr int ss = s0 % RS;
    T raw[NT], q[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      raw[t] = !EXACT || reach(y) ? qr[t][ss] : T(0);
      q[t] = corrected(raw[t], gfr[ss], t);
    }
    const T gvm1 = gvr[s4], gv0 = gvr[s5], gvp1 = gvr[s0];   // gv at y-edges y-4, y-3, y-2
    T out = T(0);
    // ---- column stage: parabolas of row y-3, y-edge e = y-3, output row r = y-4
    if constexpr (STAGE >= 1) {
      const T gce = gcw[s3];
      Parab<T> a[NT], b[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        a[t] = parabola<T, RECON>(qw[t][s1], qw[t][s2], qw[t][s3], qw[t][s4], qw[t][s5], gv0, gvp1, gce, mt0);
        b[t] = parabola<T, RECON>(xw[t][s1], xw[t][s2], xw[t][s3], xw[t][s4], xw[t][s5], gv0, gvp1, gce, mt0);
      }
      if constexpr (STAGE >= 2) {
        const T v = vr[ss];
        const bool vup = vwind_sign(ss) >= T(0);
        const T cy = v * dtdy;
        T g1[NT], g2[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          g1[t] = upwind<T>(vup, pa_A[t], pa_B[t], pa_q6[t], a[t], cy, gv0, mt0);     // times dt/dy
          g2[t] = upwind<T>(vup, pb_A[t], pb_B[t], pb_q6[t], b[t], cy, gv0, mt0);
        }
        if constexpr (STAGE >= 3) {
          const T gcm = gcw[s2];               // row y-4
          T d1 = T(0), d2 = T(0);
          if (split != 1) {
            d1 = gv0 * cy;
            d2 = gvm1 * (vprev * dtdy);
          }
          const T rg = rgw[s2];
          const T ur = uw[s2];
          const bool urp = wind_sign(s2) >= T(0);
          const T gur = guw[s2];
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            const T dG1 = g1prev[t] - g1[t];
            const T qy = split_update<T>(split, qw[t][s2], dG1, gcm, rg, d1, d2);
            T qy_m2, qy_m1, qy_p1, qy_p2;
            xwindow<T, XPULL>(qy, am2, am1, ap1, ap2, qy_m2, qy_m1, qy_p1, qy_p2);
            const Parab<T> c = parabola<T, RECON>(qy_m2, qy_m1, qy, qy_p1, qy_p2, gur, from_right(gur), gcm, mt0);
            const T f2 = upwind_x<T>(urp, c, ur * dtdx, gur, from_right(gur), mt0);
            const T dF2 = f2 - from_right(f2);
            const T dG2 = g2prev[t] - g2[t];
            // div = -(dF2 + dG2) / (dt g);  Q - dt div = Q + (dF2 + dG2) / g
            const T flx = dF2 + dG2;
            // the correction total stays out of the stored state: S(n+1) = S(n) + flux
            if constexpr (EXACT) {
              out = (qw[t][s2] + flx * rg) - T(cf[t]) * gcm;
              if (out_lane) part[t] -= (double)flx;   // used with MF-PR only
            } else if (out_lane) {
              st_at(Qo[t], (uint32_t)((y - 4) * ld + x), sw[t][s2] + flx * rg);
              part[t] -= (double)flx;
            }
            if (fside[0]) {   // MF-AF: cube-edge fluxes of output row r = y - 4 (f2: left edge of the lane's cell)
              const int r = y - 4;
              if (w_lane) st_at(fside[t], (uint32_t)(1 * N + r - NG), f2);
              if (e_lane) st_at(fside[t], (uint32_t)(0 * N + r - NG), f2);
              if (out_lane && r == NG) st_at(fside[t], (uint32_t)(3 * N + x - NG), g2prev[t]);        // edge j0
              if (out_lane && r == iend - 1) st_at(fside[t], (uint32_t)(2 * N + x - NG), g2[t]);      // edge jend
            }
          }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          g1prev[t] = g1[t];
          g2prev[t] = g2[t];
        }
        vprev = v;
      }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        pa_A[t] = a[t].rA; pa_B[t] = a[t].rB; pa_q6[t] = a[t].q6;
        pb_A[t] = b[t].rA; pb_B[t] = b[t].rB; pb_q6[t] = b[t].q6;
      }
    }
    // ---- row stage: row y.  Qx of the march's last row, ylast, is read by no stencil (the
    // last parabola row is ylast-3, whose window ends at ylast-1).  FP32, bound by issue, skips
    // it by a wave-uniform branch (2-3% of the launch).  FP64 does not: the branch changes the
    // code layout, and the one-solve and multi-solve kernels then round alike only when
    // multiply-adds are fused per expression, which costs the latency-bound FP64 kernel 1.5-2%
    if (!std::is_same<T, float>::value || STAGE < 3 || y < ylast) {
      const T u = uw[s0], gu0 = guw[s0], gc0 = gcw[s0];
      const bool up = wind_sign(s0) >= T(0);
      const T gup1 = from_right(gu0);
      T c1 = T(0), c2 = T(0);
      if (split != 1) {
        c1 = gup1 * (from_right(u) * dtdx);
        c2 = gu0 * (u * dtdx);
      }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        T q_m2, q_m1, q_p1, q_p2;
        xwindow<T, XPULL>(q[t], am2, am1, ap1, ap2, q_m2, q_m1, q_p1, q_p2);
        const Parab<T> c = parabola<T, RECON>(q_m2, q_m1, q[t], q_p1, q_p2, gu0, gup1, gc0, mt0);
        const T f1 = upwind_x<T>(up, c, u * dtdx, gu0, gup1, mt0);
        const T dF1 = f1 - from_right(f1);
        xw[t][s0] = split_update<T>(split, q[t], dF1, gc0, rgw[s0], c1, c2);
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      qw[t][s0] = q[t];
      if constexpr (std::is_same<T, float>::value) sw[t][s0] = raw[t];
    }
    // rows past ylast are read in bounds (clamped rows, masked Q) and never used
    load<false, sl(K, PF)>(y + PF);
    // EXACT: the output row's store, after the loads and issued by every lane (the first rows,
    // which have no output row yet, issue it too, dropped)
    if constexpr (EXACT)
      st_buf(qo_buf, STAGE >= 3 && out_lane ? (uint32_t)((y - 4) * ld + x) * (uint32_t)sizeof(T) : kNoStore, out);
  }

  template <int... J>
  __device__ __forceinline__ void first_rows(std::integer_sequence<int, J...>) {
    (load<true, J>(tj - 3 + J), ...);
  }

  // The first PF rows' inputs, requested before the step counter is known: Q from both state
  // buffers (Q = q0, Qa = q1), the rest from the step-invariant tables and winds.
  __device__ __forceinline__ void start() {
#pragma unroll
    for (int k = 0; k < RD; ++k) gcw[k] = rgw[k] = uw[k] = usw[k] = guw[k] = gvr[k] = T(0);
#pragma unroll
    for (int k = 0; k < RS; ++k) gfr[k] = vr[k] = vsr[k] = T(0);
    vprev = T(0);   // gv is complete from y = tj-1 on; the column stage starts at tj+2
#pragma unroll
    for (int t = 0; t < NT; ++t) {
#pragma unroll
      for (int k = 0; k < RD; ++k) qw[t][k] = xw[t][k] = sw[t][k] = T(0);
#pragma unroll
      for (int k = 0; k < RS; ++k) qr[t][k] = qar[t][k] = T(0);
      g1prev[t] = g2prev[t] = T(0);
      pa_A[t] = pa_B[t] = pa_q6[t] = pb_A[t] = pb_B[t] = pb_q6[t] = T(0);
      part[t] = 0.0;
    }
    // y0 = tj - 3 has phase 0; ylast - y0 = ny + 6 >= PF: every initial row exists
    first_rows(std::make_integer_sequence<int, PF>());
  }

  // The step's input buffer is known: keep the first rows' Q from it and read the rest there.
  __device__ __forceinline__ void resolve(bool odd) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
#pragma unroll
      for (int k = 0; k < PF; ++k) qr[t][k % RS] = odd ? qar[t][k % RS] : qr[t][k % RS];
      Q[t] = odd ? Qa[t] : Q[t];
    }
  }

  // row y0 + K0 + j in phase (K0 + j) % RD
  template <int K0, int STAGE, int J>
  __device__ __forceinline__ void row(int y) {
    body<(K0 + J) % RD, STAGE>(y + J);
  }
  template <int K0, int... J>
  __device__ __forceinline__ void full_rows(int y, std::integer_sequence<int, J...>) {
    (row<K0, 3, J>(y), ...);
  }
  template <int K0, int... J>
  __device__ __forceinline__ void last_rows(int y, std::integer_sequence<int, J...>) {
    ((y + J <= ylast ? row<K0, 3, J>(y) : void()), ...);
  }

  __device__ __forceinline__ void run() {
    // rows y0 .. y0+6 (ylast - y0 = ny + 6 >= 7: they all exist), the column stage starting
    int y = tj - 3;
    row<0, 0, 0>(y); row<0, 0, 1>(y); row<0, 0, 2>(y); row<0, 0, 3>(y);
    row<0, 0, 4>(y); row<0, 1, 5>(y); row<0, 2, 6>(y);
    // rows y0+7 on: the full step, RD rows per iteration
    for (y += 7; y + RD - 1 <= ylast; y += RD) full_rows<7>(y, std::make_integer_sequence<int, RD>());
    last_rows<7>(y, std::make_integer_sequence<int, RD - 1>());
  }
};

// One work item: NT tracers tr, tr+1, .. of one (panel, strip, segment).
template <typename T, int RECON, bool RK2, int SPLIT, bool WIND, bool XPULL, int NT = 1>
__device__ __forceinline__ void step_item(const Solve &S, const int4 wk) {
  using Mt = March<T, RECON, RK2, SPLIT, WIND, XPULL, NT>;
  const int lane = threadIdx.x;
  const csadv_solve_desc &D = S.d;
  const int p = wk.y & 7, tr = wk.y >> 3;    // panel, (first) tracer
  const int N = D.N;
  const int ti = NG + wk.z * S.strip_w, tj = NG + wk.w * S.seg_l;
  const int iend = N + NG;
  const int nx = min(S.strip_w, iend - ti), ny = min(S.seg_l, iend - tj);
  const int item = (p * S.segs + wk.w) * S.strips + wk.z;
  const bool mfpr = D.mf == 3;
  Mt m;
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
  m.N = N;
  m.w_lane = ti == NG && lane == 3;
  m.e_lane = ti + nx == iend && lane == 3 + nx;
#pragma unroll
  for (int t = 0; t < NT; ++t)
    m.fside[t] = D.mf == 2 ? reinterpret_cast<T *>(D.work) + (int64_t)((tr + t) * 6 + p) * 4 * N : nullptr;
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
  int64_t off[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    off[t] = (tr + t) * S.tstride + p * S.pstride;
    m.Q[t] = reinterpret_cast<const T *>(D.q0) + off[t];
    m.Qa[t] = reinterpret_cast<const T *>(D.q1) + off[t];
  }
  int step;
  if constexpr (WIND) {
    // the winds of the first rows need the step's coefficient: the step counter comes first
    step = D.ctl[2];
    const int row = min(max(step, 0), D.max_steps - 1);   // as k_wind_avg (one basis field)
    m.ws = D.coef[row * 2];
    m.WU = D.wbu + p * S.pstride_u;
    m.WV = D.wbv + p * S.pstride_v;
    m.start();
  } else {
    m.start();
    // read after the first rows are requested: a wait for any scalar load waits for all of them
    // (they return out of order), so reading these first would hold back the rows' loads
    step = D.ctl[2];
  }
  double c0[NT], c1[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    c0[t] = mfpr ? D.corr[tr + t] : 0.0;
    c1[t] = mfpr ? D.corr[D.ntr + tr + t] : 0.0;
    m.off[t] = (std::is_same<T, float>::value && D.offset) ? T(D.offset[tr + t]) : T(0);
  }
  if (step < 0 || (D.stop > 0 && step >= D.stop)) return;   // solve halted or finished
  const bool odd = step & 1;                  // the state of step n lives in q[n & 1]
  m.resolve(odd);
#pragma unroll
  for (int t = 0; t < NT; ++t) m.Qo[t] = reinterpret_cast<T *>(odd ? D.q0 : D.q1) + off[t];
  if constexpr (Mt::EXACT)
    m.qo_buf = buffer_of(reinterpret_cast<T *>(odd ? D.q0 : D.q1) + off[0], (uint32_t)((int64_t)m.M * m.ld * sizeof(T)));
#pragma unroll
  for (int t = 0; t < NT; ++t) m.cf[t] = odd ? c1[t] : c0[t];
  m.run();
  if (mfpr) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const double part = warp_sum(m.part[t]) / D.dt;   // sum of div g over the wavefront's cells
      if (lane == 0) D.partials[(int64_t)(tr + t) * S.part_per_tracer + item] = part;
    }
  }
  if (item == 0 && tr == 0 && lane == 0) {   // hand the step over to the next kernels of the stream
    D.ctl[1] = step + 1;
    if (mfpr) D.ctl[3] = 1;
  }
}

// Tracers per work item of the multi-solve step kernel: two in FP32 (the last item of a solve
// with an odd tracer count holds one), one in FP64.  Measured against one tracer per item
// (same box, 64 seeded solves with 256 tracers, 16-row tiles): 565 -> 554 us per launch.  The
// one-solve kernel keeps one tracer per item: its wavefronts drop from 5 to 3 per SIMD with
// two (94 -> 129 VGPRs), and N=768 with 4 tracers ran 67.7 -> 69.0 us at its best tile; with
// two and DPP windows (4 per SIMD) its whole step gained nothing.
template <typename T> constexpr int kStepGroup = std::is_same<T, float>::value ? 2 : 1;

// One item of NT = kStepGroup<T> tracers, or of one where the solve has no second tracer left
// (a wave-uniform choice).
template <typename T, int RECON, bool RK2, int SPLIT, bool WIND, bool XPULL>
__device__ __forceinline__ void step_group(const Solve &S, const int4 wk) {
  if constexpr (kStepGroup<T> == 2) {
    if ((wk.y >> 3) + 1 < S.d.ntr) step_item<T, RECON, RK2, SPLIT, WIND, XPULL, 2>(S, wk);
    else step_item<T, RECON, RK2, SPLIT, WIND, XPULL, 1>(S, wk);
  } else {
    step_item<T, RECON, RK2, SPLIT, WIND, XPULL, 1>(S, wk);
  }
}

template <typename T, int RECON, bool RK2, int SPLIT, bool WIND>
__global__ __launch_bounds__(WAVE) void k_step(const Solve *__restrict__ solves, const int4 *__restrict__ work) {
  const int4 wk = work[blockIdx.x];
  if (wk.x < 0) return;              // padding item of the XCD-banded work list
  step_group<T, RECON, RK2, SPLIT, WIND, false>(solves[wk.x], wk);
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
