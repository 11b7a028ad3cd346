// The fused step kernel (k_step, k_step1) and its launch: both directional passes of the
// split, PPM edges, upwind fluxes, update and mass-fixer partials of one (solve, panel,
// tracer, column strip, row segment) per wavefront (src/discrete_operators.py:18-101,
// src/flux.py, src/reconstruction_1d.py).  Included by the step_<dtype>_r<recon>.hip
// translation units, one per precision and reconstruction, which instantiate launch_step_r.
#pragma once
#include "csadv_internal.hpp"

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
  T off;                     // FP32: the tracer's state offset (Q = S + C gfix + off); FP64: 0
  // MF-AF: the second pass's fluxes through the cube edges go to fside ([4][N]: E, W, N, S of
  // this panel and tracer), for k_af_fix to replace them by the cube-edge averages; null otherwise.
  // w_lane / e_lane: the lanes on the panel's west / east cube edge (first / last strip)
  T *fside;
  bool w_lane, e_lane;
  int N;
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
    if constexpr (std::is_same<T, float>::value) r.q = (r.q + off) + T(cf) * r.gf;
    else r.q = r.q + T(cf) * r.gf;
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
          if (fside) {   // MF-AF: cube-edge fluxes of output row r = y - 4 (f2: left edge of the lane's cell)
            const int r = y - 4;
            if (w_lane) st_at(fside, (uint32_t)(1 * N + r - NG), f2);
            if (e_lane) st_at(fside, (uint32_t)(0 * N + r - NG), f2);
            if (out_lane && r == NG) st_at(fside, (uint32_t)(3 * N + x - NG), g2prev);        // edge j0
            if (out_lane && r == iend - 1) st_at(fside, (uint32_t)(2 * N + x - NG), g2);      // edge jend
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
  m.N = N;
  m.w_lane = ti == NG && lane == 3;
  m.e_lane = ti + nx == iend && lane == 3 + nx;
  m.fside = D.mf == 2 ? reinterpret_cast<T *>(D.work) + (int64_t)(tr * 6 + p) * 4 * N : nullptr;
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
    const int row = min(max(step, 0), D.max_steps - 1);   // as k_wind_avg (one basis field)
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
  m.off = (std::is_same<T, float>::value && D.offset) ? T(D.offset[tr]) : T(0);
  if (step < 0 || (D.stop > 0 && step >= D.stop)) return;   // solve halted or finished
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
