// libcsadv: MI355X-native dimension-split finite-volume advection on the cubed sphere.
//
// Kernels (all launched on the caller's stream, none allocates):
//   k_ghost_edges / k_ghost_corners   duo-grid ghost fill (src/interpolation.py:154-315)
//   k_step<T,RECON,TX,TY>             fused per-tile step: both directional passes of the
//                                     split, PPM edges, upwind fluxes, update, mass-fixer
//                                     partial sums (src/discrete_operators.py:18-101,
//                                     src/flux.py, src/reconstruction_1d.py)
//   k_step_reduce                     deterministic MF-PR sum -> pending correction
//   k_wind_e2c / k_wind_c2e           duo-grid vector ghost fill (src/interpolation.py:347-532)
//   k_wind_avg                        time-averaged winds (src/averaged_velocity.py:14-62)
//   k_mass / k_norms / k_finalize     diagnostics and state materialisation
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
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
  const int c = F.cur ? *F.cur : 0;
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
      const double cf = F.corr ? F.corr[f] : 0.0;
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
  const int c = F.cur ? *F.cur : 0;
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
    const double cf = F.corr ? F.corr[f] : 0.0;
    double acc = 0.0;
#pragma unroll
    for (int l = 0; l < 4; ++l) acc += ((double)Q[src[l]] + cf * gs[l]) * w[l];
    Q[tgt] = (T)acc;
  }
}

// ------------------------------------------------------------------ fused step
// Work item = (solve, panel, tile_x, tile_y).  One block computes TX x TY interior cells of
// every tracer: region of Q with a 3-cell halo in LDS, x-pass and y-pass fluxes of Q, the
// split half-steps Qx / Qy, the swapped second pass, the update and the MF-PR partial sum.
template <typename T, int RECON, int TX, int TY>
__global__ __launch_bounds__(BLOCK) void k_step(const Solve *solves, const int4 *work) {
  constexpr int RW = TX + 6, RH = TY + 6;          // region with the 3-cell halo
  constexpr int FW = (TX + 6 > TX + 1 ? TX + 6 : TX + 1);
  __shared__ T sQ[RH][RW + 1];
  __shared__ T sA[RH][TX + 1];                      // Qx rows tj-3 .. tj+TY+2, interior columns
  __shared__ T sB[TY][RW + 1];                      // Qy interior rows, columns ti-3 .. ti+TX+2
  __shared__ T sF[(RH > TY + 1 ? RH : TY + 1) * (FW + 1)];
  __shared__ double sred[BLOCK / 64];

  const int4 wk = work[blockIdx.x];
  const Solve &S = solves[wk.x];
  const csadv_solve_desc &D = S.d;
  const int p = wk.y;
  const int N = D.N, ld = D.ld, ldu = D.ldu;
  const int ti = NG + wk.z * TX, tj = NG + wk.w * TY;
  const int nx = min(TX, N + NG - ti), ny = min(TY, N + NG - tj);
  const int cur = D.ctl[0];
  const T *Qin = reinterpret_cast<const T *>(cur ? D.q1 : D.q0);
  T *Qout = reinterpret_cast<T *>(cur ? D.q0 : D.q1);
  const T *U = reinterpret_cast<const T *>(D.uavg) + p * S.pstride_u;
  const T *V = reinterpret_cast<const T *>(D.vavg) + p * S.pstride_v;
  const T *US = D.usgn ? reinterpret_cast<const T *>(D.usgn) + p * S.pstride_u : U;
  const T *VS = D.vsgn ? reinterpret_cast<const T *>(D.vsgn) + p * S.pstride_v : V;
  const T *gc = reinterpret_cast<const T *>(D.gc);
  const T *gu = reinterpret_cast<const T *>(D.gu);
  const T *gv = reinterpret_cast<const T *>(D.gv);
  const T *rgc = reinterpret_cast<const T *>(D.rgc);
  const T dt = T(D.dt), rdx = T(1.0 / D.dx), rdy = T(1.0 / D.dy);
  const T dtdx = T(D.dt / D.dx), dtdy = T(D.dt / D.dy);
  const bool mt0 = D.mt == 1;
  const int split = D.split;
  const bool mfpr = D.mf == 3;
  const int tid = threadIdx.x;
  const int x0 = ti - 3, y0 = tj - 3;               // region origin (global indices)
  const int M = S.M;
  const int tile_id = (p * S.tiles_y + wk.w) * S.tiles_x + wk.z;

  for (int tr = 0; tr < D.ntr; ++tr) {
    const T *Q = Qin + tr * S.tstride + p * S.pstride;
    const double cf = D.corr ? D.corr[tr] : 0.0;

    // 1. region load; interior cells carry the pending mass-fixer correction
    for (int idx = tid; idx < RH * RW; idx += BLOCK) {
      const int r = idx / RW, cidx = idx - r * RW;
      const int gj = y0 + r, gi = x0 + cidx;
      T val = T(0);
      if (gi < M && gj < M) {
        val = Q[(int64_t)gj * ld + gi];
        if (cf != 0.0 && gi >= NG && gi < N + NG && gj >= NG && gj < N + NG)
          val = (T)((double)val + cf * D.gc64[(int64_t)gj * ld + gi]);
      }
      sQ[r][cidx] = val;
    }
    __syncthreads();

    // 2. x-pass 1: fluxes of Q at edges ti..ti+nx of every region row
    for (int idx = tid; idx < RH * (TX + 1); idx += BLOCK) {
      const int r = idx / (TX + 1), e = idx - r * (TX + 1);
      const int gj = y0 + r, ge = ti + e;
      T f = T(0);
      if (e <= nx && gj < M) {
        const T u = U[(int64_t)gj * ldu + ge];
        const bool up = US[(int64_t)gj * ldu + ge] >= T(0);
        const int cgl = up ? ge - 1 : ge;               // upwind cell (global)
        f = upwind_flux<T, RECON>(&sQ[r][cgl - x0], 1, up, (u * dt) * rdx, u,
                                  gu[(int64_t)gj * ldu + cgl], gu[(int64_t)gj * ldu + cgl + 1],
                                  gc[(int64_t)gj * ld + cgl], mt0, gu[(int64_t)gj * ldu + ge]);
      }
      sF[r * (FW + 1) + e] = f;
    }
    __syncthreads();

    // 3. Qx = split half-step in x (interior columns, every region row)
    for (int idx = tid; idx < RH * TX; idx += BLOCK) {
      const int r = idx / TX, c = idx - r * TX;
      const int gj = y0 + r, gi = ti + c;
      T v = T(0);
      if (c < nx && gj < M) {
        const T q = sQ[r][c + 3];
        const T dF = (-(sF[r * (FW + 1) + c + 1] - sF[r * (FW + 1) + c]) * dt) * rdx;
        T c1 = T(0), c2 = T(0);
        if (split != 1) {
          c1 = gu[(int64_t)gj * ldu + gi + 1] * ((U[(int64_t)gj * ldu + gi + 1] * dt) * rdx);
          c2 = gu[(int64_t)gj * ldu + gi] * ((U[(int64_t)gj * ldu + gi] * dt) * rdx);
        }
        v = split_update<T>(split, q, dF, gc[(int64_t)gj * ld + gi], rgc[(int64_t)gj * ld + gi], c1, c2);
      }
      sA[r][c] = v;
    }
    __syncthreads();

    // 4. y-pass 1: fluxes of Q at y-edges tj..tj+ny of every region column
    for (int idx = tid; idx < (TY + 1) * RW; idx += BLOCK) {
      const int e = idx / RW, c = idx - e * RW;
      const int ge = tj + e, gi = x0 + c;
      T f = T(0);
      if (e <= ny && gi < M) {
        const T v = V[(int64_t)ge * ld + gi];
        const bool up = VS[(int64_t)ge * ld + gi] >= T(0);
        const int cgl = up ? ge - 1 : ge;
        f = upwind_flux<T, RECON>(&sQ[cgl - y0][c], RW + 1, up, (v * dt) * rdy, v,
                                  gv[(int64_t)cgl * ld + gi], gv[(int64_t)(cgl + 1) * ld + gi],
                                  gc[(int64_t)cgl * ld + gi], mt0, gv[(int64_t)ge * ld + gi]);
      }
      sF[e * (FW + 1) + c] = f;
    }
    __syncthreads();

    // 5. Qy = split half-step in y (interior rows, every region column)
    for (int idx = tid; idx < TY * RW; idx += BLOCK) {
      const int r = idx / RW, c = idx - r * RW;
      const int gj = tj + r, gi = x0 + c;
      T v = T(0);
      if (r < ny && gi < M) {
        const T q = sQ[r + 3][c];
        const T dG = (-(sF[(r + 1) * (FW + 1) + c] - sF[r * (FW + 1) + c]) * dt) * rdy;
        T c1 = T(0), c2 = T(0);
        if (split != 1) {
          c1 = gv[(int64_t)(gj + 1) * ld + gi] * ((V[(int64_t)(gj + 1) * ld + gi] * dt) * rdy);
          c2 = gv[(int64_t)gj * ld + gi] * ((V[(int64_t)gj * ld + gi] * dt) * rdy);
        }
        v = split_update<T>(split, q, dG, gc[(int64_t)gj * ld + gi], rgc[(int64_t)gj * ld + gi], c1, c2);
      }
      sB[r][c] = v;
    }
    __syncthreads();

    // 6. x-pass 2: fluxes of Qy at edges ti..ti+nx of the interior rows
    for (int idx = tid; idx < TY * (TX + 1); idx += BLOCK) {
      const int r = idx / (TX + 1), e = idx - r * (TX + 1);
      const int gj = tj + r, ge = ti + e;
      T f = T(0);
      if (r < ny && e <= nx) {
        const T u = U[(int64_t)gj * ldu + ge];
        const bool up = US[(int64_t)gj * ldu + ge] >= T(0);
        const int cgl = up ? ge - 1 : ge;
        f = upwind_flux<T, RECON>(&sB[r][cgl - x0], 1, up, (u * dt) * rdx, u,
                                  gu[(int64_t)gj * ldu + cgl], gu[(int64_t)gj * ldu + cgl + 1],
                                  gc[(int64_t)gj * ld + cgl], mt0, gu[(int64_t)gj * ldu + ge]);
      }
      sF[r * (FW + 1) + e] = f;
    }
    __syncthreads();

    // 7. keep dF2 of the owned cells in registers
    constexpr int PER = (TX * TY + BLOCK - 1) / BLOCK;
    T dF2[PER];
#pragma unroll
    for (int s = 0; s < PER; ++s) {
      const int idx = tid + s * BLOCK;
      const int r = idx / TX, c = idx - r * TX;
      dF2[s] = (idx < TX * TY) ? (-(sF[r * (FW + 1) + c + 1] - sF[r * (FW + 1) + c]) * dt) * rdx : T(0);
    }
    __syncthreads();

    // 8. y-pass 2: fluxes of Qx at y-edges tj..tj+ny of the interior columns
    for (int idx = tid; idx < (TY + 1) * TX; idx += BLOCK) {
      const int e = idx / TX, c = idx - e * TX;
      const int ge = tj + e, gi = ti + c;
      T f = T(0);
      if (e <= ny && c < nx) {
        const T v = V[(int64_t)ge * ld + gi];
        const bool up = VS[(int64_t)ge * ld + gi] >= T(0);
        const int cgl = up ? ge - 1 : ge;
        f = upwind_flux<T, RECON>(&sA[cgl - y0][c], TX + 1, up, (v * dt) * rdy, v,
                                  gv[(int64_t)cgl * ld + gi], gv[(int64_t)(cgl + 1) * ld + gi],
                                  gc[(int64_t)cgl * ld + gi], mt0, gv[(int64_t)ge * ld + gi]);
      }
      sF[e * (FW + 1) + c] = f;
    }
    __syncthreads();

    // 9. divergence, update, mass-fixer partial sum (src/discrete_operators.py:95-101,
    //    src/advection_timestep.py:43)
    double part = 0.0;
    T *Qo = Qout + tr * S.tstride + p * S.pstride;
#pragma unroll
    for (int s = 0; s < PER; ++s) {
      const int idx = tid + s * BLOCK;
      const int r = idx / TX, c = idx - r * TX;
      if (idx < TX * TY && r < ny && c < nx) {
        const int gj = tj + r, gi = ti + c;
        const T dG2 = (-(sF[(r + 1) * (FW + 1) + c] - sF[r * (FW + 1) + c]) * dt) * rdy;
        const T g = gc[(int64_t)gj * ld + gi];
        const T div = -(dF2[s] + dG2) / (dt * g);
        Qo[(int64_t)gj * ld + gi] = sQ[r + 3][c + 3] - dt * div;
        if (mfpr) part += (double)div * (double)g;
      }
    }
    if (mfpr) {
      part = warp_sum(part);
      if ((tid & 63) == 0) sred[tid >> 6] = part;
      __syncthreads();
      if (tid == 0) {
        double s = 0.0;
        for (int w = 0; w < BLOCK / 64; ++w) s += sred[w];
        D.partials[(int64_t)tr * S.part_per_tracer + tile_id] = s;
      }
    }
    __syncthreads();
  }
}

// One block per solve: fixed-order sums of the tile partials -> pending correction
// c = dt * m0 / a2 (Q^{n+1} = Q* + c g, src/discrete_operators.py:98-101); then flip the
// current buffer and advance the step counter.
__global__ void k_step_reduce(Solve *solves) {
  Solve &S = solves[blockIdx.x];
  csadv_solve_desc &D = S.d;
  __shared__ double red[BLOCK];
  if (D.mf == 3) {
    for (int tr = 0; tr < D.ntr; ++tr) {
      double s = 0.0;
      const double *P = D.partials + (int64_t)tr * S.part_per_tracer;
      for (int i = threadIdx.x; i < S.part_per_tracer; i += BLOCK) s += P[i];
      red[threadIdx.x] = s;
      __syncthreads();
      for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
      }
      if (threadIdx.x == 0) D.corr[tr] = D.dt * red[0] / D.a2;
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    D.ctl[0] ^= 1;
    D.ctl[1] += 1;
  }
}

// Fold pending corrections into the current buffer (interior only) and clear them.
template <typename T>
__global__ void k_finalize(Solve *solves) {
  Solve &S = solves[blockIdx.y];
  csadv_solve_desc &D = S.d;
  if (D.mf != 3) return;
  const int N = D.N, ld = D.ld;
  T *Q = reinterpret_cast<T *>(D.ctl[0] ? D.q1 : D.q0);
  const int64_t tot = (int64_t)6 * N * N;
  for (int tr = 0; tr < D.ntr; ++tr) {
    const double cf = D.corr[tr];
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

__global__ void k_clear_corr(Solve *solves) {
  csadv_solve_desc &D = solves[blockIdx.x].d;
  if (D.mf == 3)
    for (int tr = threadIdx.x; tr < D.ntr; tr += blockDim.x) D.corr[tr] = 0.0;
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
__global__ void k_wind_avg(Solve *solves) {
  Solve &S = solves[blockIdx.z];
  const csadv_solve_desc &D = S.d;
  const int N = D.N, M = S.M, ld = D.ld, ldu = D.ldu, K = D.nbasis;
  const int p = blockIdx.y;
  const int i0 = NG, iend = N + NG;
  const int row = D.unsteady ? min(D.ctl[1], D.max_steps - 1) : 0;
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
    double acc = 0.0;
    for (int m = 0; m < K; ++m) acc += s[m] * W[m * bs + off];
    return acc;
  };
  const int ne = N + 1;   // edges i0 .. iend
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < 2 * ne * M; idx += gridDim.x * blockDim.x) {
    const bool xdir = idx < ne * M;
    const int id2 = xdir ? idx : idx - ne * M;
    const int e = id2 % ne + i0, c = id2 / ne;      // edge index along the flux direction, row/col
    const int64_t bs = xdir ? bs_u : bs_v;
    const double *W = xdir ? Wu : Wv;
    const int64_t off = xdir ? (int64_t)c * ldu + e : (int64_t)e * ld + c;
    const int64_t step = xdir ? 1 : ld;
    const double rd = xdir ? 1.0 / D.dx : 1.0 / D.dy;
    const double u = (K == 1 && s1[0] == 1.0) ? W[off] : comb(W, bs, off, s1);
    double avg = u;
    if (rk2) {
      auto ui = [&](int64_t o) {
        const double un = (K == 1 && s1[0] == 1.0) ? W[o] : comb(W, bs, o, s1);
        const double uo = (K == 1 && s0[0] == 1.0) ? W[o] : comb(W, bs, o, s0);
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
  const T *Q = reinterpret_cast<const T *>(D.ctl[0] ? D.q1 : D.q0) + tracer * S.tstride;
  const double cf = D.corr ? D.corr[tracer] : 0.0;
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
  const T *Q = reinterpret_cast<const T *>(D.ctl[0] ? D.q1 : D.q0) + tracer * S.tstride;
  const double cf = D.corr ? D.corr[tracer] : 0.0;
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
  double *red = nullptr;          // small reduction scratch (mass / norms)
  std::vector<int> cur;           // host copy of the current-buffer index per solve
  std::vector<hipEvent_t> ev;     // profiling event pairs (start, end) around the step kernel
  int ev_used = 0;
  std::string err;
};

namespace {

struct TileShape { int tx, ty; };
constexpr TileShape kTiles[] = {{64, 16}, {32, 32}, {64, 8}, {128, 8}, {32, 16}};
constexpr int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);
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
  if (d.et != 3) { msg = "only et=3 (ET-DG) is implemented by this library"; return CSADV_E_ARG; }
  if (d.mf != 1 && d.mf != 3) { msg = "only mf=1 (none) and mf=3 (MF-PR) are implemented"; return CSADV_E_ARG; }
  if (d.nbasis < 1 || d.nbasis > 8) { msg = "nbasis must be 1..8"; return CSADV_E_RANGE; }
  if (d.unsteady && d.max_steps < 1) { msg = "unsteady winds need a coefficient table"; return CSADV_E_ARG; }
  if (!d.q0 || !d.q1 || !d.uavg || !d.vavg || !d.gc || !d.gu || !d.gv || !d.rgc || !d.gc64 || !d.lagw ||
      !d.lagk || !d.orient || !d.wbu || !d.wbv || !d.coef || !d.ctl) {
    msg = "null device pointer in descriptor";
    return CSADV_E_ARG;
  }
  if (d.mf == 3 && (!d.corr || !d.partials)) { msg = "MF-PR needs corr and partials"; return CSADV_E_ARG; }
  if (!(d.dt > 0.0) || !(d.dx > 0.0) || !(d.dy > 0.0)) { msg = "dt, dx, dy must be positive"; return CSADV_E_ARG; }
  return CSADV_OK;
}

Solve derive(const csadv_solve_desc &d, int tile) {
  Solve s{};
  s.d = d;
  s.M = d.N + 2 * NG;
  const TileShape ts = kTiles[tile];
  s.tiles_x = (d.N + ts.tx - 1) / ts.tx;
  s.tiles_y = (d.N + ts.ty - 1) / ts.ty;
  s.tiles = s.tiles_x * s.tiles_y;
  s.pstride = (int64_t)s.M * d.ld;
  s.tstride = 6 * s.pstride;
  s.pstride_u = (int64_t)s.M * d.ldu;
  s.pstride_v = (int64_t)(s.M + 1) * d.ld;
  s.part_per_tracer = 6 * s.tiles;
  return s;
}

template <typename T, int R, int TX, int TY>
hipError_t launch_step_t(csadv_t *h, hipStream_t st) {
  hipLaunchKernelGGL((k_step<T, R, TX, TY>), dim3(h->nwork), dim3(BLOCK), 0, st, h->dev, h->work);
  return hipGetLastError();
}

template <typename T, int R>
hipError_t launch_step_r(csadv_t *h, hipStream_t st) {
  switch (h->tile) {
    case 0: return launch_step_t<T, R, 64, 16>(h, st);
    case 1: return launch_step_t<T, R, 32, 32>(h, st);
    case 2: return launch_step_t<T, R, 64, 8>(h, st);
    case 3: return launch_step_t<T, R, 128, 8>(h, st);
    default: return launch_step_t<T, R, 32, 16>(h, st);
  }
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
  const int blocks = std::min(1024, ((maxN + 1) * (maxN + 8) * 2 + BLOCK - 1) / BLOCK);
  if (h->dtype == CSADV_F64)
    hipLaunchKernelGGL(k_wind_avg<double>, dim3(blocks, 6, ns), dim3(BLOCK), 0, st, h->dev);
  else
    hipLaunchKernelGGL(k_wind_avg<float>, dim3(blocks, 6, ns), dim3(BLOCK), 0, st, h->dev);
  HIPCHK(h, hipGetLastError());
  return CSADV_OK;
}

}  // namespace

extern "C" {

int csadv_abi_version(void) { return CSADV_ABI_VERSION; }

int csadv_partials_per_tracer(int32_t N, int32_t tile) {
  if (tile < 0 || tile >= kNumTiles || N < 1) return CSADV_E_ARG;
  const TileShape ts = kTiles[tile];
  return 6 * ((N + ts.tx - 1) / ts.tx) * ((N + ts.ty - 1) / ts.ty);
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
  for (int s = 0; s < nsolve; ++s) {
    std::string msg;
    int rc = validate(solves[s], msg);
    if (rc == CSADV_OK && (solves[s].dtype != h->dtype || solves[s].recon != h->recon)) {
      rc = CSADV_E_ARG;
      msg = "all solves of one handle share dtype and recon";
    }
    if (rc != CSADV_OK) {
      // the handle is returned so that csadv_last_error() can explain the failure
      h->err = "solve " + std::to_string(s) + ": " + msg;
      *out = h;
      return rc;
    }
    h->desc.push_back(solves[s]);
  }
  h->cur.assign(nsolve, 0);
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
  h->host.clear();
  std::vector<int4> work;
  std::vector<FieldSet> sets;
  for (int s = 0; s < nsolve; ++s) {
    std::string msg;
    int rc = validate(solves[s], msg);
    if (rc != CSADV_OK) return fail(h, rc, "solve " + std::to_string(s) + ": " + msg);
    h->desc[s] = solves[s];
    Solve S = derive(solves[s], h->tile);
    h->host.push_back(S);
    for (int p = 0; p < 6; ++p)
      for (int ty = 0; ty < S.tiles_y; ++ty)
        for (int tx = 0; tx < S.tiles_x; ++tx) work.push_back(make_int4(s, p, tx, ty));
    FieldSet F{};
    F.base[0] = solves[s].q0;
    F.base[1] = solves[s].q1;
    F.cur = solves[s].ctl;
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
  if ((int)work.size() != h->nwork) {
    if (h->work) (void)hipFree(h->work);
    h->work = nullptr;
    HIPCHK(h, hipMalloc(&h->work, sizeof(int4) * work.size()));
    h->nwork = (int)work.size();
  }
  HIPCHK(h, hipMemcpyAsync(h->work, work.data(), sizeof(int4) * work.size(), hipMemcpyHostToDevice, st));
  HIPCHK(h, hipMemcpyAsync(h->dev, h->host.data(), sizeof(Solve) * nsolve, hipMemcpyHostToDevice, st));
  HIPCHK(h, hipMemcpyAsync(h->qsets, sets.data(), sizeof(FieldSet) * nsolve, hipMemcpyHostToDevice, st));
  HIPCHK(h, hipStreamSynchronize(st));
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
  return ghost_fill(h, (hipStream_t)stream);
}

int csadv_step(csadv_t *h, int32_t nsteps, void *stream) {
  if (!h || !h->dev) return CSADV_E_STATE;
  if (nsteps < 0) return fail(h, CSADV_E_ARG, "nsteps < 0");
  hipStream_t st = (hipStream_t)stream;
  const int ns = (int)h->desc.size();
  for (int k = 0; k < nsteps; ++k) {
    int rc = ghost_fill(h, st);
    if (rc) return rc;
    rc = wind_avg(h, st, false);
    if (rc) return rc;
    const bool prof = 2 * (h->ev_used + 1) <= (int)h->ev.size();
    if (prof) HIPCHK(h, hipEventRecord(h->ev[2 * h->ev_used], st));
    hipError_t e = h->dtype == CSADV_F64 ? launch_step_dt<double>(h, st) : launch_step_dt<float>(h, st);
    if (e != hipSuccess) return fail(h, CSADV_E_HIP, std::string("step kernel: ") + hipGetErrorString(e));
    if (prof) {
      HIPCHK(h, hipEventRecord(h->ev[2 * h->ev_used + 1], st));
      ++h->ev_used;
    }
    hipLaunchKernelGGL(k_step_reduce, dim3(ns), dim3(BLOCK), 0, st, h->dev);
    HIPCHK(h, hipGetLastError());
  }
  for (int s = 0; s < ns; ++s) h->cur[s] = (h->cur[s] + nsteps) & 1;
  return CSADV_OK;
}

int csadv_finalize(csadv_t *h, void *stream) {
  if (!h || !h->dev) return CSADV_E_STATE;
  hipStream_t st = (hipStream_t)stream;
  const int ns = (int)h->desc.size();
  if (h->dtype == CSADV_F64)
    hipLaunchKernelGGL(k_finalize<double>, dim3(256, ns), dim3(BLOCK), 0, st, h->dev);
  else
    hipLaunchKernelGGL(k_finalize<float>, dim3(256, ns), dim3(BLOCK), 0, st, h->dev);
  hipLaunchKernelGGL(k_clear_corr, dim3(ns), dim3(64), 0, st, h->dev);
  HIPCHK(h, hipGetLastError());
  return CSADV_OK;
}

int csadv_mass(csadv_t *h, int32_t solve, int32_t tracer, double *out, void *stream) {
  if (!h || !h->dev) return CSADV_E_STATE;
  if (solve < 0 || solve >= (int)h->desc.size() || tracer < 0 || tracer >= h->desc[solve].ntr || !out)
    return fail(h, CSADV_E_ARG, "mass: bad solve / tracer / output");
  hipStream_t st = (hipStream_t)stream;
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

int csadv_norms(csadv_t *h, int32_t solve, int32_t tracer, const double *qexact, double *out, void *stream) {
  if (!h || !h->dev) return CSADV_E_STATE;
  if (solve < 0 || solve >= (int)h->desc.size() || tracer < 0 || tracer >= h->desc[solve].ntr || !out || !qexact)
    return fail(h, CSADV_E_ARG, "norms: bad solve / tracer / pointer");
  hipStream_t st = (hipStream_t)stream;
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
  if (!h || solve < 0 || solve >= (int)h->cur.size()) return CSADV_E_ARG;
  return h->cur[solve];
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

const char *csadv_last_error(csadv_t *h) { return h ? h->err.c_str() : "null handle"; }

void csadv_destroy(csadv_t *h) {
  if (!h) return;
  if (h->dev) (void)hipFree(h->dev);
  if (h->work) (void)hipFree(h->work);
  if (h->qsets) (void)hipFree(h->qsets);
  if (h->wsets) (void)hipFree(h->wsets);
  if (h->red) (void)hipFree(h->red);
  for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
  delete h;
}

}  // extern "C"
