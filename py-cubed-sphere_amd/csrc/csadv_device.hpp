// Device-side helpers shared by the csadv kernels: descriptor layout, PPM edge
// reconstructions and the upwind flux.  gfx950 (CDNA4) only; 64-wide wavefronts.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "csadv.h"
#include "csadv_plan.hpp"

namespace csadv {

// A set of FP64 or dtype fields that receive a duo-grid ghost fill.
struct FieldSet {
  void *base[2];
  const int32_t *cur;        // device current-buffer index, or null (base[0])
  int32_t nfield, N, M, ld;
  int64_t fstride, pstride;
  const double *corr;        // per-field pending correction (null = none)
  const double *gc64;        // [M][ld] metric, used with corr
  const double *lagw;
  const int32_t *lagk;
  const int32_t *orient;
};

// Neighbour cell of the halo strip (panel p, side s) at depth d and along-edge index r.
// orient row = {q, ib, id, ir, jb, jd, jr}: i = ib + id*d + ir*r, j = jb + jd*d + jr*r.
__device__ __forceinline__ void halo_src(const int32_t *o, int d, int r, int &q, int &i, int &j) {
  q = o[0];
  i = o[1] + o[2] * d + o[3] * r;
  j = o[4] + o[5] * d + o[6] * r;
}

// Side of the neighbour across (p, side) and the sign relating the two axes across the edge.
// o = orient row {q, ib, id, ir, jb, jd, jr}; sides 0 E, 1 W, 2 N, 3 S.
__device__ __forceinline__ void across(const int32_t *o, int side, int &sq, int &sign) {
  const int id = o[2], jd = o[5];
  sq = (id > 0) + (id == 0) * (2 + (jd > 0));       // id > 0: W; id < 0: E; id == 0: jd > 0 S, else N
  const int out = (side == 0 || side == 2) ? 1 : -1;
  const int in = (sq == 1 || sq == 3) ? 1 : -1;
  sign = out * in;
}

// Ghost target of (side, depth d, along-edge k); sides 0 E, 1 W, 2 N, 3 S.
// Written with selects, not a switch: where the side differs across a wavefront, hipcc
// (ROCm 7.2, gfx950) lowered the four-way switch so that the default arm's lanes kept a
// stale register for one of the two results.
__device__ __forceinline__ void ghost_target(int side, int d, int k, int N, int &i, int &j) {
  const int g = (side == 0 || side == 2) ? N + NG + d : NG - 1 - d;   // index across the side
  const bool xside = side < 2;
  i = xside ? g : k;
  j = xside ? k : g;
}

template <typename T> __device__ __forceinline__ T sgn(T x) { return (x > T(0)) ? T(1) : ((x < T(0)) ? T(-1) : T(0)); }

// PPM edge values (q_L, q_R) of one cell from its 5-cell window; src/reconstruction_1d.py:26-193.
//   RECON 1 PPM-0 (CW84 eq. 1.9), 2 PPM-CW84 (monotone), 3 PPM-PL07 (eqs. 41-42), 4 PPM-L04 (B1-B4)
template <typename T, int RECON>
__device__ __forceinline__ void recon_edges(T m2, T m1, T c0, T p1, T p2, T &qL, T &qR) {
  if constexpr (RECON == 3) {
    const T a1 = T(2.0 / 60.0), a2 = T(-13.0 / 60.0), a3 = T(47.0 / 60.0), a4 = T(27.0 / 60.0), a5 = T(-3.0 / 60.0);
    qR = a1 * m2 + a2 * m1 + a3 * c0 + a4 * p1 + a5 * p2;
    qL = a5 * m2 + a4 * m1 + a3 * c0 + a2 * p1 + a1 * p2;
  } else if constexpr (RECON == 1) {
    const T s7 = T(7.0 / 12.0);
    qR = s7 * (p1 + c0) - (p2 + m1) / T(12.0);
    qL = s7 * (c0 + m1) - (p1 + m2) / T(12.0);
  } else if constexpr (RECON == 2) {
    // limited slopes at c-1, c, c+1 (CW84 eqs. 1.7-1.8)
    auto slope = [](T a, T b, T c) {
      T d0 = T(0.5) * (c - a), d1 = T(2.0) * (c - b), d2 = T(2.0) * (b - a);
      T d = fmin(fabs(d0), fabs(d1));
      d = fmin(d, fabs(d2)) * sgn(d0);
      return ((c - b) * (b - a) > T(0)) ? d : T(0);
    };
    T sm = slope(m2, m1, c0), s0 = slope(m1, c0, p1), sp = slope(c0, p1, p2);
    T R = T(0.5) * (p1 + c0) - (sp - s0) / T(6.0);
    T L = T(0.5) * (c0 + m1) - (s0 - sm) / T(6.0);
    T dq = R - L, q6 = T(6) * c0 - T(3) * (R + L);
    if ((R - c0) * (c0 - L) <= T(0)) { R = c0; L = c0; }     // local extremum: flatten (eq. 1.10)
    bool over = fabs(dq) < fabs(q6);
    T dd = R - L, mm = c0 - T(0.5) * (R + L);
    bool left = dd * mm > (dd * dd) / T(6.0);
    bool right = -(dd * dd) / T(6.0) > dd * mm;
    if (over && left) L = T(3.0) * c0 - T(2.0) * R;
    if (over && right) R = T(3.0) * c0 - T(2.0) * L;
    qL = L; qR = R;
  } else {
    // monotone slope (L04 eq. B1) at c-1, c, c+1
    auto mono = [](T a, T b, T c) {
      T d = T(0.25) * (c - a);
      T mx = fmax(fmax(a, b), c), mn = fmin(fmin(a, b), c);
      return fmin(fmin(fabs(d), mx - b), b - mn) * sgn(d);
    };
    T dm = mono(m2, m1, c0), d0 = mono(m1, c0, p1), dp = mono(c0, p1, p2);
    T R = T(0.5) * (p1 + c0) - (dp - d0) / T(3.0);
    T L = T(0.5) * (c0 + m1) - (d0 - dm) / T(3.0);
    T s = sgn(T(2.0) * d0);
    qL = c0 - fmin(T(2.0) * fabs(d0), fabs(L - c0)) * s;   // eq. B3
    qR = c0 + fmin(T(2.0) * fabs(d0), fabs(R - c0)) * s;   // eq. B4
  }
}

// The parabola of one cell as the upwind flux uses it (src/flux.py:28-44): edge values
// from the 5-cell window, with the MT-0 metric applied (gl / gr at the cell's left and right
// edges, gcell at its centre).  With dq = q_R - q_L and q6 = 6q - 3(q_R + q_L), it is kept in
// the scaled form the flux below evaluates fastest: with a = 3q - (q_R + q_L),
//   rB = a - q_R = (q6 - dq) / 2,   lB = a - q_L = (q6 + dq) / 2,   q6' = a - q = q6 / 3.
// An edge takes it from the cell on its left when the wind there is non-negative
// (A = q_R, B = rB) and from the cell on its right otherwise (A = q_L, B = -lB).  Each cell's
// parabola is formed once and an edge only selects three values from one side.
template <typename T>
struct Parab {
  T rA, rB, lA, lB, q6;
};

// The parabola from a cell's edge values (q_L, q_R) and mean c0.
template <typename T>
__device__ __forceinline__ Parab<T> parab_of(T qL, T qR, T c0, T gl, T gr, T gcell, bool mt0) {
  T q = c0;
  if (mt0) { qL = qL * gl; qR = qR * gr; q = q * gcell; }
  const T a = T(3) * q - (qR + qL);
  return {qR, a - qR, qL, a - qL, a - q};
}

template <typename T, int RECON>
__device__ __forceinline__ Parab<T> parabola(T m2, T m1, T c0, T p1, T p2, T gl, T gr, T gcell, bool mt0) {
  T qL, qR;
  recon_edges<T, RECON>(m2, m1, c0, p1, p2, qL, qR);
  return parab_of<T>(qL, qR, c0, gl, gr, gcell, mt0);
}

// Upwind PPM flux through one edge (src/flux.py:48-70) from the parabolas of the cells on its
// left (L) and right (R): f = q_R + (c/2)(q6 - dq) - q6 c^2/3 from the left cell when up,
// f = q_L - (c/2)(q6 + dq) - q6 c^2/3 from the right cell otherwise; F = f u_avg, and
// MT-PL07 multiplies by the edge metric ge.  Returned already times dt/dx, as F dt/dx =
// f c (c = u_avg dt/dx, the Courant number), so the flux difference of a cell is the
// difference of two returned values.  In the scaled form of the parabola,
// f = A + c (B - q6' c): three operations with the final product by c.
template <typename T>
__device__ __forceinline__ T upwind(bool up, T LrA, T LrB, T Lq6, const Parab<T> &R, T cc, T ge, bool mt0) {
  const T A = up ? LrA : R.lA, B = up ? LrB : -R.lB, q6 = up ? Lq6 : R.q6;
  T f = (A + cc * (B - q6 * cc)) * cc;
  if (!mt0) f = f * ge;
  return f;
}

// Whole-wavefront DPP lane shifts (one v_mov_b32_dpp per dword, no LDS round trip):
// from_left(v) is v of lane-1, from_right(v) is v of lane+1; the end lane without a source
// (lane 0, resp. lane 63) reads zero.  The step kernel only ever stores results of lanes
// whose stencils stay at least two lanes away from either end, so the end values never
// reach a stored cell.
constexpr int kDppWaveShl1 = 0x130;   // lane i <- lane i+1
constexpr int kDppWaveShr1 = 0x138;   // lane i <- lane i-1
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <typename T> __device__ __forceinline__ T from_left(T v) { return dpp_move<kDppWaveShr1>(v); }
template <typename T> __device__ __forceinline__ T from_right(T v) { return dpp_move<kDppWaveShl1>(v); }

// Lane pull through the LDS crossbar (ds_bpermute_b32: no LDS memory, no VALU slot): each lane
// reads v of the lane whose byte address (4 * lane) it passes; the address wraps modulo 64 lanes.
__device__ __forceinline__ float lane_pull(float v, int addr) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}
__device__ __forceinline__ double lane_pull(double v, int addr) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)b);
  const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// The 5-cell window of every lane's cell along the wavefront (x): neighbours at -2, -1, +1, +2.
// PULL: through the crossbar (am2 / am1 / ap1 / ap2: byte addresses of the lanes two and one to
// the left and right, mod 64), which frees four VALU slots per window and adds the crossbar's
// latency; else by DPP shifts.  End lanes read zeros (DPP) or wrapped lanes (crossbar); no
// stored cell's stencil reaches them.
template <typename T, bool PULL>
__device__ __forceinline__ void xwindow(T v, int am2, int am1, int ap1, int ap2, T &m2, T &m1, T &p1, T &p2) {
  if constexpr (PULL) {
    m2 = lane_pull(v, am2); m1 = lane_pull(v, am1);
    p1 = lane_pull(v, ap1); p2 = lane_pull(v, ap2);
  } else {
    m1 = from_left(v); p1 = from_right(v);
    m2 = from_left(m1); p2 = from_right(p1);
  }
}

// Flux through the left edge of each lane's cell along the wavefront (x), lane = cell: every
// lane forms the flux its cell sends through its right edge when it is the upwind cell there
// (c_r = c of the right neighbour) and through its left edge when it is the upwind cell there;
// the left edge then takes the left neighbour's right-edge value when the wind there is
// non-negative and the lane's own value otherwise.  One shift of the result replaces shifting
// the left cell's three parabola values (the same arithmetic as upwind(), so the same bits).
// gl, gr: edge metrics at the left and right edges (MT-PL07 only).
template <typename T>
__device__ __forceinline__ T upwind_x(bool up, const Parab<T> &c, T cc, T gl, T gr, bool mt0) {
  const T ccr = from_right(cc);
  T fr = (c.rA + ccr * (c.rB - c.q6 * ccr)) * ccr;
  T fl = (c.lA + cc * (-c.lB - c.q6 * cc)) * cc;
  if (!mt0) { fr = fr * gr; fl = fl * gl; }
  const T from_l = from_left(fr);
  return up ? from_l : fl;
}

// Element idx of an array whose base is wave-uniform: the byte offset stays 32-bit so that
// the load takes the scalar-base + vector-offset form (callers keep idx * sizeof(T) < 4 GiB).
// Every array handed to the library is device global memory, so the access is made in the
// global address space (a generic pointer would compile to a flat access).
#define CSADV_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ T ld_at(const T *base, uint32_t idx) {
  const char CSADV_GLOBAL *b = (const char CSADV_GLOBAL *)(base);
  return *(const T CSADV_GLOBAL *)(b + idx * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st_at(T *base, uint32_t idx, T v) {
  char CSADV_GLOBAL *b = (char CSADV_GLOBAL *)(base);
  *(T CSADV_GLOBAL *)(b + idx * (uint32_t)sizeof(T)) = v;
}

// Stores that every lane issues, with no branch around them: a lane with nothing to store passes
// an offset past the buffer's end, and the buffer's range check drops its write.  Used where a
// branch would make the count of outstanding memory operations uncertain: loads and stores share
// one counter, and a wait for a load past an optional store has to assume the store was not
// issued, waiting for that much more than it needs.  The buffer covers `bytes` bytes from `base`
// (< 2^31); kNoStore is past the end of any such buffer.
constexpr uint32_t kNoStore = 0x80000000u;
constexpr int kBufferDword3 = 0x00020000;   // raw (untyped) 32-bit elements, no swizzle
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(void *base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)bytes, kBufferDword3);
}
typedef unsigned int csadv_u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st_buf(__amdgpu_buffer_rsrc_t r, uint32_t byte_off, double v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(csadv_u32x2, v), r, byte_off, 0, 0);
}
__device__ __forceinline__ void st_buf(__amdgpu_buffer_rsrc_t r, uint32_t byte_off, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, byte_off, 0, 0);
}

// The wind of one edge from a one-field basis W and its coefficient s at the step's time
// (src/averaged_velocity.py:14-62 with RK1: the wind at t_n), in FP64, in one expression shared by
// the wind kernel and the step kernel so that both give the same bits (s = 1 gives w, with a
// zero of either sign as +0).
__device__ __forceinline__ double wind_of_basis(double s, double w) { return fma(s, w, 0.0); }

// Split update of one cell after the first directional pass (src/discrete_operators.py:49-73).
//   split 1 AVLT: (gQ + dF/2)/g ; 2 L04: (gQ + dF/2 + (c1-c2) Q/2)/g ; 3 PL07: (Q + (Q+dF)/(1-(c1-c2)))/2
template <typename T>
__device__ __forceinline__ T split_update(int split, T Q, T dF, T g, T rg, T c1, T c2) {
  if (split == 1) return (Q * g + T(0.5) * dF) * rg;
  if (split == 2) return (Q * g + T(0.5) * dF + T(0.5) * (c1 - c2) * Q) * rg;
  return T(0.5) * (Q + (Q + dF) / (T(1.0) - (c1 - c2)));
}

// MF-AF (src/edges_treatment.py:231-278) on boundary cells of a new state: F holds the second
// pass's fluxes through the 24 panel sides ([panel][E, W, N, S][N], scaled by dt/dx as the step
// forms them); each is replaced by the average of the two panels' fluxes through that cube edge,
// 0.5 f + 0.5 (sign f'), with the neighbour's side and sign from the halo orientation table, by
// adding the difference to the boundary cells: Q += (f - f_avg) / g at the east and north sides
// (their fluxes leave the cell there) and (f_avg - f) / g at the west and south sides.  Item idx
// is one boundary cell of a panel (4 (N - 1) of them per panel), so a corner cell takes both of
// its sides' differences in one fixed order; run is the step that produced the state.
template <typename T>
__device__ __forceinline__ T af_average(const T *F, const int32_t *orient, int N, int p, int side, int k) {
  const T own = F[(p * 4 + side) * N + k];
  const int32_t *o = orient + (p * 4 + side) * 7;
  int q, i, j, sq, sign;
  halo_src(o, 0, k + NG, q, i, j);
  across(o, side, sq, sign);
  const T nb = F[(q * 4 + sq) * N + (sq < 2 ? j : i) - NG];
  return T(0.5) * own + T(0.5) * (sign > 0 ? nb : -nb);
}

template <typename T>
__device__ __forceinline__ void af_fix_cell(const Solve &S, int tr, int run, const T *F, int idx) {
  const csadv_solve_desc &D = S.d;
  const int N = D.N, ld = D.ld, i0 = NG, iend = N + NG;
  const int ring = 4 * (N - 1);
  if (idx >= 6 * ring) return;
  const int p = idx / ring, c = idx - p * ring;
  // ring order: south row, north row, west column, east column (corners in the rows)
  int i, j;
  if (c < N) { i = i0 + c; j = i0; }
  else if (c < 2 * N) { i = i0 + c - N; j = iend - 1; }
  else if (c < 3 * N - 2) { i = i0; j = i0 + 1 + c - 2 * N; }
  else { i = iend - 1; j = i0 + 1 + c - (3 * N - 2); }
  T d = T(0);
  if (i == iend - 1) d += F[(p * 4 + 0) * N + j - NG] - af_average(F, D.orient, N, p, 0, j - NG);
  if (i == i0) d += af_average(F, D.orient, N, p, 1, j - NG) - F[(p * 4 + 1) * N + j - NG];
  if (j == iend - 1) d += F[(p * 4 + 2) * N + i - NG] - af_average(F, D.orient, N, p, 2, i - NG);
  if (j == i0) d += af_average(F, D.orient, N, p, 3, i - NG) - F[(p * 4 + 3) * N + i - NG];
  T *Q = reinterpret_cast<T *>(((run + 1) & 1) ? D.q1 : D.q0) + tr * S.tstride + p * S.pstride;
  const T rg = reinterpret_cast<const T *>(D.rgc)[(int64_t)j * ld + i];
  Q[(int64_t)j * ld + i] += d * rg;
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
  static_assert(sizeof(Solve) == 432, "one word in each 64-byte line of the descriptor");
  const char *b = reinterpret_cast<const char *>(&S);
  int w[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) __builtin_memcpy(&w[k], b + (k < 7 ? 64 * k : 428), 4);
  asm("" : "=s"(v) : "0"(v), "s"(w[0]), "s"(w[1]), "s"(w[2]), "s"(w[3]), "s"(w[4]), "s"(w[5]), "s"(w[6]), "s"(w[7]));
  return v;
}

__device__ __forceinline__ double warp_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

}  // namespace csadv
