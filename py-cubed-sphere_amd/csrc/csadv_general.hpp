// General step: the options the fused step kernel does not cover -- ET-S72 / ET-PL07
// (adjacent-panel copy halo, PL07 edge extrapolation with cube-edge averaging) and MF-AF
// (cube-edge flux averaging) -- as a sequence of simple kernels over whole arrays, in the
// order of src/discrete_operators.py:18-101:
//
//   copy halo of Q (ET-S72/PL07; ET-DG uses the duo-grid fill)   src/interpolation.py:320-340
//   x / y reconstruction of Q (in the flux; for ET-PL07 also beside the panel sides) src/reconstruction_1d.py:26-384
//   [ET-PL07] edge extrapolation, cube-edge averaging, ghosts    src/edges_treatment.py:82-178
//   x / y upwind fluxes                                          src/flux.py:20-128
//   F / G operators and the split: Qx, Qy                        src/discrete_operators.py:31-73
//   [ET-S72/PL07] copy halo of Qx / Qy                           src/discrete_operators.py:76-78
//   second pass on (Qy, Qx), [MF-AF] flux averaging              src/discrete_operators.py:81-86
//   divergence, [MF-PR] projection, update                       src/discrete_operators.py:95-101
//
// Every cross-panel step reads the halo orientation table: the neighbour cell next to the
// ghost of (panel p, side s) at depth 0 is the neighbour's boundary cell; the direction in
// which depth grows names the neighbour's side, and the product of our outward and its
// inward axis sign says whether edge values swap (q_L <-> q_R) and fluxes change sign.
// MF-PR is applied in the step itself here (no pending correction), as the reference does.
//
// Included by general_f64.hip and general_f32.hip, which instantiate enqueue_general_t; kernels
// take the solve and the tracer.

#pragma once
#include "csadv_internal.hpp"

namespace gp {

template <typename T>
struct View {
  const Solve *S;
  T *qin, *qout, *qx, *qy, *pxl, *pxr, *pyl, *pyr, *fx, *fy, *div, *strip, *avg;
  double *red;
  int N, M, ld, ldu, i0, iend;
  int64_t pstride, pstride_u, pstride_v, cell;
};

// The four edge-value arrays by slot: 0 x q_L, 1 x q_R, 2 y q_L, 3 y q_R.  They are
// consecutive in the work buffer, so a slot computed per lane picks one without a branch
// (a per-lane multiway choice between arrays is kept out of these kernels: hipcc for gfx950
// mis-lowered a divergent switch of that kind, leaving one arm's result unset).
template <typename T>
__device__ __forceinline__ T *parab(const View<T> &v, int slot) { return v.pxl + slot * v.cell; }

template <typename T>
__device__ __forceinline__ View<T> view(const Solve *solves, int s, int tr) {
  const Solve &S = solves[s];
  const csadv_solve_desc &D = S.d;
  View<T> v;
  v.S = &S;
  v.N = D.N; v.M = S.M; v.ld = D.ld; v.ldu = D.ldu;
  v.i0 = NG; v.iend = D.N + NG;
  v.pstride = S.pstride; v.pstride_u = S.pstride_u; v.pstride_v = S.pstride_v;
  const int steps = D.ctl[1];
  v.cell = work_layout(D.N, D.ld, D.ldu).cell;
  T *q0 = reinterpret_cast<T *>(D.q0) + tr * S.tstride, *q1 = reinterpret_cast<T *>(D.q1) + tr * S.tstride;
  v.qin = (steps & 1) ? q1 : q0;
  v.qout = (steps & 1) ? q0 : q1;
  const WorkLayout w = work_layout(D.N, D.ld, D.ldu);
  T *b = reinterpret_cast<T *>(D.work);
  v.qx = b + w.o_qx; v.qy = b + w.o_qy;
  v.pxl = b + w.o_pxl; v.pxr = b + w.o_pxr; v.pyl = b + w.o_pyl; v.pyr = b + w.o_pyr;
  v.fx = b + w.o_fx; v.fy = b + w.o_fy; v.div = b + w.o_div;
  v.strip = b + w.o_strip; v.avg = b + w.o_avg;
  v.red = reinterpret_cast<double *>(b + w.o_red);
  return v;
}

// finished its own step count, or halted by a diagnostics call whose blow-up limit was exceeded
__device__ __forceinline__ bool finished(const csadv_solve_desc &D) {
  return (D.stop > 0 && D.ctl[1] >= D.stop) || D.ctl[0] != 0;
}

// cell (i, j) of panel p in a cell array
__device__ __forceinline__ int64_t cidx(const int64_t ps, int ld, int p, int j, int i) { return p * ps + (int64_t)j * ld + i; }

// ---------------------------------------------------------------- copy halo
// Gather: every ghost (p, side, depth d, along r) of the target fields reads the neighbour
// cell named by the orientation table, from field A or B; all reads before any write.
//   mode 0: the step's Q halo (both from Q);
//   mode 1: between the passes: E/W ghosts of Qy read Qx for panels 4, 5 and Qy otherwise;
//           N/S ghosts of Qx read Qy for panels 1, 3 and Qx otherwise (src/halo_data.py:191-400).
template <typename T>
__global__ void k_gp_copy_gather(const Solve *solves, int s, int tr, int mode) {
  const View<T> v = view<T>(solves, s, tr);
  const csadv_solve_desc &D = v.S->d;
  if (finished(D)) return;
  const unsigned M = v.M, total = 96u * M;
  for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const unsigned r = t % M, u = t / M;              // u = (p * 4 + side) * 4 + d
    const int d = u & 3, side = (u >> 2) & 3, p = u >> 4;
    int q, i, j;
    halo_src(D.orient + (p * 4 + side) * 7, d, r, q, i, j);
    const T *ew = (p == 4 || p == 5) ? v.qx : v.qy;
    const T *ns = (p == 1 || p == 3) ? v.qy : v.qx;
    const T *src = mode == 0 ? v.qin : (side < 2 ? ew : ns);
    v.strip[t] = src[cidx(v.pstride, v.ld, q, j, i)];
  }
}

// Scatter of the gathered strips into their target fields: E/W ghosts go to Qy (or Q), N/S
// ghosts to Qx (or Q). In mode 0 both land in Q and overlap in the corner blocks, where the
// reference's order (E/W first, then N/S) leaves the N/S values: E/W lanes skip those rows.
// Qx and Qy are consecutive in the work buffer, so the field is an offset from Qx.
template <typename T>
__global__ void k_gp_copy_scatter(const Solve *solves, int s, int tr, int mode) {
  const View<T> v = view<T>(solves, s, tr);
  const csadv_solve_desc &D = v.S->d;
  if (finished(D)) return;
  const unsigned M = v.M, total = 96u * M;
  for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const unsigned r = t % M, u = t / M;              // u = (p * 4 + side) * 4 + d, as gathered
    const int d = u & 3, side = (u >> 2) & 3, p = u >> 4;
    if (mode == 0 && side < 2 && (r < (unsigned)NG || r >= (unsigned)(v.N + NG))) continue;
    int ti, tj;
    ghost_target(side, d, r, v.N, ti, tj);
    T *dst = mode == 0 ? v.qin : v.qx + (side < 2 ? v.cell : 0);
    dst[cidx(v.pstride, v.ld, p, tj, ti)] = v.strip[t];
  }
}

// ---------------------------------------------------------------- ET-PL07 edge values
// One thread per (panel, side, interior position k): the PPM edge values
// (src/reconstruction_1d.py:26-384) of the three cells inside the side along the side's
// axis -- the only cells whose arrays the flux reads; it reconstructs every other cell
// itself -- then the PL07 edge extrapolation (eqs. 47 and 49; src/edges_treatment.py:82-114)
// over them. x-sides use field fx_field, y-sides fy_field (0 = Q, 1 = Qx, 2 = Qy).
template <typename T, int RECON>
__global__ void k_gp_extrap(const Solve *solves, int s, int tr, int fx_field, int fy_field) {
  const View<T> v = view<T>(solves, s, tr);
  const csadv_solve_desc &D = v.S->d;
  if (finished(D)) return;
  const T *Fx = fx_field == 0 ? v.qin : (fx_field == 1 ? v.qx : v.qy);
  const T *Fy = fy_field == 0 ? v.qin : (fy_field == 1 ? v.qx : v.qy);
  const unsigned N = v.N, total = 24u * N;
  const int i0 = v.i0, iend = v.iend;
  for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const unsigned u = t / N;
    const int k = (int)(t % N) + i0;
    const int side = u & 3, p = u >> 2;
    const bool xdir = side < 2, low = side & 1;   // W and S are the low ends (cells i0, i0+1)
    const T *F = xdir ? Fx : Fy;
    const int st = xdir ? 1 : v.ld;
    const int dir = low ? st : -st;               // from the boundary cell into the panel
    const int64_t c0 = (xdir ? cidx(v.pstride, v.ld, p, k, 0) : cidx(v.pstride, v.ld, p, 0, k)) +
                       (int64_t)(low ? i0 : iend - 1) * st;
    // A: the edge on the panel side (q_L at the low end, q_R at the high end); B: the other
    T *A = parab(v, (xdir ? 0 : 2) + (low ? 0 : 1)), *B = parab(v, (xdir ? 0 : 2) + (low ? 1 : 0));
    T a2 = T(0);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      const int64_t o = c0 + (int64_t)m * dir;
      T qL, qR;
      recon_edges<T, RECON>(F[o - 2 * st], F[o - st], F[o], F[o + st], F[o + 2 * st], qL, qR);
      A[o] = low ? qL : qR;
      B[o] = low ? qR : qL;
      if (m == 2) a2 = low ? qL : qR;
    }
    const T Q0 = F[c0], Q1 = F[c0 + dir], Q2 = F[c0 + 2 * dir];
    A[c0] = T(1.5) * Q0 - T(0.5) * Q1;
    const T e = (T(3.0) * Q0 + T(11.0) * Q1 - T(2.0) * (Q2 - Q0)) / T(14.0);
    B[c0] = e;
    A[c0 + dir] = e;
    if (RECON == 3) B[c0 + dir] = a2;
  }
}

// The edge value of the boundary cell of (p, side) facing the cube edge, along position k:
// E: q_R of cell iend-1 (x), W: q_L of cell i0 (x), N: q_R of cell jend-1 (y), S: q_L of j0 (y).
// The slot of that edge is side ^ 1 (E 1, W 0, N 3, S 2).
template <typename T>
__device__ __forceinline__ int64_t boundary_cell(const View<T> &v, int p, int side, int k) {
  const bool xdir = side < 2, high = side == 0 || side == 2;
  const int b = high ? v.iend - 1 : v.i0;
  return xdir ? cidx(v.pstride, v.ld, p, k, b) : cidx(v.pstride, v.ld, p, b, k);
}

// Cube-edge averaging of the facing edge values (src/edges_treatment.py:31-76): the averages
// of all 24 sides are gathered into avg (entry (p * 4 + side) * N + k - i0); k_gp_edge_ghosts
// writes them back.
template <typename T>
__global__ void k_gp_edge_avg(const Solve *solves, int s, int tr) {
  const View<T> v = view<T>(solves, s, tr);
  const csadv_solve_desc &D = v.S->d;
  if (finished(D)) return;
  const unsigned N = v.N, total = 24u * N;
  for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const unsigned u = t / N;
    const int k = (int)(t % N) + v.i0;
    const int side = u & 3, p = u >> 2;
    const T *own = parab(v, side ^ 1);
    const int64_t oi = boundary_cell(v, p, side, k);
    const int32_t *o = D.orient + (p * 4 + side) * 7;
    int q, i, j, sq, sign;
    halo_src(o, 0, k, q, i, j);
    across(o, side, sq, sign);
    const T *nb = parab(v, sq ^ 1);
    v.avg[t] = (own[oi] + nb[cidx(v.pstride, v.ld, q, j, i)]) * T(0.5);
  }
}

// The averages written back to the boundary cells' facing edges, and the ghost-cell
// parabolas (src/edges_treatment.py:119-178): the ghost cell next to (p, side) takes the
// neighbour's boundary-cell edge values, swapped when the axes point opposite ways. The
// neighbour's facing value is its average, read from avg; its other value (slot sq) is one
// that no lane here writes: a boundary cell's facing slots are written only for its own
// sides, and with N >= 8 no cell lies on two opposite sides.
template <typename T>
__global__ void k_gp_edge_ghosts(const Solve *solves, int s, int tr) {
  const View<T> v = view<T>(solves, s, tr);
  const csadv_solve_desc &D = v.S->d;
  if (finished(D)) return;
  const unsigned N = v.N, total = 24u * N;
  for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const unsigned u = t / N;
    const int k = (int)(t % N) + v.i0;
    const int side = u & 3, p = u >> 2;
    parab(v, side ^ 1)[boundary_cell(v, p, side, k)] = v.avg[t];
    const int32_t *o = D.orient + (p * 4 + side) * 7;
    int q, i, j, sq, sign;
    halo_src(o, 0, k, q, i, j);
    across(o, side, sq, sign);
    const int64_t ni = cidx(v.pstride, v.ld, q, j, i);
    const T other = parab(v, sq)[ni];
    const T facing = v.avg[(q * 4 + sq) * (int)N + (sq < 2 ? j : i) - v.i0];
    const bool facing_r = sq == 0 || sq == 2;        // E, N face with q_R; W, S with q_L
    const T l = facing_r ? other : facing, r = facing_r ? facing : other;
    const bool xdir = side < 2;
    const int g = (side == 0 || side == 2) ? v.iend : v.i0 - 1;
    const int64_t gi = xdir ? cidx(v.pstride, v.ld, p, k, g) : cidx(v.pstride, v.ld, p, g, k);
    T *L = parab(v, xdir ? 0 : 2), *R = parab(v, xdir ? 1 : 3);
    L[gi] = sign > 0 ? l : r;
    R[gi] = sign > 0 ? r : l;
  }
}

// ---------------------------------------------------------------- fluxes
// Upwind PPM fluxes at x-edges i0..iend of every row (dir 0, from field fx_field) and at
// y-edges j0..jend of every column (dir 1, from field fy_field) from the reconstruction of
// the field (src/flux.py:20-128). One launch does both directions: blocks [0, nbx) take the
// x-edges, the rest the y-edges; the two read their own inputs and write separate arrays.
// The upwind cell's edge values are reconstructed here from F (the same recon_edges as
// k_gp_extrap). Without FUSE (ET-PL07, whose edge treatment runs between the reconstruction
// and the flux), the cells next to each panel side on the interior rows -- the three inside
// and the ghost cell, whose edge values the treatment wrote -- read them from the arrays.
template <typename T, int RECON, bool FUSE>
__global__ void k_gp_flux(const Solve *solves, int s, int tr, int fx_field, int fy_field, int nbx) {
  const View<T> v = view<T>(solves, s, tr);
  const csadv_solve_desc &D = v.S->d;
  if (finished(D)) return;
  const int dir = (int)blockIdx.x >= nbx;
  const int field = dir == 0 ? fx_field : fy_field;
  const unsigned b = dir == 0 ? blockIdx.x : blockIdx.x - nbx, nb = dir == 0 ? nbx : gridDim.x - nbx;
  const T *F = field == 0 ? v.qin : (field == 1 ? v.qx : v.qy);
  const unsigned M = v.M, ne = v.N + 1, total = 6u * M * ne;
  const bool mt0 = D.mt == 1;
  const T *gc = reinterpret_cast<const T *>(D.gc);
  // flux at edge item t (and its offset in the flux array)
  auto flux_at = [&](unsigned t, int64_t &fo) -> T {
    // consecutive threads take consecutive columns in both directions (coalesced)
    const unsigned w = dir == 0 ? ne : M;
    const unsigned u = t / w, r = t % w;
    const int e = (int)(dir == 0 ? r : u % ne) + v.i0;     // edge along dir
    const int a = dir == 0 ? (int)(u % M) : (int)r;        // the other index
    const int p = dir == 0 ? (int)(u / M) : (int)(u / ne);
    T uav, usg, ge_l, ge_r, gce, ge;
    int64_t cl, cr;
    const T *ge_arr;
    int ge_st;
    int64_t ge_base;
    if (dir == 0) {
      const T *U = reinterpret_cast<const T *>(D.uavg) + p * v.pstride_u;
      const T *US = D.usgn ? reinterpret_cast<const T *>(D.usgn) + p * v.pstride_u : U;
      uav = U[(int64_t)a * v.ldu + e];
      usg = US[(int64_t)a * v.ldu + e];
      cl = cidx(v.pstride, v.ld, p, a, e - 1);
      cr = cl + 1;
      ge_arr = reinterpret_cast<const T *>(D.gu);
      ge_base = (int64_t)a * v.ldu;
      ge_st = 1;
      fo = p * v.pstride_u + (int64_t)a * v.ldu + e;
    } else {
      const T *V = reinterpret_cast<const T *>(D.vavg) + p * v.pstride_v;
      const T *VS = D.vsgn ? reinterpret_cast<const T *>(D.vsgn) + p * v.pstride_v : V;
      uav = V[(int64_t)e * v.ld + a];
      usg = VS[(int64_t)e * v.ld + a];
      cl = cidx(v.pstride, v.ld, p, e - 1, a);
      cr = cl + v.ld;
      ge_arr = reinterpret_cast<const T *>(D.gv);
      ge_base = a;
      ge_st = v.ld;
      fo = p * v.pstride_v + (int64_t)e * v.ld + a;
    }
    const bool up = usg >= T(0);
    const int64_t c = up ? cl : cr;
    const int ec = up ? e - 1 : e;             // upwind cell index along dir
    T qL, qR, q = F[c];
    const int64_t st = dir == 0 ? 1 : v.ld;
    recon_edges<T, RECON>(F[c - 2 * st], F[c - st], q, F[c + st], F[c + 2 * st], qL, qR);
    if constexpr (!FUSE) {
      if (a >= v.i0 && a < v.iend && (ec <= v.i0 + 1 || ec >= v.iend - 2)) {
        const T *L = dir == 0 ? v.pxl : v.pyl, *R = dir == 0 ? v.pxr : v.pyr;
        qL = L[c];
        qR = R[c];
      }
    }
    ge_l = ge_arr[ge_base + (int64_t)ec * ge_st];
    ge_r = ge_arr[ge_base + (int64_t)(ec + 1) * ge_st];
    gce = dir == 0 ? gc[(int64_t)a * v.ld + ec] : gc[(int64_t)ec * v.ld + a];
    ge = ge_arr[ge_base + (int64_t)e * ge_st];
    if (mt0) { qL = qL * ge_l; qR = qR * ge_r; q = q * gce; }
    const T dq = qR - qL;
    const T q6 = T(6) * q - T(3) * (qR + qL);
    const T cc = (uav * T(D.dt)) / T(dir == 0 ? D.dx : D.dy);
    T f = up ? (qR + cc * T(0.5) * (q6 - dq) - q6 * cc * cc / T(3.0))
             : (qL - cc * T(0.5) * (q6 + dq) - q6 * cc * cc / T(3.0));
    f = f * uav;
    if (!mt0) f = f * ge;
    return f;
  };
  // Two edge items per lane and iteration (t and t + stride), both computed before either is
  // stored so that their loads are in flight together; a lane without a second item
  // recomputes its first and stores it once.
  T *out = dir == 0 ? v.fx : v.fy;
  const unsigned stride = nb * blockDim.x;
  for (unsigned t = b * blockDim.x + threadIdx.x; t < total; t += 2 * stride) {
    const unsigned t2 = t + stride;
    const bool has2 = t2 < total;
    int64_t fo1, fo2;
    const T f1 = flux_at(t, fo1);
    const T f2 = flux_at(has2 ? t2 : t, fo2);
    out[fo1] = f1;
    if (has2) out[fo2] = f2;
  }
}

// MF-AF (src/edges_treatment.py:231-278): each cube-edge flux becomes 0.5 f + 0.5 (sign f'),
// f' the neighbour's flux on the same edge. All 24 sides are gathered into avg (entry
// (p * 4 + side) * N + k - i0); k_gp_div_update reads them there in place of the fluxes.
template <typename T>
__global__ void k_gp_flux_avg(const Solve *solves, int s, int tr) {
  const View<T> v = view<T>(solves, s, tr);
  const csadv_solve_desc &D = v.S->d;
  if (finished(D)) return;
  const unsigned N = v.N, total = 24u * N;
  // flux on the edge of side (p, side) at position a: a row of the x-edge array (E, W) or a
  // column of the y-edge array (N, S); both offsets are formed and one is selected
  auto fidx = [&](int p, int side, int a) -> T * {
    const int e = (side == 0 || side == 2) ? v.iend : v.i0;
    const int64_t ox = p * v.pstride_u + (int64_t)a * v.ldu + e;
    const int64_t oy = p * v.pstride_v + (int64_t)e * v.ld + a;
    return side < 2 ? v.fx + ox : v.fy + oy;
  };
  for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const unsigned u = t / N;
    const int k = (int)(t % N) + v.i0;
    const int side = u & 3, p = u >> 2;
    const T *own = fidx(p, side, k);
    const int32_t *o = D.orient + (p * 4 + side) * 7;
    int q, i, j, sq, sign;
    halo_src(o, 0, k, q, i, j);
    across(o, side, sq, sign);
    const T nb = *fidx(q, sq, sq < 2 ? j : i);
    v.avg[t] = T(0.5) * *own + T(0.5) * (sign > 0 ? nb : -nb);
  }
}

// ---------------------------------------------------------------- split, divergence, update
// dF of cell i (x) / j (y) from the fluxes: -(f[c+1] - f[c]) * dt / dx inside [i0, iend), else 0
// (the edges read exist for every cell of the panel: i + 1 <= M < ldu, j + 1 <= M rows of M + 1)
template <typename T>
__device__ __forceinline__ T dflux(T f1, T f0, T dt, T dx) {
  const T d = -(f1 - f0);
  return d * dt / dx;
}
template <typename T>
__device__ __forceinline__ T dflux_x(const View<T> &v, const T *fx, const csadv_solve_desc &D, int p, int j, int i) {
  const T *f = fx + p * v.pstride_u + (int64_t)j * v.ldu;
  const T r = dflux(f[i + 1], f[i], T(D.dt), T(D.dx));
  return (i >= v.i0 && i < v.iend) ? r : T(0);
}
template <typename T>
__device__ __forceinline__ T dflux_y(const View<T> &v, const T *fy, const csadv_solve_desc &D, int p, int j, int i) {
  const T *f = fy + p * v.pstride_v + i;
  const T r = dflux(f[(int64_t)(j + 1) * v.ld], f[(int64_t)j * v.ld], T(D.dt), T(D.dy));
  return (j >= v.i0 && j < v.iend) ? r : T(0);
}

// Qx (WHICH 0) or Qy (WHICH 1) of cell (p, j, i) from Q and the first pass's fluxes
// (src/discrete_operators.py:31-73).
template <typename T, int WHICH>
__device__ __forceinline__ T qsplit(const View<T> &v, const csadv_solve_desc &D, int p, int j, int i) {
  const T *gc = reinterpret_cast<const T *>(D.gc);
  const T Q = v.qin[cidx(v.pstride, v.ld, p, j, i)], g = gc[(int64_t)j * v.ld + i];
  const T dF = WHICH == 0 ? dflux_x(v, v.fx, D, p, j, i) : dflux_y(v, v.fy, D, p, j, i);
  if (D.split == 1) {
    const T gQ = Q * g;
    return (gQ + T(0.5) * dF) / g;
  }
  const T dt = T(D.dt);                    // cx = u dt / dx (src/cfl.py:10-19)
  T c1, c2;
  if (WHICH == 0) {
    const T *U = reinterpret_cast<const T *>(D.uavg) + p * v.pstride_u, *gu = reinterpret_cast<const T *>(D.gu);
    const T dx = T(D.dx);
    const T cx1 = (U[(int64_t)j * v.ldu + i + 1] * dt) / dx, cx2 = (U[(int64_t)j * v.ldu + i] * dt) / dx;
    c1 = gu[(int64_t)j * v.ldu + i + 1] * cx1;
    c2 = gu[(int64_t)j * v.ldu + i] * cx2;
  } else {
    const T *V = reinterpret_cast<const T *>(D.vavg) + p * v.pstride_v, *gv = reinterpret_cast<const T *>(D.gv);
    const T dy = T(D.dy);
    const T cy1 = (V[(int64_t)(j + 1) * v.ld + i] * dt) / dy, cy2 = (V[(int64_t)j * v.ld + i] * dt) / dy;
    c1 = gv[(int64_t)(j + 1) * v.ld + i] * cy1;
    c2 = gv[(int64_t)j * v.ld + i] * cy2;
  }
  if (D.split == 2) {
    const T gQ = Q * g;
    return (gQ + T(0.5) * dF + T(0.5) * (c1 - c2) * Q) / g;
  }
  return T(0.5) * (Q + (Q + dF) / (T(1.0) - (c1 - c2)));
}

// Qx, Qy of every cell
template <typename T>
__global__ void k_gp_split(const Solve *solves, int s, int tr) {
  const View<T> v = view<T>(solves, s, tr);
  const csadv_solve_desc &D = v.S->d;
  if (finished(D)) return;
  const unsigned M = v.M, total = 6u * M * M;
  // two cells per lane (t and t + stride), both computed before any store
  const unsigned stride = gridDim.x * blockDim.x;
  for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += 2 * stride) {
    const unsigned t2 = t + stride < total ? t + stride : t;
    const unsigned u = t / M, u2 = t2 / M;
    const int i = t % M, j = u % M, p = u / M;
    const int i2 = t2 % M, j2 = u2 % M, p2 = u2 / M;
    const T Qx = qsplit<T, 0>(v, D, p, j, i), Qy = qsplit<T, 1>(v, D, p, j, i);
    const T Qx2 = qsplit<T, 0>(v, D, p2, j2, i2), Qy2 = qsplit<T, 1>(v, D, p2, j2, i2);
    const int64_t c = cidx(v.pstride, v.ld, p, j, i), c2 = cidx(v.pstride, v.ld, p2, j2, i2);
    v.qx[c] = Qx;
    v.qy[c] = Qy;
    if (t2 != t) {
      v.qx[c2] = Qx2;
      v.qy[c2] = Qy2;
    }
  }
}

// div = -(dF + dG) / (dt g) on the interior and, for MF-PR, per-block FP64 sums of div g
template <typename T>
__global__ void k_gp_div(const Solve *solves, int s, int tr) {
  __shared__ double red[BLOCK];
  const View<T> v = view<T>(solves, s, tr);
  const csadv_solve_desc &D = v.S->d;
  if (finished(D)) return;
  const unsigned N = v.N, total = 6u * N * N;
  const T *gc = reinterpret_cast<const T *>(D.gc);
  double acc = 0.0;
  for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const unsigned u = t / N;
    const int i = (int)(t % N) + v.i0, j = (int)(u % N) + v.i0, p = u / N;
    const T g = gc[(int64_t)j * v.ld + i];
    const T dv = -(dflux_x(v, v.fx, D, p, j, i) + dflux_y(v, v.fy, D, p, j, i)) / (T(D.dt) * g);
    v.div[cidx(v.pstride, v.ld, p, j, i)] = dv;
    acc += (double)dv * (double)g;
  }
  if (D.mf != 3) return;
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = BLOCK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) v.red[blockIdx.x] = red[0];
}

// Q^{n+1} = Q^n - dt div on the interior (MF-PR: div - g m0 / a2), every other cell copied
// so that the ghost cells persist across steps as in the reference's single array.
template <typename T>
__global__ void k_gp_update(const Solve *solves, int s, int tr) {
  __shared__ double red[BLOCK];
  const View<T> v = view<T>(solves, s, tr);
  const csadv_solve_desc &D = v.S->d;
  if (finished(D)) return;
  double m0 = 0.0;
  if (D.mf == 3) {
    red[threadIdx.x] = 0.0;
    for (int b = threadIdx.x; b < kRedGen; b += BLOCK) red[threadIdx.x] += v.red[b];
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    m0 = red[0];
  }
  const unsigned M = v.M, total = 6u * M * M;
  const T *gc = reinterpret_cast<const T *>(D.gc);
  const bool pr = D.mf == 3;
  for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const unsigned u = t / M;
    const int i = t % M, j = u % M, p = u / M;
    const int64_t c = cidx(v.pstride, v.ld, p, j, i);
    const bool interior = i >= v.i0 && i < v.iend && j >= v.i0 && j < v.iend;
    const T q = v.qin[c];
    T dv = v.div[c];                       // every cell of div exists; only interior ones are used
    if (pr) dv = dv - (gc[(int64_t)j * v.ld + i] * T(m0)) / T(D.a2);
    v.qout[c] = interior ? q - T(D.dt) * dv : q;
  }
}

// Without MF-PR no reduction separates the divergence from the update: one pass forms div of
// each interior cell in a register and writes Q^{n+1} = Q^n - dt div (the same operations as
// k_gp_div then k_gp_update); every other cell is copied. AVG (MF-AF): the fluxes on the
// panel sides are the cube-edge averages that k_gp_flux_avg gathered (side order E, W, N, S,
// N entries per side), read in place of the flux arrays' entries there.
template <typename T, bool AVG>
__global__ void k_gp_div_update(const Solve *solves, int s, int tr) {
  const View<T> v = view<T>(solves, s, tr);
  const csadv_solve_desc &D = v.S->d;
  if (finished(D)) return;
  const unsigned M = v.M, total = 6u * M * M;
  const T *gc = reinterpret_cast<const T *>(D.gc);
  // new value of cell item t (and its offset)
  auto next_q = [&](unsigned t, int64_t &c) -> T {
    const unsigned u = t / M;
    const int i = t % M, j = u % M, p = u / M;
    c = cidx(v.pstride, v.ld, p, j, i);
    const bool interior = i >= v.i0 && i < v.iend && j >= v.i0 && j < v.iend;
    const T q = v.qin[c];
    if (interior) {
      const T *fxr = v.fx + p * v.pstride_u + (int64_t)j * v.ldu, *fyc = v.fy + p * v.pstride_v + i;
      T fw = fxr[i], fe = fxr[i + 1], fs = fyc[(int64_t)j * v.ld], fn = fyc[(int64_t)(j + 1) * v.ld];
      if constexpr (AVG) {
        const T *side = v.avg + (int64_t)p * 4 * v.N;
        if (i + 1 == v.iend) fe = side[0 * v.N + j - v.i0];
        if (i == v.i0) fw = side[1 * v.N + j - v.i0];
        if (j + 1 == v.iend) fn = side[2 * v.N + i - v.i0];
        if (j == v.i0) fs = side[3 * v.N + i - v.i0];
      }
      const T g = gc[(int64_t)j * v.ld + i];
      const T dFx = dflux(fe, fw, T(D.dt), T(D.dx)), dFy = dflux(fn, fs, T(D.dt), T(D.dy));
      const T dv = -(dFx + dFy) / (T(D.dt) * g);
      return q - T(D.dt) * dv;
    }
    return q;
  };
  // two cells per lane (t and t + stride), both computed before any store
  const unsigned stride = gridDim.x * blockDim.x;
  for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += 2 * stride) {
    const bool has2 = t + stride < total;
    int64_t c1, c2;
    const T q1 = next_q(t, c1), q2 = next_q(has2 ? t + stride : t, c2);
    v.qout[c1] = q1;
    if (has2) v.qout[c2] = q2;
  }
}

}  // namespace gp


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
        hipLaunchKernelGGL((gp::k_gp_div<T>), dim3(kRedGen), dim3(BLOCK), 0, st, h->dev, s, tr);
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
    launch_gp_advance(h, st, s);
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
