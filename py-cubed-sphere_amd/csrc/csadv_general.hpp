// General step: ET-S72 and ET-PL07, whose halo is the adjacent panels' own cells (the copy halo
// of src/interpolation.py:320-340 and src/halo_data.py) rather than the duo-grid interpolation,
// as two marched passes per tracer (src/discrete_operators.py:18-101):
//
//   pass 1: x- and y-fluxes of Q, the split half-steps Qx and Qy of every interior cell;
//   pass 2: x-fluxes of Qy, y-fluxes of Qx, and the update of Q (MF-0, MF-AF), or the
//           divergence and its partial sums (MF-PR);
//   MF-AF: k_gp_af_fix puts the cube-edge flux averages into the boundary cells;
//   MF-PR: k_gp_prsum sums the partials, k_gp_update applies the projection and the update.
//
// Each pass marches one wavefront down a column strip as the fused step does (lane = column,
// the 3-cell halo in the outer lanes, x windows by DPP, register rings for the y windows).  No
// ghost cell is stored: a lane in a ghost column (interior rows) or a row in a ghost row
// (interior columns) reads the neighbour panel's cell at the same depth through the halo
// orientation table, from the field the copy halo would have copied there -- Q in pass 1; in
// pass 2 the E/W ghosts of Qy read the neighbour's Qx on panels 4 and 5 and its Qy otherwise,
// the N/S ghosts of Qx read its Qy on panels 1 and 3 and its Qx otherwise
// (src/discrete_operators.py:76-78).  That field is always the one the neighbour's own pass
// reconstructs across the shared edge, so the ET-PL07 edge treatment needs nothing from the
// neighbour but those ghost values:
//   * the extrapolated edge values (eqs. 47 and 49, src/edges_treatment.py:82-114) of the three
//     cells inside each panel side, from the cells themselves;
//   * the cube-edge average of the two panels' extrapolations (src/edges_treatment.py:31-76),
//     the neighbour's from its cells at depths 0 and 1 = our ghosts;
//   * the ghost cell's parabola (src/edges_treatment.py:119-178): the neighbour's boundary cell's
//     edge values, the average facing us and its own eq. 49 value beyond, from ghosts 0..2.
// Those override the plain PPM edge values of the cells i0-1 .. i0+1 and iend-2 .. iend along
// each direction (interior rows / columns only), as the reference's arrays hold them when its
// flux reads them.
//
// Included by the general_<dtype>_s<split>.hip translation units, one per precision and
// splitting, which instantiate enqueue_general_split.

#pragma once
#include <utility>

#include "csadv_internal.hpp"

namespace gp {

// Element offsets of one lane's values of a field along the march: a + y * b in the row regime
// of row y (0: south ghost rows, 1: interior rows, 2: north ghost rows); bit k of ok says whether
// regime k has a source at this lane.  Regimes are wave-uniform, so one select per row picks.
struct Lines {
  int32_t a0, b0, a1, b1, a2, b2;
  uint32_t ok;
};

// The lines of column x of panel p: interior columns read the panel's own cells (at own_off) in
// interior rows and, with rows, the neighbour's cells of the N/S ghost rows (at ns_off); ghost
// columns (depth < 3), with cols, read the neighbour's cells of the E/W ghosts in interior rows
// (at ew_off).  Ghost columns in ghost rows (the corners) have no source.  ps: panel stride.
__device__ __forceinline__ Lines lane_lines(const csadv_solve_desc &D, int ps, int p, int x, int own_off, bool cols,
                                            int ew_off, bool rows, int ns_off) {
  const int ld = D.ld, i0 = NG, iend = D.N + NG;
  Lines L{0, 0, 0, 0, 0, 0, 0u};
  if (x >= i0 && x < iend) {
    L.a1 = own_off + p * ps + x;
    L.b1 = ld;
    L.ok = 2u;
    if (rows) {
      // (q, i, j) = halo_src(o, d, x): linear in the depth d, which is i0 - 1 - y in the south
      // ghost rows and y - iend in the north ones
      const int32_t *os = D.orient + (p * 4 + 3) * 7, *on = D.orient + (p * 4 + 2) * 7;
      const int bs = ns_off + os[0] * ps + (os[4] + os[6] * x) * ld + os[1] + os[3] * x, ss = os[5] * ld + os[2];
      const int bn = ns_off + on[0] * ps + (on[4] + on[6] * x) * ld + on[1] + on[3] * x, sn = on[5] * ld + on[2];
      L.a0 = bs + (i0 - 1) * ss;
      L.b0 = -ss;
      L.a2 = bn - iend * sn;
      L.b2 = sn;
      L.ok = 7u;
    }
  } else if (cols && x >= i0 - 3 && x < iend + 3) {
    // (q, i, j) = halo_src(o, d, y) of the W (x < i0) or E side: linear in the row y
    const bool west = x < i0;
    const int d = west ? i0 - 1 - x : x - iend;
    const int32_t *o = D.orient + (p * 4 + (west ? 1 : 0)) * 7;
    L.a1 = ew_off + o[0] * ps + (o[4] + o[5] * d) * ld + o[1] + o[2] * d;
    L.b1 = o[6] * ld + o[3];
    L.ok = 2u;
  }
  return L;
}

// ET-PL07 pieces (src/edges_treatment.py:82-114, 31-76): the one-sided extrapolation of a
// boundary cell c to the panel side from its neighbour n inside the panel (eq. 47), the edge
// value between cells 0 and 1 from cells 0, 1, 2 counted from the side (eq. 49), and the
// cube-edge average of our and the neighbour panel's extrapolation.
template <typename T> __device__ __forceinline__ T pl_ext(T c, T n) { return T(1.5) * c - T(0.5) * n; }
template <typename T> __device__ __forceinline__ T pl_edge(T c0, T c1, T c2) {
  // times the rounded 1/14 rather than divided by 14: within an ulp of the quotient, and a
  // multiply where a division takes about ten instructions
  return (T(3.0) * c0 + T(11.0) * c1 - T(2.0) * (c2 - c0)) * T(1.0 / 14.0);
}
template <typename T> __device__ __forceinline__ T pl_avg(T own, T nb) { return (own + nb) * T(0.5); }

// PL07 edge values (src/reconstruction_1d.py:196-229, eqs. 41-42): q_L of the cell whose window
// is (m2 .. p2) -- the RECON 3 override of q_R next to a panel side takes it from the next cell.
template <typename T> __device__ __forceinline__ T pl_left(T m2, T m1, T c0, T p1, T p2) {
  return T(-3.0 / 60.0) * m2 + T(27.0 / 60.0) * m1 + T(47.0 / 60.0) * c0 + T(-13.0 / 60.0) * p1 + T(2.0 / 60.0) * p2;
}
template <typename T> __device__ __forceinline__ T pl_right(T m2, T m1, T c0, T p1, T p2) {
  return T(2.0 / 60.0) * m2 + T(-13.0 / 60.0) * m1 + T(47.0 / 60.0) * c0 + T(27.0 / 60.0) * p1 + T(-3.0 / 60.0) * p2;
}

// One wavefront of a pass: strip columns ti .. ti+nx-1 (lane l <-> column ti-3+l), rows
// tj .. tj+ny-1.  Iteration y (rows y0 = tj-3 .. ylast = tj+ny+3):
//   * column stage at y-edge y-3 from the ring of the y-field (rows y-6 .. y-1): the
//     parabola of row y-3, the y-flux through edge y-3, then for row y-4 the y-split Qy
//     (pass 1) or the update (pass 2, with the x-flux difference the row stage kept);
//   * row stage for row y, in the segment's own rows tj .. tj+ny-1: x-window of the x-field by
//     DPP, parabolas, x-fluxes (every lane the flux through its cell's left edge), then Qx
//     (pass 1) or the x-flux difference (pass 2).
// The inputs of row y+PF are requested at the end of iteration y, PF = 2 rows ahead.  Loads and
// stores share one counter, and a wait for a load also waits for everything issued before it,
// so the march keeps that count exact: every per-row value sits in a ring one slot deeper than
// its lifetime (RD), which keeps the register allocator from rotating values still in flight
// at the loop's back-edge; every lane issues every load (a lane with no source in a row regime
// reads element 0 and masks the value where the row is used); the row's stores follow the next
// row's loads and every lane issues them (a lane with nothing to store passes an offset past
// the buffer, which drops it).  The field offsets advance by one row stride per row and are
// recomputed only where the row regime changes.
// SPLIT is the splitting (1 AVLT, 2 L04, 3 PL07), with the metric treatment it implies (MT-0
// with AVLT and L04, MT-PL07 with PL07); a compile-time constant, as in the fused step.
template <typename T, int RECON, bool RK2, int PASS, int SPLIT>
struct GMarch {
  static constexpr int PF = 2;
  static constexpr int split = SPLIT;
  static constexpr bool mt0 = SPLIT != 3;
  static constexpr int RD = PF + 5;
  // ring slot of row y + j in phase K (the phase of row y is (y - y0) mod RD)
  static constexpr int sl(int K, int j) { return ((K + j) % RD + RD) % RD; }
  // ---- per wavefront
  const T *X;              // pass 1: Q of the tracer; pass 2: Qx of the tracer (Qy = Qx + cell)
  const T *Qs;             // pass 2: Q of the tracer (the update's input)
  // pass 1: Qx (out) and Qy (outy) of the tracer; pass 2: the new Q, or with MF-PR its div
  // (out).  Each covers one cell array, under 2^31 bytes up to the largest N csadv_create
  // accepts for the general step, so that kNoStore lies past every one of them.
  __amdgpu_buffer_rsrc_t out, outy;
  bool Div;                // pass 2: out holds div (MF-PR; the update follows later)
  T *fside;                // pass 2 with MF-AF: cube-edge fluxes of the panel, [E, W, N, S][N]
  const T *U, *US, *V, *VS, *gc, *gu, *gv, *rgc;
  Lines lf, lg;            // lines of the row stage's field and (pass 2) the column stage's
  bool okf[3], okg[3];     // per row regime: the lane has a source of that field there
  int ld, ldu, N, i0, iend, x, xc, xe, tj, tend, ylast, cell, own;
  bool pl07, sp_x, out_lane, w_lane, e_lane;
  bool lo_g, lo_b, lo_1, hi_1, hi_b, hi_g;   // column x is i0-1, i0, i0+1, iend-2, iend-1, iend
  T dtdx, dtdy, dtc;
  // ---- the next row to load: its row and regime, the element offsets of the fields and of Q
  // there (of, og, oq) and the fields' row strides in that regime (bf, bg)
  int lrow, lreg;
  int of, og, oq, bf, bg;
  // ---- rings (slot of row y+j in phase K: sl(K, j))
  T fr[RD], gr[RD];        // loaded, unmasked: the row stage's field, (pass 2) the column stage's
  T vr[RD], vsr[RD];       // y-wind and its upwind-side wind at y-edge y-3 of the row's iteration
  T fw[RD];                // the column stage's field (pass 1: Q, pass 2: Qx), masked
  T dfw[RD];               // pass 2: x-flux difference of the rows
  T qsw[RD];               // pass 2: Q of the rows
  T gcw[RD], rgw[RD], gvr[RD];   // metric, 1/metric of the rows; gv at y-edge y-2 in row y's slot
  T uw[RD], usw[RD], guw[RD];    // x-edge wind, upwind-side wind, gu of the rows
  T gprev, vprev, pa_A, pa_B, pa_q6;
  double part;             // pass 2 with MF-PR: sum of div g over the wavefront's cells

  // row regime of row y: 0 south ghost rows, 1 interior, 2 north ghost rows (wave-uniform)
  __device__ __forceinline__ int regime(int y) const { return y < i0 ? 0 : (y < iend ? 1 : 2); }
  // the line (a, b) of regime reg (a lane with no source there has a = b = 0: element 0)
  __device__ __forceinline__ void line(const Lines &L, int reg, int &a, int &b) const {
    // (the members are read into values first: a select between member lvalues would keep the
    // march's state in scratch memory)
    const int a0 = L.a0, a1 = L.a1, a2 = L.a2, b0 = L.b0, b1 = L.b1, b2 = L.b2;
    a = reg == 0 ? a0 : (reg == 1 ? a1 : a2);
    b = reg == 0 ? b0 : (reg == 1 ? b1 : b2);
  }
  __device__ __forceinline__ bool has(const bool *ok, int reg) const {
    const bool k0 = ok[0], k1 = ok[1], k2 = ok[2];
    return reg == 0 ? k0 : (reg == 1 ? k1 : k2);
  }
  // one value of a field at row y, zero where the lane has no source (the first parabola row of
  // a segment next to the north side reads one row before the march)
  __device__ __forceinline__ T field(const Lines &L, const bool *ok, int y) const {
    int a, b;
    const int reg = regime(y);
    line(L, reg, a, b);
    const T v = ld_at(X, (uint32_t)(a + y * b));
    return has(ok, reg) ? v : T(0);
  }

  // Inputs of row y into ring slot S.
  template <int S>
  __device__ __forceinline__ void load(int y) {
    y = min(y, ylast);       // rows past the last are requested in bounds and never used
    if (y != lrow) {         // (wave-uniform) advance the offsets one row, or set them anew
      const int reg = regime(y);
      if (reg == lreg && y == lrow + 1) {
        of += bf;
        if constexpr (PASS == 2) og += bg;
        oq += ld;
      } else {
        int a, b;
        line(lf, reg, a, b);
        of = a + y * b;
        bf = b;
        if constexpr (PASS == 2) {
          line(lg, reg, a, b);
          og = a + y * b;
          bg = b;
        }
        oq = own + y * ld;
      }
      lrow = y;
      lreg = reg;
    }
    const uint32_t o = (uint32_t)(y * ld + xc), ou = (uint32_t)(y * ldu + xe);
    uw[S] = ld_at(U, ou);
    if constexpr (RK2) usw[S] = ld_at(US, ou);
    guw[S] = ld_at(gu, ou);
    gcw[S] = ld_at(gc, o);
    rgw[S] = ld_at(rgc, o);
    fr[S] = ld_at(X, (uint32_t)of);
    if constexpr (PASS == 2) {
      gr[S] = ld_at(X, (uint32_t)og);
      qsw[S] = ld_at(Qs, (uint32_t)oq);   // (at the clamped column: every lane reads in bounds)
    }
    const int ec = max(y - 3, 0);
    vr[S] = ld_at(V, (uint32_t)(ec * ld + xc));
    if constexpr (RK2) vsr[S] = ld_at(VS, (uint32_t)(ec * ld + xc));
    gvr[S] = ld_at(gv, (uint32_t)((ec + 1) * ld + xc));
  }

  // ET-PL07 along x: the lane's edge values when its column is next to a panel side (the
  // values every lane forms are exchanged by DPP; all lanes take part).
  __device__ __forceinline__ void pl07_x(T m2, T m1, T c, T p1, T p2, T &qL, T &qR) const {
    const T xl = pl_ext(c, p1), xr = pl_ext(c, m1);       // extrapolations to the left / right side
    const T eR = pl_edge(c, p1, p2), eL = pl_edge(c, m1, m2);
    const T alo = pl_avg(xl, from_left(xr));               // at column i0: own + neighbour's
    const T ahi = pl_avg(xr, from_right(xl));              // at column iend-1
    const T alo_g = from_right(alo), ahi_g = from_left(ahi);
    const T eR_l = from_left(eR), eL_r = from_right(eL);
    const T nL = from_right(qL), nR = from_left(qR);       // plain values of the next cells
    T L = qL, R = qR;
    if (lo_g) { L = eL; R = alo_g; }
    if (lo_b) { L = alo; R = eR; }
    if (lo_1) { L = eR_l; if (RECON == 3) R = nL; }
    if (hi_1) { if (RECON == 3) L = nR; R = eL_r; }
    if (hi_b) { L = eL; R = ahi; }
    if (hi_g) { L = ahi_g; R = eR; }
    qL = L;
    qR = R;
  }

  // ET-PL07 along y for parabola row r (wave-uniform): f3..p3 = the y-field at rows r-3 .. r+3.
  __device__ __forceinline__ void pl07_y(int r, T m3, T m2, T m1, T c, T p1, T p2, T p3, T &qL, T &qR) const {
    if (r == i0 - 1) { qL = pl_edge(c, m1, m2); qR = pl_avg(pl_ext(p1, p2), pl_ext(c, m1)); }
    else if (r == i0) { qL = pl_avg(pl_ext(c, p1), pl_ext(m1, m2)); qR = pl_edge(c, p1, p2); }
    else if (r == i0 + 1) { qL = pl_edge(m1, c, p1); if (RECON == 3) qR = pl_left(m1, c, p1, p2, p3); }
    else if (r == iend - 2) { if (RECON == 3) qL = pl_right(m3, m2, m1, c, p1); qR = pl_edge(p1, c, m1); }
    else if (r == iend - 1) { qL = pl_edge(c, m1, m2); qR = pl_avg(pl_ext(c, m1), pl_ext(p1, p2)); }
    else if (r == iend) { qL = pl_avg(pl_ext(m1, m2), pl_ext(c, p1)); qR = pl_edge(c, p1, p2); }
  }

  // Iteration y in phase K; STAGE as in the fused step: parabolas of row y-3 from y = tj+2 on
  // (>= 1), fluxes through y-edge y-3 from tj+3 on (>= 2), row y-4's result from tj+4 on (3).
  template <int K, int STAGE>
  __device__ __forceinline__ void body(const int y) {
    // slots of rows y, y-5 .. y-1 and y-6
    constexpr int s0 = sl(K, 0), s1 = sl(K, -5), s2 = sl(K, -4), s3 = sl(K, -3), s4 = sl(K, -2), s5 = sl(K, -1),
                  s6 = sl(K, -6);
    const int reg = regime(y);
    const T f = has(okf, reg) ? fr[s0] : T(0);               // row y of the row stage's field
    T cur = f;                                                // row y of the column stage's field
    if constexpr (PASS == 2) cur = has(okg, reg) ? gr[s0] : T(0);
    const T gvm1 = gvr[s4], gv0 = gvr[s5], gvp1 = gvr[s0];   // gv at y-edges y-4, y-3, y-2
    T res = T(0), qx = T(0);                                  // the row's results, stored below
    // ---- column stage: parabola of row y-3, y-edge y-3, result row y-4
    if constexpr (STAGE >= 1) {
      T qL, qR;
      recon_edges<T, RECON>(fw[s1], fw[s2], fw[s3], fw[s4], fw[s5], qL, qR);
      const int r = y - 3;
      T m3 = fw[s6];
      if constexpr (STAGE == 1 && RECON == 3) {
        // the first parabola row of a segment one row short of the north side (iend - 2): its
        // q_L override reads row r - 3, one before the march's first row, which no slot holds
        if (pl07 && r == iend - 2) m3 = PASS == 1 ? field(lf, okf, r - 3) : field(lg, okg, r - 3);
      }
      if (pl07 && ((r >= i0 - 1 && r <= i0 + 1) || (r >= iend - 2 && r <= iend)))
        pl07_y(r, m3, fw[s1], fw[s2], fw[s3], fw[s4], fw[s5], cur, qL, qR);
      const Parab<T> a = parab_of<T>(qL, qR, fw[s3], gv0, gvp1, gcw[s3], mt0);
      if constexpr (STAGE >= 2) {
        const T v = vr[s0];
        const bool vup = (RK2 ? vsr[s0] : v) >= T(0);
        const T cy = v * dtdy;
        const T g = upwind<T>(vup, pa_A, pa_B, pa_q6, a, cy, gv0, mt0);     // times dt/dy
        if constexpr (STAGE >= 3) {
          const int ro = y - 4;
          const T dG = gprev - g;
          if constexpr (PASS == 1) {
            T d1 = T(0), d2 = T(0);
            if (split != 1) {
              d1 = gv0 * cy;
              d2 = gvm1 * (vprev * dtdy);
            }
            res = split_update<T>(split, fw[s2], dG, gcw[s2], rgw[s2], d1, d2);   // Qy
          } else {
            // div = -(dF + dG) / (dt g);  Q - dt div = Q + (dF + dG) / g
            const T flx = dfw[s2] + dG;
            if (Div) {
              res = -flx / (dtc * gcw[s2]);
              if (out_lane) part += (double)res * (double)gcw[s2];
            } else {
              res = qsw[s2] + flx * rgw[s2];
            }
            if (fside && out_lane) {   // MF-AF: the y-fluxes through the S (edge i0) and N (edge iend) sides
              if (ro == i0) st_at(fside, (uint32_t)(3 * N + x - i0), gprev);
              if (ro == iend - 1) st_at(fside, (uint32_t)(2 * N + x - i0), g);
            }
          }
        }
        gprev = g;
        vprev = v;
      }
      pa_A = a.rA; pa_B = a.rB; pa_q6 = a.q6;
    }
    // ---- row stage: row y.  Its result (Qx, or the x-flux difference) is used only in the
    // segment's own rows: the halo rows of the march feed the column stage alone, since each pass
    // reads the other direction's half-step from the work buffer, not from this march.  So the
    // ny + 7 iterations of a segment run it ny times (a wave-uniform branch).
    const bool row_out = y >= tj && y < tend;
    if (row_out) {
      T m2, m1, p1, p2;
      xwindow<T, false>(f, 0, 0, 0, 0, m2, m1, p1, p2);
      T qL, qR;
      recon_edges<T, RECON>(m2, m1, f, p1, p2, qL, qR);
      if (sp_x) pl07_x(m2, m1, f, p1, p2, qL, qR);
      const T gu0 = guw[s0], gup1 = from_right(gu0), u = uw[s0];
      const bool up = (RK2 ? usw[s0] : u) >= T(0);
      const Parab<T> c = parab_of<T>(qL, qR, f, gu0, gup1, gcw[s0], mt0);
      const T f1 = upwind_x<T>(up, c, u * dtdx, gu0, gup1, mt0);     // left edge of the lane's cell
      const T dF = f1 - from_right(f1);
      if constexpr (PASS == 1) {
        T c1 = T(0), c2 = T(0);
        if (split != 1) {
          c1 = gup1 * (from_right(u) * dtdx);
          c2 = gu0 * (u * dtdx);
        }
        qx = split_update<T>(split, f, dF, gcw[s0], rgw[s0], c1, c2);
      } else {
        dfw[s0] = dF;
        if (fside) {               // MF-AF: the x-fluxes through the W (edge i0) and E (iend) sides
          if (w_lane) st_at(fside, (uint32_t)(1 * N + y - i0), f1);
          if (e_lane) st_at(fside, (uint32_t)(0 * N + y - i0), f1);
        }
      }
    }
    // row y replaces row y-RD in the ring of the column stage's field
    fw[s0] = cur;
    load<sl(K, PF)>(y + PF);
    // the row's stores (byte offsets within out; the first rows issue them too, dropped)
    constexpr uint32_t sz = sizeof(T);
    const int ro = y - 4;
    if constexpr (PASS == 1) {
      st_buf(outy, STAGE >= 3 && out_lane ? (uint32_t)(own + ro * ld) * sz : kNoStore, res);   // Qy, row y-4
      st_buf(out, out_lane && row_out ? (uint32_t)(own + y * ld) * sz : kNoStore, qx);               // Qx, row y
    } else {
      st_buf(out, STAGE >= 3 && out_lane ? (uint32_t)(own + ro * ld) * sz : kNoStore, res);
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
#pragma unroll
    for (int k = 0; k < RD; ++k)
      fr[k] = gr[k] = vr[k] = vsr[k] = fw[k] = dfw[k] = qsw[k] = gcw[k] = rgw[k] = gvr[k] = uw[k] = usw[k] = guw[k] = T(0);
    gprev = vprev = pa_A = pa_B = pa_q6 = T(0);
    part = 0.0;
    lrow = -1000;
    lreg = -1;
    of = og = oq = bf = bg = 0;
    int y = tj - 3;
    load<0>(y);
    load<1>(y + 1);
    // rows y0 .. y0+6 (ylast - y0 = ny + 6 >= 7: they all exist), the column stage starting
    row<0, 0, 0>(y); row<0, 0, 1>(y); row<0, 0, 2>(y); row<0, 0, 3>(y);
    row<0, 0, 4>(y); row<0, 1, 5>(y); row<0, 2, 6>(y);
    // rows y0+7 on: RD rows per iteration
    for (y += 7; y + RD - 1 <= ylast; y += RD) full_rows<7>(y, std::make_integer_sequence<int, RD>());
    last_rows<7>(y, std::make_integer_sequence<int, RD - 1>());
  }
};

// Work item b of one tracer's pass: band x = b % kXcds (the XCD that runs it) holds the row
// segments sg with sg * kXcds / segs == x, each listing its (panel, strip) items; past the
// band's end: padding.  Block 0 is always an item (band 0 starts at segment 0).
__device__ __forceinline__ bool gen_item(const Solve &S, int b, int &p, int &sx, int &sg) {
  const int x = b % kXcds, i = b / kXcds;
  const int lo = (x * S.segs + kXcds - 1) / kXcds, hi = ((x + 1) * S.segs + kXcds - 1) / kXcds;
  const int per = 6 * S.strips;
  if (i >= (hi - lo) * per) return false;
  const int r = i % per;
  p = r / S.strips;
  sx = r % S.strips;
  sg = lo + i / per;
  return true;
}

// The step a pass belongs to: pass 1 reads the steps completed (ctl[1]) and hands it on in
// ctl[2] (-1 when the solve is halted); the later kernels of the step read ctl[2], and the last
// one of the solve's last tracer advances ctl[1], so that no kernel reads a word it writes.
__device__ __forceinline__ bool gp_running(const csadv_solve_desc &D, int run) {
  return run >= 0 && !(D.stop > 0 && run >= D.stop);
}

template <typename T, int RECON, bool RK2, int PASS, int SPLIT>
__device__ __forceinline__ void gp_item(const Solve &S, int tr, int p, int sx, int sg, bool last) {
  const csadv_solve_desc &D = S.d;
  const int lane = threadIdx.x;
  int run;
  if constexpr (PASS == 1) {
    run = D.ctl[0] ? -1 : D.ctl[1];
    if (blockIdx.x == 0 && lane == 0) D.ctl[2] = run;
  } else {
    run = D.ctl[2];
  }
  if (!gp_running(D, run)) return;
  const int N = D.N, ld = D.ld, iend = N + NG;
  const int ti = NG + sx * S.strip_w, tj = NG + sg * S.seg_l;
  const int nx = min(S.strip_w, iend - ti), ny = min(S.seg_l, iend - tj);
  const int ps = (int)S.pstride;
  const WorkLayout w = work_layout(N, ld, D.ldu);
  T *work = reinterpret_cast<T *>(D.work);
  const bool odd = run & 1;                   // the state of step n lives in q[n & 1]
  const int64_t toff = tr * S.tstride;
  T *qin = reinterpret_cast<T *>(odd ? D.q1 : D.q0) + toff, *qout = reinterpret_cast<T *>(odd ? D.q0 : D.q1) + toff;
  GMarch<T, RECON, RK2, PASS, SPLIT> m;
  m.ld = ld;
  m.ldu = D.ldu;
  m.N = N;
  m.i0 = NG;
  m.iend = iend;
  m.x = ti - 3 + lane;
  m.xc = min(m.x, S.M - 1);
  m.xe = min(m.x, S.M);
  m.tj = tj;
  m.tend = tj + ny;
  m.ylast = tj + ny + 3;
  m.cell = (int)w.cell;
  // (the clamped column: the lanes past the array's last column read Q in bounds)
  m.own = p * ps + m.xc;
  m.pl07 = D.et == 2;
  m.out_lane = lane >= 3 && lane < 3 + nx;
  m.w_lane = ti == NG && lane == 3;
  m.e_lane = ti + nx == iend && lane == 3 + nx;
  m.lo_g = m.x == NG - 1;
  m.lo_b = m.x == NG;
  m.lo_1 = m.x == NG + 1;
  m.hi_1 = m.x == iend - 2;
  m.hi_b = m.x == iend - 1;
  m.hi_g = m.x == iend;
  m.sp_x = m.pl07 && (ti - 3 <= NG + 1 || ti + nx + 2 >= iend - 2);
  m.dtdx = T(D.dt / D.dx);
  m.dtdy = T(D.dt / D.dy);
  m.dtc = T(D.dt);
  m.U = reinterpret_cast<const T *>(D.uavg) + p * S.pstride_u;
  m.V = reinterpret_cast<const T *>(D.vavg) + p * S.pstride_v;
  m.US = RK2 ? reinterpret_cast<const T *>(D.usgn) + p * S.pstride_u : m.U;
  m.VS = RK2 ? reinterpret_cast<const T *>(D.vsgn) + p * S.pstride_v : m.V;
  m.gc = reinterpret_cast<const T *>(D.gc);
  m.gu = reinterpret_cast<const T *>(D.gu);
  m.gv = reinterpret_cast<const T *>(D.gv);
  m.rgc = reinterpret_cast<const T *>(D.rgc);
  m.Qs = qin;
  m.Div = false;
  m.fside = nullptr;
  const int cell = m.cell;
  const uint32_t cell_bytes = (uint32_t)(w.cell * sizeof(T));
  m.lg = Lines{0, 0, 0, 0, 0, 0, 0u};
  if constexpr (PASS == 1) {
    m.X = qin;
    m.out = buffer_of(work + w.o_qx, cell_bytes);
    m.outy = buffer_of(work + w.o_qy, cell_bytes);
    m.lf = lane_lines(D, ps, p, m.x, 0, true, 0, true, 0);
  } else {
    m.X = work + w.o_qx;
    m.Div = D.mf == 3;
    m.out = buffer_of(m.Div ? work + w.o_div : qout, cell_bytes);
    if (D.mf == 2) m.fside = work + w.o_side + p * 4 * N;
    // the row stage's Qy (own at + cell; E/W ghosts from the neighbour's Qx on panels 4, 5)
    m.lf = lane_lines(D, ps, p, m.x, cell, true, (p == 4 || p == 5) ? 0 : cell, false, 0);
    // the column stage's Qx (N/S ghosts from the neighbour's Qy on panels 1, 3)
    m.lg = lane_lines(D, ps, p, m.x, 0, false, 0, true, (p == 1 || p == 3) ? cell : 0);
  }
  for (int k = 0; k < 3; ++k) {
    m.okf[k] = (m.lf.ok >> k) & 1u;
    m.okg[k] = (m.lg.ok >> k) & 1u;
  }
  m.run();
  if constexpr (PASS == 2) {
    if (D.mf == 3) {
      const double part = warp_sum(m.part);
      double *red = reinterpret_cast<double *>(work + w.o_red);
      if (lane == 0) red[(p * S.segs + sg) * S.strips + sx] = part;
    }
    if (last && blockIdx.x == 0 && lane == 0) D.ctl[1] = run + 1;
  }
}

// One tracer's pass of one solve; the descriptor travels in the kernel arguments.  last: this
// pass ends the step (pass 2 of the last tracer with neither MF-AF nor MF-PR after it).
template <typename T, int RECON, bool RK2, int PASS, int SPLIT>
__global__ __launch_bounds__(WAVE) void k_gp_pass(const Solve one, int tr, int last) {
  int p, sx, sg;
  if (!gen_item(one, touch_solve(one, blockIdx.x), p, sx, sg)) return;
  gp_item<T, RECON, RK2, PASS, SPLIT>(one, tr, p, sx, sg, last != 0);
}

// MF-AF: the cube-edge averages of the second pass's fluxes into the boundary cells of the new
// state (af_fix_cell).  grid: boundary cells of the 6 panels.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_gp_af_fix(const Solve one, int tr, int last) {
  const csadv_solve_desc &D = one.d;
  const int run = D.ctl[2];
  if (!gp_running(D, run)) return;
  if (last && blockIdx.x == 0 && threadIdx.x == 0) D.ctl[1] = run + 1;
  const WorkLayout w = work_layout(D.N, D.ld, D.ldu);
  af_fix_cell<T>(one, tr, run, reinterpret_cast<const T *>(D.work) + w.o_side, blockIdx.x * BLOCK + threadIdx.x);
}

// MF-PR: the sum of div g over all cells, from the nparts wavefront partials in a fixed order
// (each lane a strided sum, the wavefronts by shuffles, then the four in index order), into the
// last double of the tail.  One block.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_gp_prsum(const Solve one, int nparts) {
  __shared__ double red[BLOCK / WAVE];
  const csadv_solve_desc &D = one.d;
  if (!gp_running(D, D.ctl[2])) return;
  const WorkLayout w = work_layout(D.N, D.ld, D.ldu);
  double *part = reinterpret_cast<double *>(reinterpret_cast<T *>(D.work) + w.o_red);
  double acc = 0.0;
  for (int k = threadIdx.x; k < nparts; k += BLOCK) acc += part[k];
  acc = warp_sum(acc);
  if ((threadIdx.x & (WAVE - 1)) == 0) red[threadIdx.x / WAVE] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[w.n_red - 1] = ((red[0] + red[1]) + red[2]) + red[3];
}

// MF-PR (src/discrete_operators.py:95-101): Q^{n+1} = Q^n - dt (div - g m0 / a2) on the
// interior, m0 the sum of div g; one thread per interior cell.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_gp_update(const Solve one, int tr, int last) {
  const csadv_solve_desc &D = one.d;
  const int run = D.ctl[2];
  if (!gp_running(D, run)) return;
  if (last && blockIdx.x == 0 && threadIdx.x == 0) D.ctl[1] = run + 1;
  const int N = D.N, ld = D.ld;
  const WorkLayout w = work_layout(N, ld, D.ldu);
  const T *work = reinterpret_cast<const T *>(D.work);
  const double m0 = reinterpret_cast<const double *>(work + w.o_red)[w.n_red - 1];
  const bool odd = run & 1;
  const int64_t toff = tr * one.tstride;
  const T *qin = reinterpret_cast<const T *>(odd ? D.q1 : D.q0) + toff;
  T *qout = reinterpret_cast<T *>(odd ? D.q0 : D.q1) + toff;
  const T *div = work + w.o_div, *gc = reinterpret_cast<const T *>(D.gc);
  const unsigned total = 6u * N * N;
  for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const unsigned u = t / N;
    const int i = (int)(t % N) + NG, j = (int)(u % N) + NG, p = u / N;
    const int64_t c = p * one.pstride + (int64_t)j * ld + i;
    const T dv = div[c] - (gc[(int64_t)j * ld + i] * T(m0)) / T(D.a2);
    qout[c] = qin[c] - T(D.dt) * dv;
  }
}

// Blocks of one tracer's pass: kXcds bands of the longest band's length.
inline int gen_blocks(const Solve &S) { return kXcds * ((S.segs + kXcds - 1) / kXcds) * 6 * S.strips; }

}  // namespace gp

// One step of every solve by the general step: per tracer the two passes, then the mass fixer's
// kernels; the last kernel of the last tracer advances the step counter.
template <typename T, int R, bool RK2, int SP>
int enqueue_general_k(csadv_t *h, hipStream_t st) {
  for (size_t s = 0; s < h->host.size(); ++s) {
    const Solve &S = h->host[s];
    const csadv_solve_desc &d = S.d;
    const int blocks = gp::gen_blocks(S);
    for (int tr = 0; tr < d.ntr; ++tr) {
      const int last = tr == d.ntr - 1;
      hipLaunchKernelGGL((gp::k_gp_pass<T, R, RK2, 1, SP>), dim3(blocks), dim3(WAVE), 0, st, S, tr, 0);
      hipLaunchKernelGGL((gp::k_gp_pass<T, R, RK2, 2, SP>), dim3(blocks), dim3(WAVE), 0, st, S, tr,
                         last && d.mf == 1);
      if (d.mf == 2) {
        const int cells = 6 * 4 * (d.N - 1);
        hipLaunchKernelGGL(gp::k_gp_af_fix<T>, dim3((cells + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, S, tr, last);
      } else if (d.mf == 3) {
        hipLaunchKernelGGL(gp::k_gp_prsum<T>, dim3(1), dim3(BLOCK), 0, st, S, 6 * S.strips * S.segs);
        hipLaunchKernelGGL(gp::k_gp_update<T>, dim3(gp_blocks((int64_t)6 * d.N * d.N)), dim3(BLOCK), 0, st, S, tr,
                           last);
      }
    }
  }
  HIPCHK(h, hipGetLastError());
  return CSADV_OK;
}

template <typename T, int R, int SP>
int enqueue_general_r(csadv_t *h, hipStream_t st) {
  return h->rk2 ? enqueue_general_k<T, R, true, SP>(h, st) : enqueue_general_k<T, R, false, SP>(h, st);
}

template <typename T, int SP>
int enqueue_general_split(csadv_t *h, hipStream_t st) {
  switch (h->recon) {
    case 1: return enqueue_general_r<T, 1, SP>(h, st);
    case 2: return enqueue_general_r<T, 2, SP>(h, st);
    case 4: return enqueue_general_r<T, 4, SP>(h, st);
    default: return enqueue_general_r<T, 3, SP>(h, st);
  }
}
