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
// the 3-cell halo in the outer lanes, x windows by DPP, a six-row ring for the y windows).  No
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
// Included by general_f64.hip and general_f32.hip, which instantiate enqueue_general_t.

#pragma once
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

template <typename T>
struct GIn {        // inputs of iteration y at the lane's column
  T f;              // row y of the field the row stage reconstructs (pass 1: Q; pass 2: Qy)
  T g;              // pass 2: row y of the field the column stage reconstructs (Qx)
  T v, vs;          // y-edge y-3: wind, upwind-side wind
};

// One wavefront of a pass: strip columns ti .. ti+nx-1 (lane l <-> column ti-3+l), rows
// tj .. tj+ny-1.  Iteration y (rows y0 = tj-3 .. ylast = tj+ny+3):
//   * column stage at y-edge y-3 from the six-row ring of the y-field (rows y-6 .. y-1): the
//     parabola of row y-3, the y-flux through edge y-3, then for row y-4 the y-split Qy
//     (pass 1) or the update (pass 2, with the x-flux difference the row stage kept);
//   * row stage for row y: x-window of the x-field by DPP, parabolas, x-fluxes (every lane the
//     flux through its cell's left edge), then Qx (pass 1) or the x-flux difference (pass 2).
// The inputs of row y+PF are requested at the end of iteration y, PF = 2 rows ahead.
template <typename T, int RECON, bool RK2, int PASS>
struct GMarch {
  static constexpr int PF = 2;
  // ---- per wavefront
  const T *X;              // pass 1: Q of the tracer; pass 2: Qx of the tracer (Qy = Qx + cell)
  const T *Qs;             // pass 2: Q of the tracer (the update's input)
  T *Out;                  // pass 1: Qx of the tracer (Qy = Qx + cell); pass 2: the new Q
  T *Div;                  // pass 2 with MF-PR: div of the tracer (the update follows later)
  T *fside;                // pass 2 with MF-AF: cube-edge fluxes of the panel, [E, W, N, S][N]
  const T *U, *US, *V, *VS, *gc, *gu, *gv, *rgc;
  Lines lf, lg;            // lines of the row stage's field and (pass 2) the column stage's
  int ld, ldu, N, i0, iend, x, xc, xe, tj, tend, ylast, cell, own, split;
  bool mt0, pl07, sp_x, out_lane, w_lane, e_lane;
  bool lo_g, lo_b, lo_1, hi_1, hi_b, hi_g;   // column x is i0-1, i0, i0+1, iend-2, iend-1, iend
  T dtdx, dtdy, dtc;
  // ---- rings: six rows (slot of row y-6+i in phase K: (K+i) % 6), winds of three rows
  T fw[6];                 // the column stage's field (pass 1: Q, pass 2: Qx), rows y-6 .. y-1
  T dfw[6];                // pass 2: x-flux difference of the rows
  T qsw[6];                // pass 2: Q of the rows (loaded straight into their slots)
  T gcw[6], rgw[6], gvr[6];   // metric, 1/metric of the rows; gv at y-edges (edge y-2 in slot K)
  T uw[3], usw[3], guw[3];    // x-edge wind, upwind-side wind, gu of row y in slot K % 3
  T gprev, vprev, pa_A, pa_B, pa_q6;
  double part;             // pass 2 with MF-PR: sum of div g over the wavefront's cells

  __device__ __forceinline__ T field(const Lines &L, int y, int reg) const {
    // (the members are read into values first: a select between member lvalues would keep the
    // march's state in scratch memory)
    const int a0 = L.a0, a1 = L.a1, a2 = L.a2, b0 = L.b0, b1 = L.b1, b2 = L.b2;
    const int a = reg == 0 ? a0 : (reg == 1 ? a1 : a2);
    const int b = reg == 0 ? b0 : (reg == 1 ? b1 : b2);
    T v = T(0);
    if ((L.ok >> reg) & 1u) v = ld_at(X, (uint32_t)(a + y * b));
    return v;
  }

  // Inputs of row y into ring slots S6 (six-row rings) and S3 (wind rings).
  template <int S6, int S3>
  __device__ __forceinline__ void load(int y, GIn<T> &r) {
    y = min(y, ylast);       // rows past the last are requested in bounds and never used
    const int reg = y < i0 ? 0 : (y < iend ? 1 : 2);
    const uint32_t o = (uint32_t)(y * ld + xc), ou = (uint32_t)(y * ldu + xe);
    uw[S3] = ld_at(U, ou);
    if constexpr (RK2) usw[S3] = ld_at(US, ou);
    guw[S3] = ld_at(gu, ou);
    gcw[S6] = ld_at(gc, o);
    rgw[S6] = ld_at(rgc, o);
    r.f = field(lf, y, reg);
    if constexpr (PASS == 2) {
      r.g = field(lg, y, reg);
      T q = T(0);
      if (out_lane && y >= tj && y < tend) q = ld_at(Qs, (uint32_t)(own + y * ld));
      qsw[S6] = q;
    }
    const int ec = max(y - 3, 0);
    r.v = ld_at(V, (uint32_t)(ec * ld + xc));
    r.vs = RK2 ? ld_at(VS, (uint32_t)(ec * ld + xc)) : r.v;
    gvr[S6] = ld_at(gv, (uint32_t)((ec + 1) * ld + xc));
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
  __device__ __forceinline__ void body(const int y, GIn<T> &in) {
    constexpr int s0 = K % 6, s1 = (K + 1) % 6, s2 = (K + 2) % 6, s3 = (K + 3) % 6, s4 = (K + 4) % 6,
                  s5 = (K + 5) % 6, w0 = K % 3;
    const T cur = PASS == 1 ? in.f : in.g;                   // row y of the column stage's field
    const T gvm1 = gvr[s4], gv0 = gvr[s5], gvp1 = gvr[s0];   // gv at y-edges y-4, y-3, y-2
    // ---- column stage: parabola of row y-3, y-edge y-3, result row y-4
    if constexpr (STAGE >= 1) {
      T qL, qR;
      recon_edges<T, RECON>(fw[s1], fw[s2], fw[s3], fw[s4], fw[s5], qL, qR);
      const int r = y - 3;
      T m3 = fw[s0];
      if constexpr (STAGE == 1 && RECON == 3) {
        // the first parabola row of a segment one row short of the north side (iend - 2): its
        // q_L override reads row r - 3, one before the march's first row, which no slot holds
        if (pl07 && r == iend - 2) m3 = field(PASS == 1 ? lf : lg, r - 3, 1);
      }
      if (pl07 && ((r >= i0 - 1 && r <= i0 + 1) || (r >= iend - 2 && r <= iend)))
        pl07_y(r, m3, fw[s1], fw[s2], fw[s3], fw[s4], fw[s5], cur, qL, qR);
      const Parab<T> a = parab_of<T>(qL, qR, fw[s3], gv0, gvp1, gcw[s3], mt0);
      if constexpr (STAGE >= 2) {
        const T v = in.v;
        const bool vup = in.vs >= T(0);
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
            const T qy = split_update<T>(split, fw[s2], dG, gcw[s2], rgw[s2], d1, d2);
            if (out_lane) st_at(Out, (uint32_t)(cell + own + ro * ld), qy);
          } else {
            // div = -(dF + dG) / (dt g);  Q - dt div = Q + (dF + dG) / g
            const T flx = dfw[s2] + dG;
            if (out_lane) {
              if (Div) {
                const T dv = -flx / (dtc * gcw[s2]);
                st_at(Div, (uint32_t)(own + ro * ld), dv);
                part += (double)dv * (double)gcw[s2];
              } else {
                st_at(Out, (uint32_t)(own + ro * ld), qsw[s2] + flx * rgw[s2]);
              }
              if (fside) {   // MF-AF: the y-fluxes through the S (edge i0) and N (edge iend) sides
                if (ro == i0) st_at(fside, (uint32_t)(3 * N + x - i0), gprev);
                if (ro == iend - 1) st_at(fside, (uint32_t)(2 * N + x - i0), g);
              }
            }
          }
        }
        gprev = g;
        vprev = v;
      }
      pa_A = a.rA; pa_B = a.rB; pa_q6 = a.q6;
    }
    // ---- row stage: row y
    const T f = in.f;
    T m2, m1, p1, p2;
    xwindow<T, false>(f, 0, 0, 0, 0, m2, m1, p1, p2);
    T qL, qR;
    recon_edges<T, RECON>(m2, m1, f, p1, p2, qL, qR);
    if (sp_x) pl07_x(m2, m1, f, p1, p2, qL, qR);
    const T gu0 = guw[w0], gup1 = from_right(gu0), u = uw[w0];
    const bool up = (RK2 ? usw[w0] : u) >= T(0);
    const Parab<T> c = parab_of<T>(qL, qR, f, gu0, gup1, gcw[s0], mt0);
    const T f1 = upwind_x<T>(up, c, u * dtdx, gu0, gup1, mt0);     // left edge of the lane's cell
    const T dF = f1 - from_right(f1);
    const bool row_out = y >= tj && y < tend;
    if constexpr (PASS == 1) {
      T c1 = T(0), c2 = T(0);
      if (split != 1) {
        c1 = gup1 * (from_right(u) * dtdx);
        c2 = gu0 * (u * dtdx);
      }
      const T qx = split_update<T>(split, f, dF, gcw[s0], rgw[s0], c1, c2);
      if (out_lane && row_out) st_at(Out, (uint32_t)(own + y * ld), qx);
    } else {
      dfw[s0] = dF;
      if (fside && row_out) {      // MF-AF: the x-fluxes through the W (edge i0) and E (iend) sides
        if (w_lane) st_at(fside, (uint32_t)(1 * N + y - i0), f1);
        if (e_lane) st_at(fside, (uint32_t)(0 * N + y - i0), f1);
      }
    }
    // row y replaces row y-6 in the ring of the column stage's field
    fw[s0] = cur;
    load<(K + PF) % 6, (K + PF) % 3>(y + PF, in);
  }

  __device__ __forceinline__ void run(GIn<T> *b) {
#pragma unroll
    for (int k = 0; k < 6; ++k) fw[k] = dfw[k] = qsw[k] = gcw[k] = rgw[k] = gvr[k] = T(0);
#pragma unroll
    for (int k = 0; k < 3; ++k) uw[k] = usw[k] = guw[k] = T(0);
    gprev = vprev = pa_A = pa_B = pa_q6 = T(0);
    part = 0.0;
    int y = tj - 3;
    load<0, 0>(y, b[0]);
    load<1, 1>(y + 1, b[1]);
    // rows y0 .. y0+6 (ylast - y0 = ny + 6 >= 7: they all exist), the column stage starting
    body<0, 0>(y, b[0]); body<1, 0>(y + 1, b[1]); body<2, 0>(y + 2, b[0]); body<3, 0>(y + 3, b[1]);
    body<4, 0>(y + 4, b[0]); body<5, 1>(y + 5, b[1]); body<0, 2>(y + 6, b[0]);
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

template <typename T, int RECON, bool RK2, int PASS>
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
  GMarch<T, RECON, RK2, PASS> m;
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
  m.own = p * ps + m.x;
  m.split = D.split;
  m.mt0 = D.split != 3;
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
  m.Div = nullptr;
  m.fside = nullptr;
  const int cell = m.cell;
  if constexpr (PASS == 1) {
    m.X = qin;
    m.Out = work + w.o_qx;
    m.lf = lane_lines(D, ps, p, m.x, 0, true, 0, true, 0);
  } else {
    m.X = work + w.o_qx;
    m.Out = qout;
    if (D.mf == 3) m.Div = work + w.o_div;
    if (D.mf == 2) m.fside = work + w.o_side + p * 4 * N;
    // the row stage's Qy (own at + cell; E/W ghosts from the neighbour's Qx on panels 4, 5)
    m.lf = lane_lines(D, ps, p, m.x, cell, true, (p == 4 || p == 5) ? 0 : cell, false, 0);
    // the column stage's Qx (N/S ghosts from the neighbour's Qy on panels 1, 3)
    m.lg = lane_lines(D, ps, p, m.x, 0, false, 0, true, (p == 1 || p == 3) ? cell : 0);
  }
  GIn<T> b[2];
  m.run(b);
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
template <typename T, int RECON, bool RK2, int PASS>
__global__ __launch_bounds__(WAVE) void k_gp_pass(const Solve one, int tr, int last) {
  int p, sx, sg;
  if (!gen_item(one, touch_solve(one, blockIdx.x), p, sx, sg)) return;
  gp_item<T, RECON, RK2, PASS>(one, tr, p, sx, sg, last != 0);
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
template <typename T, int R, bool RK2>
int enqueue_general_k(csadv_t *h, hipStream_t st) {
  for (size_t s = 0; s < h->host.size(); ++s) {
    const Solve &S = h->host[s];
    const csadv_solve_desc &d = S.d;
    const int blocks = gp::gen_blocks(S);
    for (int tr = 0; tr < d.ntr; ++tr) {
      const int last = tr == d.ntr - 1;
      hipLaunchKernelGGL((gp::k_gp_pass<T, R, RK2, 1>), dim3(blocks), dim3(WAVE), 0, st, S, tr, 0);
      hipLaunchKernelGGL((gp::k_gp_pass<T, R, RK2, 2>), dim3(blocks), dim3(WAVE), 0, st, S, tr,
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

template <typename T, int R>
int enqueue_general_r(csadv_t *h, hipStream_t st) {
  return h->rk2 ? enqueue_general_k<T, R, true>(h, st) : enqueue_general_k<T, R, false>(h, st);
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
