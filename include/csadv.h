/*
 * csadv.h -- C ABI of libcsadv.so, the MI355X-native dimension-split finite-volume
 * advection step on the gnomonic cubed sphere (PPM reconstruction, upwind fluxes,
 * operator splitting, duo-grid ghost cells).
 *
 * Every state buffer is device memory owned by the caller (PyTorch tensors on the
 * host side); the library keeps non-owning pointers to it and never frees it.  The
 * library allocates only its own small descriptor tables, in csadv_create; a step
 * neither allocates nor synchronises the stream.  All entry points return
 * 0 on success and a negative CSADV_E* code on failure; the message of the last
 * failure of a handle is returned by csadv_last_error().  Handles are not
 * thread-safe; independent handles are.
 *
 * Layouts (elements; one solve with N cells per panel edge, M = N + 8, halo 4):
 *   tracer state  q[b] : [ntr][6][M][ld]     i (x) fastest, j (y), panel, tracer
 *   x-edge winds       : [6][M][ldu]         e = 0..M along x, ldu >= M + 1
 *   y-edge winds       : [6][M + 1][ld]      e = 0..M along y
 *   metric  gc / rgc / gfix : [M][ld]   gu : [M][ldu]   gv : [M + 1][ld]  (panel invariant)
 * The reference stores the same values as numpy arrays indexed [i, j, panel].
 */
#ifndef CSADV_H
#define CSADV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CSADV_ABI_VERSION 7

enum csadv_dtype { CSADV_F64 = 0, CSADV_F32 = 1 };

enum csadv_status {
  CSADV_OK = 0,
  CSADV_E_ARG = -1,      /* invalid argument / option pair            */
  CSADV_E_HIP = -2,      /* a HIP runtime call failed                 */
  CSADV_E_STATE = -3,    /* call out of order (e.g. step before bind) */
  CSADV_E_RANGE = -4     /* size beyond a documented limit            */
};

/* Per-solve description.  Option codes are exactly those of par/advection.par
 * (src/configuration.py:90-158, decoded in src/advection_ic.py:85-149):
 *   recon 1 PPM-0 | 2 PPM-CW84 | 3 PPM-PL07 | 4 PPM-L04
 *   dp    1 RK1   | 2 RK2
 *   split 1 AVLT  | 2 L04      | 3 PL07
 *   et    1 ET-S72 | 2 ET-PL07 | 3 ET-DG
 *   mt    1 MT-0  | 2 MT-PL07   (valid pairs: split 1|2 with mt 1, split 3 with mt 2)
 *   mf    1 none  | 2 MF-AF     | 3 MF-PR
 * ET-DG runs the fused step kernel (with MF-AF: plus a fix-up of the cube-edge cells);
 * ET-S72 and ET-PL07 run the general step (two marched passes per tracer, the copy halo
 * read from the adjacent panels in the passes).  Both need the work buffer where noted
 * below.  All solves of one handle take the same one of the two.                          */
typedef struct csadv_solve_desc {
  int32_t N, ntr, dtype;
  int32_t ld, ldu;                 /* row strides (elements) of cell / x-edge arrays   */
  int32_t recon, dp, split, et, mt, mf;
  int32_t nbasis;                  /* wind basis fields (>= 1)                          */
  int32_t unsteady;                /* 1: winds recombined from the basis every step     */
  int32_t max_steps;               /* rows of the coefficient table (unsteady only)     */
  double dt, dx, dy;
  double a2;                       /* sum of gc^2 over the interior of the 6 panels    */
  /* device buffers */
  void *q0, *q1;                   /* ping-pong tracer state (dtype)                    */
  void *uavg, *vavg;               /* time-averaged contravariant winds (dtype)         */
  void *usgn, *vsgn;               /* contravariant winds fixing the upwind side (dtype)*/
  const void *gc, *gu, *gv, *rgc;  /* metric tables (dtype)                             */
  const void *gfix;                /* fused step: the metric the MF-PR correction total
                                      multiplies (dtype): gc inside the panel, sum_l w_l
                                      gc(src_l) of the ghost recipe at the 3 ghost layers (the
                                      same on every panel); the per-step ghost fill stores the
                                      uncorrected sums and the step adds C gfix on read     */
  const double *gc64;              /* metric at centres, FP64 (mass fixer, mass)        */
  const double *lagw;              /* Lagrange weights [4][M][4], FP64                  */
  const int32_t *lagk;             /* stencil starts   [4][M]                            */
  const int32_t *orient;           /* halo orientation [6][4][7]                        */
  double *wbu, *wbv;               /* wind basis, FP64: [nbasis][6][M][ldu] / [nbasis][6][M+1][ld] */
  const double *conv_pc;           /* [4][6][M][ld]   ex.e_lon, ex.e_lat, ey.e_lon, ey.e_lat       */
  const double *conv_pu;           /* [5][6][M][ldu]  same four and the determinant at x-edges     */
  const double *conv_pv;           /* [5][6][M+1][ld] same at y-edges                               */
  double *wpc;                     /* scratch [nbasis][2][6][M][ld]: lat-lon winds at centres      */
  int32_t n_strip, n_corner;       /* ghost recipes: edge-strip and corner cells          */
  const int32_t *strip_tgt;        /* [n_strip] target offsets inside one tracer           */
  const int32_t *strip_src;        /* [n_strip][4] source offsets inside one tracer        */
  const double *strip_w;           /* [n_strip][4] weights                                  */
  const int32_t *corner_tgt;       /* [n_corner]                                            */
  const int32_t *corner_src;       /* [n_corner][16]                                        */
  const double *corner_w;          /* [n_corner][16]                                        */
  const double *coef;              /* [max_steps][2][nbasis] basis coefficients (t_{n-1}, t_{n-2}) */
  double *corr;                    /* [2][ntr] MF-PR correction totals C (FP64), slot by step parity: Q = S + C gfix */
  double *partials;                /* scratch, >= ntr * csadv_partials_per_tracer()     */
  int32_t *ctl;                    /* [4] device control words, zero at init: halt flag
                                      (csadv_diagnostics), steps completed, running step,
                                      mass-fixer partials pending                          */
  int32_t tile;                    /* step tile shape of this solve (csadv_tile_rows);
                                      -1 = the handle's tile.  The per-wavefront mass-fixer
                                      partial sums follow the tile, so a solve keeps the same
                                      bits alone and in any batch when its tile is fixed here */
  int32_t stop;                    /* > 0: the step kernels leave this solve untouched once it
                                      has completed `stop` steps (a batch can then run a fixed
                                      number of steps while its solves end at their own step);
                                      0 = no limit                                           */
  void *work;                      /* general step (ET-S72, ET-PL07): scratch of csadv_work_bytes()
                                      bytes (the general step takes N up to 6680 in FP64:
                                      CSADV_E_RANGE past that); fused step with
                                      MF-AF: ntr * 24 * N elements of the dtype (the second
                                      pass's cube-edge fluxes); else unused */
  double *offset;                  /* fused step, FP32: [ntr] constants (exact in FP32) the stored
                                      state is kept from, Q = S + C gfix + offset, so that
                                      increments far below an ulp of Q (on a flat background)
                                      are not rounded away; NULL = 0.  csadv_finalize folds them
                                      into the state and zeroes them.  NULL for the general step */
} csadv_solve_desc;

typedef struct csadv_s csadv_t;

/* Create a handle for nsolve solves (1 = the reference's single run; > 1 = batched form).
 * tile selects the step kernel's tile shape (0 = built-in default).
 * Replaces: adv_simulation_par construction + option validation (src/advection_ic.py:21-203,
 * src/discrete_operators.py:48-82). */
int csadv_create(const csadv_solve_desc *solves, int32_t nsolve, int32_t tile, csadv_t **out);

/* Number of FP64 partial sums the step writes per tracer (size of desc.partials / ntr). */
int csadv_partials_per_tracer(int32_t N, int32_t tile);

/* Bytes of the work buffer the general step (ET-S72, ET-PL07) needs for one solve of
 * size N and dtype (written to *bytes).  No counterpart in the reference (its scratch arrays
 * are members of the simulation object, src/cs_datastruct.py:635-684). */
int csadv_work_bytes(int32_t N, int32_t dtype, int64_t *bytes);

/* Rows per wavefront segment of a tile shape (tile = 0 .. count-1), or CSADV_E_ARG past the
 * last shape: the tile ids an autotuner may try.  No counterpart in the reference. */
int csadv_tile_rows(int32_t tile);

/* Refresh the device copy of the descriptors after the caller changed pointers. */
int csadv_update(csadv_t *h, const csadv_solve_desc *solves, int32_t nsolve, void *stream);

/* Wind preparation from the FP64 basis fields whose interior edges are set:
 * duo-grid vector ghost fill of every basis field (src/edges_treatment.py:296-347 with
 * src/interpolation.py:347-532), then the time-averaged winds of step 1
 * (src/averaged_velocity.py:14-62).  Call once after the basis interiors are written. */
int csadv_wind_init(csadv_t *h, void *stream);

/* Scalar duo-grid ghost fill of every tracer of the current state buffer
 * (src/edges_treatment.py:284-290 -> src/interpolation.py:154-315). */
int csadv_ghost_fill(csadv_t *h, void *stream);

/* Advance every solve nsteps steps: ghost fill, winds (unsteady), split fluxes, update,
 * mass fixer.  Replaces adv_time_step + update_adv (src/advection_timestep.py:19-75)
 * called nsteps times.  Graph-capturable (no host synchronisation, no allocation). */
int csadv_step(csadv_t *h, int32_t nsteps, void *stream);

/* Replay steps from a HIP graph: steps_per_graph > 0 makes csadv_step capture that many
 * consecutive steps (all their kernels) once, on a private capture stream, and launch the
 * instantiated graph on the caller's stream for every whole group of steps; remaining
 * steps, and every step while csadv_profile pairs are armed, run as plain launches.
 * Results are bitwise identical to plain launches.  0 (the default) switches graphs off
 * and releases the graph.  No counterpart in the reference (its step is a Python call). */
int csadv_set_graph(csadv_t *h, int32_t steps_per_graph);

/* Fold the mass-fixer correction totals of one solve (solve = -1: every solve) into its
 * current state buffer so that it holds Q exactly as the reference would after the last
 * step.  During stepping the stored interior values S stay uncorrected (Q = S + C g with an
 * FP64 total C per tracer), so that an FP32 state does not round the small per-step
 * corrections away.  Other solves of the handle are left untouched, so reading one result never
 * changes the arithmetic of the solves still running. */
int csadv_finalize(csadv_t *h, int32_t solve, void *stream);

/* Total mass sum_int Q g dx dy of one tracer with a compensated FP64 sum
 * (src/diagnostics.py:14-26): csadv_diagnostics, then a copy to host out[0].  Synchronises. */
int csadv_mass(csadv_t *h, int32_t solve, int32_t tracer, double *out, void *stream);

/* Error norms of one tracer against qexact (device, [6][N][N] FP64, i fastest):
 * host out = {Linf, L1, L2} as src/errors.py:99-113 (unweighted means).  Synchronises. */
int csadv_norms(csadv_t *h, int32_t solve, int32_t tracer, const double *qexact, double *out, void *stream);

/* Diagnostics of tracers tracer0 .. tracer0 + ntracers - 1 of one solve into device memory,
 * without host synchronisation or allocation (a graph may capture it): out ([ntracers][4] device
 * doubles) = {mass, Linf, L1, L2} per tracer -- the compensated FP64 mass of
 * src/diagnostics.py:14-26 and the norms of src/errors.py:99-113 against qexact (device
 * [ntracers][6][N][N] FP64), or of the field itself when qexact is NULL.  Reductions run over
 * wavefront shuffles in a fixed order (the same bits on every call).  blowup_limit > 0: a Linf above it
 * halts the solve -- its step kernels leave it untouched from then on (src/output.py:44-47 stops
 * the run when Linf > 100); csadv_status reports the halt.  Replaces mass_computation +
 * compute_errors + the blow-up test of output_adv. */
int csadv_diagnostics(csadv_t *h, int32_t solve, int32_t tracer0, int32_t ntracers, const double *qexact,
                      double blowup_limit, double *out, void *stream);

/* csadv_diagnostics with its block partials in a fixed device array of the library instead of the
 * handle's scratch, so that the call allocates nothing and needs no scratch of its own: at most 16
 * tracers per call (CSADV_E_RANGE past that; the array holds 256 KiB on each GPU until the process
 * ends).  Same kernels, same results as csadv_diagnostics.  Calls that overlap on one GPU (on two
 * streams at once) share the array and race.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc). */
int csadv_diagnostics_noalloc(csadv_t *h, int32_t solve, int32_t tracer0, int32_t ntracers, const double *qexact,
                              double blowup_limit, double *out, void *stream);

/* Steps completed and the halt flag of one solve, read from the device (synchronises the stream);
 * the host's count of completed steps (csadv_current) follows the device after a halt.  No
 * counterpart in the reference (its loop exits on a blow-up). */
int csadv_status(csadv_t *h, int32_t solve, int32_t *steps_done, int32_t *halted, void *stream);

/* Nearest-cell remap of one tracer onto the plotting lat-lon grid (src/interpolation.py:38-149,
 * src/cs_datastruct.py:569-582): point (a, b) is lon = linspace(-pi, pi, nlon)[a],
 * lat = linspace(-pi/2, pi/2, nlat)[b]; its panel is the face of its largest Cartesian component
 * (a later face wins a tie, as the reference's mask order gives), its cell the floor of the
 * inverse equiangular coordinates.  field == NULL remaps the tracer's state (with the mass-fixer
 * total); otherwise field is a device FP64 [6][N][N] interior array (i fastest).  Device outputs,
 * each [nlon][nlat] and each optional (NULL to skip): out (FP64 values, the reference's Q_ll),
 * ix, jy, panel (int32, the reference's ll2cs tables).  nlat, nlon >= 2, nlat * nlon <= 2^31. */
int csadv_remap_latlon(csadv_t *h, int32_t solve, int32_t tracer, const double *field, int32_t nlat, int32_t nlon,
                       double *out, int32_t *ix, int32_t *jy, int32_t *panel, void *stream);

/* Index (0 or 1) of the buffer holding the current state of a solve (host-tracked). */
int csadv_current(csadv_t *h, int32_t solve);

/* Measurement hook: capacity > 0 creates that many HIP event pairs and makes csadv_step
 * record one pair around each launch of the step kernel (until the pairs run out);
 * capacity = 0 releases them.  csadv_profile_read waits for the recorded pairs and
 * returns the summed kernel time (ms) and the number of launches, then resets. */
int csadv_profile(csadv_t *h, int32_t capacity);
int csadv_profile_read(csadv_t *h, double *total_ms, int32_t *launches);

/* Mean duration in microseconds of the step kernel alone, over reps (1..4096) back-to-back
 * launches replayed from one HIP graph between one event pair on `stream`.  The repeats
 * redo the same step (same input, same output), so the state afterwards equals that of one
 * launch.  Fused step only (CSADV_E_STATE for a general-step handle).  No counterpart in the
 * reference: measurement of the kernel's roofline. */
int csadv_time_step_kernel(csadv_t *h, int32_t reps, void *stream, double *us_per_launch);

/* Mean duration in microseconds of the step kernel over reps rounds of one launch per handle
 * (handles[0 .. nhandles-1], reps * nhandles <= 8192), replayed from one HIP graph between one
 * event pair on `stream`.  With handles over distinct state buffers whose total exceeds the
 * last-level cache (256 MB on MI355X), each launch streams its state from HBM, where
 * csadv_time_step_kernel's repeats of one step find it in cache.  Each launch redoes its handle's
 * current step.  Fused step only.  No counterpart in the reference: measurement. */
int csadv_time_step_kernel_rotating(csadv_t *const *handles, int32_t nhandles, int32_t reps, void *stream,
                                    double *us_per_launch);

const char *csadv_last_error(csadv_t *h);
void csadv_destroy(csadv_t *h);
int csadv_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif /* CSADV_H */
