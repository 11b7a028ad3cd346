// Internal interface of libcsadv's translation units: the handle and the functions one unit
// calls in another.  Not part of the public ABI (include/csadv.h).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <type_traits>
#include <vector>

#include "csadv_device.hpp"

using namespace csadv;

struct csadv_s {
  std::vector<csadv_solve_desc> desc;
  std::vector<Solve> host;
  Solve *dev = nullptr;           // device descriptors (caller-visible scratch is not used)
  int4 *work = nullptr;           // step work items
  int4 *gwork = nullptr;          // ghost-fill work items of a multi-solve handle (ghost_list)
  FieldSet *qsets = nullptr;      // ghost-fill sets for the tracers
  FieldSet *wsets = nullptr;      // ghost-fill sets for the lat-lon centre winds
  int nwork = 0;
  int ngwork = 0;
  int tile = 0;
  int dtype = 0;
  int recon = 3;
  bool rk2 = false;               // dp = 2 (separate upwind-side winds)
  int split = 1;                  // splitting of every solve (its metric treatment follows)
  bool general = false;           // ET-S72 / ET-PL07: the general step
  bool mfaf = false;              // a fused-step solve with MF-AF: k_af_fix after the step kernel
  bool wind1 = false;             // every solve RK1 with a one-field time-dependent wind: the
                                  // step kernel forms its winds (no per-step k_wind_avg)
  double *red = nullptr;          // diagnostics scratch: partials and outputs (diag_scratch_words)
  int red_tracers = 1;            // tracers the diagnostics scratch holds (the largest ntr)
  std::vector<int> done;          // host copy of the steps completed per solve (ctl[1])
  std::vector<hipEvent_t> ev;     // profiling event pairs (start, end) around the step kernel
  int ev_used = 0;
  int graph_steps = 0;            // steps per captured graph (0 = plain launches)
  hipStream_t cap = nullptr;      // private stream used only for capture
  hipGraphExec_t gexec = nullptr; // instantiated graph of graph_steps steps
  std::string err;
};

// Record a failure on the handle (its message for csadv_last_error) and return the code.
int fail(csadv_t *h, int code, const std::string &msg);

#define HIPCHK(h, call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail((h), CSADV_E_HIP, std::string(#call " failed: ") + hipGetErrorString(e_)); \
  } while (0)

// Blocks of a grid-stride whole-array kernel over `total` items.
inline int gp_blocks(int64_t total) { return (int)std::min<int64_t>(4096, (total + BLOCK - 1) / BLOCK); }

// Doubles of diagnostics scratch for `ntracers` tracers: kDiagBlocks partials of kDiagWords
// per tracer, then 4 outputs per tracer.
inline int64_t diag_scratch_words(int ntracers) {
  return (int64_t)ntracers * (kDiagBlocks * kDiagWords + 4);
}

// Full 4-layer duo-grid ghost fill of every tracer's current buffer (csadv.hip).
int ghost_fill(csadv_t *h, hipStream_t st);

// The fused step kernel of every solve for reconstruction R (step_<dtype>_r<R>.hip).
template <typename T, int R>
hipError_t launch_step_r(csadv_t *h, hipStream_t st);

// One general step of every solve with splitting SP (general_<dtype>_s<SP>.hip), and by the
// handle's splitting.
template <typename T, int SP>
int enqueue_general_split(csadv_t *h, hipStream_t st);
template <typename T>
inline int enqueue_general_t(csadv_t *h, hipStream_t st) {
  if (h->split == 2) return enqueue_general_split<T, 2>(h, st);
  if (h->split == 3) return enqueue_general_split<T, 3>(h, st);
  return enqueue_general_split<T, 1>(h, st);
}
