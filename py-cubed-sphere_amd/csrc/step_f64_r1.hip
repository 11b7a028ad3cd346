// The fused step kernels of FP64 with PPM-0 (one translation unit per precision and
// reconstruction, so that the library's kernels compile in parallel).
#include "csadv_step.hpp"

template hipError_t launch_step_r<double, 1>(csadv_t *h, hipStream_t st);
