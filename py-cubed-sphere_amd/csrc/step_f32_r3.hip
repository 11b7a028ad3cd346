// The fused step kernels of FP32 with PPM-PL07 (one translation unit per precision and
// reconstruction, so that the library's kernels compile in parallel).
#include "csadv_step.hpp"

template hipError_t launch_step_r<float, 3>(csadv_t *h, hipStream_t st);
