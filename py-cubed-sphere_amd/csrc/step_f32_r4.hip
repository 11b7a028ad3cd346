// The fused step kernels of FP32 with PPM-L04 (one translation unit per precision and
// reconstruction, so that the library's kernels compile in parallel).
#include "csadv_step.hpp"

template hipError_t launch_step_r<float, 4>(csadv_t *h, hipStream_t st);
