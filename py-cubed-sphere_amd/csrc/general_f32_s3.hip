// The general step (ET-S72, ET-PL07, MF-AF) of FP32 with the PL07 splitting: one translation unit
// per precision and splitting, so that the library's kernels compile in parallel.
#include "csadv_general.hpp"

template int enqueue_general_split<float, 3>(csadv_t *h, hipStream_t st);
