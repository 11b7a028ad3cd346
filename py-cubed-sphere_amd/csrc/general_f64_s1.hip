// The general step (ET-S72, ET-PL07, MF-AF) of FP64 with the AVLT splitting: one translation unit
// per precision and splitting, so that the library's kernels compile in parallel.
#include "csadv_general.hpp"

template int enqueue_general_split<double, 1>(csadv_t *h, hipStream_t st);
