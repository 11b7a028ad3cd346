// The general step (ET-S72, ET-PL07, MF-AF) of FP32: one translation unit per precision.
#include "csadv_general.hpp"

template int enqueue_general_t<float>(csadv_t *h, hipStream_t st);
