// FFT pass kernels and launchers, FP32 instantiation.
#define OSC_REAL float
#include "fft_dispatch.inl"
