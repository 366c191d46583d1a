// FFT pass kernels and launchers, FP64 instantiation.
#define OSC_REAL double
#include "fft_dispatch.inl"
