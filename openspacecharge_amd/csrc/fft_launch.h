// Host-side launchers of the FFT pass kernels (defined per precision in
// fft_f64.hip / fft_f32.hip).
#pragma once
#include "fft_kernels.h"

namespace osc {

// lengths with a compile-time plan
#define OSC_FOR_EACH_L(X)                                                                   \
  X(16) X(18) X(24) X(32) X(36) X(48) X(54) X(64) X(72) X(96) X(108) X(128) X(144) X(162)   \
      X(192) X(216) X(256) X(288) X(324) X(384) X(432) X(486) X(512) X(576) X(648) X(768)  \
      X(864) X(972) X(1024)

bool fft_has_plan(int L);

// Online tuner of the strided passes' tile width (AxisTiles).  For each
// (precision, L, pass kind, log2 of the line count) the first calls run the
// distinct candidate widths in turn, each bracketed by two events on the
// launch stream; once every candidate has kTuneSamples completed timings
// (read without waiting, hipEventQuery) the fastest is used from then on.
// Widths do not change results, so the calls that measure are real work.
// fft_tune_mode: 1 tune (default), 0 the default width, 2 + v width v.
// With OSC_FFT_TUNE_FILE set, decisions are read from and appended to it.
class FftTune {
 public:
  FftTune(int prec, int L, int kind, int64_t lines, const bool distinct[3], hipStream_t s);
  int variant() const { return var_; }
  void done(hipStream_t s);

 private:
  int key_ = -1;
  int var_ = 1;
  hipEvent_t a_ = nullptr, b_ = nullptr;
};
void fft_tune_mode(int mode);
// keys decided, keys still measuring
void fft_tune_stats(int64_t* decided, int64_t* measuring);
// true when every prime factor of L is <= 23 and the generic engine fits LDS
bool fft_len_supported(int L);
GenericPlan generic_plan(int L);

template <typename T>
void launch_axis(int L, int dir, const cpx<T>* in, cpx<T>* out, const AxisIO& a, int g1, int g2,
                 const cpx<T>* tw, hipStream_t s);
template <typename T>
void launch_xr2c(int L, const double* in, cpx<T>* out, const R2CIO& a, const cpx<T>* tw,
                 hipStream_t s);
// x pass over the corner difference of a primitive F (lengths with a plan only)
template <typename T>
void launch_xr2c_diff(int L, const double* F, cpx<T>* out, const R2CIO& a, const CornerDiff& cd,
                      const cpx<T>* tw, hipStream_t s);
template <typename T>
void launch_xc2r(int L, const cpx<T>* in, const C2RIO<T>& a, int nout, const cpx<T>* tw,
                 hipStream_t s);
// true when L has a plan and k_zconv's 32-bit per-thread offsets cover the
// strides of a (ZinvIO::sz, o_sz)
template <typename T>
bool zconv_fits(int L, const ZinvIO<T>& a);
template <typename T>
void launch_zinv(int L, const cpx<T>* rho, cpx<T>* out, const ZinvIO<T>& a, int Ly, int nout,
                 const cpx<T>* tw, hipStream_t s);

}  // namespace osc
