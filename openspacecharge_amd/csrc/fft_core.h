// LDS-staged Stockham FFT engine for gfx950.
//
// A workgroup transforms BX lines of length L.  Each line is handled by TPL
// threads; every thread keeps E = L/TPL complex values in registers and the
// radix stages exchange data through LDS (one buffer, read-all / barrier /
// write-all per stage).  Plans (radix sequences) are compile-time for the
// lengths in FftPlan; any other length uses FftGeneric (runtime radices, LDS
// double buffer).  Both engines present the same interface:
//
//   load(f)   f(b, p) -> value of line b at position p (stage-0 positions)
//   run(lds, tw)
//   store(f)  f(b, p, value) for the thread's output positions (natural order)
//   dump(lds) writes the thread's outputs into LDS at natural positions
//             (then LDS index lidx(b, p) holds X[p] of line b after a barrier)
//
// Direction DIR = -1: X[k] = sum_j x[j] exp(-2 pi i jk/L); DIR = +1 the
// conjugate kernel; both unnormalised.  Twiddle table tw[m] = exp(-2 pi i m/L).
#pragma once
#include "osc_common.h"

namespace osc {

// ---------------------------------------------------------------------------
// small DFT butterflies
// ---------------------------------------------------------------------------
template <int R, int DIR, typename T>
__device__ __forceinline__ void dft(cpx<T> (&a)[R]);

template <int DIR, typename T>
__device__ __forceinline__ void dft2(cpx<T>& a, cpx<T>& b) {
  cpx<T> t = a;
  a = add(t, b);
  b = sub(t, b);
}

template <>
__device__ __forceinline__ void dft<2, -1, double>(cpx<double> (&a)[2]) { dft2<-1>(a[0], a[1]); }
template <>
__device__ __forceinline__ void dft<2, 1, double>(cpx<double> (&a)[2]) { dft2<1>(a[0], a[1]); }
template <>
__device__ __forceinline__ void dft<2, -1, float>(cpx<float> (&a)[2]) { dft2<-1>(a[0], a[1]); }
template <>
__device__ __forceinline__ void dft<2, 1, float>(cpx<float> (&a)[2]) { dft2<1>(a[0], a[1]); }

template <int DIR, typename T>
__device__ __forceinline__ void dft3v(cpx<T>& a0, cpx<T>& a1, cpx<T>& a2) {
  const T s3 = T(0.86602540378443864676372317075294);  // sqrt(3)/2
  cpx<T> t = add(a1, a2);
  cpx<T> m = sub(a0, scale(t, T(0.5)));
  cpx<T> d = scale(sub(a1, a2), s3);
  cpx<T> s = mul_i<DIR>(d);
  a0 = add(a0, t);
  a1 = add(m, s);
  a2 = sub(m, s);
}

template <int DIR, typename T>
__device__ __forceinline__ void dft4v(cpx<T>& a0, cpx<T>& a1, cpx<T>& a2, cpx<T>& a3) {
  cpx<T> s02 = add(a0, a2), d02 = sub(a0, a2);
  cpx<T> s13 = add(a1, a3), d13 = mul_i<DIR>(sub(a1, a3));
  a0 = add(s02, s13);
  a2 = sub(s02, s13);
  a1 = add(d02, d13);
  a3 = sub(d02, d13);
}

template <int DIR, typename T>
struct Dft {
  static __device__ __forceinline__ void r3(cpx<T> (&a)[3]) { dft3v<DIR>(a[0], a[1], a[2]); }
  static __device__ __forceinline__ void r4(cpx<T> (&a)[4]) { dft4v<DIR>(a[0], a[1], a[2], a[3]); }
  // radix 6 = radix-2 over two radix-3 (decimation in time)
  static __device__ __forceinline__ void r6(cpx<T> (&a)[6]) {
    const T h = T(0.5), s3 = T(0.86602540378443864676372317075294);
    cpx<T> e0 = a[0], e1 = a[2], e2 = a[4], o0 = a[1], o1 = a[3], o2 = a[5];
    dft3v<DIR>(e0, e1, e2);
    dft3v<DIR>(o0, o1, o2);
    // w6^1 = (1/2, DIR*sqrt3/2), w6^2 = (-1/2, DIR*sqrt3/2)
    cpx<T> w1 = mk(h, T(DIR) * s3), w2 = mk(-h, T(DIR) * s3);
    cpx<T> t1 = mul(o1, w1), t2 = mul(o2, w2);
    a[0] = add(e0, o0);
    a[3] = sub(e0, o0);
    a[1] = add(e1, t1);
    a[4] = sub(e1, t1);
    a[2] = add(e2, t2);
    a[5] = sub(e2, t2);
  }
  // radix 8 = radix-2 over two radix-4 (decimation in time)
  static __device__ __forceinline__ void r8(cpx<T> (&a)[8]) {
    const T r2 = T(0.70710678118654752440084436210485);
    cpx<T> e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
    cpx<T> o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
    dft4v<DIR>(e0, e1, e2, e3);
    dft4v<DIR>(o0, o1, o2, o3);
    // w8^1 = r2*(1 + DIR i), w8^2 = DIR i, w8^3 = r2*(-1 + DIR i)
    cpx<T> t1 = scale(add(o1, mul_i<DIR>(o1)), r2);
    cpx<T> t2 = mul_i<DIR>(o2);
    cpx<T> t3 = scale(sub(mul_i<DIR>(o3), o3), r2);
    a[0] = add(e0, o0);
    a[4] = sub(e0, o0);
    a[1] = add(e1, t1);
    a[5] = sub(e1, t1);
    a[2] = add(e2, t2);
    a[6] = sub(e2, t2);
    a[3] = add(e3, t3);
    a[7] = sub(e3, t3);
  }
};

template <int R, int DIR, typename T>
__device__ __forceinline__ void butterfly(cpx<T> (&a)[R]) {
  if constexpr (R == 2) {
    dft2<DIR>(a[0], a[1]);
  } else if constexpr (R == 3) {
    Dft<DIR, T>::r3(a);
  } else if constexpr (R == 4) {
    Dft<DIR, T>::r4(a);
  } else if constexpr (R == 6) {
    Dft<DIR, T>::r6(a);
  } else if constexpr (R == 8) {
    Dft<DIR, T>::r8(a);
  } else {
    static_assert(R == 2, "unsupported radix");
  }
}

// ---------------------------------------------------------------------------
// compile-time plans
// ---------------------------------------------------------------------------
constexpr int cgcd(int a, int b) { return b == 0 ? a : cgcd(b, a % b); }

template <int L>
struct FftPlan {
  static constexpr int ns = 0;  // no specialised plan
  static constexpr int r[5] = {1, 1, 1, 1, 1};
};

#define OSC_PLAN(LL, NS, R0, R1, R2, R3, R4)             \
  template <>                                            \
  struct FftPlan<LL> {                                   \
    static constexpr int ns = NS;                        \
    static constexpr int r[5] = {R0, R1, R2, R3, R4};    \
  };

// L = 2n for the doubled grids of n = 2^a 3^b (16 <= L <= 1024)
OSC_PLAN(16, 2, 4, 4, 1, 1, 1)
OSC_PLAN(18, 2, 6, 3, 1, 1, 1)
OSC_PLAN(24, 2, 6, 4, 1, 1, 1)
OSC_PLAN(32, 2, 8, 4, 1, 1, 1)
OSC_PLAN(36, 2, 6, 6, 1, 1, 1)
OSC_PLAN(48, 3, 4, 4, 3, 1, 1)
OSC_PLAN(54, 3, 6, 3, 3, 1, 1)
OSC_PLAN(64, 2, 8, 8, 1, 1, 1)
OSC_PLAN(72, 3, 6, 6, 2, 1, 1)
OSC_PLAN(96, 3, 6, 4, 4, 1, 1)
OSC_PLAN(108, 3, 6, 6, 3, 1, 1)
OSC_PLAN(128, 3, 8, 4, 4, 1, 1)
OSC_PLAN(144, 3, 6, 6, 4, 1, 1)
OSC_PLAN(162, 4, 6, 3, 3, 3, 1)
OSC_PLAN(192, 4, 4, 4, 4, 3, 1)
OSC_PLAN(216, 3, 6, 6, 6, 1, 1)
OSC_PLAN(256, 3, 8, 8, 4, 1, 1)
OSC_PLAN(288, 4, 4, 4, 6, 3, 1)
OSC_PLAN(324, 4, 6, 6, 3, 3, 1)
OSC_PLAN(384, 4, 4, 4, 4, 6, 1)
OSC_PLAN(432, 4, 6, 6, 6, 2, 1)
OSC_PLAN(486, 5, 6, 3, 3, 3, 3)
OSC_PLAN(512, 3, 8, 8, 8, 1, 1)
OSC_PLAN(576, 4, 4, 4, 6, 6, 1)
OSC_PLAN(648, 4, 6, 6, 6, 3, 1)
OSC_PLAN(768, 4, 8, 4, 4, 6, 1)
OSC_PLAN(864, 4, 6, 6, 6, 4, 1)
OSC_PLAN(972, 5, 6, 6, 3, 3, 3)
OSC_PLAN(1024, 4, 8, 8, 4, 4, 1)
#undef OSC_PLAN

template <int L>
struct PlanInfo {
  using P = FftPlan<L>;
  static constexpr int ns = P::ns;
  static constexpr int tpl_raw() {
    int g = 0;
    for (int s = 0; s < ns; ++s) g = cgcd(g, L / P::r[s]);
    return g;
  }
  static constexpr int TPL = ns ? tpl_raw() : 1;
  static constexpr int E = L / TPL;
  static constexpr int ns_before(int s) {
    int p = 1;
    for (int i = 0; i < s; ++i) p *= P::r[i];
    return p;
  }
};

// LDS padding: one element per bank row (256 B) to spread strided accesses
template <typename T>
struct LdsPad {
  static constexpr int shift = sizeof(cpx<T>) == 16 ? 4 : 5;
  static __host__ __device__ constexpr int idx(int i) { return i + (i >> shift); }
  static __host__ __device__ constexpr int size(int n) { return n + (n >> shift) + 1; }
};

// Workgroups are dealt to the 8 XCDs round-robin in dispatch order
// (workgroup w runs on XCD w % 8), and each XCD has its own L2.  A strided
// pass reads rows of BX adjacent columns -- 32 or 64 bytes, less than a
// 128-byte line, at the long lengths -- so workgroups of neighbouring columns
// read the same lines; dealt round-robin, each XCD fetches each line from
// HBM.  xcd_block maps dispatch index w of n to a logical workgroup such that
// every XCD runs one contiguous range of logical workgroups, in order:
// neighbours then share one L2.  A bijection of [0, n).
__device__ __forceinline__ uint32_t xcd_block(uint32_t w, uint32_t n) {
  constexpr uint32_t kXcd = 8;
  const uint32_t x = w % kXcd, i = w / kXcd, per = n / kXcd, rem = n % kXcd;
  return x * per + (x < rem ? x : rem) + i;
}

// ---------------------------------------------------------------------------
// compile-time engine
// ---------------------------------------------------------------------------
// PFAST: thread index runs along positions first (contiguous lines, x axis);
// otherwise along lines first (strided axes, coalesced over the tile lines).
template <typename T, int L, int DIR, int BX, bool PFAST>
struct FftTile {
  using PI = PlanInfo<L>;
  using P = FftPlan<L>;
  static constexpr int TPL = PI::TPL;
  static constexpr int E = PI::E;
  static constexpr int NT = BX * TPL;
  static constexpr int NS = PI::ns;
  static constexpr int R0 = P::r[0];
  static constexpr int RL = P::r[NS - 1];
  static constexpr int LDS = LdsPad<T>::size(BX * L);

  int b, t;
  cpx<T> v[E];

  __device__ __forceinline__ FftTile() {
    const int tid = threadIdx.x;
    if (PFAST) {
      t = tid % TPL;
      b = tid / TPL;
    } else {
      b = tid % BX;
      t = tid / BX;
    }
  }
  static __device__ __forceinline__ int lidx(int b, int p) {
    return LdsPad<T>::idx(PFAST ? b * L + p : p * BX + b);
  }
  __device__ __forceinline__ bool active() const { return b < BX && t < TPL; }

  template <class F>
  __device__ __forceinline__ void load(F&& f) {
#pragma unroll
    for (int q = 0; q < E / R0; ++q)
#pragma unroll
      for (int r = 0; r < R0; ++r) v[q * R0 + r] = f(b, t + q * TPL + r * (L / R0));
  }
  template <class F>
  __device__ __forceinline__ void store(F&& f) const {
#pragma unroll
    for (int q = 0; q < E / RL; ++q)
#pragma unroll
      for (int r = 0; r < RL; ++r) f(b, t + q * TPL + r * (L / RL), v[q * RL + r]);
  }
  // load/store with each position split as p = t + c: f receives c, a
  // compile-time constant once the loops are unrolled (a multiple of TPL)
  template <class F>
  __device__ __forceinline__ void load_c(F&& f) {
#pragma unroll
    for (int q = 0; q < E / R0; ++q)
#pragma unroll
      for (int r = 0; r < R0; ++r) v[q * R0 + r] = f(q * TPL + r * (L / R0));
  }
  template <class F>
  __device__ __forceinline__ void store_c(F&& f) const {
#pragma unroll
    for (int q = 0; q < E / RL; ++q)
#pragma unroll
      for (int r = 0; r < RL; ++r) f(q * TPL + r * (L / RL), v[q * RL + r]);
  }
  __device__ __forceinline__ void dump(cpx<T>* lds) const {
#pragma unroll
    for (int q = 0; q < E / RL; ++q)
#pragma unroll
      for (int r = 0; r < RL; ++r) lds[lidx(b, t + q * TPL + r * (L / RL))] = v[q * RL + r];
  }

  template <int S>
  __device__ __forceinline__ void stage(cpx<T>* lds, const cpx<T>* __restrict__ tw) {
    constexpr int R = P::r[S];
    constexpr int Ns = PI::ns_before(S);
    constexpr int NB = E / R;
    constexpr int TWS = L / (Ns * R);
    if constexpr (S > 0) {
#pragma unroll
      for (int q = 0; q < NB; ++q)
#pragma unroll
        for (int r = 0; r < R; ++r) v[q * R + r] = lds[lidx(b, t + q * TPL + r * (L / R))];
      if constexpr (S < NS - 1) __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int j = t + q * TPL;
      cpx<T> a[R];
#pragma unroll
      for (int r = 0; r < R; ++r) a[r] = v[q * R + r];
      const int k = j % Ns;
      if constexpr (S > 0) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cpx<T> w = tw[unsigned(k * r * TWS)];  // unsigned: scalar-base address form
          if (DIR > 0) w.y = -w.y;
          a[r] = mul(a[r], w);
        }
      }
      butterfly<R, DIR>(a);
      if constexpr (S < NS - 1) {
        const int od = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) lds[lidx(b, od + r * Ns)] = a[r];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) v[q * R + r] = a[r];
      }
    }
    if constexpr (S < NS - 1) {
      __syncthreads();
      stage<S + 1>(lds, tw);
    }
  }
  __device__ __forceinline__ void run(cpx<T>* lds, const cpx<T>* __restrict__ tw) { stage<0>(lds, tw); }
};

// ---------------------------------------------------------------------------
// runtime-radix fallback (any length whose prime factors are <= 23)
// ---------------------------------------------------------------------------
struct GenericPlan {
  int L;
  int ns;
  int r[24];
};

// One line per group of GTPL threads; LDS holds two buffers of L per line.
template <typename T, int DIR, int BX, int GTPL>
struct FftGeneric {
  static constexpr int NT = BX * GTPL;
  int b, t;
  int L;
  __device__ __forceinline__ FftGeneric(int L_) : L(L_) {
    b = threadIdx.x / GTPL;
    t = threadIdx.x % GTPL;
  }
  __device__ __forceinline__ bool active() const { return b < BX; }
  __device__ __forceinline__ int lidx(int bb, int p, int buf) const {
    return LdsPad<T>::idx((buf * BX + bb) * L + p);
  }
  static __host__ __device__ int lds_size(int L) { return LdsPad<T>::size(2 * BX * L); }

  cpx<T>* lds_ = nullptr;
  int cur = 0;

  template <class F>
  __device__ __forceinline__ void load(cpx<T>* lds, F&& f) {
    lds_ = lds;
    cur = 0;
    for (int p = t; p < L; p += GTPL) lds[lidx(b, p, 0)] = f(b, p);
  }
  __device__ void run(const GenericPlan& pl, const cpx<T>* __restrict__ tw) {
    __syncthreads();
    int Ns = 1;
    for (int s = 0; s < pl.ns; ++s) {
      const int R = pl.r[s];
      const int nbf = L / R;
      const int tws = L / (Ns * R);
      for (int j = t; j < nbf; j += GTPL) {
        cpx<T> a[23], y[23];
        const int k = j % Ns;
        for (int r = 0; r < R; ++r) {
          cpx<T> vv = lds_[lidx(b, j + r * nbf, cur)];
          if (r > 0 && k > 0) {
            cpx<T> w = tw[k * r * tws];
            if (DIR > 0) w.y = -w.y;
            vv = mul(vv, w);
          }
          a[r] = vv;
        }
        // direct DFT of size R: y[q] = sum_r a[r] W_R^{rq}, W_R^m = tw[m * L/R]
        for (int q = 0; q < R; ++q) {
          cpx<T> acc = a[0];
          for (int r = 1; r < R; ++r) {
            cpx<T> w = tw[((r * q) % R) * nbf];
            if (DIR > 0) w.y = -w.y;
            acc = add(acc, mul(a[r], w));
          }
          y[q] = acc;
        }
        const int od = (j / Ns) * Ns * R + k;
        for (int r = 0; r < R; ++r) lds_[lidx(b, od + r * Ns, cur ^ 1)] = y[r];
      }
      cur ^= 1;
      Ns *= R;
      __syncthreads();
    }
  }
  template <class F>
  __device__ __forceinline__ void store(F&& f) const {
    for (int p = t; p < L; p += GTPL) f(b, p, lds_[lidx(b, p, cur)]);
  }
  // natural-order result already in LDS buffer `cur`
  __device__ __forceinline__ int result_buf() const { return cur; }
};

}  // namespace osc
