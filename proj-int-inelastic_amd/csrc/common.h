// Shared device building blocks: complex arithmetic, the 64-entry twiddle
// table, the in-register radix-2 FFT of one pencil, and the spherical-mean
// special functions (J0 in 2D, sin(x)/x in 3D).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ib {

template <typename T>
struct cx {
  T re, im;
};

template <typename T>
__host__ __device__ __forceinline__ cx<T> mk(T r, T i) { return cx<T>{r, i}; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator*(T s, cx<T> a) { return {s * a.re, s * a.im}; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> conj(cx<T> a) { return {a.re, -a.im}; }

// cos / sin of 2 pi k / 64, k = 0..63, correctly rounded in the first octant
// and completed by exact symmetries (so cos(pi/2) is exactly 0).
__device__ constexpr double TW64[64][2] = {
    {0x1.0000000000000p+0, 0x0.0p+0},  // k=0
    {0x1.fd88da3d12526p-1, 0x1.917a6bc29b42cp-4},  // k=1
    {0x1.f6297cff75cb0p-1, 0x1.8f8b83c69a60ap-3},  // k=2
    {0x1.e9f4156c62ddap-1, 0x1.294062ed59f05p-2},  // k=3
    {0x1.d906bcf328d46p-1, 0x1.87de2a6aea963p-2},  // k=4
    {0x1.c38b2f180bdb1p-1, 0x1.e2b5d3806f63bp-2},  // k=5
    {0x1.a9b66290ea1a3p-1, 0x1.1c73b39ae68c8p-1},  // k=6
    {0x1.8bc806b151741p-1, 0x1.44cf325091dd6p-1},  // k=7
    {0x1.6a09e667f3bcdp-1, 0x1.6a09e667f3bcdp-1},  // k=8
    {0x1.44cf325091dd6p-1, 0x1.8bc806b151741p-1},  // k=9
    {0x1.1c73b39ae68c8p-1, 0x1.a9b66290ea1a3p-1},  // k=10
    {0x1.e2b5d3806f63bp-2, 0x1.c38b2f180bdb1p-1},  // k=11
    {0x1.87de2a6aea963p-2, 0x1.d906bcf328d46p-1},  // k=12
    {0x1.294062ed59f05p-2, 0x1.e9f4156c62ddap-1},  // k=13
    {0x1.8f8b83c69a60ap-3, 0x1.f6297cff75cb0p-1},  // k=14
    {0x1.917a6bc29b42cp-4, 0x1.fd88da3d12526p-1},  // k=15
    {0x0.0p+0, 0x1.0000000000000p+0},  // k=16
    {-0x1.917a6bc29b42cp-4, 0x1.fd88da3d12526p-1},  // k=17
    {-0x1.8f8b83c69a60ap-3, 0x1.f6297cff75cb0p-1},  // k=18
    {-0x1.294062ed59f05p-2, 0x1.e9f4156c62ddap-1},  // k=19
    {-0x1.87de2a6aea963p-2, 0x1.d906bcf328d46p-1},  // k=20
    {-0x1.e2b5d3806f63bp-2, 0x1.c38b2f180bdb1p-1},  // k=21
    {-0x1.1c73b39ae68c8p-1, 0x1.a9b66290ea1a3p-1},  // k=22
    {-0x1.44cf325091dd6p-1, 0x1.8bc806b151741p-1},  // k=23
    {-0x1.6a09e667f3bcdp-1, 0x1.6a09e667f3bcdp-1},  // k=24
    {-0x1.8bc806b151741p-1, 0x1.44cf325091dd6p-1},  // k=25
    {-0x1.a9b66290ea1a3p-1, 0x1.1c73b39ae68c8p-1},  // k=26
    {-0x1.c38b2f180bdb1p-1, 0x1.e2b5d3806f63bp-2},  // k=27
    {-0x1.d906bcf328d46p-1, 0x1.87de2a6aea963p-2},  // k=28
    {-0x1.e9f4156c62ddap-1, 0x1.294062ed59f05p-2},  // k=29
    {-0x1.f6297cff75cb0p-1, 0x1.8f8b83c69a60ap-3},  // k=30
    {-0x1.fd88da3d12526p-1, 0x1.917a6bc29b42cp-4},  // k=31
    {-0x1.0000000000000p+0, 0x0.0p+0},  // k=32
    {-0x1.fd88da3d12526p-1, -0x1.917a6bc29b42cp-4},  // k=33
    {-0x1.f6297cff75cb0p-1, -0x1.8f8b83c69a60ap-3},  // k=34
    {-0x1.e9f4156c62ddap-1, -0x1.294062ed59f05p-2},  // k=35
    {-0x1.d906bcf328d46p-1, -0x1.87de2a6aea963p-2},  // k=36
    {-0x1.c38b2f180bdb1p-1, -0x1.e2b5d3806f63bp-2},  // k=37
    {-0x1.a9b66290ea1a3p-1, -0x1.1c73b39ae68c8p-1},  // k=38
    {-0x1.8bc806b151741p-1, -0x1.44cf325091dd6p-1},  // k=39
    {-0x1.6a09e667f3bcdp-1, -0x1.6a09e667f3bcdp-1},  // k=40
    {-0x1.44cf325091dd6p-1, -0x1.8bc806b151741p-1},  // k=41
    {-0x1.1c73b39ae68c8p-1, -0x1.a9b66290ea1a3p-1},  // k=42
    {-0x1.e2b5d3806f63bp-2, -0x1.c38b2f180bdb1p-1},  // k=43
    {-0x1.87de2a6aea963p-2, -0x1.d906bcf328d46p-1},  // k=44
    {-0x1.294062ed59f05p-2, -0x1.e9f4156c62ddap-1},  // k=45
    {-0x1.8f8b83c69a60ap-3, -0x1.f6297cff75cb0p-1},  // k=46
    {-0x1.917a6bc29b42cp-4, -0x1.fd88da3d12526p-1},  // k=47
    {0x0.0p+0, -0x1.0000000000000p+0},  // k=48
    {0x1.917a6bc29b42cp-4, -0x1.fd88da3d12526p-1},  // k=49
    {0x1.8f8b83c69a60ap-3, -0x1.f6297cff75cb0p-1},  // k=50
    {0x1.294062ed59f05p-2, -0x1.e9f4156c62ddap-1},  // k=51
    {0x1.87de2a6aea963p-2, -0x1.d906bcf328d46p-1},  // k=52
    {0x1.e2b5d3806f63bp-2, -0x1.c38b2f180bdb1p-1},  // k=53
    {0x1.1c73b39ae68c8p-1, -0x1.a9b66290ea1a3p-1},  // k=54
    {0x1.44cf325091dd6p-1, -0x1.8bc806b151741p-1},  // k=55
    {0x1.6a09e667f3bcdp-1, -0x1.6a09e667f3bcdp-1},  // k=56
    {0x1.8bc806b151741p-1, -0x1.44cf325091dd6p-1},  // k=57
    {0x1.a9b66290ea1a3p-1, -0x1.1c73b39ae68c8p-1},  // k=58
    {0x1.c38b2f180bdb1p-1, -0x1.e2b5d3806f63bp-2},  // k=59
    {0x1.d906bcf328d46p-1, -0x1.87de2a6aea963p-2},  // k=60
    {0x1.e9f4156c62ddap-1, -0x1.294062ed59f05p-2},  // k=61
    {0x1.f6297cff75cb0p-1, -0x1.8f8b83c69a60ap-3},  // k=62
    {0x1.fd88da3d12526p-1, -0x1.917a6bc29b42cp-4},  // k=63
};

// Multiply v by exp(DIR * 2 pi i * ti / 64).  ti is a compile-time constant
// after unrolling, so the special cases fold away.
template <int DIR, typename T>
__device__ __forceinline__ cx<T> twiddle(cx<T> v, int ti) {
  if (ti == 0) return v;
  if (ti == 16) return (DIR > 0) ? cx<T>{-v.im, v.re} : cx<T>{v.im, -v.re};
  if (ti == 32) return cx<T>{-v.re, -v.im};
  if (ti == 48) return (DIR > 0) ? cx<T>{v.im, -v.re} : cx<T>{-v.im, v.re};
  const T c = (T)TW64[ti][0];
  const T s = (T)(DIR > 0 ? TW64[ti][1] : -TW64[ti][1]);
  return cx<T>{v.re * c - v.im * s, v.re * s + v.im * c};
}

template <int N>
__host__ __device__ constexpr int bitrev(int i) {
  int r = 0;
  for (int b = 1; b < N; b <<= 1) {
    r = (r << 1) | (i & 1);
    i >>= 1;
  }
  return r;
}

// In-register iterative radix-2 DIT FFT of length N (N | 64), natural order in
// and out.  DIR = -1: forward exp(-2 pi i k x / N); DIR = +1: inverse,
// unnormalised.
template <int N, int DIR, typename T>
__device__ __forceinline__ void fft_reg(cx<T>* a) {
  static_assert(N >= 2 && N <= 64 && (N & (N - 1)) == 0, "N must be a power of two <= 64");
  cx<T> t[N];
#pragma unroll
  for (int i = 0; i < N; ++i) t[i] = a[bitrev<N>(i)];
#pragma unroll
  for (int len = 2; len <= N; len <<= 1) {
    const int half = len >> 1;
    const int step = 64 / len;
#pragma unroll
    for (int i = 0; i < N; i += len) {
#pragma unroll
      for (int k = 0; k < half; ++k) {
        const cx<T> u = t[i + k];
        const int ti = k * step;
        if (ti % 16 != 0) {
          // u + w v by FMAs, then u - w v = 2 u - (u + w v): 6 operations per
          // butterfly instead of 8 (Linzer & Feig 1993)
          const cx<T> x = t[i + k + half];
          cx<T> a;
          if (ti == 8 || ti == 24 || ti == 40 || ti == 56) {
            const T h = (T)0.70710678118654752440;
            const cx<T> w = twiddle<DIR>(cx<T>{(T)1, (T)0}, ti);
            // w = h (w.re/h + i w.im/h) with components +-1
            const T sr = w.re > 0 ? (T)1 : (T)-1, si = w.im > 0 ? (T)1 : (T)-1;
            a.re = fma(h, sr * x.re - si * x.im, u.re);
            a.im = fma(h, si * x.re + sr * x.im, u.im);
          } else {
            const T c = (T)TW64[ti][0];
            const T s = (T)(DIR > 0 ? TW64[ti][1] : -TW64[ti][1]);
            a.re = fma(x.re, c, fma(-x.im, s, u.re));
            a.im = fma(x.re, s, fma(x.im, c, u.im));
          }
          t[i + k] = a;
          t[i + k + half] = cx<T>{fma((T)2, u.re, -a.re), fma((T)2, u.im, -a.im)};
          continue;
        }
        const cx<T> v = twiddle<DIR>(t[i + k + half], ti);
        t[i + k] = u + v;
        t[i + k + half] = u - v;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) a[i] = t[i];
}

// exp(i pi t) in FP32 for t in half turns: t reduced to [-1, 1] in FP64, then the FP32
// sincospi of the reduced value (as accurate as the FP32 rounding of an FP64 sincos, and
// without the large-argument reduction of sincos, whose table costs registers and scratch).
__device__ __forceinline__ cx<float> cispi_f32(double t) {
  t -= 2.0 * rint(0.5 * t);
  float sn, cs;
  sincospif((float)t, &sn, &cs);
  return cx<float>{cs, sn};
}

// The separable phase factors of the 2D FP32 forms in half turns, one expression for the
// plan's tables and the kernels' fallbacks (identical values either way):
// E = exp(-i pi r_j sigma kf / L), G = exp(i a_j r_j sigma kf).
__device__ __forceinline__ cx<float> phase2_f32(int which, double rj, double aj, double s, double kf, double L) {
  return cispi_f32(which == 0 ? -(rj * s * kf) / L : (aj * rj * s * kf) / M_PI);
}

// Integer frequency of FFT-ordered index i: 0..N/2-1, then -N/2..-1.
__host__ __device__ __forceinline__ int kfreq(int i, int N) { return i < N / 2 ? i : i - N; }

// J0(x) by the midpoint rule on J0(x) = (1/pi) int_0^pi cos(x sin t) dt.  The
// integrand is periodic and entire, so the K-node rule converges geometrically
// (Trefethen & Weideman, SIAM Review 56, 2014); its error is
// 2 sum_q (+-) J_{2qK}(x), below 1e-17 for 2K >= 1.5 x + 20.  K = 128 covers
// x <= 150.  The integrand is symmetric about pi/2, so K/2 cosines suffice.
__host__ __device__ inline double j0_midpoint(double x) {
  const int K = 128;
  double s = 0.0;
  for (int m = 0; m < K / 2; ++m) {
    const double t = M_PI * (m + 0.5) / K;
    s += cos(x * sin(t));
  }
  return 2.0 * s / K;
}

// sin(x)/x with the removable singularity (numpy.sinc(x/pi)).
__host__ __device__ inline double sinc_x(double x) { return x == 0.0 ? 1.0 : sin(x) / x; }

__host__ __device__ inline double sp_func(int dim, double x) {
  return dim == 2 ? j0_midpoint(x) : sinc_x(x);
}

// Workgroup id -> (cell, partial index jj) with nrs workgroups per cell.  xg = 1: the nrs
// workgroups of cell c are numbered i = c (mod 8), so they run on one XCD (the dispatcher
// gives workgroup i to XCD i mod 8) and share the cell's fhat and f in its L2; the grid has
// ceil(batch / 8) * 8 * nrs workgroups.  xg = 0 (batches under 8 cells): consecutive
// numbering, so a small batch's workgroups spread over all XCDs; grid batch * nrs.
__device__ __forceinline__ void cell_of_block(long long id, int nrs, int xg, long long& cell, int& jj) {
  if (xg) {
    jj = (int)((id / 8) % nrs);
    cell = (id / (8LL * nrs)) * 8 + id % 8;
  } else {
    jj = (int)(id % nrs);
    cell = id / nrs;
  }
}

// grid of a per-(cell, partial) launch, numbered as cell_of_block reads it
static inline long long cell_grid(long long batch, int nrs, int xg) {
  return xg ? (batch + 7) / 8 * 8 * nrs : batch * nrs;
}

}  // namespace ib
