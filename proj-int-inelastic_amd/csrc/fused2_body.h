// Device code of the fused C2C pair loops: k_fused2 (FP32 2D 16^2, 32^2), k_fusedc3 (FP32
// 3D 16^3) and k_fusedw (FP32 2D 64^2, FP64 2D 32^2) -- arguments and per-workgroup bodies.
// Included by collide_fused2.hip (the kernels and their host drivers) and collide_group.hip
// (the grouped launch of the fused shapes).  The bodies take their LDS arrays as references
// to arrays at fixed LDS addresses (see fused3_body.h).
#pragma once
#include <type_traits>

#include "common.h"

namespace ib {

struct F2Args {
  const cx<float>* fh;  // [cell][N][N], fhat / n
  const cx<float>* ephase;  // [pair][2][N] or null
  const cx<float>* gphase;  // [param][pair][2][N] or null
  const float* f32;     // [cell][N][N]
  cx<float>* part;      // [cell][wpc][N][N]
  const int* pidx;
  long long batch;
  int wpc;
  const double* sigma;
  const double* rj;
  const double* wsig;
  const double* crad;
  const double* a;
  const float* spg;
  int nr, m, nq;
  double L;
};

// REAL: the 32^2 form works on the half spectrum kx = 0 .. N/2 of real data (16 threads per
// quadrature pair, fused2r_body); 16^2 keeps the C2C form (N threads per pair).
template <int N>
struct F2Shape {
  static constexpr int THREADS = N >= 32 ? 128 : 256;
  static constexpr bool REAL = N == 32;
  static constexpr int TPP = REAL ? N / 2 : N;  // threads per quadrature pair
  static constexpr int G = THREADS / TPP;       // pairs in flight per workgroup
};

// Elements of one pair's half spectrum in the real form (16 rows of 32).
constexpr int kF2rStride = 16 * 32;

// LDS of one k_fused2 workgroup.  C2C (16^2): Wg the pairs' work arrays, Tab the phase
// tables, Fhs / Fs the staged fhat / n and f.  Real form (32^2): Wg the pairs' half spectra
// (kF2rStride elements each), Tab the y-axis E and G factors, Fhs one packed row per pair.
template <int N>
struct F2L {
  static constexpr int G = F2Shape<N>::G, PAD = N + 1, SN = N < 32 ? N : 1;
  static constexpr bool REAL = F2Shape<N>::REAL;
  using Wg = std::conditional_t<REAL, cx<float>[G * kF2rStride], cx<float>[G][N][PAD]>;
  using Tab = cx<float>[G][REAL ? 2 : 4][N];
  using Fhs = std::conditional_t<REAL, cx<float>[G][N], cx<float>[SN][SN + 1]>;
  using Fs = float[SN][SN + 1];
};

// Separable phase factor of quadrature pair (j, mm) on axis ax at frequency index k:
// E = exp(-i pi r_j sigma_ax kf / L) (which = 0) or G = exp(i a_j r_j sigma_ax kf) of
// parameter set ps (which = 1), from the plan's tables when it has them, else by the
// expression the tables are built with (phase2_f32: identical values either way).
__device__ __forceinline__ cx<float> f2_factor(const F2Args& A, int which, int ax, int pair, int j, int mm,
                                               int ps, int k, int N) {
  if (which == 0 && A.ephase) return A.ephase[((long long)pair * 2 + ax) * N + k];
  if (which == 1 && A.gphase) return A.gphase[(((long long)ps * A.nr * A.m + pair) * 2 + ax) * N + k];
  return phase2_f32(which, A.rj[j], which ? A.a[ps * A.nr + j] : 0.0, A.sigma[mm * 2 + ax], (double)kfreq(k, N),
                    A.L);
}

// Orders a wave's LDS accesses around an exchange between lanes of one 16-lane group.
__device__ __forceinline__ void f2r_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Real-data form of the 2D 32^2 pair loop (FP32).  f is real, so with E and G replaced by
// their Hermitian parts E_H, G_H (the Nyquist entry of each 1D factor by its real part; at
// kx = ky = 16 the real part of the product) h = f * ifft(E_H fhat) is real and its
// spectrum Hermitian: a pair needs the rows kx = 0 .. 16 only.  Rows 0 and 16 are real after
// the y transform and travel packed as row 0 (row 0 + i row 16).  The anti-Hermitian parts
// live on the Nyquist lines and are added by k_f2_nyq.  One 16-lane group per pair:
//   A  lane t: row t of E_H fhat (lane 0: rows 0 + i 16), inverse along y
//   B  lane t: columns y = 2t, 2t + 1 as one complex pencil (Hermitian extension in x),
//      inverse along x, * f, forward along x, unpack the two columns' rows 0 .. 16
//   C  lane t: row t forward along y, G_H * (.) and the sum; the packed row's pair
//      of rows is split by all lanes, two frequencies each
// The group's lanes share a wave, so its LDS exchanges need no workgroup barrier.  The
// unpack of B keeps a factor 2, taken out by the 1/2 in every G weight.  Element (kx, y) of
// a pair's half spectrum sits at kx * 32 + (y ^ kx): the 16 lanes reading or writing their
// rows at one y touch 16 distinct 8-byte bank pairs, and the column pairs (y, y + 1) of B
// stay one aligned 16-byte block (its halves swapped on odd rows), read and written as
// one 128-bit access.  At the end the workgroup's pairs are summed (fixed order) and the
// partial is written as the full Hermitian spectrum.
__device__ __forceinline__ void fused2r_body(const F2Args& A, const long long cell, const int w, const int tid,
                                             cx<float> (&Wf)[F2Shape<32>::G * kF2rStride],
                                             cx<float> (&tab)[F2Shape<32>::G][2][32],
                                             cx<float> (&S16)[F2Shape<32>::G][32]) {
  constexpr int N = 32, H = 16, G = F2Shape<32>::G, THREADS = F2Shape<32>::THREADS;
  const int g = tid >> 4, t = tid & 15;
  const int ps = A.pidx ? A.pidx[cell] : 0;
  const cx<float>* fhc = A.fh + cell * N * N;
  const float* fc = A.f32 + cell * N * N;
  cx<float>* Wr = Wf + g * kF2rStride;
  const int npairs = A.nr * A.m;
  const int stride = A.wpc * G;
  const int rounds = (npairs + stride - 1) / stride;
  cx<float> acc[N];
#pragma unroll
  for (int k = 0; k < N; ++k) acc[k] = cx<float>{0.f, 0.f};
  cx<float> accp[2] = {cx<float>{0.f, 0.f}, cx<float>{0.f, 0.f}};  // packed row at ky = t, t + 16
  cx<float> v[N];
  for (int r = 0; r < rounds; ++r) {
    const int pair = w * G + g + r * stride;
    if (pair < npairs) {
      const int j = pair / A.m, mm = pair % A.m;
      // fhat / n and f are re-read every pair (from L2) through an opaque offset: hoisted
      // out of the loop they would hold 130 registers across it
      int zo = 0;
      asm volatile("" : "+v"(zo));
      const cx<float>* fhz = fhc + zo;
      const float* fz = fc + zo;
      // the lane's LDS and table offsets (row swizzle, sp index) likewise from an opaque copy
      // of its row, recomputed each pair instead of held in registers across the loop
      int kxt = t;
      asm volatile("" : "+v"(kxt));
      // y-axis factors of E and G in LDS; x-axis factors at kx = t and 16 in registers
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        tab[g][0][t + h * H] = f2_factor(A, 0, 1, pair, j, mm, ps, t + h * H, N);
        tab[g][1][t + h * H] = f2_factor(A, 1, 1, pair, j, mm, ps, t + h * H, N);
      }
      const cx<float> ext = f2_factor(A, 0, 0, pair, j, mm, ps, kxt, N);
      const cx<float> ex16 = f2_factor(A, 0, 0, pair, j, mm, ps, H, N);
      const cx<float> gxt = f2_factor(A, 1, 0, pair, j, mm, ps, kxt, N);
      const cx<float> gx16 = f2_factor(A, 1, 0, pair, j, mm, ps, H, N);
      f2r_sync();
      // row 16 of E_H fhat times i, at ky = t, t + 16, for lane 0's packed row
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int k = t + h * H;
        const cx<float> ey = tab[g][0][k];
        const cx<float> e16 = k == H ? cx<float>{(ex16 * ey).re, 0.f} : ex16.re * ey;
        const cx<float> x = e16 * fhz[H * N + k];
        S16[g][k] = cx<float>{-x.im, x.re};
      }
      // A: row t of E_H fhat (kx = t is not the Nyquist row: ky = 16 takes Re e_y(16))
      const cx<float>* fr = fhz + kxt * N;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        cx<float> ey = tab[g][0][k];
        if (k == H) ey = cx<float>{ey.re, 0.f};
        v[k] = (ext * ey) * fr[k];
      }
      f2r_sync();
      if (t == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) v[k] = v[k] + S16[g][k];
      }
      fft_reg<N, +1>(v);
      int ta = t;
      asm volatile("" : "+v"(ta));
#pragma unroll
      for (int k = 0; k < N; ++k) Wr[ta * N + (k ^ ta)] = v[k];
      f2r_sync();
      // B: columns y0 = 2t, y1 = 2t + 1; Z(kx) = X_y0(kx) + i X_y1(kx) over kx = 0 .. 31
      {
        int tb = t;
        asm volatile("" : "+v"(tb));
        const int y0 = 2 * tb;
        {
          const cx<float> a = Wr[y0], b = Wr[y0 + 1];  // packed row 0: (Re, Im) = rows 0, 16
          v[0] = cx<float>{a.re, b.re};
          v[H] = cx<float>{a.im, b.im};
        }
#pragma unroll
        for (int kx = 1; kx < H; ++kx) {
          const float4 q = *reinterpret_cast<const float4*>(&Wr[kx * N + ((y0 ^ kx) & ~1)]);
          const cx<float> a = (kx & 1) ? cx<float>{q.z, q.w} : cx<float>{q.x, q.y};
          const cx<float> b = (kx & 1) ? cx<float>{q.x, q.y} : cx<float>{q.z, q.w};
          v[kx] = cx<float>{a.re - b.im, a.im + b.re};
          v[N - kx] = cx<float>{a.re + b.im, b.re - a.im};
        }
        fft_reg<N, +1>(v);
        const float2* fcol = reinterpret_cast<const float2*>(fz + y0);
#pragma unroll
        for (int x = 0; x < N; ++x) {
          const float2 fv = fcol[x * (N / 2)];
          v[x] = cx<float>{fv.x * v[x].re, fv.y * v[x].im};
        }
        fft_reg<N, -1>(v);
        // unpack (times 2): H0 = U + conj U(-kx), H1 = (U - conj U(-kx)) / i; addresses
        // recomputed (opaque copy) rather than held across the transforms
        int tw = t;
        asm volatile("" : "+v"(tw));
        const int y0w = 2 * tw;
        *reinterpret_cast<float4*>(&Wr[y0w]) =
            make_float4(2.f * v[0].re, 2.f * v[H].re, 2.f * v[0].im, 2.f * v[H].im);
#pragma unroll
        for (int kx = 1; kx < H; ++kx) {
          const cx<float> U = v[kx], Um = v[N - kx];
          const cx<float> h0{U.re + Um.re, U.im - Um.im}, h1{U.im + Um.im, Um.re - U.re};
          *reinterpret_cast<float4*>(&Wr[kx * N + ((y0w ^ kx) & ~1)]) =
              (kx & 1) ? make_float4(h1.re, h1.im, h0.re, h0.im) : make_float4(h0.re, h0.im, h1.re, h1.im);
        }
      }
      f2r_sync();
      // C: row t forward along y, G_H * (.) and the sum over the group's pairs
      int tc = t;
      asm volatile("" : "+v"(tc));
#pragma unroll
      for (int k = 0; k < N; ++k) v[k] = Wr[tc * N + (k ^ tc)];
      fft_reg<N, -1>(v);
      const float wj = (float)(0.5 * A.crad[ps * A.nr + j] * A.wsig[mm]);
      const float* sp = A.spg + ((long long)ps * A.nr + j) * A.nq;
#pragma unroll
      for (int k0 = 0; k0 < N; k0 += 8) {
#pragma unroll
        for (int k = k0; k < k0 + 8; ++k) {
          const int kk = kfreq(k, N);
          cx<float> gy = tab[g][1][k];
          if (k == H) gy = cx<float>{gy.re, 0.f};
          const cx<float> Gk = (wj * sp[tc * tc + kk * kk]) * (gxt * gy);
          acc[k] = acc[k] + Gk * v[k];
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      // the packed row: lane 0 parks its transform, every lane splits two frequencies
      if (t == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) S16[g][k] = v[k];
      }
      f2r_sync();
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int k = t + h * H, km = (N - k) & (N - 1), kk = kfreq(k, N);
        const cx<float> Z = S16[g][k], Zm = S16[g][km];
        const cx<float> gy = tab[g][1][k];
        const cx<float> g0 = (wj * sp[kk * kk]) * (k == H ? cx<float>{gy.re, 0.f} : gy);
        const cx<float> g16 = (wj * sp[H * H + kk * kk]) * (k == H ? cx<float>{(gx16 * gy).re, 0.f} : gx16.re * gy);
        const cx<float> al = 0.5f * (g0 + g16), be = 0.5f * (g0 - g16);
        accp[h] = accp[h] + al * Z + be * conj(Zm);
      }
      f2r_sync();
    }
  }
  // sum over the workgroup's pairs: rows kx >= 1 from acc, the packed row from accp
  __syncthreads();
  if (t > 0) {
#pragma unroll
    for (int k = 0; k < N; ++k) Wr[t * N + (k ^ t)] = acc[k];
  }
  Wr[t] = accp[0];
  Wr[t + H] = accp[1];
  __syncthreads();
  auto half = [&](int kx, int ky) {
    cx<float> s{0.f, 0.f};
#pragma unroll
    for (int gg = 0; gg < G; ++gg) s = s + Wf[gg * kF2rStride + kx * N + (ky ^ kx)];
    return s;
  };
  cx<float>* out = A.part + (cell * A.wpc + w) * N * N;
  for (int e = tid; e < N * N; e += THREADS) {
    const int kx = e / N, ky = e % N, kym = (N - ky) & (N - 1);
    cx<float> s;
    if (kx == 0 || kx == H) {
      const cx<float> P = half(0, ky), Pm = half(0, kym);
      s = kx == 0 ? cx<float>{0.5f * (P.re + Pm.re), 0.5f * (P.im - Pm.im)}
                  : cx<float>{0.5f * (P.im + Pm.im), 0.5f * (Pm.re - P.re)};
    } else if (kx < H) {
      s = half(kx, ky);
    } else {
      s = conj(half(N - kx, kym));
    }
    out[e] = s;
  }
}

// The rounds of one workgroup: cell `cell` (< A.batch), workgroup w of the cell's wpc,
// thread tid.  Every workgroup runs the same number of rounds (the last ones of a
// workgroup may have no pair: their barriers only), so that the grouped launch can run
// several of them in one 512-thread workgroup.
template <int N>
__device__ __forceinline__ void fused2_body(const F2Args& A, const long long cell, const int w, const int tid,
                                            typename F2L<N>::Wg& Wg, typename F2L<N>::Tab& tab,
                                            typename F2L<N>::Fhs& fhs, typename F2L<N>::Fs& fs) {
  if constexpr (F2Shape<N>::REAL) {
    (void)fs;
    fused2r_body(A, cell, w, tid, Wg, tab, fhs);
  } else {
  constexpr int THREADS = F2Shape<N>::THREADS, G = F2Shape<N>::G, PAD = N + 1;
  // 16^2 stages the cell's fhat / n and f in LDS; 32^2 reads them from L2, as
  // staging would cost the fourth workgroup per CU (8% slower measured; at 16^2
  // the LDS staging is 12% faster and does not change the occupancy)
  constexpr bool STAGE = N < 32;
  constexpr int SN = STAGE ? N : 1;
  // tab: E axis 0, E axis 1, G axis 0, G axis 1
  const int g = tid / N, l = tid % N;
  const int ps = A.pidx ? A.pidx[cell] : 0;
  const cx<float>* fhr = A.fh + cell * N * N + (long long)l * N;  // row l of fhat / n
  const float* fcl = A.f32 + cell * N * N + l;                     // column l of f
  if constexpr (STAGE) {
    for (int e = tid; e < N * N; e += THREADS) {
      fhs[e / N][e % N] = A.fh[cell * N * N + e];
      fs[e / N][e % N] = A.f32[cell * N * N + e];
    }
  }
  const int npairs = A.nr * A.m;
  const int stride = A.wpc * G;
  const int rounds = (npairs + stride - 1) / stride;
  const int kl = kfreq(l, N);
  cx<float> acc[N];
#pragma unroll
  for (int k = 0; k < N; ++k) acc[k] = cx<float>{0.f, 0.f};
  __syncthreads();
  cx<float> v[N];
  for (int r = 0; r < rounds; ++r) {
    const int pair = w * G + g + r * stride;
    const bool act = pair < npairs;
    const int j = act ? pair / A.m : 0, mm = act ? pair % A.m : 0;
    if (act) {
      if (A.ephase) {
        const cx<float>* e = A.ephase + (long long)pair * 2 * N;
        tab[g][0][l] = e[l];
        tab[g][1][l] = e[N + l];
      } else {
#pragma unroll
        for (int ax = 0; ax < 2; ++ax)
          tab[g][ax][l] = phase2_f32(0, A.rj[j], 0.0, A.sigma[mm * 2 + ax], (double)kl, A.L);
      }
      if (A.gphase) {
        const cx<float>* gt = A.gphase + ((long long)ps * npairs + pair) * 2 * N;
        tab[g][2][l] = gt[l];
        tab[g][3][l] = gt[N + l];
      } else {
#pragma unroll
        for (int ax = 0; ax < 2; ++ax)
          tab[g][2 + ax][l] = phase2_f32(1, A.rj[j], A.a[ps * A.nr + j], A.sigma[mm * 2 + ax], (double)kl, A.L);
      }
    }
    __syncthreads();
    if (act) {  // A
      const cx<float> e0 = tab[g][0][l];
#pragma unroll
      for (int k = 0; k < N; ++k)
        v[k] = (e0 * tab[g][1][k]) * (STAGE ? fhs[l % SN][k % SN] : fhr[k]);
      fft_reg<N, +1>(v);
#pragma unroll
      for (int k = 0; k < N; ++k) Wg[g][l][k] = v[k];
    }
    __syncthreads();
    if (act) {  // B
#pragma unroll
      for (int i = 0; i < N; ++i) v[i] = Wg[g][i][l];
      fft_reg<N, +1>(v);
#pragma unroll
      for (int i = 0; i < N; ++i) v[i] = (STAGE ? fs[i % SN][l % SN] : fcl[i * N]) * v[i];
      fft_reg<N, -1>(v);
#pragma unroll
      for (int i = 0; i < N; ++i) Wg[g][i][l] = v[i];
    }
    __syncthreads();
    if (act) {  // C
#pragma unroll
      for (int k = 0; k < N; ++k) v[k] = Wg[g][l][k];
      fft_reg<N, -1>(v);
      const float wj = (float)(A.crad[ps * A.nr + j] * A.wsig[mm]);
      const float* sp = A.spg + ((long long)ps * A.nr + j) * A.nq;
      const cx<float> g0 = tab[g][2][l];
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const int kk = kfreq(k, N);
        const cx<float> Gk = (wj * sp[kl * kl + kk * kk]) * (g0 * tab[g][3][k]);
        acc[k] = acc[k] + Gk * v[k];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < N; ++k) Wg[g][l][k] = acc[k];
  __syncthreads();
  cx<float>* out = A.part + (cell * A.wpc + w) * N * N;
  for (int e = tid; e < N * N; e += THREADS) {
    cx<float> s{0.f, 0.f};
#pragma unroll
    for (int gg = 0; gg < G; ++gg) s = s + Wg[gg][e / N][e % N];
    out[e] = s;
  }
  }
}

// Orders a wave's LDS accesses around an exchange between its own lanes: a wave's
// LDS instructions execute in program order, so only the compiler must be kept
// from moving them.
__device__ __forceinline__ void wave_sync2() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

namespace c3 {
constexpr int N = 16;
constexpr int PAD = N + 1;
constexpr int THREADS = N * N;
constexpr long long NPTS = (long long)N * N * N;
}  // namespace c3

struct C3Args {
  const cx<float>* fh;  // [cell][N][N][N], fhat / n
  const float* f32;     // [cell][N][N][N]
  cx<float>* part;      // [cell][nr][ns][N][N][N]
  const int* pidx;
  long long batch;
  const double* sigma;
  const double* rj;
  const double* wsig;
  const double* crad;
  const double* a;
  const float* spg;
  int nr, m, nq;
  int ns;  // workgroups per (cell, j), each over a contiguous range of the angles
  int xg;  // workgroup numbering (cell_of_block)
  double L;
};

// Pair schedule (2 workgroup barriers per quadrature pair): plane kx of the work
// array is written and read in passes A, B, D and E only by the 16 threads (kx, *),
// which share a wave, so those exchanges need no barrier; the x pass C reads and
// writes every plane.  W[x][y][z] rows are padded to N + 1, so every pass reads
// and writes 16 lanes of distinct banks.
// One workgroup's loop: cell `cell` (< A.batch), partial index jj (radial node jj / ns,
// angle range jj % ns), thread tid.
__device__ __forceinline__ void fusedc3_body(const C3Args& A, const long long cell, const int jj, const int tid,
                                             cx<float> (&W)[c3::N][c3::N][c3::PAD],
                                             cx<float> (&tE)[2][3][c3::N], cx<float> (&tG)[2][3][c3::N],
                                             float (&sp)[3 * (c3::N / 2) * (c3::N / 2) + 1]) {
  using namespace c3;
  const int nrs = A.nr * A.ns;
  const int j = jj / A.ns, asp = jj % A.ns;
  const int m0 = (asp * A.m) / A.ns, mc = ((asp + 1) * A.m) / A.ns - m0;  // angles m0 .. m0 + mc - 1
  const int ps = A.pidx ? A.pidx[cell] : 0;
  for (int q = tid; q < A.nq; q += THREADS) sp[q] = A.spg[((long long)ps * A.nr + j) * A.nq + q];
  const double aj_pi = A.a[ps * A.nr + j] * A.rj[j] / M_PI;
  const double rpl_pi = A.rj[j] / A.L;  // pi r_j / L in half turns
  // pencil role: the two fixed indices (r0, r1) of the thread's pencil in every pass
  const int r0 = tid / N, r1 = tid % N;
  cx<float> fhv[N];  // fhat along z at (kx, ky) = (r0, r1)
  float fv[N];       // f along x at (y, z) = (r0, r1)
  {
    const cx<float>* fh = A.fh + cell * NPTS + (long long)(r0 * N + r1) * N;
    const float* fc = A.f32 + cell * NPTS + r0 * N + r1;
#pragma unroll
    for (int k = 0; k < N; ++k) fhv[k] = fh[k];
#pragma unroll
    for (int x = 0; x < N; ++x) fv[x] = fc[x * N * N];
  }
  cx<float> acc[N];
#pragma unroll
  for (int k = 0; k < N; ++k) acc[k] = cx<float>{0.f, 0.f};
  // phase factors of pair mm along each axis (the separable factors of
  // inelastic_collisions.py:97-105), FP32 sincospi of the FP64-reduced angle
  auto make_tables = [&](int mm, int b) {
    if (tid < 6 * N) {
      const int which = tid / (3 * N), ax = (tid % (3 * N)) / N, k = tid % N;
      const double s = A.sigma[mm * 3 + ax];
      const double kf = (double)kfreq(k, N);
      // the angle in half turns, reduced to [-1, 1] in FP64; the FP32 sincospi of the
      // reduced value is as accurate as the FP32 rounding of an FP64 sincos
      double t = which == 0 ? -(rpl_pi * s * kf) : (aj_pi * s * kf);
      t -= 2.0 * rint(0.5 * t);
      float sn, cs;
      sincospif((float)t, &sn, &cs);
      (which ? tG : tE)[b][ax][k] = cx<float>{cs, sn};
    }
  };
  make_tables(m0, 0);
  __syncthreads();
  cx<float> v[N];
  for (int t = 0; t < mc; ++t) {
    const int mm = m0 + t, b = t & 1;
    // A: E * fhat along z at (kx, ky) = (r0, r1), inverse along z
    {
      const cx<float> row = tE[b][0][r0] * tE[b][1][r1];
#pragma unroll
      for (int k = 0; k < N; ++k) v[k] = (row * tE[b][2][k]) * fhv[k];
      fft_reg<N, +1>(v);
#pragma unroll
      for (int z = 0; z < N; ++z) W[r0][r1][z] = v[z];
    }
    wave_sync2();  // plane kx = r0 belongs to the 16 threads (r0, *) of one wave
    // B: inverse along y at (kx, z) = (r0, r1)
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int k = bitrev<N>(i);
      v[k] = W[r0][k][r1];
    }
    fft_reg<N, +1>(v);
#pragma unroll
    for (int y = 0; y < N; ++y) W[r0][y][r1] = v[y];
    __syncthreads();
    // C: along x at (y, z) = (r0, r1): inverse, * f, forward
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int k = bitrev<N>(i);
      v[k] = W[k][r0][r1];
    }
    fft_reg<N, +1>(v);
#pragma unroll
    for (int x = 0; x < N; ++x) v[x] = fv[x] * v[x];
    fft_reg<N, -1>(v);
#pragma unroll
    for (int k = 0; k < N; ++k) W[k][r0][r1] = v[k];
    // the next pair's phase tables, into the other buffer: its last readers (phases A
    // and E of the previous pair) are behind the barrier above
    if (t + 1 < mc) make_tables(mm + 1, b ^ 1);
    __syncthreads();
    // D: forward along y at (kx, z) = (r0, r1)
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int k = bitrev<N>(i);
      v[k] = W[r0][k][r1];
    }
    fft_reg<N, -1>(v);
#pragma unroll
    for (int k = 0; k < N; ++k) W[r0][k][r1] = v[k];
    wave_sync2();  // again within the plane's wave
    // E: forward along z at (kx, ky) = (r0, r1), G * (.) and the sum over angles
    // (the thread's own row: the next pair's phase A may overwrite it right after)
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int k = bitrev<N>(i);
      v[k] = W[r0][r1][k];
    }
    fft_reg<N, -1>(v);
    const cx<float> grow = (float)A.wsig[mm] * (tG[b][0][r0] * tG[b][1][r1]);
#pragma unroll
    for (int k = 0; k < N; ++k) acc[k] = acc[k] + (grow * tG[b][2][k]) * v[k];
  }
  // radial weight and sp(a_j r_j |k|), then this (cell, j)'s partial spectrum
  const float cj = (float)A.crad[ps * A.nr + j];
  const int kx = kfreq(r0, N), ky = kfreq(r1, N);
  cx<float>* out = A.part + ((long long)cell * nrs + jj) * NPTS + (long long)(r0 * N + r1) * N;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int kz = kfreq(k, N);
    out[k] = (cj * sp[kx * kx + ky * ky + kz * kz]) * acc[k];
  }
}

// One half of an N-point transform (DIR = +1 inverse, -1 forward) of a pencil
// a[0..N-1], computed by lane h of the pencil's two lanes: on entry r[n] =
// a[n] + (1 - 2h) a[n + N/2] (the first radix-2 decimation-in-frequency stage);
// r[n] is multiplied by tw[n] = 1 (h = 0) or exp(2 pi i n / N) (h = 1, conjugated
// for DIR = -1), read from LDS so that the twiddle constants take no registers,
// and an N/2-point FFT leaves X[2k + h] in r[k].
template <int DIR, int H, typename T>
__device__ __forceinline__ void half_dif(cx<T> (&r)[H], const cx<T>* tw) {
#pragma unroll
  for (int n = 1; n < H; ++n) {
    const cx<T> w = tw[n];
    r[n] = DIR > 0 ? r[n] * w : r[n] * conj(w);
  }
  fft_reg<H, DIR>(r);
}

template <typename T>
struct CWArgs {
  const cx<T>* fh;      // [cell][N][N], fhat / n
  const T* f;           // [cell][N][N]
  const cx<T>* ephase;  // [pair][2][N] or null
  const cx<T>* gphase;  // [param][pair][2][N] or null
  cx<T>* part;          // [cell][nr][ns][N][N]
  const int* pidx;
  long long batch;
  const double* sigma;
  const double* rj;
  const double* wsig;
  const double* crad;
  const double* a;
  const T* spg;
  int nr, m, nq;
  int ns;  // workgroups per (cell, j), each over a contiguous range of the angles
  int xg;  // workgroup numbering (cell_of_block)
  double L;
};

// 2D grids whose pencil would not fit in one lane's registers next to the sum over
// angles (FP32 64^2, FP64 32^2): one 2N-thread workgroup per (cell, radial node j)
// loops over the M angles with the N x N complex work array in LDS; each pass
// gives a pencil to two adjacent lanes (half_dif), so the sum over angles of a
// row's N frequencies takes N/2 complex registers per lane.  Per pair: A rows
// (E * fhat from L2, inverse along axis 1); B columns (inverse along axis 0, * f,
// forward along axis 0; the column's two lanes are in one wave); C rows (forward
// along axis 1, G * (.) and the sum).  Two workgroup barriers per pair.
//
// The work array is unpadded, element (r, c) at r N + (c ^ g(r)).  The lanes (p, h) of a
// wave access (p, 2k + h) in phase A, (2k + h, p) at the end of phase B and (n, p), (p, n)
// in the reads; the row permutation g puts every group of lanes served in one LDS cycle on
// distinct banks in all four patterns (a padded row of N + 1 elements puts lanes (p, 1) and
// (p + 1, 0) on one bank in both write patterns).
//   8-byte elements (FP32): g(r) = 2 floor(r / 2) ^ (3N/8) (r mod 2).
//   16-byte elements (FP64), for gfx950's 16-byte LDS accesses as measured: a read serves
//   16 consecutive lanes per cycle, each group of 8 lanes only from one 128-byte half of
//   the banks, the 16 lanes' distinct addresses on distinct 16-byte bank groups; a write
//   serves 8 lanes per cycle on distinct bank groups.  g(r) = 4 (r & 1) | (r & 2) |
//   2 (r & 4): the 4 rows read by a group of 8 lanes in phase C lie in one half, the 8 rows
//   of 16 lanes on distinct groups; g(r) / 2 is distinct over 4 rows (phase A writes) and
//   g(2k + 1) ^ g(2k) = 4 (phase B writes).
template <typename T, int N>
__device__ __forceinline__ constexpr int wsw(int r) {
  if constexpr (sizeof(cx<T>) == 16)
    return ((r & 1) << 2) | (r & 2) | ((r & 4) << 1);
  else
    return (2 * (r >> 1)) ^ ((3 * N / 8) * (r & 1));
}

// SPLIT = false is the one-range form (ns = 1): its FP64 instance keeps two waves per SIMD,
// which the split form's extra scalar values would cost it.
// One workgroup's loop: cell `cell` (< A.batch), partial index jj (radial node jj / ns,
// angle range jj % ns), thread tid.  twd: ones, then exp(2 pi i n / N); its second row
// starts one element later than H, so that the two lanes of a pencil (reading entry n of
// row 0 and of row 1 together) use different banks.
template <typename T, int N, bool SPLIT>
__device__ __forceinline__ void fusedw_body(const CWArgs<T>& A, const long long cell, const int jj, const int tid,
                                            cx<T> (&W)[N * N], cx<T> (&tE)[2][2][N], cx<T> (&tG)[2][2][N],
                                            cx<T> (&twd)[2][N / 2 + 1]) {
  constexpr int H = N / 2, THREADS = 2 * N;
  // wsw(2k + h) = wsw(2k) ^ h DH for both permutations
  constexpr int DH = wsw<T, N>(1) ^ wsw<T, N>(0);
  static_assert((wsw<T, N>(3) ^ wsw<T, N>(2)) == DH && (wsw<T, N>(2 * H - 1) ^ wsw<T, N>(2 * H - 2)) == DH,
                "row permutation: wsw(2k + 1) ^ wsw(2k) must not depend on k");
  constexpr long long NPTS = (long long)N * N;
  const int ns = SPLIT ? A.ns : 1;
  const int nrs = A.nr * ns;
  const int j = jj / ns, asp = jj % ns;
  const int m0 = SPLIT ? (asp * A.m) / ns : 0;  // angles m0 .. mend - 1
  const int mend = SPLIT ? ((asp + 1) * A.m) / ns : A.m;
  const int ps = A.pidx ? A.pidx[cell] : 0;
  if (tid < N) {
    const int n = tid % H;
    twd[tid / H][n] = tid < H ? cx<T>{(T)1, (T)0}
                              : cx<T>{(T)TW64[n * (64 / N)][0], (T)TW64[n * (64 / N)][1]};
  }
  const int p = tid >> 1, h = tid & 1;  // pencil and half
  const cx<T>* tw = twd[h];
  // swizzled offsets: row p's element 2k + h is at p N + (2k ^ hg); row 2k + h's element p is
  // at h N + 2k N + (wsw(2k) ^ ph).  Each phase takes its offsets from an opaque copy, so the
  // compiler forms the loop-invariant addresses next to their accesses instead of holding
  // all of them in registers across the pair loop.
  const int hg0 = h ^ wsw<T, N>(p);
  const int ph0 = p ^ (DH * h);
  auto opaque = [](int v) {
    asm volatile("" : "+v"(v));
    return v;
  };
  const T sg = h ? (T)-1 : (T)1;
  const double aj = A.a[ps * A.nr + j] * A.rj[j];
  const double rpl = M_PI * A.rj[j] / A.L;
  const int npairs = A.nr * A.m;
  const cx<T>* fh = A.fh + cell * NPTS;
  const T* fc = A.f + cell * NPTS;
  // phase factors of pair (j, mm) along each axis: from the plan's tables when it
  // has them, else FP64 sincos (inelastic_collisions.py:97-105)
  auto load_tables = [&](int mm, int b) {
    const int pair = j * A.m + mm;
    for (int e = tid; e < 4 * N; e += THREADS) {
      const int which = e / (2 * N), ax = (e / N) % 2, k = e % N;
      cx<T> val;
      if (which == 0 && A.ephase) {
        val = A.ephase[((long long)pair * 2 + ax) * N + k];
      } else if (which == 1 && A.gphase) {
        val = A.gphase[(((long long)ps * npairs + pair) * 2 + ax) * N + k];
      } else if constexpr (std::is_same<T, float>::value) {
        val = phase2_f32(which, A.rj[j], A.a[ps * A.nr + j], A.sigma[mm * 2 + ax], (double)kfreq(k, N), A.L);
      } else {
        const double s = A.sigma[mm * 2 + ax];
        const double kf = (double)kfreq(k, N);
        const double th = which == 0 ? -(rpl * s * kf) : (aj * s * kf);
        double sn, cs;
        sincos(th, &sn, &cs);
        val = cx<T>{(T)cs, (T)sn};
      }
      (which ? tG : tE)[b][ax][k] = val;
    }
  };
  auto split_row = [&](cx<T> (&r)[H], const cx<T>& a0, const cx<T>& a1, int n) {
    r[n] = cx<T>{a0.re + sg * a1.re, a0.im + sg * a1.im};
  };
  // FP32 with the plan's tables: the next pair's entries of this thread (one E and
  // one G value: the loop of load_tables visits e = tid and e = tid + 2N) are loaded
  // at the start of phase B and stored to LDS at its end, so their L2 latency
  // overlaps B (3.5% at 64^2; no gain in FP64)
  const bool pre = sizeof(T) == 4 && A.ephase && A.gphase;
  const int tax = (tid / N) % 2, tk = tid % N;
  cx<T> nt[2];
  auto fetch = [&](int mm) {
    const int pair = j * A.m + mm;
    nt[0] = A.ephase[((long long)pair * 2 + tax) * N + tk];
    nt[1] = A.gphase[(((long long)ps * npairs + pair) * 2 + tax) * N + tk];
  };
  cx<T> acc[H];
#pragma unroll
  for (int k = 0; k < H; ++k) acc[k] = cx<T>{(T)0, (T)0};
  load_tables(m0, 0);
  __syncthreads();
  cx<T> r[H];
  for (int mm = m0; mm < mend; ++mm) {
    const int b = (mm - m0) & 1;
    // A: row p = kx: E * fhat, inverse along axis 1 (ky -> y); the row factor
    // E_x(kx) is constant along the row and applied to the outputs
    {
      const cx<T>* fr = fh + (long long)p * N;
#pragma unroll
      for (int n = 0; n < H; ++n) {
        split_row(r, tE[b][1][n] * fr[n], tE[b][1][n + H] * fr[n + H], n);
        if ((n & 7) == 7) __builtin_amdgcn_sched_barrier(0);  // 16 loads in flight at a time
      }
      half_dif<+1>(r, tw);
      const cx<T> ex = tE[b][0][p];
      const int hg = opaque(hg0), rowp = opaque(p * N);
#pragma unroll
      for (int k = 0; k < H; ++k) W[rowp + ((2 * k) ^ hg)] = ex * r[k];
    }
    __syncthreads();
    // B: column p = y: inverse along axis 0 (kx -> x), * f, forward along axis 0;
    // the next pair's phase tables (their buffer's last readers, phases A and C of
    // the previous pair, are behind the barrier above)
    {
      if (pre && mm + 1 < mend) fetch(mm + 1);
      int pc = opaque(p);
#pragma unroll
      for (int n = 0; n < H; ++n)
        split_row(r, W[n * N + (pc ^ wsw<T, N>(n))], W[(n + H) * N + (pc ^ wsw<T, N>(n + H))], n);
      half_dif<+1>(r, tw);
      // 32-bit offset made opaque per pair, so the addresses are formed next to
      // their loads instead of being held in registers across the pair loop
      int toff = h * N + p;
      asm volatile("" : "+v"(toff));
#pragma unroll
      for (int k = 0; k < H; ++k) {
        r[k] = fc[2 * k * N + toff] * r[k];
        if ((k & 7) == 7) __builtin_amdgcn_sched_barrier(0);
      }
      wave_sync2();  // both lanes of the column have read it
      int ph = opaque(ph0), rowh = opaque(h * N);
#pragma unroll
      for (int k = 0; k < H; ++k) W[rowh + 2 * k * N + (wsw<T, N>(2 * k) ^ ph)] = r[k];
      wave_sync2();  // ... and written their halves
      pc = opaque(p);
#pragma unroll
      for (int n = 0; n < H; ++n)
        split_row(r, W[n * N + (pc ^ wsw<T, N>(n))], W[(n + H) * N + (pc ^ wsw<T, N>(n + H))], n);
      half_dif<-1>(r, tw);
      wave_sync2();
      ph = opaque(ph0);
      rowh = opaque(h * N);
#pragma unroll
      for (int k = 0; k < H; ++k) W[rowh + 2 * k * N + (wsw<T, N>(2 * k) ^ ph)] = r[k];
    }
    if (mm + 1 < mend) {
      if (pre) {
        tE[b ^ 1][tax][tk] = nt[0];
        tG[b ^ 1][tax][tk] = nt[1];
      } else {
        load_tables(mm + 1, b ^ 1);
      }
    }
    __syncthreads();
    // C: row p = kx: forward along axis 1 (y -> ky), G * (.), sum over angles (the
    // row's own two lanes are the only readers, so the next pair's phase A may
    // overwrite it without a barrier)
    {
      const int gp = opaque(wsw<T, N>(p)), rowp = opaque(p * N);
#pragma unroll
      for (int n = 0; n < H; ++n) split_row(r, W[rowp + (n ^ gp)], W[rowp + ((n + H) ^ gp)], n);
      half_dif<-1>(r, tw);
      const cx<T> gx = (T)A.wsig[mm] * tG[b][0][p];
#pragma unroll
      for (int k = 0; k < H; ++k) acc[k] = acc[k] + (gx * tG[b][1][2 * k + h]) * r[k];
      wave_sync2();
    }
  }
  // radial weight and sp(a_j r_j |k|), then this (cell, j, angle range)'s partial spectrum
  const T cj = (T)A.crad[ps * A.nr + j];
  const T* sp = A.spg + ((long long)ps * A.nr + j) * A.nq;
  const int kx = kfreq(p, N);
  cx<T>* out = A.part + ((long long)cell * nrs + jj) * NPTS + (long long)p * N;
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const int ky = kfreq(2 * k + h, N);
    out[2 * k + h] = (cj * sp[kx * kx + ky * ky]) * acc[k];
  }
}

}  // namespace ib
