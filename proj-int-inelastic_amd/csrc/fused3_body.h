// Device code of the fused FP32 3D 32^3 pair loop (k_fused3): constants, arguments,
// LDS layout and the per-workgroup body.  Included by collide_fused.hip (the kernel
// and its host driver) and collide_group.hip (the grouped launch of the fused shapes).
#pragma once
#include "common.h"

namespace ib {

namespace f3 {
constexpr int N = 32;
constexpr int NPL = 16;  // packed kx planes
constexpr int PAD = N + 1;
constexpr int THREADS = 512;
constexpr long long NPTS = (long long)N * N * N;
constexpr long long PLANE = (long long)N * N;
}  // namespace f3

struct F3Args {
  const cx<float>* fh;  // [cell][17][32][32], kx = 0..16, scaled by 1/n
  const float* f32;     // [cell][32][32][32]
  cx<float>* part;      // [cell][nr][kp][16][32][32]: sum (kp = 1), sum and compensation (kp = 2)
  const int* pidx;
  long long batch;
  long long cell0;
  const double* sigma;
  const double* rj;
  const double* wsig;
  const double* crad;
  const double* a;
  const float* spg;
  int nr, m, nq;
  int ns;  // workgroups per (cell, j), each over a contiguous range of the angles
  int xg;  // 1: the workgroups of a cell on one XCD (batches of 8 cells or more)
  double L;
};

__device__ __forceinline__ cx<float> mulf(cx<float> a, cx<float> b) { return a * b; }

// s += x, Kahan-compensated (c carries the lost low-order part) or plain
template <bool KAHAN>
__device__ __forceinline__ void acc_add(cx<float>& s, cx<float>& c, cx<float> x) {
  if constexpr (KAHAN) {
    const float yr = x.re - c.re;
    const float tr = s.re + yr;
    c.re = (tr - s.re) - yr;
    s.re = tr;
    const float yi = x.im - c.im;
    const float ti = s.im + yi;
    c.im = (ti - s.im) - yi;
    s.im = ti;
  } else {
    (void)c;
    s = s + x;
  }
}

// A complex Kahan compensation kept between pairs as two bf16 halves of one
// 32-bit word: the high 16 bits of each FP32 component (truncation).  The term is
// the rounding error of the running sum, so dropping 2^-8 of it per step leaves
// the compensated sum's error near (2 + M / 2^8) ulp instead of the plain sum's
// (M - 1) ulp.
__device__ __forceinline__ unsigned pack_bf16x2(cx<float> c) {
  // bytes 3, 2 of c.im over bytes 3, 2 of c.re
  return __builtin_amdgcn_perm(__float_as_uint(c.im), __float_as_uint(c.re), 0x07060302u);
}
__device__ __forceinline__ cx<float> unpack_bf16x2(unsigned b) {
  return cx<float>{__uint_as_float(b << 16), __uint_as_float(b & 0xffff0000u)};
}

// Hermitian part of a separable factor phi_x(kx) phi_y(ky) phi_z(kz): every
// factor is Hermitian except its Nyquist entry, so the Hermitian part is
// Re(product of the Nyquist factors) * (product of the others).  A row fixes
// (kx, ky); sym_at() adds kz.
struct SymRow {
  cx<float> nyq, rest, row;
};

__device__ __forceinline__ SymRow sym_row(cx<float> fx, bool nx, cx<float> fy, bool ny) {
  const cx<float> one{1.f, 0.f};
  const cx<float> nyq = (nx ? fx : one) * (ny ? fy : one);
  const cx<float> rest = (nx ? one : fx) * (ny ? one : fy);
  return SymRow{nyq, rest, nyq.re * rest};
}

__device__ __forceinline__ cx<float> sym_at(const SymRow& r, cx<float> fz, bool nz) {
  return nz ? (r.nyq * fz).re * r.rest : r.row * fz;
}

// Orders a wave's LDS writes before its later LDS reads of other lanes' data: a
// wave's LDS instructions execute in program order, so only the compiler must be
// kept from moving them.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Pair schedule (2 workgroup barriers per quadrature pair).  Wave w owns the
// kx planes 2w and 2w + 1 in the kz and ky passes, so A -> B and D -> E exchange
// data inside a wave only:
//   A  E * fhat and the inverse kz transform (pencil (p, ky) in registers); the
//      packed plane p = 0 adds i * X16 staged in phase C of the previous pair
//   B  inverse ky                                             -- barrier --
//   C  along x (C2R, * f, R2C); side work of all threads: G * (.) and the sum
//      for the packed plane of the previous pair (from SP with its (-ky, -kz)
//      mirror), the phase tables two pairs ahead, X16 of the next pair
//                                                             -- barrier --
//   D  forward ky
//   E  forward kz, then G * (.) and the sum (p >= 1, registers) or the packed
//      plane parked in SP (p = 0)
// The phase tables are four-buffered (pair mm uses buffer mm mod 4), so phase C
// fills one that no phase of the same barrier interval reads.
// The pair loop of one workgroup: cell `cell` (< A.batch) of the launch, partial index jj
// (radial node j = jj / ns, angle range jj % ns), thread tid.  Shared by k_fused3 and the
// grouped launch of the fused shapes (collide_group.hip).  The LDS arrays come in as
// references to arrays at fixed LDS addresses (k_fused3's own __shared__ arrays, or fixed
// offsets of the grouped kernel's LDS bytes): those addresses fold into the LDS
// instructions' offsets, where members of one __shared__ struct or union cost the kernel
// 70 more spilled registers.
template <bool KAHAN>
__device__ __forceinline__ void fused3_body(const F3Args& A, const long long cell, const int jj, const int tid,
                                            cx<float> (&W)[f3::NPL][f3::N][f3::PAD],
                                            cx<float> (&tE)[4][3][f3::N], cx<float> (&tG)[4][3][f3::N],
                                            float (&sp)[3 * (f3::N / 2) * (f3::N / 2) + 1],
                                            cx<float> (&S16)[f3::N][f3::PAD], cx<float> (&SP)[f3::N][f3::PAD]) {
  using namespace f3;
  constexpr int KP = KAHAN ? 2 : 1;
  const int nrs = A.nr * A.ns;
  const int j = jj / A.ns, asp = jj % A.ns;
  const int m0 = (asp * A.m) / A.ns, mc = ((asp + 1) * A.m) / A.ns - m0;  // angles m0 .. m0 + mc - 1
  const int ps = A.pidx ? A.pidx[A.cell0 + cell] : 0;
  for (int q = tid; q < A.nq; q += THREADS) sp[q] = A.spg[((long long)ps * A.nr + j) * A.nq + q];
  const double aj = A.a[ps * A.nr + j] * A.rj[j];
  const double crad_j = A.crad[ps * A.nr + j];
  const double rpl_pi = A.rj[j] / A.L;       // pi r_j / L in half turns
  const double aj_pi = aj / M_PI;
  const int p = tid / N, l = tid % N;                 // plane-pass role
  const cx<float>* fh = A.fh + (A.cell0 + cell) * 17 * PLANE;
  const float* fc = A.f32 + (A.cell0 + cell) * NPTS;
  const int kyf = kfreq(l, N);
  // Kahan: the compensation of the kz-pencil sums (bf16 pairs) lives in this
  // workgroup's compensation plane of the partial buffer, laid out [k][thread] so
  // that every access of a wave is one contiguous 256-byte run, read and rewritten
  // once per pair in phase E: 64 KiB per workgroup stay in L2, and the 64
  // registers a register-held FP32 compensation needs stay out of the pair loop
  // (register-held, it spilled 444 bytes per lane and ran at 5.5k cell-evals/s;
  // FP32 in this buffer, 6.8k; bf16, 8.9k)
  unsigned* compb =
      (unsigned*)(A.part + ((A.cell0 + cell) * nrs + jj) * KP * NPL * PLANE + NPL * PLANE);

  cx<float> acc[N];
  cx<float> acc2[2], cmp2[2];
#pragma unroll
  for (int k = 0; k < N; ++k) acc[k] = cx<float>{0.f, 0.f};
  acc2[0] = acc2[1] = cmp2[0] = cmp2[1] = cx<float>{0.f, 0.f};

  // The side tasks recompute their indices from an opaque copy of tid each pair,
  // so that none of their addresses occupies a register across the pair loop.
  auto make_tables = [&](int mm, int b) {
    int tq = tid;
    asm volatile("" : "+v"(tq));
    if (tq < 6 * N) {
      const int which = tq / (3 * N), ax = (tq % (3 * N)) / N, k = tq % N;
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
  // fhat on the kx = 16 plane at this thread's two points (the same for every pair)
  cx<float> f16[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) f16[h] = fh[16 * PLANE + tid + h * THREADS];
  auto make_x16 = [&](int b) {
    int tq = tid;
    asm volatile("" : "+v"(tq));
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int e = tq + h * THREADS, ky = e / N, kz = e % N;
      const SymRow r16 = sym_row(tE[b][0][N / 2], true, tE[b][1][ky], ky == N / 2);
      S16[ky][kz] = sym_at(r16, tE[b][2][kz], kz == N / 2) * f16[h];
    }
  };
  auto packed_acc = [&](int b, float wjp) {
    int tq = tid;
    asm volatile("" : "+v"(tq));
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int e = tq + h * THREADS, ky = e / N, kz = e % N;
      const cx<float> P = SP[ky][kz], Pm = SP[(N - ky) % N][(N - kz) % N];
      const cx<float> h0{0.5f * (P.re + Pm.re), 0.5f * (P.im - Pm.im)};
      const cx<float> h16{0.5f * (P.im + Pm.im), 0.5f * (Pm.re - P.re)};
      const int kyq = kfreq(ky, N), kzq = kfreq(kz, N);
      const int qq = kyq * kyq + kzq * kzq;
      const bool ny = ky == N / 2, nz = kz == N / 2;
      const SymRow r0 = sym_row(tG[b][0][0], false, tG[b][1][ky], ny);
      const SymRow r16 = sym_row(tG[b][0][N / 2], true, tG[b][1][ky], ny);
      const cx<float> G0 = (wjp * sp[qq]) * sym_at(r0, tG[b][2][kz], nz);
      const cx<float> G16 = (wjp * sp[qq + (N / 2) * (N / 2)]) * sym_at(r16, tG[b][2][kz], nz);
      const cx<float> t0 = G0 * h0, t1 = G16 * h16;
      acc_add<KAHAN>(acc2[h], cmp2[h], cx<float>{t0.re - t1.im, t0.im + t1.re});
    }
  };

  make_tables(m0, 0);
  if (mc > 1) make_tables(m0 + 1, 1);
  __syncthreads();
  make_x16(0);
  __syncthreads();

  float wj_prev = 0.f;
  cx<float> v[N];
  // fhat along this thread's kz pencil is the same for every pair: phase E
  // reloads it into the pencil registers as they free up, so the L2 latency
  // overlaps the sum over angles
  const cx<float>* fp = fh + (long long)p * PLANE + (long long)l * N;
#pragma unroll
  for (int k = 0; k < N; ++k) v[k] = fp[k];
  for (int t = 0; t < mc; ++t) {
    const int mm = m0 + t, b = t & 3;
    const float wj = (float)(crad_j * A.wsig[mm]);
    const float ws_half = (float)(0.5 * A.wsig[mm]);
    // A: E * fhat, inverse along kz
    {
      const SymRow rr = sym_row(tE[b][0][p], false, tE[b][1][l], l == N / 2);
#pragma unroll
      for (int k = 0; k < N; ++k) v[k] = sym_at(rr, tE[b][2][k], k == N / 2) * v[k];
      if (p == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const cx<float> x = S16[l][k];
          v[k] = cx<float>{v[k].re - x.im, v[k].im + x.re};
        }
      }
      fft_reg<N, +1>(v);
#pragma unroll
      for (int k = 0; k < N; ++k) W[p][l][k] = v[k];
    }
    wave_sync();  // planes p = 2w, 2w + 1 belong to wave w: phase B reads only them
    // B: inverse along ky
#pragma unroll
    for (int i = 0; i < N; ++i) {  // in the order the radix-2 stages consume them
      const int k = bitrev<N>(i);
      v[k] = W[p][k][l];
    }
    fft_reg<N, +1>(v);
#pragma unroll
    for (int k = 0; k < N; ++k) W[p][k][l] = v[k];
    __syncthreads();
    // C: along x, two real pencils (xy, xz), (xy, xz + 16) as one complex FFT
    // (16 lanes of a row touch 16 consecutive complex values: no LDS bank conflicts)
    {
      // roles from an opaque copy of tid, so the pencils' LDS addresses are formed
      // here each pair instead of occupying registers across the pair loop
      int tq = tid;
      asm volatile("" : "+v"(tq));
      const int xy = tq / 16, xz = tq % 16;
      cx<float> Z[N];
      {
        const cx<float> a0 = W[0][xy][xz], b0 = W[0][xy][xz + 16];
        Z[0] = cx<float>{a0.re, b0.re};
        Z[N / 2] = cx<float>{a0.im, b0.im};
      }
#pragma unroll
      for (int i = 1; i < N / 2; ++i) {
        const int k = bitrev<N / 2>(i);
        const cx<float> a = W[k][xy][xz], bb = W[k][xy][xz + 16];
        Z[k] = cx<float>{a.re - bb.im, a.im + bb.re};
        Z[N - k] = cx<float>{a.re + bb.im, bb.re - a.im};
      }
      // 32-bit per-thread offset, made opaque per pair so the 32 addresses are
      // formed next to their loads instead of being hoisted out of the loop
      int toff = xy * N + xz;
      asm volatile("" : "+v"(toff));
      // f along the two pencils, in two halves of 16 x: the first half is loaded
      // before the inverse transform (its latency hides behind the FFT), the
      // second before the first is used
      float fa[N], fb[N];
#pragma unroll
      for (int x = 0; x < N / 2; ++x) {
        fa[x] = fc[x * (int)PLANE + toff];
        fb[x] = fc[x * (int)PLANE + toff + 16];
      }
      __builtin_amdgcn_sched_barrier(0);
      fft_reg<N, +1>(Z);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int x = N / 2; x < N; ++x) {
        fa[x] = fc[x * (int)PLANE + toff];
        fb[x] = fc[x * (int)PLANE + toff + 16];
      }
#pragma unroll
      for (int x = 0; x < N / 2; ++x) Z[x] = cx<float>{Z[x].re * fa[x], Z[x].im * fb[x]};
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int x = N / 2; x < N; ++x) Z[x] = cx<float>{Z[x].re * fa[x], Z[x].im * fb[x]};
      __builtin_amdgcn_sched_barrier(0);
      fft_reg<N, -1>(Z);
      W[0][xy][xz] = cx<float>{Z[0].re, Z[N / 2].re};
      W[0][xy][xz + 16] = cx<float>{Z[0].im, Z[N / 2].im};
#pragma unroll
      for (int k = 1; k < N / 2; ++k) {
        const cx<float> zk = Z[k], zm = Z[N - k];
        // planes 1..15 carry twice their value; phase E folds the 1/2 into its weight
        W[k][xy][xz] = cx<float>{zk.re + zm.re, zk.im - zm.im};
        W[k][xy][xz + 16] = cx<float>{zk.im + zm.im, zm.re - zk.re};
      }
    }
    // side work: the packed plane of the previous pair, the phase tables two pairs
    // ahead and X16 of the next pair (tables are four-buffered, so none of these
    // touches a buffer another phase of this barrier interval reads)
    if (t > 0) packed_acc((t - 1) & 3, wj_prev);
    if (t + 2 < mc) make_tables(mm + 2, (t + 2) & 3);
    if (t + 1 < mc) make_x16((t + 1) & 3);
    __syncthreads();
    // D: forward along ky
#pragma unroll
    for (int i = 0; i < N; ++i) {  // in the order the radix-2 stages consume them
      const int k = bitrev<N>(i);
      v[k] = W[p][k][l];
    }
    fft_reg<N, -1>(v);
#pragma unroll
    for (int k = 0; k < N; ++k) W[p][k][l] = v[k];
    wave_sync();
    // E: forward along kz, G * (.) and the sum over angles
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int k = bitrev<N>(i);
      v[k] = W[p][l][k];
    }
    unsigned cv[KAHAN ? N : 1];
    int to = tid;  // Kahan: opaque per pair, so the 32 addresses are formed next to their accesses
    if constexpr (KAHAN) asm volatile("" : "+v"(to));
    fft_reg<N, -1>(v);
    if constexpr (KAHAN) {  // loaded after the transform (fewer live registers)
      if (p != 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) cv[k] = t > 0 ? compb[k * THREADS + to] : 0u;
      }
    }
    if (p == 0) {
#pragma unroll
      for (int k = 0; k < N; ++k) {
        SP[l][k] = v[k];
        v[k] = fp[k];
      }
    } else {
      SymRow rr = sym_row(tG[b][0][p], false, tG[b][1][l], l == N / 2);
      rr.rest = ws_half * rr.rest;
      rr.row = ws_half * rr.row;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const cx<float> x = sym_at(rr, tG[b][2][k], k == N / 2) * v[k];
        if constexpr (KAHAN) {
          cx<float> c = unpack_bf16x2(cv[k]);
          acc_add<true>(acc[k], c, x);
          compb[k * THREADS + to] = pack_bf16x2(c);
        } else {
          acc[k] = acc[k] + x;
        }
        v[k] = fp[k];
      }
    }
    wj_prev = wj;
  }
  __syncthreads();
  packed_acc((mc - 1) & 3, wj_prev);

  cx<float>* out = A.part + ((A.cell0 + cell) * nrs + jj) * KP * NPL * PLANE;
  cx<float> cfin[KAHAN ? N : 1];
  if constexpr (KAHAN) {
    if (p != 0) {
#pragma unroll
      for (int k = 0; k < N; ++k) cfin[k] = unpack_bf16x2(compb[k * THREADS + tid]);
    }
    __syncthreads();  // every thread holds its compensation before the plane is rewritten
  }
  if (p != 0) {
    const float cj = (float)crad_j;
    const int q0 = p * p + kyf * kyf;
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const int kzf = kfreq(k, N);
      const float g = cj * sp[q0 + kzf * kzf];
      out[(long long)p * PLANE + (long long)l * N + k] = g * acc[k];
      if constexpr (KAHAN) {
        const cx<float> c = cfin[k];
        out[NPL * PLANE + (long long)p * PLANE + (long long)l * N + k] = -g * c;
      }
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int e = tid + h * THREADS;
    out[e] = acc2[h];
    if constexpr (KAHAN) out[NPL * PLANE + e] = cx<float>{-cmp2[h].re, -cmp2[h].im};
  }
}

}  // namespace ib
