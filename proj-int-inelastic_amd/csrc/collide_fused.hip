// Fused FP32 form of Q(f, f) for 3D 32^3 grids (the 3D velocity batch).
//
// Same operator as reference src/collision/inelastic_collisions.py:208-227,
// computed with real-data transforms: the Hermitian parts E_H, G_H of the phase
// factors make h = f * ifft(E_H fhat) real, so each 32^3 transform needs only the
// kx in [0, 16] half spectrum; the kx = 0 and kx = 16 planes (real after the x
// transform) are packed into one complex plane.  The anti-Hermitian parts E_A, G_A
// live on the three Nyquist planes only; their term G_A * fft(f * ifft(E_A fhat))
// is computed exactly by k_f3_nyq (collide_nyq.hip) and added in FP64 before the
// final transform, so the sum is the reference's C2C operator.
//
// Work decomposition: one 512-thread workgroup per (cell, radial node j)
// loops over the M angles.  The 16 packed kx-planes (32 x 33 padded complex
// FP32 = 132 KiB) stay in LDS for the whole pair; the pass along kz is done
// with the pencil in registers and fused with E * fhat on the way in and with
// G * (.) plus the sum over angles on the way out; the
// pass along x turns two real pencils into one complex FFT.  Workgroups of
// one cell are numbered so they land on the same XCD (the dispatcher assigns
// workgroup i to XCD i mod 8), so fhat and f of a cell are read from one L2.
// sp(a_j r_j |k|) and the radial weight do not depend on the angle, so the
// sum over angles runs on phase * transform only and is scaled once at the
// end.  Per-(cell, j) sums go to a partial buffer; the FP64 reduce
// over j, the final inverse transform, the loss term and the Laplacian run in
// FP64 (collide_plain.hip kernels), as does the prologue fhat = fft(f).
#include <algorithm>

#include "common.h"
#include "plan.h"

// The sum over the M angles inside a workgroup is plain FP32 (plan precision
// IB_PREC_FP32) or Kahan-compensated FP32 (IB_PREC_FP32_KAHAN, compensation kept
// between pairs as bf16 in an L2-resident buffer); the sum over radial nodes is
// FP64 either way (k_f3_reduce).

namespace ib {

// FP64 staged-engine pieces (collide_plain.hip)
int staged_prologue_fp64(Plan* P, const double* f, long long nb, void* fhat64, void* work);
int staged_epilogue_fp64(Plan* P, const double* f, double* q, long long nb, void* acc64,
                         void* fhat64, void* work, int with_loss, double hb, const int* pidx);
// fused C2C engines: FP32 2D and 3D 16^3, FP64 2D 32^2 (collide_fused2.hip)
int collide_fused2(Plan* P, const double* f, double* q, long long batch, const int* pidx,
                   int with_loss, double hb, hipStream_t st, int* handled);
// exact anti-Hermitian (Nyquist-plane) term of the real-data form (collide_nyq.hip)
size_t nyq_scratch_bytes(int nparts);
int launch_f3_nyq(Plan* P, const cx<float>* xp, const float* f32, cx<float>* out, cx<double>* acc64,
                  long long nb, const int* pidx, hipStream_t st);

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
template <bool KAHAN>
__global__ __launch_bounds__(512, 1) void k_fused3(F3Args A) {
  using namespace f3;
  constexpr int KP = KAHAN ? 2 : 1;
  __shared__ cx<float> W[NPL][N][PAD];
  __shared__ cx<float> tE[4][3][N];
  __shared__ cx<float> tG[4][3][N];
  __shared__ float sp[3 * (N / 2) * (N / 2) + 1];
  __shared__ cx<float> S16[N][PAD];
  __shared__ cx<float> SP[N][PAD];
  const int tid = threadIdx.x;
  const long long id = blockIdx.x;
  const int nrs = A.nr * A.ns;
  long long cell;
  int jj;  // partial index: radial node j, angle range asp
  cell_of_block(id, nrs, A.xg, cell, jj);
  const int j = jj / A.ns, asp = jj % A.ns;
  if (cell >= A.batch) return;  // uniform over the workgroup
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

// fhat (FP64, full spectrum) -> kx = 0..16 planes in FP32, scaled by 1/n; f -> FP32;
// the three Nyquist planes of fhat / n, [cell][3][32][32]: kx = 16 (ky, kz), ky = 16
// (kx, kz), kz = 16 (kx, ky), for the anti-Hermitian term.
__global__ void k_f3_pack(const cx<double>* __restrict__ fhat, const double* __restrict__ f,
                          cx<float>* __restrict__ fh32, float* __restrict__ f32,
                          cx<float>* __restrict__ xp, long long nb) {
  using namespace f3;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per = NPTS;
  if (gid >= nb * per) return;
  const long long c = gid / per, i = gid % per;
  f32[gid] = (float)f[gid];
  const int kx = (int)(i / PLANE), ky = (int)((i / N) % N), kz = (int)(i % N);
  const cx<double> v = fhat[gid];
  const double s = 1.0 / (double)NPTS;
  const cx<float> vs{(float)(v.re * s), (float)(v.im * s)};
  if (kx <= N / 2) fh32[c * 17 * PLANE + i] = vs;
  cx<float>* xc = xp + c * 3 * PLANE;
  if (kx == N / 2) xc[ky * N + kz] = vs;
  if (ky == N / 2) xc[PLANE + kx * N + kz] = vs;
  if (kz == N / 2) xc[2 * PLANE + kx * N + ky] = vs;
}

// Sum the per-j partials in FP64 and expand the packed half spectrum to the
// full FP64 spectrum (Hermitian completion), as the FP64 epilogue expects.  A
// thread owns one point of the stored planes (kx = 1..15, or the packed plane)
// and writes it together with its conjugate mirror, so every partial is read once.
__global__ void k_f3_reduce(const cx<float>* __restrict__ part, cx<double>* __restrict__ acc64,
                            long long nb, int nr, int kp) {
  using namespace f3;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nb * NPL * PLANE) return;
  const long long c = gid / (NPL * PLANE), i = gid % (NPL * PLANE);
  const int pl = (int)(i / PLANE), ky = (int)((i / N) % N), kz = (int)(i % N);
  auto sum_at = [&](int yy, int zz) {
    double re = 0.0, im = 0.0;
    const cx<float>* b = part + c * nr * kp * NPL * PLANE + (long long)pl * PLANE + (long long)yy * N + zz;
    for (int j = 0; j < nr; ++j) {
      const cx<float> hi = b[(long long)j * kp * NPL * PLANE];
      re += (double)hi.re;
      im += (double)hi.im;
      if (kp == 2) {
        const cx<float> lo = b[(long long)j * kp * NPL * PLANE + NPL * PLANE];
        re += (double)lo.re;
        im += (double)lo.im;
      }
    }
    return cx<double>{re, im};
  };
  const int my = (N - ky) % N, mz = (N - kz) % N;
  cx<double>* out = acc64 + c * NPTS;
  const cx<double> P = sum_at(ky, kz);
  if (pl == 0) {  // packed kx = 0 (re) and kx = 16 (im) planes
    const cx<double> Pm = sum_at(my, mz);
    out[(long long)ky * N + kz] = cx<double>{0.5 * (P.re + Pm.re), 0.5 * (P.im - Pm.im)};
    out[(N / 2) * PLANE + (long long)ky * N + kz] = cx<double>{0.5 * (P.im + Pm.im), 0.5 * (Pm.re - P.re)};
  } else {
    out[(long long)pl * PLANE + (long long)ky * N + kz] = P;
    out[(long long)(N - pl) * PLANE + (long long)my * N + mz] = cx<double>{P.re, -P.im};
  }
}

// which fused engine runs a plan: 0 none, 1 the 3D 32^3 real-data form here, 2 collide_fused2
static int fused_kind(const Plan* P) {
  if ((P->precision == IB_PREC_FP32 && (P->dim == 2 || P->n == 16)) ||
      (P->precision == IB_PREC_FP64 && P->dim == 2 && P->n == 32))
    return 2;
  return (P->dim == 3 && P->n == 32 && P->precision == IB_PREC_FP32) ? 1 : 0;
}

// per-cell scratch of the 3D 32^3 form: FP64 f-hat, accumulator and two work arrays; FP32
// partial planes per radial node (twice with Kahan compensation), FP32 f-hat and f, the
// Nyquist planes and the Nyquist term's partials
static long long f3_sizes(const Plan* P, long long batch, ScratchSizes* z, size_t* part_bytes,
                          size_t* fh32_bytes) {
  using namespace f3;
  const size_t b64 = (size_t)NPTS * sizeof(cx<double>);
  const int kp = P->kahan ? 2 : 1;
  const int parts = P->nr * P->splits();  // partial planes per cell: one per (j, angle range)
  *part_bytes = (size_t)parts * kp * NPL * PLANE * sizeof(cx<float>);
  *fh32_bytes = (size_t)17 * PLANE * sizeof(cx<float>) + (size_t)NPTS * sizeof(float);
  const size_t aux_cell = *part_bytes + *fh32_bytes + nyq_scratch_bytes(parts);
  const long long sub = fused_sub(P, batch, 256, 4 * b64 + aux_cell);
  z->fhat = sub * b64;
  z->acc = sub * b64;
  z->work = 2 * sub * b64;
  z->aux = sub * aux_cell;
  return sub;
}

int fused_sizes(const Plan* P, long long batch, ScratchSizes* z) {
  const int k = fused_kind(P);
  if (k == 2) return fused2_sizes(P, batch, z);
  if (k == 0) return 0;
  size_t pb, fb;
  f3_sizes(P, batch, z, &pb, &fb);
  return 1;
}

int collide_fused(Plan* P, const double* f, double* q, long long batch, const int* pidx,
                  int with_loss, double hb, hipStream_t st, int* handled) {
  using namespace f3;
  *handled = 0;
  const int kind = fused_kind(P);
  if (kind == 2) return collide_fused2(P, f, q, batch, pidx, with_loss, hb, st, handled);
  if (kind == 0) return IB_OK;
  *handled = 1;
  ScratchSizes z;
  size_t part_bytes, fh32_bytes;
  const long long sub = f3_sizes(P, batch, &z, &part_bytes, &fh32_bytes);
  const int kp = P->kahan ? 2 : 1;
  int rc;
  if ((rc = P->reserve(z))) return rc;
  cx<float>* part = (cx<float>*)P->aux.ptr;
  cx<float>* fh32 = (cx<float>*)((char*)P->aux.ptr + sub * part_bytes);
  float* f32 = (float*)((char*)fh32 + sub * 17 * PLANE * sizeof(cx<float>));
  cx<float>* xp = (cx<float>*)(f32 + sub * NPTS);          // sub x 3 planes
  cx<float>* nyq_out = xp + sub * 3 * PLANE;               // sub x nr x splits x 3 planes
  for (long long c0 = 0; c0 < batch; c0 += sub) {
    const long long nb = std::min(sub, batch - c0);
    const double* fc = f + c0 * NPTS;
    if ((rc = staged_prologue_fp64(P, fc, nb, P->fhat.ptr, P->work.ptr))) return rc;
    const long long tot = nb * NPTS;
    hipLaunchKernelGGL(k_f3_pack, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                       (const cx<double>*)P->fhat.ptr, fc, fh32, f32, xp, nb);
    IB_HIP(hipGetLastError());
    F3Args A{};
    A.fh = fh32;
    A.f32 = f32;
    A.part = part;
    A.pidx = pidx;
    A.batch = nb;
    A.cell0 = 0;
    A.sigma = P->d_sigma;
    A.rj = P->d_rj;
    A.wsig = P->d_wsig;
    A.crad = P->d_crad;
    A.a = P->d_a;
    A.spg = (const float*)P->d_spg;
    A.nr = P->nr;
    A.m = P->m;
    A.nq = P->nq;
    A.ns = P->splits();
    A.xg = nb >= 8;
    A.L = P->L;
    // pidx is indexed from the sub-batch start
    A.pidx = pidx ? pidx + c0 : nullptr;
    if ((rc = P->timed_begin(st))) return rc;
    const unsigned grid = (unsigned)cell_grid(nb, P->nr * A.ns, A.xg);
    if (P->kahan)
      hipLaunchKernelGGL(k_fused3<true>, dim3(grid), dim3(THREADS), 0, st, A);
    else
      hipLaunchKernelGGL(k_fused3<false>, dim3(grid), dim3(THREADS), 0, st, A);
    IB_HIP(hipGetLastError());
    P->timed_end(st, nb);
    const long long half = nb * NPL * PLANE;
    hipLaunchKernelGGL(k_f3_reduce, dim3((unsigned)((half + 255) / 256)), dim3(256), 0, st, part,
                       (cx<double>*)P->acc.ptr, nb, P->nr * A.ns, kp);
    IB_HIP(hipGetLastError());
    if ((rc = launch_f3_nyq(P, xp, f32, nyq_out, (cx<double>*)P->acc.ptr, nb, A.pidx, st))) return rc;
    if ((rc = staged_epilogue_fp64(P, fc, q + c0 * NPTS, nb, P->acc.ptr, P->fhat.ptr, P->work.ptr,
                                   with_loss, hb, pidx ? pidx + c0 : nullptr)))
      return rc;
  }
  return IB_OK;
}

}  // namespace ib
