// Fused FP32 C2C forms of Q(f, f): 2D grids (N = 16, 32) and the 3D 16^3 grid.
//
// Same operator as reference src/collision/inelastic_collisions.py:208-227 with
// C2C transforms (the reference's own semantics: the complex work array of a
// quadrature pair fits in LDS, so no real-data reduction is needed, and grids
// this coarse can carry Nyquist content the real-data form would symmetrise).
//
// 2D: a workgroup serves one cell and a strided subset of its (j, m)
// quadrature pairs with G = THREADS / N concurrent pairs, one N-thread group
// per pair:
//   A  row pencils: E * fhat, inverse along axis 1
//   B  column pencils: inverse along axis 0, * f, forward along axis 0
//   C  row pencils: forward along axis 1, * G, sum over the group's pairs in
//      registers
// At 16^2 fhat / n and f of the cell are staged in LDS once per workgroup; at
// 32^2 each group reads its row of fhat / n and its column of f from L2.  The
// groups' sums are added in LDS, each workgroup writes an FP32 partial.
//
// 3D 16^3: one 256-thread workgroup per (cell, radial node j) loops over the
// M angles with the 16^3 complex work array in LDS; a thread owns one pencil
// in each of the five passes (E * fhat and inverse z; inverse y; inverse x,
// * f, forward x; forward y; forward z, * G and the sum over angles).  The
// thread's fhat pencil and f pencil stay in registers for the whole loop, and
// sp(a_j r_j |k|) and the radial weight are applied once at the end.
//
// In both forms the FP64 sum over workgroups (radial nodes), the final inverse
// transform, the loss term and the Laplacian run in FP64 (collide_plain.hip),
// as does the prologue fhat = fft(f).
#include <algorithm>

#include "common.h"
#include "edge2_body.h"
#include "group.h"
#include "plan.h"

namespace ib {

int staged_prologue_fp64(Plan* P, const double* f, long long nb, void* fhat64, void* work);
int staged_epilogue_fp64(Plan* P, const double* f, double* q, long long nb, void* acc64,
                         void* fhat64, void* work, int with_loss, double hb, const int* pidx);

template <int N>
__global__ __launch_bounds__(F2Shape<N>::THREADS, 2) void k_fused2(F2Args A) {
  __shared__ __align__(16) typename F2L<N>::Wg Wg;
  __shared__ typename F2L<N>::Tab tab;
  __shared__ typename F2L<N>::Fhs fhs;
  __shared__ typename F2L<N>::Fs fs;
  const long long cell = blockIdx.x / A.wpc;
  if (cell >= A.batch) return;  // uniform over the workgroup
  fused2_body<N>(A, cell, (int)(blockIdx.x % A.wpc), threadIdx.x, Wg, tab, fhs, fs);
}

// Nyquist-line term of the real-data 2D 32^2 form.  The anti-Hermitian parts of E and G
// live on the lines kx = 16 and ky = 16: E_A(16, ky) = i s_x h_y(ky) (ky != 16),
// E_A(kx, 16) = i s_y h_x(kx) (kx != 16), E_A(16, 16) = i Im(e_x(16) e_y(16)), with
// s_d = Im e_d(16) and h_d = e_d with its Nyquist entry replaced by the real part.  Hence
// ifft(E_A fhat) = i beta, beta(x, y) = s_x (-1)^x P1(y) + s_y (-1)^y P2(x), where
// P1 + i P2 = ifft(h_y fhat(16, .) + i h_x fhat(., 16)) (both real).  fft(f i beta) on the
// line kx = 16 is i fft(R1), R1(y) = sum_x (-1)^x f beta, and on ky = 16 i fft(R2),
// R2(x) = sum_y (-1)^y f beta; the sums split into the angle-independent row and column
// sums of f and two 32 x 32 products with (-1)^x f, (-1)^y f.  The term G_A * i fft(.) is
// -w sp(|k|) Im(...) * fft(.) on the lines; it is Hermitian, and Re ifft keeps
// G_H fft(f ifft(E_H fhat)) + G_A fft(f ifft(E_A fhat)), the reference's C2C result
// (inelastic_collisions.py:210-227).  One 64-lane workgroup per (cell, w) takes the chunks
// of 64 quadrature pairs q = w, w + wpc, ..., one pair per lane, sums its lanes' line terms
// in lane order and adds them to partial (cell, w) (written by k_fused2 before); a cell's
// result does not depend on the batch.
// The factors exp(i pi t kf) of one axis at kf = 0 .. 15: kf = 1 and 8 directly, the others
// as products (at most 7 steps from a direct value); kf = -k is the conjugate.
struct NyqAxis {
  cx<float> c1, c8;
  __device__ __forceinline__ void init(double t) {
    c1 = cispi_f32(t);
    c8 = cispi_f32(8.0 * t);
  }
};

// The value of the other lane of a lane pair (lane ^ 1, DPP quad_perm [1, 0, 3, 2]).
__device__ __forceinline__ float xch1(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ cx<float> xch1(cx<float> v) { return cx<float>{xch1(v.re), xch1(v.im)}; }

__global__ __launch_bounds__(128) void k_f2_nyq(F2Args A) {
  constexpr int N = 32, H = 16, NL = N + H;  // line points: kx = 16 (all ky), ky = 16 (kx < 16)
  __shared__ __align__(16) float fs[N][N];
  __shared__ float fsum[2][N];  // [0][y] = sum_x f, [1][x] = sum_y f
  __shared__ cx<float> lin[2][N];  // fhat / n on kx = 16 ([ky]) and on ky = 16 ([kx])
  __shared__ cx<float> red[NL][65];
  const long long cell = blockIdx.x / A.wpc;
  const int w = (int)(blockIdx.x % A.wpc);
  if (cell >= A.batch || w * 64 >= A.nr * A.m) return;  // uniform: no chunk for this partial
  const int tid = threadIdx.x, slot = tid >> 1, h = tid & 1;  // pair slot, lane of the pair
  const cx<float>* fhc = A.fh + cell * N * N;
  const float* fc = A.f32 + cell * N * N;
  for (int e = tid; e < N * N; e += 128) fs[e / N][e % N] = fc[e];
  if (tid < 64) lin[tid >> 5][tid & 31] = tid < 32 ? fhc[H * N + tid] : fhc[(tid - 32) * N + H];
  __syncthreads();
  if (tid < 64) {
    float s = 0.f;
    const int q = tid & 31;
#pragma unroll 8
    for (int i = 0; i < N; ++i) s += tid < 32 ? fs[i][q] : fs[q][i];
    fsum[tid >> 5][q] = s;
  }
  __syncthreads();
  const int ps = A.pidx ? A.pidx[cell] : 0;
  const int npairs = A.nr * A.m;
  const float sg = h ? -1.f : 1.f;  // (-1)^n of this lane's indices n = 2k + h
  // the lane's line terms summed over its pairs in registers: ky = 8h + m (and, lane 1,
  // ky = 16) on kx = 16, kx = 2k + h on ky = 16
  cx<float> lrow[H / 2], lcol[H / 2];
  float lnyq = 0.f;
#pragma unroll
  for (int m = 0; m < H / 2; ++m) lrow[m] = cx<float>{0.f, 0.f};
#pragma unroll
  for (int k = 0; k < H / 2; ++k) lcol[k] = cx<float>{0.f, 0.f};
  // chunks of 64 pairs, one pair per lane pair: chunk q goes to partial q mod wpc (in chunk order)
  for (int q = w; q * 64 < npairs; q += A.wpc) {
    const int pair = q * 64 + slot;
    if (pair >= npairs) continue;  // uniform over the lane pair
    const int j = pair / A.m, mm = pair % A.m;
    // LDS reads through an opaque offset, so that the compiler does not hoist the
    // loop-invariant reads of f out of the pair loop
    int zo = 0;
    asm volatile("" : "+v"(zo));
    const float* fsz = &fs[0][0] + zo;
    const float* fsumz = &fsum[0][0] + zo;
    const cx<float>* linz = &lin[0][0] + zo;
    const double rl = A.rj[j] / A.L, ar = A.a[ps * A.nr + j] * A.rj[j] / M_PI;
    const double sgx = A.sigma[mm * 2], sgy = A.sigma[mm * 2 + 1];
    // P1 + i P2 = ifft(h_y fhat(16, .) + i h_x fhat(., 16)); E = exp(-i pi r sigma kf / L)
    NyqAxis ex, ey;
    ex.init(-rl * sgx);
    ey.init(-rl * sgy);
    const cx<float> ex16 = cispi_f32(16.0 * rl * sgx), ey16 = cispi_f32(16.0 * rl * sgy);  // kf = -16
    cx<float> u[N];
    {
      cx<float> hx{1.f, 0.f}, hy{1.f, 0.f};
#pragma unroll
      for (int k = 0; k <= H; ++k) {
        if (k == 8) {
          hx = ex.c8;
          hy = ey.c8;
        } else if (k == H) {
          hx = cx<float>{ex16.re, 0.f};
          hy = cx<float>{ey16.re, 0.f};
        } else if (k > 0) {
          hx = hx * ex.c1;
          hy = hy * ey.c1;
        }
        const cx<float> a = hy * linz[k], b = hx * linz[N + k];
        u[k] = cx<float>{a.re - b.im, a.im + b.re};
        if (k > 0 && k < H) {
          const cx<float> am = conj(hy) * linz[N - k], bm = conj(hx) * linz[2 * N - k];
          u[N - k] = cx<float>{am.re - bm.im, am.im + bm.re};
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    // inverse transform split over the lane pair: lane h gets the outputs 2k + h
    // (r[n] = (u[n] + (-1)^h u[n + 16]) w32^{n h}, then 16 points)
    cx<float> r[H];
#pragma unroll
    for (int n = 0; n < H; ++n) {
      const cx<float> d{u[n].re + sg * u[n + H].re, u[n].im + sg * u[n + H].im};
      const cx<float> tw = h ? cx<float>{(float)TW64[2 * n][0], (float)TW64[2 * n][1]} : cx<float>{1.f, 0.f};
      r[n] = d * tw;
    }
    fft_reg<H, +1>(r);  // r[k].re = P1(2k + h), r[k].im = P2(2k + h)
    float p1[N];        // P1 at every y (the pair's two lanes' values)
#pragma unroll
    for (int k = 0; k < H; ++k) {
      const float o = xch1(r[k].re);
      p1[2 * k] = h ? o : r[k].re;
      p1[2 * k + 1] = h ? r[k].re : o;
    }
    // the lane's rows x = 2k + h: sum_x (-1)^x f P2 (partial, all y) and sum_y (-1)^y f P1
    float t1[N], t2[H];
#pragma unroll
    for (int i = 0; i < N; ++i) t1[i] = 0.f;
#pragma unroll
    for (int k = 0; k < H; ++k) {
      t2[k] = 0.f;
      const float* row = fsz + (2 * k + h) * N;
      const float p2 = r[k].im;
#pragma unroll
      for (int y0 = 0; y0 < N; y0 += 4) {
        const float4 qv = *reinterpret_cast<const float4*>(row + y0);
        const float fv[4] = {qv.x, qv.y, qv.z, qv.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int y = y0 + e;
          t1[y] = fmaf(fv[e], p2, t1[y]);
          t2[k] = fmaf((y & 1) ? -fv[e] : fv[e], p1[y], t2[k]);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    const float sx = ex16.im, sy = ey16.im;
    cx<float> z[H];
#pragma unroll
    for (int k = 0; k < H; ++k) {
      // sum_x (-1)^x f P2 at y = 2k + h: this lane's rows (sign (-1)^h) plus the other lane's
      const float mine = sg * (h ? t1[2 * k + 1] : t1[2 * k]);
      const float send = sg * (h ? t1[2 * k] : t1[2 * k + 1]);
      const float T1 = mine + xch1(send);
      const int n = 2 * k + h;
      const float r1 = fmaf(sx * r[k].re, fsumz[n], sy * sg * T1);
      const float r2 = fmaf(sy * r[k].im, fsumz[N + n], sx * sg * t2[k]);
      z[k] = cx<float>{r1, r2};
    }
    fft_reg<H, -1>(z);  // Y_h(k') of this lane's samples n = 2k + h
    // X(k') = Y_0 + w32^{-k'} Y_1, X(k' + 16) = Y_0 - w32^{-k'} Y_1: every lane forms all 32
    cx<float> X[N];
#pragma unroll
    for (int k = 0; k < H; ++k) {
      const cx<float> o = xch1(z[k]);
      const cx<float> y0 = h ? o : z[k], y1 = h ? z[k] : o;
      const cx<float> t = twiddle<-1>(y1, 2 * k);
      X[k] = y0 + t;
      X[k + H] = y0 - t;
    }
    // line weights: -w sp(|k|) Im(.) of the anti-Hermitian part of G = exp(i a r sigma kf);
    // lane 0 takes ky = 0 .. 15 and lane 1 ky = 16 .. 31 on kx = 16, lane h the kx = 2k + h
    // on ky = 16 (each candidate formed with fixed registers, the lane's chosen by select)
    const float wj = (float)(A.crad[ps * A.nr + j] * A.wsig[mm]);
    const float* sp = A.spg + ((long long)ps * A.nr + j) * A.nq;
    NyqAxis gx, gy;
    gx.init(ar * sgx);
    gy.init(ar * sgy);
    const cx<float> gx16 = cispi_f32(-16.0 * ar * sgx), gy16 = cispi_f32(-16.0 * ar * sgy);
    // kx = 16: the term at ky and at -ky are conjugates (R1 is real), so lane h forms
    // ky = 8h + m (m = 0 .. 7), lane 1 also ky = 16; the mirrors are written at the end
    cx<float> hgx{1.f, 0.f}, hgx1{1.f, 0.f};
    cx<float> hgy = h ? gy.c8 : cx<float>{1.f, 0.f};
#pragma unroll
    for (int m = 0; m < H / 2; ++m) {
      if (m > 0) hgy = hgy * gy.c1;  // h_y(8h + m)
      // both candidates through opaque copies: a select between two array elements would
      // otherwise become a lane-dependent register index (the array moved to scratch)
      cx<float> z1 = X[m + 8], z0 = X[m], m1 = X[N - 8 - m], m0 = X[(N - m) & (N - 1)];
      asm volatile("" : "+v"(z1.re), "+v"(z1.im), "+v"(z0.re), "+v"(z0.im));
      asm volatile("" : "+v"(m1.re), "+v"(m1.im), "+v"(m0.re), "+v"(m0.im));
      const cx<float> Z = h ? z1 : z0, Zm = h ? m1 : m0;
      const cx<float> F1{0.5f * (Z.re + Zm.re), 0.5f * (Z.im - Zm.im)};  // (X + conj X(-ky)) / 2
      const int ky = 8 * h + m;
      lrow[m] = lrow[m] + (-wj * sp[H * H + ky * ky]) * ((gx16.im * hgy) * F1);
      __builtin_amdgcn_sched_barrier(0);
    }
    lnyq = lnyq + (-wj * sp[2 * H * H] * (gx16 * gy16).im * X[H].re);  // ky = 16 (used by lane 1)
#pragma unroll
    for (int k = 0; k < H / 2; ++k) {
      // ky = 16, kx = 2k + h: F2 = (X - conj X(-kx)) / (2 i), weight Im(g_y(16)) h_x(kx)
      if (k > 0) hgx = (k == 4) ? gx.c8 : hgx * gx.c1 * gx.c1;  // h_x(2k)
      hgx1 = (2 * k + 1 == 8) ? gx.c8 : hgx * gx.c1;           // h_x(2k + 1)
      const int k0 = 2 * k, k1 = 2 * k + 1;
      const cx<float> Z0 = X[k0], Z0m = X[(N - k0) & (N - 1)], Z1 = X[k1], Z1m = X[N - k1];
      const cx<float> F0{0.5f * (Z0.im + Z0m.im), 0.5f * (Z0m.re - Z0.re)};
      const cx<float> F1{0.5f * (Z1.im + Z1m.im), 0.5f * (Z1m.re - Z1.re)};
      const cx<float> v2 = h ? (gy16.im * hgx1) * F1 : (gy16.im * hgx) * F0;
      const int kx = 2 * k + h;
      lcol[k] = lcol[k] + (-wj * sp[kx * kx + H * H]) * v2;
      __builtin_amdgcn_sched_barrier(0);
    }
  }
#pragma unroll
  for (int m = 0; m < H / 2; ++m) {
    const int ky = 8 * h + m;
    red[ky][slot] = lrow[m];
    if (ky > 0) red[N - ky][slot] = conj(lrow[m]);
  }
  if (h) red[H][slot] = cx<float>{lnyq, 0.f};
#pragma unroll
  for (int k = 0; k < H / 2; ++k) red[N + 2 * k + h][slot] = lcol[k];
  __syncthreads();
  if (tid < NL) {
    cx<float> s{0.f, 0.f};
    for (int l = 0; l < 64; ++l) s = s + red[tid][l];
    cx<float>* out = A.part + (cell * A.wpc + w) * N * N;
    if (tid < N) {
      out[H * N + tid] = out[H * N + tid] + s;
    } else {
      const int kx = tid - N;
      out[kx * N + H] = out[kx * N + H] + s;
      if (kx > 0) out[(N - kx) * N + H] = out[(N - kx) * N + H] + conj(s);
    }
  }
}

// Phase tables of every quadrature pair (j, m), axis a and frequency index k:
// E = exp(-i pi r_j sigma_a kf / L) and, per parameter set, G = exp(i a_j r_j sigma_a kf)
// (the separable factors of inelastic_collisions.py:97-105): FP64 sincos for FP64 plans,
// phase2_f32 (the FP32 kernels' fallback expression) for FP32 plans.
template <typename T>
__global__ void k_phase_tables2(cx<T>* ephase, cx<T>* gphase, const double* rj, const double* sigma,
                                const double* a, int nr, int m, int n, int n_param, double L) {
  const long long per = (long long)nr * m * 2 * n;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = per * (1 + (gphase ? n_param : 0));
  if (gid >= total) return;
  const long long t = gid % per;
  const int k = (int)(t % n), ax = (int)((t / n) % 2);
  const long long pair = t / (2LL * n);
  const int j = (int)(pair / m), mm = (int)(pair % m);
  const double kf = (double)kfreq(k, n), s = sigma[mm * 2 + ax];
  const int ps = gid < per ? 0 : (int)(gid / per - 1);
  cx<T> v;
  if constexpr (std::is_same<T, float>::value) {  // the FP32 kernels' own expression
    v = phase2_f32(gid < per ? 0 : 1, rj[j], gid < per ? 0.0 : a[ps * nr + j], s, kf, L);
  } else {
    double sn, cs;
    sincos(gid < per ? -(M_PI * rj[j] * s * kf) / L : a[ps * nr + j] * rj[j] * s * kf, &sn, &cs);
    v = cx<T>{(T)cs, (T)sn};
  }
  if (gid < per)
    ephase[t] = v;
  else
    gphase[gid - per] = v;
}

int launch_phase_tables2(Plan* P, hipStream_t st) {
  const long long per = (long long)P->nr * P->m * 2 * P->n;
  const long long total = per * (1 + (P->d_gphase ? P->n_param : 0));
  const dim3 grid((unsigned)((total + 255) / 256));
  if (P->precision == IB_PREC_FP64)
    hipLaunchKernelGGL(k_phase_tables2<double>, grid, dim3(256), 0, st, (cx<double>*)P->d_ephase,
                       (cx<double>*)P->d_gphase, P->d_rj, P->d_sigma, P->d_a, P->nr, P->m, P->n,
                       P->n_param, P->L);
  else
    hipLaunchKernelGGL(k_phase_tables2<float>, grid, dim3(256), 0, st, (cx<float>*)P->d_ephase,
                       (cx<float>*)P->d_gphase, P->d_rj, P->d_sigma, P->d_a, P->nr, P->m, P->n,
                       P->n_param, P->L);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

// fhat / n and f in the fused kernels' type.
template <typename T>
__global__ void k_f2_pack(const cx<double>* __restrict__ fhat, const double* __restrict__ f,
                          cx<T>* __restrict__ fh32, T* __restrict__ f32, long long total,
                          double inv_n) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  f32[gid] = (T)f[gid];
  const cx<double> v = fhat[gid];
  fh32[gid] = cx<T>{(T)(v.re * inv_n), (T)(v.im * inv_n)};
}

// FP64 sum of each cell's wpc partial spectra, in partial order.
template <typename T>
__global__ void k_f2_reduce(const cx<T>* __restrict__ part, cx<double>* __restrict__ acc64,
                            long long nb, int wpc, long long npts) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nb * npts) return;
  const long long c = gid / npts, e = gid % npts;
  double re = 0.0, im = 0.0;
  for (int w = 0; w < wpc; ++w) {
    const cx<T> x = part[(c * wpc + w) * npts + e];
    re += (double)x.re;
    im += (double)x.im;
  }
  acc64[gid] = cx<double>{re, im};
}

__global__ __launch_bounds__(c3::THREADS) void k_fusedc3(C3Args A) {
  using namespace c3;
  __shared__ cx<float> W[N][N][PAD];
  __shared__ cx<float> tE[2][3][N];
  __shared__ cx<float> tG[2][3][N];
  __shared__ float sp[3 * (N / 2) * (N / 2) + 1];
  long long cell;
  int jj;
  cell_of_block(blockIdx.x, A.nr * A.ns, A.xg, cell, jj);
  if (cell >= A.batch) return;  // uniform over the workgroup
  fusedc3_body(A, cell, jj, threadIdx.x, W, tE, tG, sp);
}

// Scratch of the fused C2C forms per cell: FP64 f-hat, accumulator and two work arrays, then
// the per-workgroup partial sums (one complex array per radial node, or per pair group), the
// working-precision f-hat and f.
template <typename T>
static size_t c2c_cell_bytes(long long npts, long long parts) {
  const size_t b64 = (size_t)npts * sizeof(cx<double>);
  return 4 * b64 + (size_t)(parts + 1) * npts * sizeof(cx<T>) + (size_t)npts * sizeof(T);
}

template <typename T>
static void c2c_layout(long long sub, long long npts, long long parts, ScratchSizes* z) {
  const size_t b64 = (size_t)npts * sizeof(cx<double>);
  z->fhat = sub * b64;
  z->acc = sub * b64;
  z->work = 2 * sub * b64;
  z->aux = sub * (c2c_cell_bytes<T>(npts, parts) - 4 * b64);
}

template <typename T>
static long long c2c_sizes(const Plan* P, long long batch, long long dflt, long long npts, long long parts,
                           ScratchSizes* z) {
  const long long sub = fused_sub(P, batch, dflt, c2c_cell_bytes<T>(npts, parts));
  c2c_layout<T>(sub, npts, parts, z);
  return sub;
}

static long long fc3_sizes(const Plan* P, long long batch, ScratchSizes* z) {
  return c2c_sizes<float>(P, batch, 1024, c3::NPTS, (long long)P->nr * P->splits(), z);
}

// workgroups per cell of k_fused2: the plan's rounds of G pairs each (by default 16 for the
// C2C 16^2 form, 8 for the real-data 32^2 form, whose rounds hold twice the pairs)
template <int N>
static int f2_wpc(const Plan* P) {
  const int per = (P->pair_rounds > 0 ? P->pair_rounds : (F2Shape<N>::REAL ? 8 : 16)) * F2Shape<N>::G;
  return std::max(1, (P->nr * P->m + per - 1) / per);
}

// aux pieces of a C2C sub-batch of `sub` cells: the partial spectra (parts per cell), then
// fhat / n and f in the kernel's type
template <typename T>
struct C2CPtrs {
  cx<T>* part;
  cx<T>* fh;
  T* f;
};

template <typename T>
static C2CPtrs<T> c2c_ptrs(const Plan* P, long long sub, long long npts, long long parts) {
  C2CPtrs<T> r;
  r.part = (cx<T>*)P->aux.ptr;
  r.fh = (cx<T>*)((char*)P->aux.ptr + (size_t)sub * parts * npts * sizeof(cx<T>));
  r.f = (T*)((char*)r.fh + (size_t)sub * npts * sizeof(cx<T>));
  return r;
}

// 2D 16^2 / 32^2 forms: prologue and epilogue in one launch each (edge2_body.h)
template <int N>
static constexpr bool edges_fused() {
  return N <= 32;
}

template <typename T, int N>
static int front2d(Plan* P, const double* fc, long long nb, const C2CPtrs<T>& b, hipStream_t st) {
  constexpr int CPW = 64 / N;
  hipLaunchKernelGGL((k_front2d<T, N>), dim3((unsigned)((nb + CPW - 1) / CPW)), dim3(64), 0, st, fc,
                     (cx<double>*)P->fhat.ptr, b.fh, b.f, nb);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

template <typename T, int N>
static int back2d(Plan* P, const double* fc, double* qc, long long nb, long long parts, const C2CPtrs<T>& b,
                  const int* pc, int with_loss, double hb, hipStream_t st) {
  BackArgs K{};
  K.f = fc;
  K.fhat = (const cx<double>*)P->fhat.ptr;
  K.q = qc;
  K.lam = P->d_lam64;
  K.pidx = pc;
  K.nq = P->nq;
  K.with_loss = with_loss;
  K.hb = hb;
  K.L = P->L;
  hipLaunchKernelGGL((k_back2d<T, N>), dim3((unsigned)nb), dim3(kBackThreads), 0, st, K, b.part, (int)parts);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

// prologue fhat = fft(f) in FP64 and the packing into the kernel's type
template <typename T>
static int c2c_front(Plan* P, const double* fc, long long nb, long long npts, const C2CPtrs<T>& b,
                     hipStream_t st) {
  int rc;
  if ((rc = staged_prologue_fp64(P, fc, nb, P->fhat.ptr, P->work.ptr))) return rc;
  const long long tot = nb * npts;
  hipLaunchKernelGGL(k_f2_pack<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                     (const cx<double>*)P->fhat.ptr, fc, b.fh, b.f, tot, 1.0 / (double)npts);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

// after the pair loop: FP64 sum of each cell's partials in partial order, the FP64 epilogue
template <typename T>
static int c2c_back(Plan* P, const double* fc, double* qc, long long nb, long long npts, long long parts,
                    const C2CPtrs<T>& b, const int* pc, int with_loss, double hb, hipStream_t st) {
  const long long tot = nb * npts;
  hipLaunchKernelGGL(k_f2_reduce<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, b.part,
                     (cx<double>*)P->acc.ptr, nb, (int)parts, npts);
  IB_HIP(hipGetLastError());
  return staged_epilogue_fp64(P, fc, qc, nb, P->acc.ptr, P->fhat.ptr, P->work.ptr, with_loss, hb, pc);
}

// plan fields every C2C argument struct carries
template <typename Args>
static void c2c_common(const Plan* P, Args& A, long long nb, const int* pc) {
  A.pidx = pc;
  A.batch = nb;
  A.sigma = P->d_sigma;
  A.rj = P->d_rj;
  A.wsig = P->d_wsig;
  A.crad = P->d_crad;
  A.a = P->d_a;
  A.nr = P->nr;
  A.m = P->m;
  A.nq = P->nq;
  A.L = P->L;
}

static C3Args c3_args(const Plan* P, const C2CPtrs<float>& b, long long nb, const int* pc) {
  C3Args A{};
  c2c_common(P, A, nb, pc);
  A.fh = b.fh;
  A.f32 = b.f;
  A.part = b.part;
  A.spg = (const float*)P->d_spg;
  A.ns = P->splits();
  A.xg = nb >= 8;
  return A;
}

template <typename T>
static CWArgs<T> cw_args(const Plan* P, const C2CPtrs<T>& b, long long nb, const int* pc) {
  CWArgs<T> A{};
  c2c_common(P, A, nb, pc);
  A.fh = b.fh;
  A.f = b.f;
  A.ephase = (const cx<T>*)P->d_ephase;
  A.gphase = (const cx<T>*)P->d_gphase;
  A.part = b.part;
  A.spg = (const T*)P->d_spg;
  A.ns = P->splits();
  // the one-range form keeps the XCD-grouped numbering (without a split, each of a small
  // batch's workgroups has a CU to itself either way)
  A.xg = A.ns > 1 ? nb >= 8 : 1;
  return A;
}

template <int N>
static F2Args f2_args(const Plan* P, const C2CPtrs<float>& b, long long nb, const int* pc) {
  F2Args A{};
  c2c_common(P, A, nb, pc);
  A.fh = b.fh;
  A.ephase = (const cx<float>*)P->d_ephase;
  A.gphase = (const cx<float>*)P->d_gphase;
  A.f32 = b.f;
  A.part = b.part;
  A.wpc = f2_wpc<N>(P);
  A.spg = (const float*)P->d_spg;
  return A;
}

static int run_fusedc3(Plan* P, const double* f, double* q, long long batch, const int* pidx,
                       int with_loss, double hb, hipStream_t st) {
  using namespace c3;
  ScratchSizes z;
  const long long sub = fc3_sizes(P, batch, &z);
  int rc;
  if ((rc = P->reserve(z))) return rc;
  const long long parts = (long long)P->nr * P->splits();
  const C2CPtrs<float> b = c2c_ptrs<float>(P, sub, NPTS, parts);
  for (long long c0 = 0; c0 < batch; c0 += sub) {
    const long long nb = std::min(sub, batch - c0);
    const double* fc = f + c0 * NPTS;
    const int* pc = pidx ? pidx + c0 : nullptr;
    if ((rc = c2c_front<float>(P, fc, nb, NPTS, b, st))) return rc;
    const C3Args A = c3_args(P, b, nb, pc);
    if ((rc = P->timed_begin(st))) return rc;
    hipLaunchKernelGGL(k_fusedc3, dim3((unsigned)cell_grid(nb, (int)parts, A.xg)), dim3(THREADS), 0, st, A);
    IB_HIP(hipGetLastError());
    P->timed_end(st, nb);
    if ((rc = c2c_back<float>(P, fc, q + c0 * NPTS, nb, NPTS, parts, b, pc, with_loss, hb, st))) return rc;
  }
  return IB_OK;
}

template <typename T, int N, bool SPLIT>
__global__ __launch_bounds__(2 * N) void k_fusedw(CWArgs<T> A) {
  __shared__ cx<T> W[N * N];
  __shared__ cx<T> tE[2][2][N];
  __shared__ cx<T> tG[2][2][N];
  __shared__ cx<T> twd[2][N / 2 + 1];
  long long cell;
  int jj;
  cell_of_block(blockIdx.x, A.nr * (SPLIT ? A.ns : 1), SPLIT ? A.xg : 1, cell, jj);
  if (cell >= A.batch) return;  // uniform over the workgroup
  fusedw_body<T, N, SPLIT>(A, cell, jj, threadIdx.x, W, tE, tG, twd);
}

template <typename T, int N>
static int run_fusedw(Plan* P, const double* f, double* q, long long batch, const int* pidx,
                      int with_loss, double hb, hipStream_t st) {
  constexpr long long NPTS = (long long)N * N;
  const int ns = P->splits();
  const long long parts = (long long)P->nr * ns;
  ScratchSizes z;
  const long long sub = c2c_sizes<T>(P, batch, 1024, NPTS, parts, &z);
  int rc;
  if ((rc = P->reserve(z))) return rc;
  const C2CPtrs<T> b = c2c_ptrs<T>(P, sub, NPTS, parts);
  for (long long c0 = 0; c0 < batch; c0 += sub) {
    const long long nb = std::min(sub, batch - c0);
    const double* fc = f + c0 * NPTS;
    const int* pc = pidx ? pidx + c0 : nullptr;
    constexpr bool E = edges_fused<N>();
    if constexpr (E)
      rc = front2d<T, N>(P, fc, nb, b, st);
    else
      rc = c2c_front<T>(P, fc, nb, NPTS, b, st);
    if (rc) return rc;
    const CWArgs<T> A = cw_args<T>(P, b, nb, pc);
    if ((rc = P->timed_begin(st))) return rc;
    const unsigned grid = (unsigned)cell_grid(nb, (int)parts, A.xg);
    if (ns > 1)
      hipLaunchKernelGGL((k_fusedw<T, N, true>), dim3(grid), dim3(2 * N), 0, st, A);
    else
      hipLaunchKernelGGL((k_fusedw<T, N, false>), dim3(grid), dim3(2 * N), 0, st, A);
    IB_HIP(hipGetLastError());
    P->timed_end(st, nb);
    if constexpr (E)
      rc = back2d<T, N>(P, fc, q + c0 * NPTS, nb, parts, b, pc, with_loss, hb, st);
    else
      rc = c2c_back<T>(P, fc, q + c0 * NPTS, nb, NPTS, parts, b, pc, with_loss, hb, st);
    if (rc) return rc;
  }
  return IB_OK;
}

template <int N>
static int run_fused2(Plan* P, const double* f, double* q, long long batch, const int* pidx,
                      int with_loss, double hb, hipStream_t st) {
  constexpr int THREADS = F2Shape<N>::THREADS;
  constexpr long long NPTS = (long long)N * N;
  // workgroups per cell fixed by the plan (its rounds of G pairs each), not by the batch
  // size: a cell's FP32 summation order, and so its result, is the same in any batch
  const int wpc = f2_wpc<N>(P);
  ScratchSizes z;
  const long long sub = c2c_sizes<float>(P, batch, 4096, NPTS, wpc, &z);
  int rc;
  if ((rc = P->reserve(z))) return rc;
  const C2CPtrs<float> b = c2c_ptrs<float>(P, sub, NPTS, wpc);
  for (long long c0 = 0; c0 < batch; c0 += sub) {
    const long long nb = std::min(sub, batch - c0);
    const double* fc = f + c0 * NPTS;
    const int* pc = pidx ? pidx + c0 : nullptr;
    if ((rc = front2d<float, N>(P, fc, nb, b, st))) return rc;
    const F2Args A = f2_args<N>(P, b, nb, pc);
    if ((rc = P->timed_begin(st))) return rc;
    hipLaunchKernelGGL(k_fused2<N>, dim3((unsigned)(nb * wpc)), dim3(THREADS), 0, st, A);
    IB_HIP(hipGetLastError());
    P->timed_end(st, nb);
    if constexpr (F2Shape<N>::REAL) {
      hipLaunchKernelGGL(k_f2_nyq, dim3((unsigned)(nb * wpc)), dim3(128), 0, st, A);
      IB_HIP(hipGetLastError());
    }
    if ((rc = back2d<float, N>(P, fc, q + c0 * NPTS, nb, wpc, b, pc, with_loss, hb, st))) return rc;
  }
  return IB_OK;
}

// Grouped launch (group.h), FP32 C2C forms: the shape's kind, points per cell and partial
// spectra per cell, or 0 when the plan has no such form
static int f2_group_kind(const Plan* P, long long* npts, long long* parts) {
  if (P->precision == IB_PREC_FP64) return -1;
  const int ns = P->splits();
  if (P->dim == 3 && P->n == 16) {
    *npts = c3::NPTS;
    *parts = (long long)P->nr * ns;
    return GK_C3;
  }
  if (P->dim != 2) return -1;
  switch (P->n) {
    case 16: *npts = 256; *parts = f2_wpc<16>(P); return GK_F2_16;
    case 32: *npts = 1024; *parts = f2_wpc<32>(P); return GK_F2_32;
    case 64: *npts = 4096; *parts = (long long)P->nr * ns; return ns > 1 ? GK_W64S : GK_W64;
    default: return -1;
  }
}

int f2_group_fits(const Plan* P, long long batch) {
  long long npts = 0, parts = 0;
  return f2_group_kind(P, &npts, &parts) >= 0 &&
         (size_t)batch * c2c_cell_bytes<float>(npts, parts) <= P->scratch_budget;
}

// The whole batch as one segment when its scratch fits the plan's budget (*ok = 0
// otherwise: the plan then runs alone).
int f2_group_front(Plan* P, const double* f, long long batch, const int* pidx, hipStream_t st,
                   GroupSeg* seg, int* ok) {
  *ok = 0;
  long long npts = 0, parts = 0;
  const int kind = f2_group_kind(P, &npts, &parts);
  if (kind < 0 || (size_t)batch * c2c_cell_bytes<float>(npts, parts) > P->scratch_budget) return IB_OK;
  ScratchSizes z;
  c2c_layout<float>(batch, npts, parts, &z);
  int rc;
  if ((rc = P->reserve(z))) return rc;
  const C2CPtrs<float> b = c2c_ptrs<float>(P, batch, npts, parts);
  if ((rc = c2c_front<float>(P, f, batch, npts, b, st))) return rc;
  seg->kind = kind;
  const int ns = P->splits();
  switch (kind) {
    case GK_C3:
      seg->c3 = c3_args(P, b, batch, pidx);
      seg->units = batch * P->nr;
      break;
    case GK_W64:
    case GK_W64S:
      seg->w = cw_args<float>(P, b, batch, pidx);
      seg->units = batch * P->nr;
      break;
    case GK_F2_16:
      seg->f2 = f2_args<16>(P, b, batch, pidx);
      seg->units = batch * seg->f2.wpc;
      break;
    default:
      seg->f2 = f2_args<32>(P, b, batch, pidx);
      seg->units = batch * seg->f2.wpc;
      break;
  }
  seg->blocks = group_blocks(kind, seg->units, ns, 0);
  *ok = 1;
  return IB_OK;
}

int f2_group_back(Plan* P, const double* f, double* q, long long batch, const int* pidx, int with_loss,
                  double hb, hipStream_t st) {
  long long npts = 0, parts = 0;
  if (f2_group_kind(P, &npts, &parts) < 0) {
    set_error("grouped fused launch: plan has no grouped form");
    return IB_ERR_ARG;
  }
  const C2CPtrs<float> b = c2c_ptrs<float>(P, batch, npts, parts);
  if (P->dim == 2 && P->n == 32) {  // the real-data form's Nyquist-line term
    const F2Args A = f2_args<32>(P, b, batch, pidx);
    hipLaunchKernelGGL(k_f2_nyq, dim3((unsigned)(batch * A.wpc)), dim3(128), 0, st, A);
    IB_HIP(hipGetLastError());
  }
  return c2c_back<float>(P, f, q, batch, npts, parts, b, pidx, with_loss, hb, st);
}

int fused2_sizes(const Plan* P, long long batch, ScratchSizes* z) {
  if (P->precision == IB_PREC_FP64) {
    if (P->dim == 2 && P->n == 32)
      return c2c_sizes<double>(P, batch, 1024, 32 * 32, (long long)P->nr * P->splits(), z), 1;
    return 0;
  }
  if (P->dim == 3) return P->n == 16 ? (fc3_sizes(P, batch, z), 1) : 0;
  switch (P->n) {
    case 16: return c2c_sizes<float>(P, batch, 4096, 16 * 16, f2_wpc<16>(P), z), 1;
    case 32: return c2c_sizes<float>(P, batch, 4096, 32 * 32, f2_wpc<32>(P), z), 1;
    case 64: return c2c_sizes<float>(P, batch, 1024, 64 * 64, (long long)P->nr * P->splits(), z), 1;
    default: return 0;
  }
}

int collide_fused2(Plan* P, const double* f, double* q, long long batch, const int* pidx,
                   int with_loss, double hb, hipStream_t st, int* handled) {
  *handled = 1;
  if (P->precision == IB_PREC_FP64) {  // FP64 fused form of 2D 32^2 (FP64 throughout)
    if (P->dim == 2 && P->n == 32) return run_fusedw<double, 32>(P, f, q, batch, pidx, with_loss, hb, st);
    *handled = 0;
    return IB_OK;
  }
  if (P->dim == 3) {
    if (P->n == 16) return run_fusedc3(P, f, q, batch, pidx, with_loss, hb, st);
    *handled = 0;
    return IB_OK;
  }
  switch (P->n) {
    case 16: return run_fused2<16>(P, f, q, batch, pidx, with_loss, hb, st);
    case 32: return run_fused2<32>(P, f, q, batch, pidx, with_loss, hb, st);
    case 64: return run_fusedw<float, 64>(P, f, q, batch, pidx, with_loss, hb, st);
    default:
      *handled = 0;
      return IB_OK;
  }
}

}  // namespace ib
