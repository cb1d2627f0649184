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
#include "plan.h"

namespace ib {

int staged_prologue_fp64(Plan* P, const double* f, long long nb, void* fhat64, void* work);
int staged_epilogue_fp64(Plan* P, const double* f, double* q, long long nb, void* acc64,
                         void* fhat64, void* work, int with_loss, double hb, const int* pidx);

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

template <int N>
struct F2Shape {
  static constexpr int THREADS = N >= 32 ? 128 : 256;
  static constexpr int G = THREADS / N;
};

template <int N>
__global__ __launch_bounds__(F2Shape<N>::THREADS) void k_fused2(F2Args A) {
  constexpr int THREADS = F2Shape<N>::THREADS, G = F2Shape<N>::G, PAD = N + 1;
  // 16^2 stages the cell's fhat / n and f in LDS; 32^2 reads them from L2, as
  // staging would cost the fourth workgroup per CU (8% slower measured; at 16^2
  // the LDS staging is 12% faster and does not change the occupancy)
  constexpr bool STAGE = N < 32;
  constexpr int SN = STAGE ? N : 1;
  __shared__ cx<float> Wg[G][N][PAD];
  __shared__ cx<float> tab[G][4][N];  // E axis 0, E axis 1, G axis 0, G axis 1
  __shared__ cx<float> fhs[SN][SN + 1];
  __shared__ float fs[SN][SN + 1];
  const int tid = threadIdx.x;
  const int g = tid / N, l = tid % N;
  const long long cell = blockIdx.x / A.wpc;
  const int w = (int)(blockIdx.x % A.wpc);
  if (cell >= A.batch) return;  // uniform over the workgroup
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
  const int rounds = npairs > w * G ? (npairs - w * G + stride - 1) / stride : 0;
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
        const double rj = A.rj[j];
#pragma unroll
        for (int ax = 0; ax < 2; ++ax) {
          double sn, cs;
          sincos(-(M_PI * rj * A.sigma[mm * 2 + ax] * (double)kl) / A.L, &sn, &cs);
          tab[g][ax][l] = cx<float>{(float)cs, (float)sn};
        }
      }
      if (A.gphase) {
        const cx<float>* gt = A.gphase + ((long long)ps * npairs + pair) * 2 * N;
        tab[g][2][l] = gt[l];
        tab[g][3][l] = gt[N + l];
      } else {
        const double aj = A.a[ps * A.nr + j] * A.rj[j];
#pragma unroll
        for (int ax = 0; ax < 2; ++ax) {
          double sn, cs;
          sincos(aj * A.sigma[mm * 2 + ax] * (double)kl, &sn, &cs);
          tab[g][2 + ax][l] = cx<float>{(float)cs, (float)sn};
        }
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

// Phase tables of every quadrature pair (j, m), axis a and frequency index k:
// E = exp(-i pi r_j sigma_a kf / L) and, per parameter set, G = exp(i a_j r_j sigma_a kf)
// (the separable factors of inelastic_collisions.py:97-105), FP64 sincos (rounded to
// FP32 for FP32 plans).
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
  double sn, cs;
  if (gid < per) {
    sincos(-(M_PI * rj[j] * s * kf) / L, &sn, &cs);
    ephase[t] = cx<T>{(T)cs, (T)sn};
  } else {
    const int ps = (int)(gid / per - 1);
    sincos(a[ps * nr + j] * rj[j] * s * kf, &sn, &cs);
    gphase[gid - per] = cx<T>{(T)cs, (T)sn};
  }
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
__global__ __launch_bounds__(c3::THREADS) void k_fusedc3(C3Args A) {
  using namespace c3;
  __shared__ cx<float> W[N][N][PAD];
  __shared__ cx<float> tE[2][3][N];
  __shared__ cx<float> tG[2][3][N];
  __shared__ float sp[3 * (N / 2) * (N / 2) + 1];
  const int tid = threadIdx.x;
  const long long id = blockIdx.x;
  const int nrs = A.nr * A.ns;
  long long cell;
  int jj;  // partial index: radial node j, angle range asp
  cell_of_block(id, nrs, A.xg, cell, jj);
  const int j = jj / A.ns, asp = jj % A.ns;
  if (cell >= A.batch) return;  // uniform over the workgroup
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

// Scratch of the fused C2C forms per cell: FP64 f-hat, accumulator and two work arrays, then
// the per-workgroup partial sums (one complex array per radial node, or per pair group), the
// working-precision f-hat and f.
template <typename T>
static long long c2c_sizes(const Plan* P, long long batch, long long dflt, long long npts, long long parts,
                           ScratchSizes* z) {
  const size_t b64 = (size_t)npts * sizeof(cx<double>);
  const size_t aux_cell = (size_t)(parts + 1) * npts * sizeof(cx<T>) + (size_t)npts * sizeof(T);
  const long long sub = fused_sub(P, batch, dflt, 4 * b64 + aux_cell);
  z->fhat = sub * b64;
  z->acc = sub * b64;
  z->work = 2 * sub * b64;
  z->aux = sub * aux_cell;
  return sub;
}

static long long fc3_sizes(const Plan* P, long long batch, ScratchSizes* z) {
  return c2c_sizes<float>(P, batch, 1024, c3::NPTS, (long long)P->nr * P->splits(), z);
}

// workgroups per cell of k_fused2: the plan's rounds of G pairs each (16 by default)
template <int N>
static int f2_wpc(const Plan* P) {
  const int per = (P->pair_rounds > 0 ? P->pair_rounds : 16) * F2Shape<N>::G;
  return std::max(1, (P->nr * P->m + per - 1) / per);
}

static int run_fusedc3(Plan* P, const double* f, double* q, long long batch, const int* pidx,
                       int with_loss, double hb, hipStream_t st) {
  using namespace c3;
  ScratchSizes z;
  const long long sub = fc3_sizes(P, batch, &z);
  int rc;
  if ((rc = P->reserve(z))) return rc;
  const int ns = P->splits();
  const size_t part_b = (size_t)sub * P->nr * ns * NPTS * sizeof(cx<float>);
  const size_t fh_b = (size_t)sub * NPTS * sizeof(cx<float>);
  cx<float>* part = (cx<float>*)P->aux.ptr;
  cx<float>* fh32 = (cx<float>*)((char*)P->aux.ptr + part_b);
  float* f32 = (float*)((char*)fh32 + fh_b);
  for (long long c0 = 0; c0 < batch; c0 += sub) {
    const long long nb = std::min(sub, batch - c0);
    const double* fc = f + c0 * NPTS;
    const int* pc = pidx ? pidx + c0 : nullptr;
    if ((rc = staged_prologue_fp64(P, fc, nb, P->fhat.ptr, P->work.ptr))) return rc;
    const long long tot = nb * NPTS;
    hipLaunchKernelGGL(k_f2_pack<float>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                       (const cx<double>*)P->fhat.ptr, fc, fh32, f32, tot, 1.0 / (double)NPTS);
    IB_HIP(hipGetLastError());
    C3Args A{};
    A.fh = fh32;
    A.f32 = f32;
    A.part = part;
    A.pidx = pc;
    A.batch = nb;
    A.sigma = P->d_sigma;
    A.rj = P->d_rj;
    A.wsig = P->d_wsig;
    A.crad = P->d_crad;
    A.a = P->d_a;
    A.spg = (const float*)P->d_spg;
    A.nr = P->nr;
    A.m = P->m;
    A.nq = P->nq;
    A.ns = ns;
    A.xg = nb >= 8;
    A.L = P->L;
    if ((rc = P->timed_begin(st))) return rc;
    hipLaunchKernelGGL(k_fusedc3, dim3((unsigned)cell_grid(nb, P->nr * ns, A.xg)), dim3(THREADS), 0, st, A);
    IB_HIP(hipGetLastError());
    P->timed_end(st, nb);
    hipLaunchKernelGGL(k_f2_reduce<float>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, part,
                       (cx<double>*)P->acc.ptr, nb, P->nr * ns, NPTS);
    IB_HIP(hipGetLastError());
    if ((rc = staged_epilogue_fp64(P, fc, q + c0 * NPTS, nb, P->acc.ptr, P->fhat.ptr, P->work.ptr,
                                   with_loss, hb, pc)))
      return rc;
  }
  return IB_OK;
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
// The work array is unpadded, element (r, c) at r N + (c ^ g(r)) with
// g(r) = 2 floor(r / 2) ^ (3N/8) (r mod 2).  The lanes (p, h) of a wave access
// (p, 2k + h) in phase A, (2k + h, p) at the end of phase B and (n, p), (p, n)
// in the reads; with this permutation of each row every group of lanes served in
// one LDS cycle touches distinct banks in all four patterns, whether the LDS
// serves 128 or 256 bytes per cycle (a padded row of N + 1 elements puts lanes
// (p, 1) and (p + 1, 0) on one bank in both write patterns).
template <int N>
__device__ __forceinline__ constexpr int wsw(int r) {
  return (2 * (r >> 1)) ^ ((3 * N / 8) * (r & 1));
}

// SPLIT = false is the one-range form (ns = 1): its FP64 instance keeps two waves per SIMD,
// which the split form's extra scalar values would cost it.
template <typename T, int N, bool SPLIT>
__global__ __launch_bounds__(2 * N) void k_fusedw(CWArgs<T> A) {
  constexpr int H = N / 2, THREADS = 2 * N, SWH = 3 * N / 8;
  constexpr long long NPTS = (long long)N * N;
  __shared__ cx<T> W[N * N];
  __shared__ cx<T> tE[2][2][N];
  __shared__ cx<T> tG[2][2][N];
  // ones, then exp(2 pi i n / N); the second row starts one element later than H, so that
  // the two lanes of a pencil (reading entry n of row 0 and of row 1 together) use different banks
  __shared__ cx<T> twd[2][H + 1];
  const int tid = threadIdx.x;
  const long long id = blockIdx.x;
  const int ns = SPLIT ? A.ns : 1;
  const int nrs = A.nr * ns;
  long long cell;
  int jj;  // partial index: radial node j, angle range asp
  cell_of_block(id, nrs, SPLIT ? A.xg : 1, cell, jj);
  const int j = jj / ns, asp = jj % ns;
  if (cell >= A.batch) return;  // uniform over the workgroup
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
  // at h N + 2k N + (2k ^ ph).  Each phase takes its offsets from an opaque copy, so the
  // compiler forms the loop-invariant addresses next to their accesses instead of holding
  // all of them in registers across the pair loop.
  const int hg0 = h ^ wsw<N>(p);
  const int ph0 = p ^ (SWH * h);
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
        split_row(r, W[n * N + (pc ^ wsw<N>(n))], W[(n + H) * N + (pc ^ wsw<N>(n + H))], n);
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
      for (int k = 0; k < H; ++k) W[rowh + 2 * k * N + ((2 * k) ^ ph)] = r[k];
      wave_sync2();  // ... and written their halves
      pc = opaque(p);
#pragma unroll
      for (int n = 0; n < H; ++n)
        split_row(r, W[n * N + (pc ^ wsw<N>(n))], W[(n + H) * N + (pc ^ wsw<N>(n + H))], n);
      half_dif<-1>(r, tw);
      wave_sync2();
      ph = opaque(ph0);
      rowh = opaque(h * N);
#pragma unroll
      for (int k = 0; k < H; ++k) W[rowh + 2 * k * N + ((2 * k) ^ ph)] = r[k];
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
      const int gp = opaque(wsw<N>(p)), rowp = opaque(p * N);
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

template <typename T, int N>
static int run_fusedw(Plan* P, const double* f, double* q, long long batch, const int* pidx,
                      int with_loss, double hb, hipStream_t st) {
  constexpr long long NPTS = (long long)N * N;
  ScratchSizes z;
  const int ns = P->splits();
  const long long sub = c2c_sizes<T>(P, batch, 1024, NPTS, (long long)P->nr * ns, &z);
  int rc;
  if ((rc = P->reserve(z))) return rc;
  const size_t part_b = (size_t)sub * P->nr * ns * NPTS * sizeof(cx<T>);
  const size_t fh_b = (size_t)sub * NPTS * sizeof(cx<T>);
  cx<T>* part = (cx<T>*)P->aux.ptr;
  cx<T>* fhn = (cx<T>*)((char*)P->aux.ptr + part_b);
  T* fT = (T*)((char*)fhn + fh_b);
  for (long long c0 = 0; c0 < batch; c0 += sub) {
    const long long nb = std::min(sub, batch - c0);
    const double* fc = f + c0 * NPTS;
    const int* pc = pidx ? pidx + c0 : nullptr;
    if ((rc = staged_prologue_fp64(P, fc, nb, P->fhat.ptr, P->work.ptr))) return rc;
    const long long tot = nb * NPTS;
    hipLaunchKernelGGL(k_f2_pack<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                       (const cx<double>*)P->fhat.ptr, fc, fhn, fT, tot, 1.0 / (double)NPTS);
    IB_HIP(hipGetLastError());
    CWArgs<T> A{};
    A.fh = fhn;
    A.f = fT;
    A.ephase = (const cx<T>*)P->d_ephase;
    A.gphase = (const cx<T>*)P->d_gphase;
    A.part = part;
    A.pidx = pc;
    A.batch = nb;
    A.sigma = P->d_sigma;
    A.rj = P->d_rj;
    A.wsig = P->d_wsig;
    A.crad = P->d_crad;
    A.a = P->d_a;
    A.spg = (const T*)P->d_spg;
    A.nr = P->nr;
    A.m = P->m;
    A.nq = P->nq;
    A.ns = ns;
    A.L = P->L;
    // the one-range form keeps the XCD-grouped numbering (without a split, each of a small
    // batch's workgroups has a CU to itself either way)
    A.xg = ns > 1 ? nb >= 8 : 1;
    if ((rc = P->timed_begin(st))) return rc;
    const unsigned grid = (unsigned)cell_grid(nb, P->nr * ns, A.xg);
    if (ns > 1)
      hipLaunchKernelGGL((k_fusedw<T, N, true>), dim3(grid), dim3(2 * N), 0, st, A);
    else
      hipLaunchKernelGGL((k_fusedw<T, N, false>), dim3(grid), dim3(2 * N), 0, st, A);
    IB_HIP(hipGetLastError());
    P->timed_end(st, nb);
    hipLaunchKernelGGL(k_f2_reduce<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, part,
                       (cx<double>*)P->acc.ptr, nb, P->nr * ns, NPTS);
    IB_HIP(hipGetLastError());
    if ((rc = staged_epilogue_fp64(P, fc, q + c0 * NPTS, nb, P->acc.ptr, P->fhat.ptr, P->work.ptr,
                                   with_loss, hb, pc)))
      return rc;
  }
  return IB_OK;
}

template <int N>
static int run_fused2(Plan* P, const double* f, double* q, long long batch, const int* pidx,
                      int with_loss, double hb, hipStream_t st) {
  constexpr int G = F2Shape<N>::G, THREADS = F2Shape<N>::THREADS;
  constexpr long long NPTS = (long long)N * N;
  const int npairs = P->nr * P->m;
  // workgroups per cell fixed by the plan (16 rounds of G pairs each), not by the
  // batch size: a cell's FP32 summation order, and so its result, is the same in
  // any batch
  const int wpc = f2_wpc<N>(P);
  ScratchSizes z;
  const long long sub = c2c_sizes<float>(P, batch, 4096, NPTS, wpc, &z);
  int rc;
  if ((rc = P->reserve(z))) return rc;
  const size_t part_b = (size_t)sub * wpc * NPTS * sizeof(cx<float>);
  const size_t fh_b = (size_t)sub * NPTS * sizeof(cx<float>);
  cx<float>* part = (cx<float>*)P->aux.ptr;
  cx<float>* fh32 = (cx<float>*)((char*)P->aux.ptr + part_b);
  float* f32 = (float*)((char*)fh32 + fh_b);
  for (long long c0 = 0; c0 < batch; c0 += sub) {
    const long long nb = std::min(sub, batch - c0);
    const double* fc = f + c0 * NPTS;
    const int* pc = pidx ? pidx + c0 : nullptr;
    if ((rc = staged_prologue_fp64(P, fc, nb, P->fhat.ptr, P->work.ptr))) return rc;
    const long long tot = nb * NPTS;
    hipLaunchKernelGGL(k_f2_pack<float>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                       (const cx<double>*)P->fhat.ptr, fc, fh32, f32, tot, 1.0 / (double)NPTS);
    IB_HIP(hipGetLastError());
    F2Args A{};
    A.fh = fh32;
    A.ephase = (const cx<float>*)P->d_ephase;
    A.gphase = (const cx<float>*)P->d_gphase;
    A.f32 = f32;
    A.part = part;
    A.pidx = pc;
    A.batch = nb;
    A.wpc = wpc;
    A.sigma = P->d_sigma;
    A.rj = P->d_rj;
    A.wsig = P->d_wsig;
    A.crad = P->d_crad;
    A.a = P->d_a;
    A.spg = (const float*)P->d_spg;
    A.nr = P->nr;
    A.m = P->m;
    A.nq = P->nq;
    A.L = P->L;
    if ((rc = P->timed_begin(st))) return rc;
    hipLaunchKernelGGL(k_fused2<N>, dim3((unsigned)(nb * wpc)), dim3(THREADS), 0, st, A);
    IB_HIP(hipGetLastError());
    P->timed_end(st, nb);
    hipLaunchKernelGGL(k_f2_reduce<float>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, part,
                       (cx<double>*)P->acc.ptr, nb, wpc, NPTS);
    IB_HIP(hipGetLastError());
    if ((rc = staged_epilogue_fp64(P, fc, q + c0 * NPTS, nb, P->acc.ptr, P->fhat.ptr, P->work.ptr,
                                   with_loss, hb, pc)))
      return rc;
  }
  return IB_OK;
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
