// Front and back of the fused 2D 16^2 / 32^2 forms in two launches instead of five: the FP64
// prologue fhat = fft(f) with the packing of the pair loop's inputs in one kernel (k_front2d),
// and the FP64 sum over the partials with the whole FP64 epilogue in another (k_back2d).  Both
// repeat the staged engine's operations in its order (collide_plain.hip: B_PRO then A_OUT_PRO,
// k_f2_pack; k_f2_reduce, A_IN_EPI then B_EPI), so the results are bit-identical to the staged
// prologue / epilogue; small batches, whose steps are bound by launches, take three launches
// per Q instead of seven.
#pragma once
#include "common.h"

namespace ib {

// One 64-thread workgroup per 64 / N cells (N = 16, 32): rows then columns of the FP64
// transform of f, written as the full FP64 spectrum and as fhat / n and f in the pair loop's
// type T.
template <typename T, int N>
__global__ __launch_bounds__(64) void k_front2d(const double* __restrict__ f, cx<double>* __restrict__ fhat64,
                                                cx<T>* __restrict__ fhT, T* __restrict__ fT, long long nb) {
  constexpr int CPW = 64 / N;
  constexpr long long NPTS = (long long)N * N;
  __shared__ cx<double> W[CPW][N][N + 1];
  const int tid = threadIdx.x, sub = tid / N, r = tid % N;
  const long long cell = (long long)blockIdx.x * CPW + sub;
  const bool active = cell < nb;
  cx<double> v[N];
  if (active) {
    const double* fr = f + cell * NPTS + (long long)r * N;
    T* fo = fT + cell * NPTS + (long long)r * N;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      v[i] = cx<double>{fr[i], 0.0};
      fo[i] = (T)fr[i];
    }
    fft_reg<N, -1>(v);
#pragma unroll
    for (int i = 0; i < N; ++i) W[sub][r][i] = v[i];
  }
  __syncthreads();
  if (active) {
    const int p = r;  // column
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = W[sub][i][p];
    fft_reg<N, -1>(v);
    const double inv_n = 1.0 / (double)NPTS;
    cx<double>* d = fhat64 + cell * NPTS + p;
    cx<T>* dT = fhT + cell * NPTS + p;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      d[(long long)i * N] = v[i];
      dT[(long long)i * N] = cx<T>{(T)(v[i].re * inv_n), (T)(v[i].im * inv_n)};
    }
  }
}

// What the epilogue needs besides the partial spectra.
struct BackArgs {
  const double* f;          // [cell][N][N] FP64 input (the loss term's f)
  const cx<double>* fhat;   // [cell][N][N] unnormalised FP64 spectrum of f
  double* q;                // [cell][N][N] output
  const double* lam;        // [param][nq] FP64 loss symbol
  const int* pidx;
  int nq, with_loss;
  double hb, L;
};

constexpr int kBackThreads = 256;

// One 256-thread workgroup per cell: the FP64 sum of the cell's nparts partial spectra
// (part[cell][w], w < nparts) in partial order, every thread a few elements with several
// partials' loads in flight, then the epilogue: gain (+ heat bath) and loss spectra
// inverse-transformed along axis 0 then axis 1, Q = Re(gain) - f Re(loss).  Re(gain) is
// staged in q itself (each row by the thread that later finishes it), so that no row of the
// loss transform is held in registers across the gain transform.
template <typename T, int N>
__global__ __launch_bounds__(kBackThreads) void k_back2d(BackArgs K, const cx<T>* __restrict__ part,
                                                         int nparts) {
  constexpr int THREADS = kBackThreads;
  constexpr long long NPTS = (long long)N * N;
  constexpr int EPT = N * N / THREADS;       // elements per thread
  constexpr int U = EPT >= 4 ? 4 : 8;        // partials loaded together
  static_assert(N * N % THREADS == 0 && THREADS >= N, "epilogue workgroup shape");
  __shared__ cx<double> W[N][N + 1];
  const long long cell = blockIdx.x;
  const int tid = threadIdx.x;
  {
    const cx<T>* pc = part + cell * nparts * NPTS + tid;
    double re[EPT], im[EPT];
#pragma unroll
    for (int k = 0; k < EPT; ++k) re[k] = im[k] = 0.0;
    int w = 0;
    for (; w + U <= nparts; w += U) {
      cx<T> x[U][EPT];
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int k = 0; k < EPT; ++k) x[u][k] = pc[(long long)(w + u) * NPTS + k * THREADS];
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int k = 0; k < EPT; ++k) {
          re[k] += (double)x[u][k].re;
          im[k] += (double)x[u][k].im;
        }
    }
    for (; w < nparts; ++w) {
#pragma unroll
      for (int k = 0; k < EPT; ++k) {
        const cx<T> x = pc[(long long)w * NPTS + k * THREADS];
        re[k] += (double)x.re;
        im[k] += (double)x.im;
      }
    }
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int e = tid + k * THREADS;
      W[e / N][e % N] = cx<double>{re[k], im[k]};
    }
  }
  __syncthreads();
  const cx<double>* fh = K.fhat + cell * NPTS;
  const double inv_n = 1.0 / (double)NPTS;
  const int p = tid;  // column, then row
  const int qrest = kfreq(p % N, N) * kfreq(p % N, N);
  double* qr = K.q + cell * NPTS + (long long)p * N;
  cx<double> v[N];
  // gain: columns, then rows into q
  if (p < N) {
    const double lap_c = -(M_PI * M_PI) / (K.L * K.L);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      cx<double> x = W[i][p];
      if (K.hb != 0.0) {
        const int ki = kfreq(i, N);
        const double lk = (double)(K.hb * lap_c * (double)(ki * ki + qrest));
        x = x + lk * fh[(long long)i * N + p];
      }
      v[i] = inv_n * x;
    }
    fft_reg<N, +1>(v);
#pragma unroll
    for (int i = 0; i < N; ++i) W[i][p] = v[i];
  }
  __syncthreads();
  if (p < N) {
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = W[p][i];
    fft_reg<N, +1>(v);
#pragma unroll
    for (int i = 0; i < N; ++i) qr[i] = v[i].re;
  }
  if (!K.with_loss) return;  // uniform over the workgroup
  __syncthreads();  // gain rows read before the loss columns overwrite W
  // loss: columns, then rows, each finishing its row of q
  if (p < N) {
    const int ps = K.pidx ? K.pidx[cell] : 0;
    const double* lam = K.lam + (long long)ps * K.nq;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const int ki = kfreq(i, N);
      v[i] = (inv_n * lam[ki * ki + qrest]) * fh[(long long)i * N + p];
    }
    fft_reg<N, +1>(v);
#pragma unroll
    for (int i = 0; i < N; ++i) W[i][p] = v[i];
  }
  __syncthreads();
  if (p < N) {
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = W[p][i];
    fft_reg<N, +1>(v);
    const double* fr = K.f + cell * NPTS + (long long)p * N;
#pragma unroll
    for (int i = 0; i < N; ++i) qr[i] = qr[i] - fr[i] * v[i].re;
  }
}

}  // namespace ib
