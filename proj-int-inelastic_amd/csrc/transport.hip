// Transport and per-step kernels, FP64.  Compiled with -ffp-contract=off so
// that every product and sum rounds exactly as the reference's numpy code.
//
//  * k_tvd1d: ghost cells (src/solver.py:59-105, src/state.py:109-122), the
//    upwind Riemann solver (src/riemann/advection_1D.py:4-21), wave limiting
//    (src/limiters/tvd.py:6-40) and the wave-propagation update
//    (src/boltzmann.py:177-205), fused into one stencil: thread (i, v) reads
//    the five cells i-2..i+2 at velocity v (coalesced over v).
//  * k_axpy, k_df_combine: the in-place updates of src/boltzmann.py:84-164.
//  * k_moments, k_project: FP64 moments (src/state.py:31-52, any d) and the
//    new conservative L2 projection of Q.
#pragma clang fp contract(off)
#include <hip/hip_runtime.h>

#include <cfloat>

#include "plan.h"

namespace ib {

__device__ __forceinline__ long long ghost_row(int J, int nx, int bc_lo, int bc_hi) {
  // fbc row J (0 .. nx + 3) -> row of f, with 2 ghost cells on each side
  if (J < 2) return bc_lo == 1 ? 0 : nx - 2 + J;
  if (J >= nx + 2) return bc_hi == 1 ? nx - 1 : J - nx - 2;
  return J - 2;
}

// np.ma true_divide masks where |num| * tiny >= |den|; masked ratios become 0.
__device__ __forceinline__ double ratio(double num, double den) {
  return (fabs(num) * DBL_MIN >= fabs(den)) ? 0.0 : num / den;
}

__device__ __forceinline__ double phi_lim(double th, int limiter) {
  if (limiter == 0) return fmax(0.0, fmin(1.0, th));  // minmod
  return fmax(0.0, fmin(fmin(0.5 * (1.0 + th), 2.0), 2.0 * th));  // MC
}

__device__ __forceinline__ double limited(double wm, double w0, double wp, double spos,
                                          int limiter) {
  const double num = spos * (w0 * wm) + (1.0 - spos) * (wp * w0);
  const double th = ratio(num, w0 * w0);
  const double ph = phi_lim(th, limiter);
  const double zero = (w0 == 0.0) ? 1.0 : 0.0;
  const double nz = 1.0 - zero;
  return w0 * zero + (ph * w0) * nz;
}

__global__ void k_tvd1d(const double* __restrict__ f, double* __restrict__ df, int nx,
                        long long nv, const double* __restrict__ v1, double dtdx, int order,
                        int limiter, int bc_lo, int bc_hi) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)nx * nv) return;
  const int i = (int)(gid / nv);
  const long long v = gid % nv;
  const int J = i + 2;
  double F[5];
#pragma unroll
  for (int t = 0; t < 5; ++t) F[t] = f[ghost_row(J - 2 + t, nx, bc_lo, bc_hi) * nv + v];
  const double s = v1[v];
  const double wJm2 = F[1] - F[0];
  const double wJm1 = F[2] - F[1];
  const double wJ = F[3] - F[2];
  const double wJp1 = F[4] - F[3];
  const double amdf = fmin(s, 0.0) * wJ;
  const double apdf = fmax(s, 0.0) * wJm1;
  double d = (-dtdx) * (amdf + apdf);
  if (order == 2) {
    const double spos = (s > 0.0) ? 1.0 : 0.0;
    const double lm1 = limited(wJm2, wJm1, wJ, spos, limiter);
    const double l0 = limited(wJm1, wJ, wJp1, spos, limiter);
    const double sabs = fabs(s);
    const double om = 1.0 - sabs * dtdx;
    const double c = (0.5 * sabs) * om;
    const double F0 = c * lm1;
    const double F1 = c * l0;
    d = d - dtdx * (F1 - F0);
  }
  df[gid] = d;
}

__global__ void k_axpy(double* __restrict__ y, const double* __restrict__ x, double a, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] = y[i] + a * x[i];
}

__global__ void k_df_combine(const double* __restrict__ adv, const double* __restrict__ q,
                             double* __restrict__ out, double dt, double kn, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double c = (dt * q[i]) / kn;
    out[i] = adv ? adv[i] + c : c;
  }
}

template <int BS>
__device__ double block_sum(double x, double* sh) {
  const int t = threadIdx.x;
  sh[t] = x;
  __syncthreads();
#pragma unroll
  for (int s = BS / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] = sh[t] + sh[t + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double vcoord(long long idx, int axis, int dim, int n, double L) {
  long long div = 1;
  for (int a = dim - 1; a > axis; --a) div *= n;
  const int i = (int)((idx / div) % n);
  const double dv = 2 * L / n;
  return -L + (i + 0.5) * dv;
}

// out[c] = [rho, m_1..m_d, E] * dv^d, one 256-thread workgroup per cell.
__global__ __launch_bounds__(256) void k_moments(const double* __restrict__ f, double* __restrict__ out,
                                                 int dim, int n, double L) {
  __shared__ double sh[256];
  const long long c = blockIdx.x;
  long long npts = 1;
  for (int a = 0; a < dim; ++a) npts *= n;
  const double* fc = f + c * npts;
  double s0 = 0, s1[3] = {0, 0, 0}, s2 = 0;
  for (long long i = threadIdx.x; i < npts; i += 256) {
    const double x = fc[i];
    s0 += x;
    double v2 = 0;
    for (int a = 0; a < dim; ++a) {
      const double va = vcoord(i, a, dim, n, L);
      s1[a] += x * va;
      v2 += va * va;
    }
    s2 += x * v2;
  }
  const double dv = 2 * L / n;
  double w = 1;
  for (int a = 0; a < dim; ++a) w *= dv;
  const double r0 = block_sum<256>(s0, sh);
  double r1[3];
  for (int a = 0; a < dim; ++a) r1[a] = block_sum<256>(s1[a], sh);
  const double r2 = block_sum<256>(s2, sh);
  if (threadIdx.x == 0) {
    double* o = out + c * (dim + 2);
    o[0] = r0 * w;
    for (int a = 0; a < dim; ++a) o[1 + a] = r1[a] * w;
    o[dim + 1] = 0.5 * r2 * w;
  }
}

// Q <- Q - mean(Q) - sum_a (<Q, v_a> / <v_a, v_a>) v_a, one workgroup per cell.
// The constraint rows {1, v_a} are mutually orthogonal on the symmetric grid,
// so the projection needs only these d + 1 reductions.
__global__ __launch_bounds__(256) void k_project(double* __restrict__ q, int dim, int n, double L) {
  __shared__ double sh[256];
  const long long c = blockIdx.x;
  long long npts = 1;
  for (int a = 0; a < dim; ++a) npts *= n;
  double* qc = q + c * npts;
  double s0 = 0, s1[3] = {0, 0, 0};
  for (long long i = threadIdx.x; i < npts; i += 256) {
    const double x = qc[i];
    s0 += x;
    for (int a = 0; a < dim; ++a) s1[a] += x * vcoord(i, a, dim, n, L);
  }
  const double r0 = block_sum<256>(s0, sh);
  double coef[3] = {0, 0, 0};
  double vv = 0;  // sum over the grid of v_a^2 = n^(d-1) sum_i v_i^2
  const double dv = 2 * L / n;
  for (int i = 0; i < n; ++i) {
    const double vi = -L + (i + 0.5) * dv;
    vv += vi * vi;
  }
  vv *= (double)(npts / n);
  for (int a = 0; a < dim; ++a) coef[a] = block_sum<256>(s1[a], sh) / vv;
  const double mean = r0 / (double)npts;
  for (long long i = threadIdx.x; i < npts; i += 256) {
    double x = qc[i] - mean;
    for (int a = 0; a < dim; ++a) x = x - coef[a] * vcoord(i, a, dim, n, L);
    qc[i] = x;
  }
}

int launch_tvd1d(const double* f, double* df, int nx, long long nv, const double* v1, double dtdx,
                 int order, int limiter, int bc_lo, int bc_hi, hipStream_t st) {
  const long long total = (long long)nx * nv;
  hipLaunchKernelGGL(k_tvd1d, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, f, df, nx,
                     nv, v1, dtdx, order, limiter, bc_lo, bc_hi);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

static unsigned grid_for(long long n) {
  long long b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

int launch_axpy(double* y, const double* x, double a, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_axpy, dim3(grid_for(n)), dim3(256), 0, st, y, x, a, n);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

int launch_df_combine(const double* adv, const double* q, double* out, double dt, double kn,
                      long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_df_combine, dim3(grid_for(n)), dim3(256), 0, st, adv, q, out, dt, kn, n);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

int launch_moments(const double* f, double* out, long long batch, int dim, int n, double L,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_moments, dim3((unsigned)batch), dim3(256), 0, st, f, out, dim, n, L);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

int launch_project(double* q, long long batch, int dim, int n, double L, hipStream_t st) {
  hipLaunchKernelGGL(k_project, dim3((unsigned)batch), dim3(256), 0, st, q, dim, n, L);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

}  // namespace ib
