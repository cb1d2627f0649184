// Transport and per-step kernels, FP64.  Compiled with -ffp-contract=off so
// that every product and sum rounds exactly as the reference's numpy code.
//
//  * k_tvd1d: ghost cells (src/solver.py:59-105, src/state.py:109-122), the
//    upwind Riemann solver (src/riemann/advection_1D.py:4-21), wave limiting
//    (src/limiters/tvd.py:6-40) and the wave-propagation update
//    (src/boltzmann.py:177-205), fused into one stencil: thread (i, v) reads
//    the five cells i-2..i+2 at velocity v (coalesced over v).  The same kernel
//    runs along either space axis of a 2D-space grid (dimension-split sum).
//  * k_axpy, k_axpy_chain, k_df_combine: the in-place updates of src/boltzmann.py:84-164.
//  * k_moments, k_project: FP64 moments (src/state.py:31-52, any d) and the
//    new conservative L2 projection of Q.
//  * k_upwind: the flux-form second-order upwind scheme with the van Leer
//    limiter (src/riemann/upwind.py:6-50), one output point per thread.
//  * k_fast_sweep: one Gauss-Seidel sweep of the steady kinetic problem
//    (src/iter_solver.py:38-62); a thread carries one velocity point through
//    the space recurrence.
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

// f, df: (outer, nx, nv) with the stencil along the middle axis; the speed of
// column v is v1[v % vmod].  accumulate: df += increment (the second sweep of a
// dimension-split step) instead of df = increment.
//
// A thread owns TVD_R consecutive cells i0 .. i0 + R - 1 of one column (o, v):
// it loads the R + 4 rows of the ghost-extended column once, forms the R + 3
// waves and the R + 1 limited waves of its interfaces (each interface's
// division done once), then writes R increments.  Workgroups cover 256
// columns of one (o, cell block); the block coordinates come from the
// workgroup index, so the per-thread index arithmetic is 32-bit.
constexpr int TVD_R = 4;
constexpr int TVD_THREADS = 256;

__global__ __launch_bounds__(TVD_THREADS) void k_tvd1d(
    const double* __restrict__ f, double* __restrict__ df, int nx, long long nv,
    const double* __restrict__ v1, long long vmod, double dtdx, int order, int limiter,
    int bc_lo, int bc_hi, int accumulate, unsigned vchunks, unsigned ichunks) {
  const unsigned b = blockIdx.x;
  const unsigned vc = b % vchunks, rest = b / vchunks;
  const int i0 = (int)(rest % ichunks) * TVD_R;
  const long long o = rest / ichunks;
  const long long v = (long long)vc * TVD_THREADS + threadIdx.x;
  if (v >= nv) return;
  const int cnt = min(TVD_R, nx - i0);
  const double* fo = f + o * nx * nv + v;
  double* dfo = df + o * nx * nv + v;
  double F[TVD_R + 4];
#pragma unroll
  for (int t = 0; t < TVD_R + 4; ++t) {
    const int tt = min(t, cnt + 3);  // rows past the column's end are not used
    F[t] = fo[ghost_row(i0 + tt, nx, bc_lo, bc_hi) * nv];
  }
  const double s = v1[v % vmod];
  double w[TVD_R + 3];
#pragma unroll
  for (int t = 0; t < TVD_R + 3; ++t) w[t] = F[t + 1] - F[t];
  const double spos = (s > 0.0) ? 1.0 : 0.0;
  double lim[TVD_R + 1];
  if (order == 2) {
#pragma unroll
    for (int u = 0; u < TVD_R + 1; ++u) lim[u] = limited(w[u], w[u + 1], w[u + 2], spos, limiter);
  }
  const double sabs = fabs(s);
  const double om = 1.0 - sabs * dtdx;
  const double c = (0.5 * sabs) * om;
#pragma unroll
  for (int r = 0; r < TVD_R; ++r) {
    if (r < cnt) {
      // cell i0 + r: waves w[r] .. w[r + 3] are W_{J-2} .. W_{J+1} (J = i + 2)
      const double amdf = fmin(s, 0.0) * w[r + 2];
      const double apdf = fmax(s, 0.0) * w[r + 1];
      double d = (-dtdx) * (amdf + apdf);
      if (order == 2) {
        const double F0 = c * lim[r];
        const double F1 = c * lim[r + 1];
        d = d - dtdx * (F1 - F0);
      }
      double* out = dfo + (long long)(i0 + r) * nv;
      *out = accumulate ? *out + d : d;
    }
  }
}

// Wave limiting alone (src/limiters/tvd.py:6-27) on (n, inner) waves: rows 1 .. n - 2 are
// limited with the speed s[((row - 1) inner + column) % smod], rows 0 and n - 1 copied.
// Out of place: every output row reads the three input rows around it.
__global__ void k_tvd_limit(const double* __restrict__ w, double* __restrict__ out, long long n,
                            long long inner, const double* __restrict__ s, long long smod,
                            int limiter) {
  const long long tot = n * inner;
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < tot;
       g += (long long)gridDim.x * blockDim.x) {
    const long long i = g / inner, j = g - i * inner;
    if (i == 0 || i == n - 1) {
      out[g] = w[g];
      continue;
    }
    const double spos = s[((i - 1) * inner + j) % smod] > 0.0 ? 1.0 : 0.0;
    out[g] = limited(w[g - inner], w[g], w[g + inner], spos, limiter);
  }
}

// Ghost arrays of src/solver.py:51-105 and src/state.py:95-122.  k_embed copies f (n0[, n1],
// inner) into the interior of fbc (n0 + 2g[, n1 + 2g], inner), or back (extract); k_ghost_axis
// fills the g ghost rows of each side of the middle axis of (outer, n + 2g, inner) in place from
// that axis's interior rows: extrapolation copies the first / last interior row, periodic wraps
// (lower ghosts = the last g interior rows, upper = the first g); bc 0 leaves a side as it is.
__global__ void k_embed(const double* __restrict__ a, double* __restrict__ b, long long n0, long long n1,
                        long long inner, int g, int pad1, int extract) {
  const long long tot = n0 * n1 * inner;
  const long long m1 = pad1 ? n1 + 2LL * g : n1;  // pad1: axis 1 is padded too (2D space)
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < tot;
       t += (long long)gridDim.x * blockDim.x) {
    const long long k = t % inner, ij = t / inner, j = ij % n1, i = ij / n1;
    const long long u = ((i + g) * m1 + (pad1 ? j + g : j)) * inner + k;
    if (extract) b[t] = a[u];
    else b[u] = a[t];
  }
}

__global__ void k_ghost_axis(double* __restrict__ a, long long outer, long long n, long long inner, int g,
                             int bc_lo, int bc_hi) {
  const long long tot = outer * 2LL * g * inner;
  const long long rows = n + 2LL * g;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < tot;
       t += (long long)gridDim.x * blockDim.x) {
    const long long k = t % inner, r = (t / inner) % (2LL * g), o = t / (inner * 2LL * g);
    const bool lower = r < g;
    const int bc = lower ? bc_lo : bc_hi;
    if (bc == 0) continue;
    long long dst, src;
    if (lower) {
      dst = r;
      src = bc == 1 ? g : n + r;            // periodic: interior row n - g + r
    } else {
      dst = n + g + (r - g);
      src = bc == 1 ? n + g - 1 : g + (r - g);
    }
    a[(o * rows + dst) * inner + k] = a[(o * rows + src) * inner + k];
  }
}

__global__ void k_axpy(double* __restrict__ y, const double* __restrict__ x, double a, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] = y[i] + a * x[i];
}

// y + a (b (c x)), multiplied in that order: the chained increments df1 = M0 df0,
// df2 = M1 df1 of the telescopic steps without materialising df1 or df2, rounding as
// the sequence of scalar-array products does (a factor of 1 is exact)
__global__ void k_axpy_chain(double* __restrict__ y, const double* __restrict__ x, double a, double b,
                             double c, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] = y[i] + a * (b * (c * x[i]));
}

__global__ void k_df_combine(const double* __restrict__ adv, const double* __restrict__ q,
                             double* __restrict__ out, double dt, double kn, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double c = (dt * q[i]) / kn;
    out[i] = adv ? adv[i] + c : c;
  }
}

// Sums K values over the workgroup: a butterfly over the 64 lanes of each wave,
// then the BS / 64 wave sums through LDS, added by every thread in the same order.
template <int BS, int K>
__device__ void block_sums(double (&x)[K], double* sh) {
  constexpr int NW = BS / 64;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int k = 0; k < K; ++k) x[k] += __shfl_xor(x[k], o, 64);
  }
  const int w = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) sh[k * NW + w] = x[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < NW; ++i) t += sh[k * NW + i];
    x[k] = t;
  }
  __syncthreads();
}

// Grid coordinates of a cell's flattened velocity index: v_i = -L + (i + 1/2) dv
// per axis, from a per-workgroup table; lg = log2 n for power-of-two n (shifts),
// -1 otherwise (32-bit division).
struct VGrid {
  int dim, n, lg;
  int npts;
  __device__ int index(int i, int a) const {
    const int rest = dim - 1 - a;
    if (lg >= 0) return (i >> (lg * rest)) & (n - 1);
    int div = 1;
    for (int t = 0; t < rest; ++t) div *= n;
    return (i / div) % n;
  }
};

constexpr int MOM_THREADS = 512;

__device__ void fill_vtab(double* vtab, int n, double L) {
  for (int i = threadIdx.x; i < n; i += MOM_THREADS) vtab[i] = -L + (i + 0.5) * (2 * L / n);
  __syncthreads();
}

// Streams the points of one cell through fn(x, i): 16-byte loads, MOM_UNROLL of
// them issued before use (enough bytes in flight per wave to cover HBM latency),
// scalar loads when the cell has an odd number of points.
constexpr int MOM_UNROLL = 4;

template <typename Fn>
__device__ __forceinline__ void for_each_point(const double* __restrict__ fc, int npts, Fn fn) {
  if ((npts & 1) == 0) {
    const double2* f2 = reinterpret_cast<const double2*>(fc);
    const int n2 = npts / 2;
    int i2 = threadIdx.x;
    for (; i2 + (MOM_UNROLL - 1) * MOM_THREADS < n2; i2 += MOM_UNROLL * MOM_THREADS) {
      double2 x[MOM_UNROLL];
#pragma unroll
      for (int u = 0; u < MOM_UNROLL; ++u) x[u] = f2[i2 + u * MOM_THREADS];
#pragma unroll
      for (int u = 0; u < MOM_UNROLL; ++u) {
        fn(x[u].x, 2 * (i2 + u * MOM_THREADS));
        fn(x[u].y, 2 * (i2 + u * MOM_THREADS) + 1);
      }
    }
    for (; i2 < n2; i2 += MOM_THREADS) {
      const double2 x = f2[i2];
      fn(x.x, 2 * i2);
      fn(x.y, 2 * i2 + 1);
    }
  } else {
    for (int i = threadIdx.x; i < npts; i += MOM_THREADS) fn(fc[i], i);
  }
}

// out[c] = [rho, m_1..m_d, E] * dv^d, one workgroup per cell.
template <int DIM>
__global__ __launch_bounds__(MOM_THREADS) void k_moments(const double* __restrict__ f,
                                                         double* __restrict__ out, VGrid g,
                                                         double L) {
  __shared__ double sh[(DIM + 2) * (MOM_THREADS / 64)];
  __shared__ double vtab[256];
  fill_vtab(vtab, g.n, L);
  const long long c = blockIdx.x;
  double s[DIM + 2];  // rho, m_1 .. m_d, sum f |v|^2
#pragma unroll
  for (int k = 0; k < DIM + 2; ++k) s[k] = 0.0;
  for_each_point(f + c * g.npts, g.npts, [&](double x, int i) {
    s[0] += x;
    double v2 = 0;
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      const double va = vtab[g.index(i, a)];
      s[1 + a] += x * va;
      v2 += va * va;
    }
    s[DIM + 1] += x * v2;
  });
  const double dv = 2 * L / g.n;
  double w = 1;
  for (int a = 0; a < DIM; ++a) w *= dv;
  block_sums<MOM_THREADS>(s, sh);
  if (threadIdx.x == 0) {
    double* o = out + c * (DIM + 2);
    for (int k = 0; k < DIM + 1; ++k) o[k] = s[k] * w;
    o[DIM + 1] = 0.5 * s[DIM + 1] * w;
  }
}

// Q <- Q - mean(Q) - sum_a (<Q, v_a> / <v_a, v_a>) v_a, one workgroup per cell.
// The constraint rows {1, v_a} are mutually orthogonal on the symmetric grid,
// so the projection needs only these d + 1 reductions.
template <int DIM>
__global__ __launch_bounds__(MOM_THREADS) void k_project(double* __restrict__ q, VGrid g,
                                                         double L) {
  __shared__ double sh[(DIM + 1) * (MOM_THREADS / 64)];
  __shared__ double vtab[256];
  fill_vtab(vtab, g.n, L);
  const long long c = blockIdx.x;
  double* qc = q + c * g.npts;
  double s[DIM + 1];  // sum q, sum q v_a
#pragma unroll
  for (int k = 0; k < DIM + 1; ++k) s[k] = 0.0;
  for_each_point(qc, g.npts, [&](double x, int i) {
    s[0] += x;
#pragma unroll
    for (int a = 0; a < DIM; ++a) s[1 + a] += x * vtab[g.index(i, a)];
  });
  block_sums<MOM_THREADS>(s, sh);
  double vv = 0;  // sum over the grid of v_a^2 = n^(d-1) sum_i v_i^2
  for (int i = 0; i < g.n; ++i) vv += vtab[i] * vtab[i];
  vv *= (double)(g.npts / g.n);
  double coef[DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a) coef[a] = s[1 + a] / vv;
  const double mean = s[0] / (double)g.npts;
  auto proj = [&](double x, int i) {
    x = x - mean;
#pragma unroll
    for (int a = 0; a < DIM; ++a) x = x - coef[a] * vtab[g.index(i, a)];
    return x;
  };
  if ((g.npts & 1) == 0) {
    double2* q2 = reinterpret_cast<double2*>(qc);
    const int n2 = g.npts / 2;
    int i2 = threadIdx.x;
    for (; i2 + (MOM_UNROLL - 1) * MOM_THREADS < n2; i2 += MOM_UNROLL * MOM_THREADS) {
      double2 x[MOM_UNROLL];
#pragma unroll
      for (int u = 0; u < MOM_UNROLL; ++u) x[u] = q2[i2 + u * MOM_THREADS];
#pragma unroll
      for (int u = 0; u < MOM_UNROLL; ++u) {
        const int i = 2 * (i2 + u * MOM_THREADS);
        q2[i2 + u * MOM_THREADS] = double2{proj(x[u].x, i), proj(x[u].y, i + 1)};
      }
    }
    for (; i2 < n2; i2 += MOM_THREADS) {
      const double2 x = q2[i2];
      q2[i2] = double2{proj(x.x, 2 * i2), proj(x.y, 2 * i2 + 1)};
    }
  } else {
    for (int i = threadIdx.x; i < g.npts; i += MOM_THREADS) qc[i] = proj(qc[i], i);
  }
}

static VGrid vgrid(int dim, int n) {
  VGrid g{dim, n, -1, 1};
  for (int a = 0; a < dim; ++a) g.npts *= n;
  if ((n & (n - 1)) == 0) {
    g.lg = 0;
    while ((1 << g.lg) < n) ++g.lg;
  }
  return g;
}

int launch_tvd1d(const double* f, double* df, long long outer, int nx, long long nv,
                 const double* v1, long long vmod, double dtdx, int order, int limiter, int bc_lo,
                 int bc_hi, int accumulate, hipStream_t st) {
  if (outer <= 0 || nx <= 0 || nv <= 0) return IB_OK;
  const long long vchunks = (nv + TVD_THREADS - 1) / TVD_THREADS;
  const long long ichunks = (nx + TVD_R - 1) / TVD_R;
  const long long blocks = vchunks * ichunks * outer;
  if (blocks > 0x7fffffffLL) {
    set_error("transport grid too large (more than 2^31 workgroups)");
    return IB_ERR_ARG;
  }
  hipLaunchKernelGGL(k_tvd1d, dim3((unsigned)blocks), dim3(TVD_THREADS), 0, st, f, df, nx, nv, v1,
                     vmod, dtdx, order, limiter, bc_lo, bc_hi, accumulate, (unsigned)vchunks,
                     (unsigned)ichunks);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

static unsigned grid_for(long long n) {
  long long b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

int launch_tvd_limit(const double* w, double* out, long long n, long long inner, const double* s,
                     long long smod, int limiter, hipStream_t st) {
  hipLaunchKernelGGL(k_tvd_limit, dim3(grid_for(n * inner)), dim3(256), 0, st, w, out, n, inner, s,
                     smod, limiter);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

int launch_embed(const double* a, double* b, long long n0, long long n1, long long inner, int g,
                 int pad1, int extract, hipStream_t st) {
  hipLaunchKernelGGL(k_embed, dim3(grid_for(n0 * n1 * inner)), dim3(256), 0, st, a, b, n0, n1, inner, g,
                     pad1, extract);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

int launch_ghost_axis(double* a, long long outer, long long n, long long inner, int g, int bc_lo,
                      int bc_hi, hipStream_t st) {
  hipLaunchKernelGGL(k_ghost_axis, dim3(grid_for(outer * 2LL * g * inner)), dim3(256), 0, st, a, outer,
                     n, inner, g, bc_lo, bc_hi);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

int launch_axpy(double* y, const double* x, double a, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_axpy, dim3(grid_for(n)), dim3(256), 0, st, y, x, a, n);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

int launch_axpy_chain(double* y, const double* x, double a, double b, double c, long long n,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_axpy_chain, dim3(grid_for(n)), dim3(256), 0, st, y, x, a, b, c, n);
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
  if (dim == 2)
    hipLaunchKernelGGL(k_moments<2>, dim3((unsigned)batch), dim3(MOM_THREADS), 0, st, f, out,
                       vgrid(dim, n), L);
  else
    hipLaunchKernelGGL(k_moments<3>, dim3((unsigned)batch), dim3(MOM_THREADS), 0, st, f, out,
                       vgrid(dim, n), L);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

int launch_project(double* q, long long batch, int dim, int n, double L, hipStream_t st) {
  if (dim == 2)
    hipLaunchKernelGGL(k_project<2>, dim3((unsigned)batch), dim3(MOM_THREADS), 0, st, q,
                       vgrid(dim, n), L);
  else
    hipLaunchKernelGGL(k_project<3>, dim3((unsigned)batch), dim3(MOM_THREADS), 0, st, q,
                       vgrid(dim, n), L);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

// ---------------------------------------------------------------------------
// Flux-form upwind scheme, src/riemann/upwind.py.  Arrays are (n, inner) with
// the stencil along axis 0.  Every expression keeps numpy's evaluation order.
constexpr double UPWIND_EPS = 1e-8;  // upwind.py:3

__device__ __forceinline__ double van_leer(double r) {  // upwind.py:16-17
  return (r + fabs(r)) / (1.0 + fabs(r));
}
// np.maximum(v, 0.) / np.minimum(v, 0.): the first operand wins ties and NaN
__device__ __forceinline__ double npmax0(double v) { return (v >= 0.0 || isnan(v)) ? v : 0.0; }
__device__ __forceinline__ double npmin0(double v) { return (v <= 0.0 || isnan(v)) ? v : 0.0; }

// F of F_p_2 (upwind.py:22-25) at row i in 1..n-2, with its speed vp
__device__ __forceinline__ double flux_p(double fm, double f0, double fp, double vp, double dx,
                                         double dt) {
  const double r = (f0 - fm) / ((fp - f0) + UPWIND_EPS);
  const double phi = van_leer(r);
  return f0 + ((0.5 * phi) * (1.0 - (vp * dt) / dx)) * (fp - f0);
}
// F of F_m_2 (upwind.py:32-35)
__device__ __forceinline__ double flux_m(double fm, double f0, double fp, double vm, double dx,
                                         double dt) {
  const double r = (fp - f0) / ((f0 - fm) + UPWIND_EPS);
  const double phi = van_leer(r);
  return f0 + ((0.5 * phi) * (-1.0 - (vm * dt) / dx)) * (f0 - fm);
}

// mode 0 grad_vf1 (out rows n-2), 1 grad_vf2, 2 F_p_2, 3 F_m_2 (out rows n-4).
// Modes 0, 1: v is (n, inner); modes 2, 3: v is the speed aligned with f[1:-1], (n-2, inner).
__global__ void k_upwind(const double* __restrict__ f, const double* __restrict__ v,
                         double* __restrict__ out, long long n, long long inner, double dx,
                         double dt, int mode) {
  const long long rows = mode == 0 ? n - 2 : n - 4;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= rows * inner) return;
  const long long o = gid / inner, j = gid % inner;
  auto F = [&](long long i) { return f[i * inner + j]; };
  auto V = [&](long long i) { return v[i * inner + j]; };
  double res;
  if (mode == 0) {  // upwind.py:47-50: diff_x_up(vp f) + diff_x_down(vm f)
    const double up = npmax0(V(o + 1)) * F(o + 1) - npmax0(V(o)) * F(o);
    const double dn = npmin0(V(o + 2)) * F(o + 2) - npmin0(V(o + 1)) * F(o + 1);
    res = up + dn;
  } else if (mode == 1) {  // upwind.py:40-44
    const double fp2 = flux_p(F(o + 1), F(o + 2), F(o + 3), npmax0(V(o + 2)), dx, dt) -
                       flux_p(F(o), F(o + 1), F(o + 2), npmax0(V(o + 1)), dx, dt);
    const double fm2 = flux_m(F(o + 2), F(o + 3), F(o + 4), npmin0(V(o + 3)), dx, dt) -
                       flux_m(F(o + 1), F(o + 2), F(o + 3), npmin0(V(o + 2)), dx, dt);
    res = npmax0(V(o + 2)) * fp2 + npmin0(V(o + 2)) * fm2;
  } else if (mode == 2) {  // F[2:-2] - F[1:-3], speed index i - 1
    res = flux_p(F(o + 1), F(o + 2), F(o + 3), V(o + 1), dx, dt) -
          flux_p(F(o), F(o + 1), F(o + 2), V(o), dx, dt);
  } else {  // F[3:-1] - F[2:-2]
    res = flux_m(F(o + 2), F(o + 3), F(o + 4), V(o + 2), dx, dt) -
          flux_m(F(o + 1), F(o + 2), F(o + 3), V(o + 1), dx, dt);
  }
  out[gid] = res;
}

int launch_upwind(const double* f, const double* v, double* out, long long n, long long inner,
                  double dx, double dt, int mode, hipStream_t st) {
  const long long total = (mode == 0 ? n - 2 : n - 4) * inner;
  if (total <= 0) return IB_OK;
  hipLaunchKernelGGL(k_upwind, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, f, v, out,
                     n, inner, dx, dt, mode);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

// ---------------------------------------------------------------------------
// Fast sweeping, src/iter_solver.py:38-62.  f, gain: (nx, nvr, nvc); rho: (nx,);
// vx: (nvr,) the v_x of each velocity row.  Rows with v_x > 0 sweep i = 1..nx-2
// forward, rows with v_x < 0 backward; rows with v_x = 0 and the end cells keep
// their values.  err (one double, cleared by k_zero1 first) receives
// max |f_new - f_old| through an integer max on the bits of the non-negative
// differences (their order equals the order of the values).
__global__ void k_zero1(double* p) { *p = 0.0; }

__global__ void k_fast_sweep(double* __restrict__ f, const double* __restrict__ gain,
                             const double* __restrict__ rho, const double* __restrict__ vx, int nx,
                             int nvr, long long nvc, double dx, double nu, double* err) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nv = (long long)nvr * nvc;
  if (gid >= nv) return;
  const double v = vx[gid / nvc];
  const double vdx = v / dx;
  double dmax = 0.0;
  if (v > 0.0) {
    double prev = f[gid];
    for (int i = 1; i < nx - 1; ++i) {
      const long long at = (long long)i * nv + gid;
      const double nw = (nu * gain[at] + vdx * prev) / (nu * rho[i] + vdx);
      const double d = fabs(nw - f[at]);
      dmax = (d > dmax || isnan(d)) ? d : dmax;
      f[at] = nw;
      prev = nw;
    }
  } else if (v < 0.0) {
    double prev = f[(long long)(nx - 1) * nv + gid];
    for (int i = nx - 2; i >= 1; --i) {
      const long long at = (long long)i * nv + gid;
      const double nw = (nu * gain[at] - vdx * prev) / (nu * rho[i] - vdx);
      const double d = fabs(nw - f[at]);
      dmax = (d > dmax || isnan(d)) ? d : dmax;
      f[at] = nw;
      prev = nw;
    }
  }
  if (dmax != 0.0) atomicMax((unsigned long long*)err, (unsigned long long)__double_as_longlong(dmax));
}

int launch_fast_sweep(double* f, const double* gain, const double* rho, const double* vx, int nx,
                      int nvr, long long nvc, double dx, double nu, double* err, hipStream_t st) {
  hipLaunchKernelGGL(k_zero1, dim3(1), dim3(1), 0, st, err);
  IB_HIP(hipGetLastError());
  const long long nv = (long long)nvr * nvc;
  hipLaunchKernelGGL(k_fast_sweep, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, st, f, gain,
                     rho, vx, nx, nvr, nvc, dx, nu, err);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

}  // namespace ib
