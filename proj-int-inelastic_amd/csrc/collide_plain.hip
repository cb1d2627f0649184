// Plain (staged) form of the fast spectral collision operator Q(f, f).
//
// Follows reference src/collision/inelastic_collisions.py:208-227 (collide):
// C2C transforms over the d velocity axes, gain = ifft(sum_{j,m} G_jm *
// fft(ifft(E_jm * fhat) * f)), loss = f * ifft(Lambda * fhat), real part at
// the end.  The sum over j of the reference's Nr loss transforms is folded
// into one transform of the combined symbol Lambda = sum_j l_j S_j (linear).
//
// Each N^d transform is split into an axis-0 pass over "pencils" (one thread
// owns the N values along axis 0 of one pencil, in registers) and a pass over
// the remaining axes ("slabs": a row in 2D, an N x N tile in LDS in 3D).  The
// work array of every (cell, quadrature pair) item goes through HBM between
// the passes; items are processed in chunks sized to a scratch budget.  The
// element-wise products (E * fhat, * f, * G and the quadrature sum) are fused
// into the passes and the E / G phase factors are generated per pair from 1D
// tables (E and G are separable in k), so no (Nr, M, n) table is stored.
//
// Precision T = double gives the plain FP64 form; T = float the plain FP32 form.
#include <algorithm>
#include <type_traits>

#include "common.h"
#include "plan.h"

namespace ib {

enum { A_IN_PAIR = 0, A_OUT_PAIR = 1, A_OUT_PRO = 2, A_IN_EPI = 3 };
enum { B_PAIR = 0, B_PRO = 1, B_EPI = 2 };

template <int N, int D>
struct Geo {
  static constexpr int NP = (D == 2) ? N : N * N;  // pencils along axis 0 (= slab size)
  static constexpr long long NPTS = (long long)N * NP;
  static constexpr int NPW = NP < 64 ? NP : 64;  // pencils handled by one sub-group
  static constexpr int SUBS = 64 / NPW;          // items per 64-thread workgroup
  static constexpr int BLK = NP / NPW;           // pencil blocks per item
};

template <typename T>
struct AArgs {
  const cx<T>* src;
  const cx<T>* src2;
  cx<T>* dst;
  cx<T>* part;  // A_OUT_PAIR with S > 1: per-(cell, split) partial sums
  long long items;
  int P, p0, first, S;
  const int* pidx;
  long long cell0;
  const double* sigma;
  const double* rj;
  const double* wsig;
  const double* crad;
  const double* a;
  const T* spg;
  const T* lam;
  int nr, m, nq, with_loss, no_gain;
  double L, hb;
};

template <typename T>
struct BArgs {
  cx<T>* w;
  const double* f;
  double* q;
  long long items;
  int P, with_loss;
};

// LDS bytes of one 64-thread block of each pass (the grouped kernel takes the
// maximum over the shapes it serves).
template <typename T, int N, int D, int MODE>
struct Axis0Lds {
  static constexpr int SUBS = Geo<N, D>::SUBS;
  static constexpr size_t tab = sizeof(cx<T>) * SUBS * D * N;
  static constexpr size_t bytes = tab + sizeof(cx<T>) * (MODE == A_OUT_PAIR ? N : 1) * 64;
};

template <typename T, int N>
struct Rest3Lds {
  static constexpr size_t bytes = sizeof(cx<T>) * (64 / N) * N * (N + 1);
};

// ---------------------------------------------------------------------------
// Axis-0 pass.  One thread per axis-0 pencil; 64-thread blocks.  (bx, by) is
// the block's position in the pass's 2D grid (pencil block, item group); lds
// points to Axis0Lds<...>::bytes of workgroup memory.
template <typename T, int N, int D, int MODE>
__device__ __forceinline__ void axis0_body(const AArgs<T>& A, int bx, long long by, char* lds) {
  using G = Geo<N, D>;
  constexpr int NP = G::NP, NPW = G::NPW, SUBS = G::SUBS;
  constexpr long long NPTS = G::NPTS;
  auto tab = reinterpret_cast<cx<T>(*)[D][N]>(lds);
  auto accs = reinterpret_cast<cx<T>(*)[64]>(lds + Axis0Lds<T, N, D, MODE>::tab);
  const int tid = threadIdx.x;
  const int sub = tid / NPW, lane = tid % NPW;
  const long long item = by * SUBS + sub;
  const int p = bx * NPW + lane;
  const bool active = item < A.items;
  int k1 = 0, k2 = 0;
  if (D == 2) {
    k1 = p;
  } else {
    k1 = p / N;
    k2 = p % N;
  }
  int qrest = kfreq(k1, N) * kfreq(k1, N);
  if (D == 3) qrest += kfreq(k2, N) * kfreq(k2, N);
  const T inv_n = (T)(1.0 / (double)NPTS);
  cx<T> v[N];

  if (MODE == A_IN_PAIR) {
    const long long cell = item / A.P;
    const int pair = A.p0 + (int)(item % A.P);
    const int j = pair / A.m, mm = pair % A.m;
    if (active) {
      for (int t = lane; t < D * N; t += NPW) {
        const int ax = t / N, k = t % N;
        const double th = -(M_PI * A.rj[j] * A.sigma[mm * D + ax] * (double)kfreq(k, N)) / A.L;
        double s, c;
        sincos(th, &s, &c);
        tab[sub][ax][k] = cx<T>{(T)c, (T)s};
      }
    }
    __syncthreads();
    if (!active) return;
    cx<T> er = tab[sub][1][k1];
    if (D == 3) er = er * tab[sub][2][k2];
    er = inv_n * er;
    const cx<T>* s = A.src + cell * NPTS + p;
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = s[(long long)i * NP] * (tab[sub][0][i] * er);
    fft_reg<N, +1>(v);
    cx<T>* d = A.dst + item * NPTS + p;
#pragma unroll
    for (int i = 0; i < N; ++i) d[(long long)i * NP] = v[i];
  } else if (MODE == A_OUT_PAIR) {
    // S > 1 splits the chunk's pairs of a cell over S items (small batches, whose
    // cells alone would leave most CUs idle); split s sums pairs [lo, hi) into its
    // own partial, and k_split_add sums the partials in split order.
    const int S = A.S > 1 ? A.S : 1;
    const long long cell = item / S;
    const int split = (int)(item % S);
    const int lo = (int)(((long long)split * A.P) / S), hi = (int)(((long long)(split + 1) * A.P) / S);
    const int tmax = (A.P + S - 1) / S;  // the same trip count in every item of the workgroup
    const int ps = (active && A.pidx) ? A.pidx[A.cell0 + cell] : 0;
    for (int t = 0; t < tmax; ++t) {
      const int pl = lo + t;
      const bool on = active && pl < hi;
      const int pair = A.p0 + (on ? pl : lo);
      const int j = pair / A.m, mm = pair % A.m;
      __syncthreads();
      if (on) {
        const double aj = A.a[ps * A.nr + j] * A.rj[j];
        for (int t = lane; t < D * N; t += NPW) {
          const int ax = t / N, k = t % N;
          const double th = aj * A.sigma[mm * D + ax] * (double)kfreq(k, N);
          double s, c;
          sincos(th, &s, &c);
          tab[sub][ax][k] = cx<T>{(T)c, (T)s};
        }
      }
      __syncthreads();
      if (!on) continue;
      cx<T> gr = tab[sub][1][k1];
      if (D == 3) gr = gr * tab[sub][2][k2];
      const T wj = (T)(A.crad[ps * A.nr + j] * A.wsig[mm]);
      const T* sp = A.spg + ((long long)ps * A.nr + j) * A.nq;
      const cx<T>* s = A.src + (cell * A.P + pl) * NPTS + p;
#pragma unroll
      for (int i = 0; i < N; ++i) v[i] = s[(long long)i * NP];
      fft_reg<N, -1>(v);
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int ki = kfreq(i, N);
        const T g = wj * sp[ki * ki + qrest];
        const cx<T> tv = (g * (tab[sub][0][i] * gr)) * v[i];
        accs[i][tid] = (t == 0) ? tv : accs[i][tid] + tv;
      }
    }
    if (!active) return;
    if (S > 1) {
      cx<T>* d = A.part + item * NPTS + p;
#pragma unroll
      for (int i = 0; i < N; ++i) d[(long long)i * NP] = accs[i][tid];
      return;
    }
    cx<T>* d = A.dst + cell * NPTS + p;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const cx<T> tv = accs[i][tid];
      d[(long long)i * NP] = A.first ? tv : d[(long long)i * NP] + tv;
    }
  } else if (MODE == A_OUT_PRO) {
    if (!active) return;
    const cx<T>* s = A.src + item * NPTS + p;
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = s[(long long)i * NP];
    fft_reg<N, -1>(v);
    cx<T>* d = A.dst + item * NPTS + p;
#pragma unroll
    for (int i = 0; i < N; ++i) d[(long long)i * NP] = v[i];
  } else {  // A_IN_EPI
    if (!active) return;
    const long long cell = item;
    const int ps = A.pidx ? A.pidx[A.cell0 + cell] : 0;
    const T* lam = A.lam + (long long)ps * A.nq;
    const cx<T>* ac = A.src + cell * NPTS + p;
    const cx<T>* fh = A.src2 + cell * NPTS + p;
    const double lap_c = -(M_PI * M_PI) / (A.L * A.L);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      cx<T> x = A.no_gain ? cx<T>{(T)0, (T)0} : ac[(long long)i * NP];
      if (A.hb != 0.0) {
        const int ki = kfreq(i, N);
        const T lk = (T)(A.hb * lap_c * (double)(ki * ki + qrest));
        x = x + lk * fh[(long long)i * NP];
      }
      v[i] = inv_n * x;
    }
    fft_reg<N, +1>(v);
    cx<T>* d = A.dst + (2 * cell) * NPTS + p;
#pragma unroll
    for (int i = 0; i < N; ++i) d[(long long)i * NP] = v[i];
    if (A.with_loss) {
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int ki = kfreq(i, N);
        v[i] = (inv_n * lam[ki * ki + qrest]) * fh[(long long)i * NP];
      }
      fft_reg<N, +1>(v);
      cx<T>* d2 = A.dst + (2 * cell + 1) * NPTS + p;
#pragma unroll
      for (int i = 0; i < N; ++i) d2[(long long)i * NP] = v[i];
    }
  }
}

template <typename T, int N, int D, int MODE>
__global__ __launch_bounds__(64) void k_axis0(AArgs<T> A) {
  __shared__ __align__(16) char lds[Axis0Lds<T, N, D, MODE>::bytes];
  axis0_body<T, N, D, MODE>(A, blockIdx.x, blockIdx.y, lds);
}

// ---------------------------------------------------------------------------
// Remaining-axes pass, 2D: one thread per row (axis 1 pencil, contiguous).
template <typename T, int N, int MODE>
__device__ __forceinline__ void rest2_body(const BArgs<T>& B, long long bx) {
  constexpr long long NPTS = (long long)N * N;
  const long long gid = bx * 64 + threadIdx.x;
  const long long item = gid / N;
  const int row = (int)(gid % N);
  if (item >= B.items) return;
  cx<T> v[N];
  if (MODE == B_PAIR) {
    const long long cell = item / B.P;
    cx<T>* w = B.w + item * NPTS + (long long)row * N;
    const double* f = B.f + cell * NPTS + (long long)row * N;
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = w[i];
    fft_reg<N, +1>(v);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = (T)f[i] * v[i];
    fft_reg<N, -1>(v);
#pragma unroll
    for (int i = 0; i < N; ++i) w[i] = v[i];
  } else if (MODE == B_PRO) {
    const double* f = B.f + item * NPTS + (long long)row * N;
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = cx<T>{(T)f[i], (T)0};
    fft_reg<N, -1>(v);
    cx<T>* w = B.w + item * NPTS + (long long)row * N;
#pragma unroll
    for (int i = 0; i < N; ++i) w[i] = v[i];
  } else {  // B_EPI
    const long long cell = item;
    const double* f = B.f + cell * NPTS + (long long)row * N;
    double* q = B.q + cell * NPTS + (long long)row * N;
    T lr[N];
    if (B.with_loss) {
      const cx<T>* w2 = B.w + (2 * cell + 1) * NPTS + (long long)row * N;
#pragma unroll
      for (int i = 0; i < N; ++i) v[i] = w2[i];
      fft_reg<N, +1>(v);
#pragma unroll
      for (int i = 0; i < N; ++i) lr[i] = v[i].re;
    }
    const cx<T>* w1 = B.w + (2 * cell) * NPTS + (long long)row * N;
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = w1[i];
    fft_reg<N, +1>(v);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double out = (double)v[i].re;
      if (B.with_loss) out = out - f[i] * (double)lr[i];
      q[i] = out;
    }
  }
}

template <typename T, int N, int MODE>
__global__ __launch_bounds__(64) void k_rest2(BArgs<T> B) {
  rest2_body<T, N, MODE>(B, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Remaining-axes pass, 3D: one N x N slab (fixed axis-0 index) per sub-group of
// N threads, staged in LDS with rows padded to N + 1 (conflict-free columns
// and rows).
template <typename T, int N, int MODE>
__device__ __forceinline__ void rest3_body(const BArgs<T>& B, long long bx, char* lds) {
  constexpr int NP = N * N;
  constexpr long long NPTS = (long long)N * NP;
  constexpr int SPW = 64 / N;
  auto tile = reinterpret_cast<cx<T>(*)[N][N + 1]>(lds);
  const int tid = threadIdx.x;
  const int s = tid / N, l = tid % N;
  const long long gs = bx * SPW + s;
  const long long item = gs / N;
  const int i0 = (int)(gs % N);
  const bool active = item < B.items;
  cx<T> v[N];
  auto col_pass = [&](auto dir_tag) {
    constexpr int DIR = decltype(dir_tag)::value;
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = tile[s][i][l];
    fft_reg<N, DIR>(v);
#pragma unroll
    for (int i = 0; i < N; ++i) tile[s][i][l] = v[i];
  };
  using FWD = std::integral_constant<int, -1>;
  using INV = std::integral_constant<int, +1>;

  if (MODE == B_PAIR) {
    const long long cell = item / B.P;
    cx<T>* w = B.w + item * NPTS + (long long)i0 * NP;
    if (active) {
#pragma unroll 4
      for (int r = 0; r < N; ++r) tile[s][r][l] = w[r * N + l];
    }
    __syncthreads();
    if (active) col_pass(INV{});
    __syncthreads();
    if (active) {
      const double* f = B.f + cell * NPTS + (long long)i0 * NP + (long long)l * N;
#pragma unroll
      for (int i = 0; i < N; ++i) v[i] = tile[s][l][i];
      fft_reg<N, +1>(v);
#pragma unroll
      for (int i = 0; i < N; ++i) v[i] = (T)f[i] * v[i];
      fft_reg<N, -1>(v);
#pragma unroll
      for (int i = 0; i < N; ++i) tile[s][l][i] = v[i];
    }
    __syncthreads();
    if (active) col_pass(FWD{});
    __syncthreads();
    if (active) {
#pragma unroll 4
      for (int r = 0; r < N; ++r) w[r * N + l] = tile[s][r][l];
    }
  } else if (MODE == B_PRO) {
    if (active) {
      const double* f = B.f + item * NPTS + (long long)i0 * NP + (long long)l * N;
#pragma unroll
      for (int i = 0; i < N; ++i) v[i] = cx<T>{(T)f[i], (T)0};
      fft_reg<N, -1>(v);
#pragma unroll
      for (int i = 0; i < N; ++i) tile[s][l][i] = v[i];
    }
    __syncthreads();
    if (active) col_pass(FWD{});
    __syncthreads();
    if (active) {
      cx<T>* w = B.w + item * NPTS + (long long)i0 * NP;
#pragma unroll 4
      for (int r = 0; r < N; ++r) w[r * N + l] = tile[s][r][l];
    }
  } else {  // B_EPI
    const long long cell = item;
    T lr[N];
    if (B.with_loss) {
      if (active) {
        const cx<T>* w2 = B.w + (2 * cell + 1) * NPTS + (long long)i0 * NP;
#pragma unroll 4
        for (int r = 0; r < N; ++r) tile[s][r][l] = w2[r * N + l];
      }
      __syncthreads();
      if (active) col_pass(INV{});
      __syncthreads();
      if (active) {
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = tile[s][l][i];
        fft_reg<N, +1>(v);
#pragma unroll
        for (int i = 0; i < N; ++i) lr[i] = v[i].re;
      }
      __syncthreads();
    }
    if (active) {
      const cx<T>* w1 = B.w + (2 * cell) * NPTS + (long long)i0 * NP;
#pragma unroll 4
      for (int r = 0; r < N; ++r) tile[s][r][l] = w1[r * N + l];
    }
    __syncthreads();
    if (active) col_pass(INV{});
    __syncthreads();
    if (active) {
#pragma unroll
      for (int i = 0; i < N; ++i) v[i] = tile[s][l][i];
      fft_reg<N, +1>(v);
      const long long off = cell * NPTS + (long long)i0 * NP + (long long)l * N;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        double out = (double)v[i].re;
        if (B.with_loss) out = out - B.f[off + i] * (double)lr[i];
        B.q[off + i] = out;
      }
    }
  }
}

template <typename T, int N, int MODE>
__global__ __launch_bounds__(64) void k_rest3(BArgs<T> B) {
  __shared__ __align__(16) char lds[Rest3Lds<T, N>::bytes];
  rest3_body<T, N, MODE>(B, blockIdx.x, lds);
}

// ---------------------------------------------------------------------------
template <typename T, int N, int D, int MODE>
static int launch_a(const AArgs<T>& A, hipStream_t st) {
  using G = Geo<N, D>;
  dim3 grid(G::BLK, (unsigned)((A.items + G::SUBS - 1) / G::SUBS));
  hipLaunchKernelGGL((k_axis0<T, N, D, MODE>), grid, dim3(64), 0, st, A);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

// acc[c] (+)= sum over s of part[c][s], in split order.
template <typename T>
__global__ void k_split_add(const cx<T>* __restrict__ part, cx<T>* __restrict__ acc, long long nb,
                            int S, long long npts, int first) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nb * npts) return;
  const long long c = gid / npts, e = gid % npts;
  cx<T> sum = part[c * S * npts + e];
  for (int s = 1; s < S; ++s) sum = sum + part[(c * S + s) * npts + e];
  acc[gid] = first ? sum : acc[gid] + sum;
}

template <typename T, int N, int D, int MODE>
static int launch_b(const BArgs<T>& B, hipStream_t st) {
  if constexpr (D == 2) {
    const long long threads = B.items * N;
    hipLaunchKernelGGL((k_rest2<T, N, MODE>), dim3((unsigned)((threads + 63) / 64)), dim3(64), 0, st,
                       B);
  } else {
    constexpr int SPW = 64 / N;
    const long long slabs = B.items * N;
    hipLaunchKernelGGL((k_rest3<T, N, MODE>), dim3((unsigned)((slabs + SPW - 1) / SPW)), dim3(64), 0,
                       st, B);
  }
  IB_HIP(hipGetLastError());
  return IB_OK;
}

#define IB_TRY(x)               \
  do {                          \
    int _r = (x);               \
    if (_r != IB_OK) return _r; \
  } while (0)

// pairs per chunk a staged sub-batch is sized for (fewer pairs per chunk re-read and
// re-write the accumulator more often: 3D 32^3 FP64 1,140 cells/s at 2 pairs per
// chunk, 1,390 at 8, 1,470 at 16, 1,508 at 32, 1,391 at 64)
constexpr long long kChunkPairs = 32;

// sub-batch of cells and pairs per chunk, bounded by the scratch budget.  Every
// chunk re-reads and re-writes the sub-batch's accumulator, so the sub-batch is
// cut until a chunk holds kChunkPairs pairs (3D 32^3 FP64: 64 cells x 32 pairs
// per GiB instead of 1024 cells x 2 pairs)
static void staged_shape(const Plan* P, long long batch, size_t arr, long long* bs_out, long long* pc_out) {
  const int npairs = P->nr * P->m;
  long long bs = batch;
  const long long max_cells = (long long)std::max<size_t>(1, P->scratch_budget / (2 * arr));
  if (bs > max_cells) bs = max_cells;
  long long pc = P->chunk_pairs > 0 ? P->chunk_pairs
                                    : (long long)std::max<size_t>(1, P->scratch_budget / (bs * arr));
  if (P->chunk_pairs == 0) {
    const long long want = std::min<long long>(npairs, kChunkPairs);
    if (pc < want) {
      bs = std::min<long long>(batch, std::max<long long>(1, (long long)(P->scratch_budget / (want * arr))));
      pc = (long long)std::max<size_t>(1, P->scratch_budget / (bs * arr));
    }
  }
  if (pc > npairs) pc = npairs;
  if (pc < 2 && npairs >= 2) pc = 2;  // the epilogue needs 2 slots per cell
  *bs_out = bs;
  *pc_out = pc;
}

// pair-sum split of a chunk of np pairs over nb cells when the cells alone give fewer than
// 512 workgroups (wg0): S partial accumulators, at most 512 MiB of them (and, in a
// static-scratch call, what the region has left)
static long long split_factor(long long wg0, long long np, long long nb, size_t arr) {
  if (wg0 >= 512 || np < 4) return 1;
  long long S = std::min<long long>((1024 + wg0 - 1) / wg0, np / 2);
  S = std::min<long long>(S, std::max<long long>(1, (512LL << 20) / (nb * (long long)arr)));
  S = std::min<long long>(S, (long long)std::min<size_t>(arena_left() / ((size_t)nb * arr), (size_t)S));
  return S;
}

template <typename T, int N, int D>
static int collide_staged(Plan* P, const double* f, double* q, long long batch, const int* pidx,
                          int with_loss, double hb, int no_gain, hipStream_t st) {
  using G = Geo<N, D>;
  constexpr long long NPTS = G::NPTS;
  const size_t arr = (size_t)NPTS * sizeof(cx<T>);
  const int npairs = P->nr * P->m;
  long long bs, pc;
  staged_shape(P, batch, arr, &bs, &pc);
  IB_TRY(P->fhat.ensure(bs * arr));
  IB_TRY(P->acc.ensure(bs * arr));
  IB_TRY(P->work.ensure((size_t)bs * (size_t)std::max<long long>(pc, 2) * arr));
  cx<T>* fh = (cx<T>*)P->fhat.ptr;
  cx<T>* ac = (cx<T>*)P->acc.ptr;
  cx<T>* wk = (cx<T>*)P->work.ptr;

  for (long long c0 = 0; c0 < batch; c0 += bs) {
    const long long nb = std::min(bs, batch - c0);
    const double* fc = f + c0 * NPTS;
    double* qc = q + c0 * NPTS;
    AArgs<T> A{};
    A.sigma = P->d_sigma;
    A.rj = P->d_rj;
    A.wsig = P->d_wsig;
    A.crad = P->d_crad;
    A.a = P->d_a;
    A.spg = (const T*)P->d_spg;
    A.lam = (const T*)P->d_lam;
    A.nr = P->nr;
    A.m = P->m;
    A.nq = P->nq;
    A.L = P->L;
    A.hb = hb;
    A.with_loss = with_loss;
    A.no_gain = no_gain;
    A.pidx = pidx;
    A.cell0 = c0;
    BArgs<T> B{};
    B.f = fc;
    B.q = qc;
    B.with_loss = with_loss;
    // prologue: fhat = fft(f)
    B.w = wk;
    B.items = nb;
    B.P = 1;
    IB_TRY((launch_b<T, N, D, B_PRO>(B, st)));
    A.src = wk;
    A.dst = fh;
    A.items = nb;
    IB_TRY((launch_a<T, N, D, A_OUT_PRO>(A, st)));
    // quadrature pairs in chunks
    for (int p0 = 0; p0 < (no_gain ? 0 : npairs); p0 += (int)pc) {
      const int np = (int)std::min<long long>(pc, npairs - p0);
      A.src = fh;
      A.dst = wk;
      A.items = nb * np;
      A.P = np;
      A.p0 = p0;
      IB_TRY((launch_a<T, N, D, A_IN_PAIR>(A, st)));
      B.w = wk;
      B.items = nb * np;
      B.P = np;
      IB_TRY(P->timed_begin(st));
      IB_TRY((launch_b<T, N, D, B_PAIR>(B, st)));
      P->timed_end(st, nb * np);
      A.src = wk;
      A.dst = ac;
      A.first = (p0 == 0);
      // split the pair sum when the cells alone give fewer than 512 workgroups
      const long long wg0 = (long long)G::BLK * ((nb + G::SUBS - 1) / G::SUBS);
      const long long S = split_factor(wg0, np, nb, arr);
      if (S > 1) {
        IB_TRY(P->aux.ensure((size_t)(nb * S) * arr));
        A.part = (cx<T>*)P->aux.ptr;
        A.S = (int)S;
        A.items = nb * S;
        IB_TRY((launch_a<T, N, D, A_OUT_PAIR>(A, st)));
        const long long tot = nb * NPTS;
        hipLaunchKernelGGL(k_split_add<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                           (const cx<T>*)A.part, ac, nb, (int)S, NPTS, (int)A.first);
        IB_HIP(hipGetLastError());
        A.S = 1;
      } else {
        A.items = nb;
        IB_TRY((launch_a<T, N, D, A_OUT_PAIR>(A, st)));
      }
    }
    // epilogue: q = Re ifft(acc + hb lap fhat) - f Re ifft(Lambda fhat)
    A.src = ac;
    A.src2 = fh;
    A.dst = wk;
    A.items = nb;
    IB_TRY((launch_a<T, N, D, A_IN_EPI>(A, st)));
    B.w = wk;
    B.items = nb;
    B.P = 1;
    IB_TRY((launch_b<T, N, D, B_EPI>(B, st)));
  }
  return IB_OK;
}

template <typename T>
static int collide_staged_dispatch(Plan* P, const double* f, double* q, long long batch,
                                   const int* pidx, int with_loss, double hb, int no_gain,
                                   hipStream_t st) {
  const int key = P->dim * 100 + P->n;
  switch (key) {
    case 208: return collide_staged<T, 8, 2>(P, f, q, batch, pidx, with_loss, hb, no_gain, st);
    case 216: return collide_staged<T, 16, 2>(P, f, q, batch, pidx, with_loss, hb, no_gain, st);
    case 232: return collide_staged<T, 32, 2>(P, f, q, batch, pidx, with_loss, hb, no_gain, st);
    case 264: return collide_staged<T, 64, 2>(P, f, q, batch, pidx, with_loss, hb, no_gain, st);
    case 308: return collide_staged<T, 8, 3>(P, f, q, batch, pidx, with_loss, hb, no_gain, st);
    case 316: return collide_staged<T, 16, 3>(P, f, q, batch, pidx, with_loss, hb, no_gain, st);
    case 332: return collide_staged<T, 32, 3>(P, f, q, batch, pidx, with_loss, hb, no_gain, st);
    default:
      set_error("unsupported (dim, n) for the staged engine");
      return IB_ERR_UNSUPPORTED;
  }
}

template <int N, int D>
static void staged_wg(long long nb, long long* wg0) {
  using G = Geo<N, D>;
  *wg0 = (long long)G::BLK * ((nb + G::SUBS - 1) / G::SUBS);
}

static long long staged_wg0(const Plan* P, long long nb) {
  long long w = 0;
  switch (P->dim * 100 + P->n) {
    case 208: staged_wg<8, 2>(nb, &w); break;
    case 216: staged_wg<16, 2>(nb, &w); break;
    case 232: staged_wg<32, 2>(nb, &w); break;
    case 264: staged_wg<64, 2>(nb, &w); break;
    case 308: staged_wg<8, 3>(nb, &w); break;
    case 316: staged_wg<16, 3>(nb, &w); break;
    case 332: staged_wg<32, 3>(nb, &w); break;
    default: w = 0;
  }
  return w;
}

// grouped chunk size: all cells of a plan in one batch
static long long group_pc(const Plan* P, long long batch, size_t arr) {
  long long pc = std::max<long long>(2, (long long)(P->scratch_budget / (batch * arr)));
  return std::min<long long>(pc, std::max(P->nr * P->m, 2));
}

int staged_sizes(const Plan* P, long long batch, int grouped, ScratchSizes* z) {
  if (batch <= 0) return IB_OK;
  const size_t arr = (size_t)P->npts * (P->precision == IB_PREC_FP64 ? sizeof(cx<double>)
                                                                     : sizeof(cx<float>));
  const int npairs = P->nr * P->m;
  if (grouped && !staged_runs_alone(P, batch)) {
    const long long pc = group_pc(P, batch, arr);
    z->fhat = batch * arr;
    z->acc = batch * arr;
    z->work = (size_t)batch * pc * arr;
    // the grouped split depends on the other segments' workgroups: bound it by this
    // segment alone
    const long long S = split_factor(staged_wg0(P, batch), std::min<long long>(pc, npairs), batch, arr);
    z->aux = S > 1 ? (size_t)(batch * S) * arr : 0;
    return IB_OK;
  }
  long long bs, pc;
  staged_shape(P, batch, arr, &bs, &pc);
  z->fhat = bs * arr;
  z->acc = bs * arr;
  z->work = (size_t)bs * (size_t)std::max<long long>(pc, 2) * arr;
  // the largest split over the sub-batch sizes (full and last) and chunk sizes (full and last)
  size_t aux = 0;
  const long long nbs[2] = {std::min(bs, batch), batch % bs ? batch % bs : bs};
  const long long nps[2] = {std::min<long long>(pc, npairs), npairs % pc ? npairs % pc : pc};
  for (long long nb : nbs)
    for (long long np : nps) {
      const long long S = split_factor(staged_wg0(P, nb), np, nb, arr);
      if (S > 1) aux = std::max(aux, (size_t)(nb * S) * arr);
    }
  z->aux = aux;
  return IB_OK;
}

// The grouped path keeps all cells of a plan in one batch, so its chunks hold
// scratch_budget / (batch x transform bytes) pairs.  When that is below
// kChunkPairs the plan runs alone through collide_staged's sub-batches instead: a
// batch this large fills the device without the other plans of the group.
bool staged_runs_alone(const Plan* P, long long batch) {
  const size_t arr = (size_t)P->npts * (P->precision == IB_PREC_FP64 ? sizeof(cx<double>)
                                                                     : sizeof(cx<float>));
  const long long want = std::min<long long>((long long)P->nr * P->m, kChunkPairs);
  return P->chunk_pairs == 0 && (double)batch * (double)want * (double)arr > (double)P->scratch_budget;
}

int collide_plain(Plan* P, const double* f, double* q, long long batch, const int* pidx,
                  int with_loss, double hb, int no_gain, hipStream_t st) {
  if (P->precision == IB_PREC_FP64)
    return collide_staged_dispatch<double>(P, f, q, batch, pidx, with_loss, hb, no_gain, st);
  return collide_staged_dispatch<float>(P, f, q, batch, pidx, with_loss, hb, no_gain, st);
}

// ---------------------------------------------------------------------------
// Grouped persistent launch: one kernel per stage serves several plans of
// different grid shapes (an uneven batch).  The segment table travels in the
// kernel arguments; each 64-thread block walks a strided range of the
// concatenated virtual blocks of all segments and dispatches on the segment's
// shape to the same pass bodies the per-shape kernels use.
constexpr int kMaxSeg = 8;

template <typename T>
struct GSeg {
  int kind;
  long long boff, blocks;
  AArgs<T> A;
  BArgs<T> B;
};

template <typename T>
struct GSegs {
  int nseg;
  long long total;
  GSeg<T> s[kMaxSeg];
};

#define IB_KINDS(X) X(8, 2) X(16, 2) X(32, 2) X(64, 2) X(8, 3) X(16, 3) X(32, 3)

template <typename T, int MODE, bool PASS_A>
__global__ __launch_bounds__(64) void k_group(GSegs<T> G) {
  extern __shared__ __align__(16) char lds[];
  for (long long vb = blockIdx.x; vb < G.total; vb += gridDim.x) {
    __syncthreads();  // the previous item's LDS readers are done
    int s = 0;
    while (s + 1 < G.nseg && G.s[s + 1].boff <= vb) ++s;
    const long long local = vb - G.s[s].boff;
    if constexpr (PASS_A) {
      switch (G.s[s].kind) {
#define IB_CASE_A(NN, DD)                                                          \
  case DD * 100 + NN: {                                                            \
    constexpr int BLK = Geo<NN, DD>::BLK;                                          \
    axis0_body<T, NN, DD, MODE>(G.s[s].A, (int)(local % BLK), local / BLK, lds);   \
  } break;
        IB_KINDS(IB_CASE_A)
#undef IB_CASE_A
        default:
          break;
      }
    } else {
      switch (G.s[s].kind) {
        case 208: rest2_body<T, 8, MODE>(G.s[s].B, local); break;
        case 216: rest2_body<T, 16, MODE>(G.s[s].B, local); break;
        case 232: rest2_body<T, 32, MODE>(G.s[s].B, local); break;
        case 264: rest2_body<T, 64, MODE>(G.s[s].B, local); break;
        case 308: rest3_body<T, 8, MODE>(G.s[s].B, local, lds); break;
        case 316: rest3_body<T, 16, MODE>(G.s[s].B, local, lds); break;
        case 332: rest3_body<T, 32, MODE>(G.s[s].B, local, lds); break;
        default: break;
      }
    }
  }
}

// virtual blocks and LDS bytes of one segment of a pass
template <typename T, int MODE, bool PASS_A>
static void seg_shape(int kind, long long items, long long* blocks, size_t* lds) {
  switch (kind) {
#define IB_SHAPE(NN, DD)                                                                 \
  case DD * 100 + NN:                                                                    \
    if (PASS_A) {                                                                        \
      using Gm = Geo<NN, DD>;                                                            \
      *blocks = (long long)Gm::BLK * ((items + Gm::SUBS - 1) / Gm::SUBS);                \
      *lds = Axis0Lds<T, NN, DD, MODE>::bytes;                                           \
    } else if (DD == 2) {                                                                \
      *blocks = (items * NN + 63) / 64;                                                  \
      *lds = 0;                                                                          \
    } else {                                                                             \
      *blocks = (items * NN + (64 / NN) - 1) / (64 / NN);                                \
      *lds = Rest3Lds<T, NN>::bytes;                                                     \
    }                                                                                    \
    break;
    IB_KINDS(IB_SHAPE)
#undef IB_SHAPE
    default:
      *blocks = 0;
      *lds = 0;
  }
}

template <typename T, int MODE, bool PASS_A>
static int launch_group(GSegs<T>& G, hipStream_t st) {
  size_t lds = 0;
  long long total = 0;
  for (int i = 0; i < G.nseg; ++i) {
    long long nb;
    size_t l;
    seg_shape<T, MODE, PASS_A>(G.s[i].kind, PASS_A ? G.s[i].A.items : G.s[i].B.items, &nb, &l);
    G.s[i].boff = total;
    G.s[i].blocks = nb;
    total += nb;
    lds = std::max(lds, l);
  }
  G.total = total;
  if (total == 0) return IB_OK;
  static size_t attr_bytes = 0;
  if (lds > 65536 && lds > attr_bytes) {
    IB_HIP(hipFuncSetAttribute((const void*)k_group<T, MODE, PASS_A>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr_bytes = lds;
  }
  const unsigned grid = (unsigned)std::min<long long>(total, 4096);
  hipLaunchKernelGGL((k_group<T, MODE, PASS_A>), dim3(grid), dim3(64), lds, st, G);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

template <typename T>
static int collide_group_t(Plan** P, int np, const double* const* f, double* const* q,
                           const long long* batch, const int* const* pidx, int with_loss,
                           double hb, hipStream_t st) {
  struct S {
    Plan* p;
    long long nb, pc;
    int npairs;
    size_t arr;
  };
  std::vector<S> segs;
  for (int i = 0; i < np; ++i) {
    if (batch[i] == 0) continue;
    const size_t arr = (size_t)P[i]->npts * sizeof(cx<T>);
    const long long pc = group_pc(P[i], batch[i], arr);
    const int npairs = P[i]->nr * P[i]->m;
    IB_TRY(P[i]->fhat.ensure(batch[i] * arr));
    IB_TRY(P[i]->acc.ensure(batch[i] * arr));
    IB_TRY(P[i]->work.ensure((size_t)batch[i] * pc * arr));
    segs.push_back(S{P[i], batch[i], pc, npairs, arr});
  }
  const int ns = (int)segs.size();
  if (ns == 0) return IB_OK;
  if (ns > kMaxSeg) {
    set_error("at most 8 plans per grouped launch");
    return IB_ERR_ARG;
  }
  // index of each non-empty segment in the caller's arrays
  std::vector<int> idx;
  for (int i = 0; i < np; ++i)
    if (batch[i] != 0) idx.push_back(i);
  auto base = [&](int k) {
    Plan* p = segs[k].p;
    GSeg<T> g{};
    g.kind = p->dim * 100 + p->n;
    AArgs<T>& A = g.A;
    A.sigma = p->d_sigma;
    A.rj = p->d_rj;
    A.wsig = p->d_wsig;
    A.crad = p->d_crad;
    A.a = p->d_a;
    A.spg = (const T*)p->d_spg;
    A.lam = (const T*)p->d_lam;
    A.nr = p->nr;
    A.m = p->m;
    A.nq = p->nq;
    A.L = p->L;
    A.hb = hb;
    A.with_loss = with_loss;
    A.pidx = pidx ? pidx[idx[k]] : nullptr;
    A.cell0 = 0;
    g.B.f = f[idx[k]];
    g.B.q = q[idx[k]];
    g.B.with_loss = with_loss;
    return g;
  };
  GSegs<T> G{};
  // prologue
  G.nseg = ns;
  for (int k = 0; k < ns; ++k) {
    G.s[k] = base(k);
    G.s[k].B.w = (cx<T>*)segs[k].p->work.ptr;
    G.s[k].B.items = segs[k].nb;
    G.s[k].B.P = 1;
  }
  IB_TRY((launch_group<T, B_PRO, false>(G, st)));
  for (int k = 0; k < ns; ++k) {
    G.s[k].A.src = (const cx<T>*)segs[k].p->work.ptr;
    G.s[k].A.dst = (cx<T>*)segs[k].p->fhat.ptr;
    G.s[k].A.items = segs[k].nb;
  }
  IB_TRY((launch_group<T, A_OUT_PRO, true>(G, st)));
  // quadrature pairs: round r serves chunk r of every plan that has one
  for (int r = 0;; ++r) {
    GSegs<T> Gi{}, Gb{}, Go{};
    int owner[kMaxSeg];  // segment of this round -> index into segs
    int live = 0;
    for (int k = 0; k < ns; ++k) {
      const long long p0 = (long long)r * segs[k].pc;
      if (p0 >= segs[k].npairs) continue;
      const int npr = (int)std::min<long long>(segs[k].pc, segs[k].npairs - p0);
      GSeg<T> g = base(k);
      g.A.src = (const cx<T>*)segs[k].p->fhat.ptr;
      g.A.dst = (cx<T>*)segs[k].p->work.ptr;
      g.A.items = segs[k].nb * npr;
      g.A.P = npr;
      g.A.p0 = (int)p0;
      Gi.s[live] = g;
      g.B.w = (cx<T>*)segs[k].p->work.ptr;
      g.B.items = segs[k].nb * npr;
      g.B.P = npr;
      Gb.s[live] = g;
      g.A.src = (const cx<T>*)segs[k].p->work.ptr;
      g.A.dst = (cx<T>*)segs[k].p->acc.ptr;
      g.A.items = segs[k].nb;
      g.A.first = (r == 0);
      Go.s[live] = g;
      owner[live] = k;
      ++live;
    }
    if (live == 0) break;
    Gi.nseg = Gb.nseg = Go.nseg = live;
    IB_TRY((launch_group<T, A_IN_PAIR, true>(Gi, st)));
    IB_TRY((launch_group<T, B_PAIR, false>(Gb, st)));
    // split the pair sums (as in collide_staged) when all segments' cells together give
    // fewer than 512 workgroups
    long long wg0 = 0;
    for (int k = 0; k < live; ++k) {
      long long nbk;
      size_t l;
      seg_shape<T, A_OUT_PAIR, true>(Go.s[k].kind, Go.s[k].A.items, &nbk, &l);
      wg0 += nbk;
    }
    int split[kMaxSeg];
    for (int k = 0; k < live; ++k) {
      split[k] = 1;
      if (wg0 >= 512) continue;
      Plan* pk = segs[owner[k]].p;
      const long long nbk = Go.s[k].A.items;
      const long long arr = pk->npts * (long long)sizeof(cx<T>);
      long long S = std::min<long long>((1024 + wg0 - 1) / wg0, Go.s[k].A.P / 2);
      S = std::min<long long>(S, std::max<long long>(1, (512LL << 20) / (nbk * arr)));
      if (S < 2) continue;
      IB_TRY(pk->aux.ensure((size_t)(nbk * S * arr)));
      Go.s[k].A.part = (cx<T>*)pk->aux.ptr;
      Go.s[k].A.S = (int)S;
      Go.s[k].A.items = nbk * S;
      split[k] = (int)S;
    }
    IB_TRY((launch_group<T, A_OUT_PAIR, true>(Go, st)));
    for (int k = 0; k < live; ++k) {
      if (split[k] == 1) continue;
      const long long nbk = Go.s[k].A.items / split[k], pts = segs[owner[k]].p->npts;
      const long long tot = nbk * pts;
      hipLaunchKernelGGL(k_split_add<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                         (const cx<T>*)Go.s[k].A.part, Go.s[k].A.dst, nbk, split[k], pts,
                         Go.s[k].A.first);
      IB_HIP(hipGetLastError());
    }
  }
  // epilogue
  for (int k = 0; k < ns; ++k) {
    G.s[k] = base(k);
    G.s[k].A.src = (const cx<T>*)segs[k].p->acc.ptr;
    G.s[k].A.src2 = (const cx<T>*)segs[k].p->fhat.ptr;
    G.s[k].A.dst = (cx<T>*)segs[k].p->work.ptr;
    G.s[k].A.items = segs[k].nb;
  }
  IB_TRY((launch_group<T, A_IN_EPI, true>(G, st)));
  for (int k = 0; k < ns; ++k) {
    G.s[k].B.w = (cx<T>*)segs[k].p->work.ptr;
    G.s[k].B.items = segs[k].nb;
    G.s[k].B.P = 1;
  }
  IB_TRY((launch_group<T, B_EPI, false>(G, st)));
  return IB_OK;
}

int collide_group(Plan** P, int np, const double* const* f, double* const* q,
                  const long long* batch, const int* const* pidx, int with_loss, double hb,
                  hipStream_t st) {
  if (P[0]->precision == IB_PREC_FP64)
    return collide_group_t<double>(P, np, f, q, batch, pidx, with_loss, hb, st);
  return collide_group_t<float>(P, np, f, q, batch, pidx, with_loss, hb, st);
}

// ---------------------------------------------------------------------------
// FP64 prologue / epilogue of the fused FP32 engines (any grid shape).
template <int N, int D>
static int prologue_fp64(Plan* P, const double* f, long long nb, void* fhat64, void* work) {
  hipStream_t st = P->stream;
  BArgs<double> B{};
  B.f = f;
  B.w = (cx<double>*)work;
  B.items = nb;
  B.P = 1;
  IB_TRY((launch_b<double, N, D, B_PRO>(B, st)));
  AArgs<double> A{};
  A.src = (const cx<double>*)work;
  A.dst = (cx<double>*)fhat64;
  A.items = nb;
  IB_TRY((launch_a<double, N, D, A_OUT_PRO>(A, st)));
  return IB_OK;
}

template <int N, int D>
static int epilogue_fp64(Plan* P, const double* f, double* q, long long nb, void* acc64,
                         void* fhat64, void* work, int with_loss, double hb, const int* pidx) {
  hipStream_t st = P->stream;
  AArgs<double> A{};
  A.src = (const cx<double>*)acc64;
  A.src2 = (const cx<double>*)fhat64;
  A.dst = (cx<double>*)work;
  A.items = nb;
  A.lam = P->d_lam64;
  A.nq = P->nq;
  A.pidx = pidx;
  A.L = P->L;
  A.hb = hb;
  A.with_loss = with_loss;
  IB_TRY((launch_a<double, N, D, A_IN_EPI>(A, st)));
  BArgs<double> B{};
  B.w = (cx<double>*)work;
  B.f = f;
  B.q = q;
  B.items = nb;
  B.P = 1;
  B.with_loss = with_loss;
  IB_TRY((launch_b<double, N, D, B_EPI>(B, st)));
  return IB_OK;
}

int staged_prologue_fp64(Plan* P, const double* f, long long nb, void* fhat64, void* work) {
  switch (P->dim * 100 + P->n) {
    case 216: return prologue_fp64<16, 2>(P, f, nb, fhat64, work);
    case 232: return prologue_fp64<32, 2>(P, f, nb, fhat64, work);
    case 264: return prologue_fp64<64, 2>(P, f, nb, fhat64, work);
    case 316: return prologue_fp64<16, 3>(P, f, nb, fhat64, work);
    case 332: return prologue_fp64<32, 3>(P, f, nb, fhat64, work);
    default:
      set_error("no FP64 prologue for this shape");
      return IB_ERR_UNSUPPORTED;
  }
}

int staged_epilogue_fp64(Plan* P, const double* f, double* q, long long nb, void* acc64,
                         void* fhat64, void* work, int with_loss, double hb, const int* pidx) {
  switch (P->dim * 100 + P->n) {
    case 216: return epilogue_fp64<16, 2>(P, f, q, nb, acc64, fhat64, work, with_loss, hb, pidx);
    case 232: return epilogue_fp64<32, 2>(P, f, q, nb, acc64, fhat64, work, with_loss, hb, pidx);
    case 264: return epilogue_fp64<64, 2>(P, f, q, nb, acc64, fhat64, work, with_loss, hb, pidx);
    case 316: return epilogue_fp64<16, 3>(P, f, q, nb, acc64, fhat64, work, with_loss, hb, pidx);
    case 332: return epilogue_fp64<32, 3>(P, f, q, nb, acc64, fhat64, work, with_loss, hb, pidx);
    default:
      set_error("no FP64 epilogue for this shape");
      return IB_ERR_UNSUPPORTED;
  }
}

// ---------------------------------------------------------------------------
// Table construction: sp(a_j r_j |k|) per (parameter set, j, |k|^2) and the
// combined loss symbol Lambda(|k|^2) = sum_j |S| r_j^(gamma+d-1) w_j sp(pi r_j |k| / L).
template <typename T>
__global__ void k_tables(T* spg, T* lam, double* lam64, const double* a, const double* rj,
                         const double* crad, int n_param, int nr, int nq, int dim, double L,
                         double S) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long ng = (long long)n_param * nr * nq;
  if (gid < ng) {
    const int q = (int)(gid % nq);
    const long long pj = gid / nq;
    const int j = (int)(pj % nr);
    const double x = (a[pj] * rj[j]) * sqrt((double)q);
    spg[gid] = (T)sp_func(dim, x);
  } else if (gid < ng + (long long)n_param * nq) {
    const long long pq = gid - ng;
    const int ps = (int)(pq / nq), q = (int)(pq % nq);
    const double kn = sqrt((double)q);
    double s = 0.0;
    for (int j = 0; j < nr; ++j) {
      const double lk = S * crad[(long long)ps * nr + j];
      s += lk * sp_func(dim, (M_PI * rj[j]) * kn / L);
    }
    lam[pq] = (T)s;
    lam64[pq] = s;
  }
}

int launch_tables(Plan* P, const double* d_wr, hipStream_t st) {
  (void)d_wr;
  const long long total = (long long)P->n_param * (P->nr + 1) * P->nq;
  const unsigned blocks = (unsigned)((total + 255) / 256);
  if (P->precision == IB_PREC_FP64)
    hipLaunchKernelGGL(k_tables<double>, dim3(blocks), dim3(256), 0, st, (double*)P->d_spg,
                       (double*)P->d_lam, P->d_lam64, P->d_a, P->d_rj, P->d_crad, P->n_param, P->nr,
                       P->nq, P->dim, P->L, P->sphere_area);
  else
    hipLaunchKernelGGL(k_tables<float>, dim3(blocks), dim3(256), 0, st, (float*)P->d_spg,
                       (float*)P->d_lam, P->d_lam64, P->d_a, P->d_rj, P->d_crad, P->n_param, P->nr,
                       P->nq, P->dim, P->L, P->sphere_area);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

}  // namespace ib
