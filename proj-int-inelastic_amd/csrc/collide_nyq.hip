// Exact Nyquist-plane term of the real-data FP32 form of Q(f, f) on 3D 32^3 grids.
//
// The reference transforms C2C (src/collision/inelastic_collisions.py:208-227):
// per quadrature pair Qhat += G * fft(f * ifft(E * fhat)), then Q = Re ifft(Qhat).
// With fhat Hermitian, split the separable phase factors into Hermitian and
// anti-Hermitian parts, E = E_H + E_A, G = G_H + G_A.  Every 1D factor is
// Hermitian except its Nyquist entry e(16) = c + i s, so E_A (and G_A) vanish
// off the three Nyquist planes kx = 16, ky = 16, kz = 16; on them
//   E_A(k) = i Im(prod_{d in S(k)} e_d(16)) prod_{d not in S(k)} e_d(k_d),
// S(k) the axes where k_d = 16.  Re ifft(.) keeps the Hermitian part of Qhat:
//   Herm(G * fft(f * ifft(E fhat))) = G_H * fft(f * ifft(E_H fhat))      (k_fused3)
//                                   + G_A * fft(f * ifft(E_A fhat))      (this file)
// The second term lives on the Nyquist planes only.  b = ifft(E_A fhat) = i beta with
//   beta(x,y,z) = s_x (-1)^x P1(y,z) + s_y (-1)^y P2(x,z) + s_z (-1)^z P3(x,y)
//                 - s_x s_y s_z fhat(16,16,16) (-1)^(x+y+z),
// P1 = ifft2(h_y h_z fhat(16,.,.)), P2 = ifft2(h_x h_z fhat(.,16,.)),
// P3 = ifft2(h_x h_y fhat(.,.,16)) (real: their inputs are 2D-Hermitian; h_d is
// e_d with its Nyquist entry replaced by c_d).  fft(f b) on the plane kx = 16 is
// i fft2(R1) with R1(y,z) = sum_x (-1)^x f beta, and likewise R2 (sum over y) on
// ky = 16, R3 (sum over z) on kz = 16.  The term at plane point k is
//   -w_sigma Im(prod_{S(k)} g_d(16)) prod_{not S(k)} g_d(k_d) V(k) * r^(g+d-1) w_j sp(a r |k|).
//
// Work decomposition: one 512-thread workgroup per (cell, radial node j) loops over
// the M angles.  Per angle: wave 0 runs the three 2D inverse transforms (P1 + i P2
// packed in one complex transform, P3 in a second) with one 32-point pencil per lane;
// all waves form u = sigma_y sigma_z f beta along x pencils (thread (z, y) and
// (z, y + 16)) and reduce it: over x in registers (R1), over z across the 32 lanes of a
// half-wave (R3: a lane exchange, then DPP mirrors), over y in registers, across the
// half-waves (a lane exchange) and across waves through LDS (R2); wave 0 runs the
// forward transforms (R1 + i R2 packed, R3); all threads weight the plane points and
// sum over angles in registers.  The per-(cell, j) plane terms go to a buffer that
// k_f3n_reduce sums over j in FP64 into the full spectrum.
#include "common.h"
#include "plan.h"

namespace ib {

namespace f3n {
constexpr int N = 32;
constexpr int H = N / 2;
constexpr int PAD = N + 1;
constexpr int THREADS = 512;
constexpr int PLANE = N * N;
constexpr int SLOTS = 3 * PLANE;        // plane points per cell, lines counted on each plane
constexpr int PER = SLOTS / THREADS;    // slots per thread
}  // namespace f3n

struct F3NArgs {
  const cx<float>* xp;  // [cell][3][32][32]: fhat/n on kx = 16 (ky, kz), ky = 16 (kx, kz), kz = 16 (kx, ky)
  const float* f32;     // [cell][32][32][32]
  cx<float>* out;       // [cell][nr][3][32][32]: per-(cell, j) plane terms (0 on repeated line points)
  const int* pidx;
  long long batch;
  const double* sigma;
  const double* rj;
  const double* wsig;
  const double* crad;
  const double* a;
  const float* spg;
  int nr, m, nq;
  double L;
};

// A slot owns the plane point once: plane 1 skips kx = 16, plane 2 skips kx = 16 and ky = 16.
__host__ __device__ __forceinline__ bool nyq_slot_owned(int plane, int a, int b) {
  return plane == 0 || (plane == 1 && a != f3n::H) || (plane == 2 && a != f3n::H && b != f3n::H);
}

// (kx, ky, kz) of a slot
__host__ __device__ __forceinline__ void nyq_slot_k(int plane, int a, int b, int* k) {
  k[0] = plane == 0 ? f3n::H : a;
  k[1] = plane == 0 ? a : (plane == 1 ? f3n::H : b);
  k[2] = plane == 2 ? f3n::H : b;
}

__device__ __forceinline__ float dpp_f(float v, int ctrl_id) {
  const int iv = __builtin_bit_cast(int, v);
  int r;
  switch (ctrl_id) {
    case 0: r = __builtin_amdgcn_mov_dpp(iv, 0x140, 0xF, 0xF, false); break;  // row_mirror: r <-> 15 - r
    case 1: r = __builtin_amdgcn_mov_dpp(iv, 0x141, 0xF, 0xF, false); break;  // row_half_mirror: r <-> 7 - r
    case 2: r = __builtin_amdgcn_mov_dpp(iv, 0x4E, 0xF, 0xF, false); break;   // quad_perm [2,3,0,1]
    default: r = __builtin_amdgcn_mov_dpp(iv, 0xB1, 0xF, 0xF, false); break;  // quad_perm [1,0,3,2]
  }
  return __builtin_bit_cast(float, r);
}

// lo(lane) + lo(lane ^ MASK) on lanes whose MASK bit is clear, hi(lane) + hi(lane ^ MASK) on
// the others (ds_bpermute; this compiler returns only the first of the two registers that
// v_permlane16/32_swap exchange, so those instructions are not used)
template <int MASK>
__device__ __forceinline__ float xchg_add(float lo, float hi, int lane) {
  asm volatile("" : "+v"(lo), "+v"(hi));  // values, not a lane-dependent load index
  const bool up = (lane & MASK) != 0;
  const float keep = up ? hi : lo;
  const float send = up ? lo : hi;
  return keep + __shfl_xor(send, MASK);
}

// Sum of v over the 32 lanes of a half-wave by recursive halving: on return v[0] of
// lane l holds sum_lanes v[l & 31].  Stage order: lane bit 4 (ds_bpermute), then
// bits 3, 2, 1, 0 by DPP (mirror partners differ in the stage's bit).
__device__ __forceinline__ void halfwave_reduce32(float* v, int lane) {
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = xchg_add<16>(v[i], v[i + 16], lane);
#pragma unroll
  for (int st = 0; st < 4; ++st) {
    const int half = 8 >> st;
    const bool up = (lane >> (3 - st)) & 1;
#pragma unroll
    for (int i = 0; i < half; ++i) {
      // opaque copies: a select between two loads of v would otherwise be folded into one
      // load at a lane-dependent index, which puts v in scratch memory
      float lo = v[i], hi = v[i + half];
      asm volatile("" : "+v"(lo), "+v"(hi));
      const float keep = up ? hi : lo;
      const float send = up ? lo : hi;
      v[i] = keep + dpp_f(send, st);
    }
  }
}

// Orders a wave's LDS writes before its later reads of other lanes' data.
__device__ __forceinline__ void nyq_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 1D factor with its Nyquist entry replaced by the real part (the Hermitian part)
__device__ __forceinline__ cx<float> hfac(const cx<float>* row, int k) {
  const cx<float> v = row[k];
  return k == f3n::H ? cx<float>{v.re, 0.f} : v;
}

__global__ __launch_bounds__(512, 2) void k_f3_nyq(F3NArgs A) {
  using namespace f3n;
  __shared__ cx<float> tE[2][3][N];
  __shared__ cx<float> tG[2][3][N];
  __shared__ cx<float> Za[N][PAD];
  __shared__ cx<float> Zb[N][PAD];
  __shared__ float RP[8][N][N];
  const int tid = threadIdx.x;
  const long long id = blockIdx.x;
  const int c8 = (int)(id % 8);
  const int j = (int)((id / 8) % A.nr);
  const long long cell = (id / (8LL * A.nr)) * 8 + c8;
  if (cell >= A.batch) return;  // uniform over the workgroup
  const int ps = A.pidx ? A.pidx[cell] : 0;
  const double aj = A.a[ps * A.nr + j] * A.rj[j];
  const double rpl_pi = A.rj[j] / A.L;
  const double aj_pi = aj / M_PI;
  const cx<float>* X1 = A.xp + cell * SLOTS;
  const cx<float>* X2 = X1 + PLANE;
  const cx<float>* X3 = X2 + PLANE;
  const float* fc = A.f32 + cell * (long long)PLANE * N;
  const float fppp = X1[H * N + H].re;  // fhat(16, 16, 16) / n (real)

  auto make_tables = [&](int mm, int b) {
    if (tid < 6 * N) {
      const int which = tid / (3 * N), ax = (tid % (3 * N)) / N, k = tid % N;
      const double s = A.sigma[mm * 3 + ax];
      const double kf = (double)kfreq(k, N);
      double t = which == 0 ? -(rpl_pi * s * kf) : (aj_pi * s * kf);
      t -= 2.0 * rint(0.5 * t);  // half turns reduced to [-1, 1] in FP64
      float sn, cs;
      sincospif((float)t, &sn, &cs);
      (which ? tG : tE)[b][ax][k] = cx<float>{cs, sn};
    }
  };

  const int lane = tid & 63, wv = tid >> 6;
  const int z = tid & 31, yq = tid >> 5;  // N^3 roles: x pencils (y, z) and (y + 16, z)
  const float sgz = (z & 1) ? -1.f : 1.f, sgy = (yq & 1) ? -1.f : 1.f;

  cx<float> acc[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) acc[i] = cx<float>{0.f, 0.f};

  make_tables(0, 0);
  __syncthreads();
  for (int mm = 0; mm < A.m; ++mm) {
    const int b = mm & 1;
    // A (wave 0): P1 + i P2 = ifft2(h_z (h_y X1 + i h_x X2)), P3 = ifft2(h_x h_y X3)
    if (tid < 64) {
      const int a = tid & 31;
      cx<float>(*Z)[PAD] = tid < 32 ? Za : Zb;
      cx<float> v[N];
      // rows in blocks of 8 points, so that at most one block's loads are in flight
      if (tid < 32) {
        const cx<float> hy = hfac(tE[b][1], a), hx = hfac(tE[b][0], a);
#pragma unroll
        for (int c0 = 0; c0 < N; c0 += 8) {
#pragma unroll
          for (int c = c0; c < c0 + 8; ++c) {
            const cx<float> t1 = hy * X1[a * N + c], t2 = hx * X2[a * N + c];
            v[c] = hfac(tE[b][2], c) * cx<float>{t1.re - t2.im, t1.im + t2.re};
          }
          __builtin_amdgcn_sched_barrier(0);
        }
      } else {
        const cx<float> hx = hfac(tE[b][0], a);
#pragma unroll
        for (int c0 = 0; c0 < N; c0 += 8) {
#pragma unroll
          for (int c = c0; c < c0 + 8; ++c) v[c] = (hx * hfac(tE[b][1], c)) * X3[a * N + c];
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      fft_reg<N, +1>(v);
#pragma unroll
      for (int c = 0; c < N; ++c) Z[a][c] = v[c];
      nyq_wave_sync();
#pragma unroll
      for (int r = 0; r < N; ++r) v[r] = Z[r][a];
      fft_reg<N, +1>(v);
#pragma unroll
      for (int r = 0; r < N; ++r) Z[r][a] = v[r];
    }
    __syncthreads();
    // B (all): u = sigma_y sigma_z f beta on the pencils (y, z), (y + 16, z); reductions
    float r1[2], r3[2];
    {
      const float sx = tE[b][0][H].im, sy = tE[b][1][H].im, sz = tE[b][2][H].im;
      const float C = sx * sy * sz * fppp;
      const float K2 = sy * sgz, K3 = sz * sgy;
      float u[2][N];
      // roles from an opaque copy of tid, so that the pencils' global and LDS addresses
      // are formed next to their loads each pair instead of occupying registers across
      // the pair loop
      int tq = tid;
      asm volatile("" : "+v"(tq));
      const int zz = tq & 31, y0 = tq >> 5, y1 = y0 + 16;
      // sigma_y sigma_z (s_x P1 - C sigma_y sigma_z) of the two pencils
      const float K1a = sgy * sgz * sx * Za[y0][zz].re - C;
      const float K1b = sgy * sgz * sx * Za[y1][zz].re - C;
      const int toff = y0 * N + zz;
      float s1a = 0.f, s1b = 0.f;
      // blocks of 8 x, so that the loads of one block are in flight together and no more
#pragma unroll
      for (int x0 = 0; x0 < N; x0 += 8) {
#pragma unroll
        for (int x = x0; x < x0 + 8; ++x) {
          const float p2 = Za[x][zz].im;
          const float ta = fmaf(K2, p2, (x & 1) ? -K1a : K1a);
          const float tb = fmaf(K2, p2, (x & 1) ? -K1b : K1b);
          u[0][x] = fc[x * PLANE + toff] * fmaf(K3, Zb[x][y0].re, ta);
          u[1][x] = fc[x * PLANE + toff + 16 * N] * fmaf(K3, Zb[x][y1].re, tb);
          s1a = (x & 1) ? s1a - u[0][x] : s1a + u[0][x];
          s1b = (x & 1) ? s1b - u[1][x] : s1b + u[1][x];
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      r1[0] = sgy * sgz * s1a;
      r1[1] = sgy * sgz * s1b;
      // R2 over y: the two pencils in registers, the half-waves by permlane32, waves via LDS
      float t2[H];
#pragma unroll
      for (int i = 0; i < H; ++i) t2[i] = xchg_add<32>(u[0][i] + u[1][i], u[0][i + H] + u[1][i + H], lane);
#pragma unroll
      for (int i = 0; i < H; ++i) RP[wv][(lane >> 5) * H + i][z] = t2[i];
      // R3 over z: across the 32 lanes of the half-wave
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        halfwave_reduce32(u[p], lane);
        r3[p] = sgy * u[p][0];
      }
    }
    __syncthreads();
    // C (all): Za <- R1 + i R2 ((y, z) and (x, z) share the index), Zb <- R3 (x, y)
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int y = yq + 16 * p;
      float s2 = 0.f;
#pragma unroll
      for (int w = 0; w < 8; ++w) s2 += RP[w][y][z];
      Za[y][z] = cx<float>{r1[p], sgz * s2};
      Zb[z][y] = cx<float>{r3[p], 0.f};
    }
    __syncthreads();
    // D (wave 0): forward 2D transforms; waves 1-3: the next angle's phase tables
    if (tid < 64) {
      const int a = tid & 31;
      cx<float>(*Z)[PAD] = tid < 32 ? Za : Zb;
      cx<float> v[N];
#pragma unroll
      for (int c = 0; c < N; ++c) v[c] = Z[a][c];
      fft_reg<N, -1>(v);
#pragma unroll
      for (int c = 0; c < N; ++c) Z[a][c] = v[c];
      nyq_wave_sync();
#pragma unroll
      for (int r = 0; r < N; ++r) v[r] = Z[r][a];
      fft_reg<N, -1>(v);
#pragma unroll
      for (int r = 0; r < N; ++r) Z[r][a] = v[r];
    } else if (mm + 1 < A.m) {
      int tq = tid - 64;
      if (tq < 6 * N) {
        const int which = tq / (3 * N), ax = (tq % (3 * N)) / N, k = tq % N;
        const double s = A.sigma[(mm + 1) * 3 + ax];
        const double kf = (double)kfreq(k, N);
        double t = which == 0 ? -(rpl_pi * s * kf) : (aj_pi * s * kf);
        t -= 2.0 * rint(0.5 * t);
        float sn, cs;
        sincospif((float)t, &sn, &cs);
        (which ? tG : tE)[b ^ 1][ax][k] = cx<float>{cs, sn};
      }
    }
    __syncthreads();
    // E (all): V on the plane points, weighted by -G_A / (i r^(g+d-1) w_j sp), summed over angles
    {
      const float ws = (float)A.wsig[mm];
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int s = tid + i * THREADS;
        const int plane = s / PLANE, a = (s / N) % N, c = s % N;
        cx<float> V;
        if (plane == 2) {
          V = Zb[a][c];
        } else {
          const cx<float> W = Za[a][c], Wm = Za[(N - a) % N][(N - c) % N];
          V = plane == 0 ? cx<float>{0.5f * (W.re + Wm.re), 0.5f * (W.im - Wm.im)}
                         : cx<float>{0.5f * (W.im + Wm.im), 0.5f * (Wm.re - W.re)};
        }
        int k[3];
        nyq_slot_k(plane, a, c, k);
        cx<float> nyq{1.f, 0.f}, rest{1.f, 0.f};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const cx<float> g = tG[b][d][k[d]];
          if (k[d] == H) nyq = nyq * g;
          else rest = rest * g;
        }
        acc[i] = acc[i] + ((ws * nyq.im) * rest) * V;
      }
    }
    __syncthreads();
  }
  const float cj = (float)A.crad[ps * A.nr + j];
  const float* sp = A.spg + ((long long)ps * A.nr + j) * A.nq;
  cx<float>* out = A.out + (cell * A.nr + j) * SLOTS;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int s = tid + i * THREADS;
    const int plane = s / PLANE, a = (s / N) % N, c = s % N;
    cx<float> v{0.f, 0.f};
    if (nyq_slot_owned(plane, a, c)) {
      int k[3];
      nyq_slot_k(plane, a, c, k);
      int q = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int kf = kfreq(k[d], N);
        q += kf * kf;
      }
      v = (-cj * sp[q]) * acc[i];
    }
    out[s] = v;
  }
}

// acc64 (full FP64 spectrum) += sum over j of the plane terms, each plane point once.
__global__ void k_f3n_reduce(const cx<float>* __restrict__ part, cx<double>* __restrict__ acc64,
                             long long nb, int nr) {
  using namespace f3n;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nb * SLOTS) return;
  const long long c = gid / SLOTS;
  const int s = (int)(gid % SLOTS);
  const int plane = s / PLANE, a = (s / N) % N, b = s % N;
  if (!nyq_slot_owned(plane, a, b)) return;
  double re = 0.0, im = 0.0;
  const cx<float>* p = part + c * nr * SLOTS + s;
  for (int j = 0; j < nr; ++j) {
    const cx<float> v = p[(long long)j * SLOTS];
    re += (double)v.re;
    im += (double)v.im;
  }
  int k[3];
  nyq_slot_k(plane, a, b, k);
  cx<double>* o = acc64 + c * (long long)PLANE * N + (long long)k[0] * PLANE + k[1] * N + k[2];
  *o = cx<double>{o->re + re, o->im + im};
}

size_t nyq_scratch_bytes(int nr) {
  return (size_t)(nr + 1) * f3n::SLOTS * sizeof(cx<float>);  // out planes per j + packed fhat planes
}

// Launch the Nyquist term for nb cells whose FP32 f (f32) and fhat/n planes (xp) are packed;
// out holds nr * SLOTS complex FP32 per cell; acc64 receives the FP64 sum.
int launch_f3_nyq(Plan* P, const cx<float>* xp, const float* f32, cx<float>* out, cx<double>* acc64,
                  long long nb, const int* pidx, hipStream_t st) {
  using namespace f3n;
  F3NArgs A{};
  A.xp = xp;
  A.f32 = f32;
  A.out = out;
  A.pidx = pidx;
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
  A.L = P->L;
  const long long groups = (nb + 7) / 8;
  hipLaunchKernelGGL(k_f3_nyq, dim3((unsigned)(groups * 8 * P->nr)), dim3(THREADS), 0, st, A);
  IB_HIP(hipGetLastError());
  const long long tot = nb * SLOTS;
  hipLaunchKernelGGL(k_f3n_reduce, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, out, acc64,
                     nb, P->nr);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

}  // namespace ib
