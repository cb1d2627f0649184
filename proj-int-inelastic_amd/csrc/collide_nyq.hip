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
// the M angles in groups of four.  Wave q of a group runs angle q's 2D inverse
// transforms (P1 + i P2 packed in one complex transform, P3 in a second) with one
// 32-point pencil per lane and leaves P1, P2, P3 as transposed real planes; all waves
// form u = sigma_y sigma_z f beta along x pencils (thread (z, y) and (z, y + 16), f held
// in registers for the whole kernel) and reduce it, angle by angle: over x in registers
// (R1), over z across the 32 lanes of a half-wave (R3) and over y across the two
// pencils, the half-waves and, through LDS, the waves (R2), by recursive halvings in a
// lane-dependent register order that needs no selects (nyq_xl0); wave q runs angle q's
// forward transforms (R1 + i R2 packed, R3); all threads weight the plane points with
// separable per-angle tables and sum over angles in registers.  The per-(cell, j) plane
// terms go to a buffer that k_f3n_reduce sums over j in FP64 into the full spectrum.
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
constexpr int G = 4;                    // angles per group
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
  int ns;  // workgroups per (cell, j), each over a contiguous range of the angles
  int xg;  // workgroup numbering (cell_of_block)
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
  // every source lane of these patterns lies in the row, so bound_ctrl never applies; with it
  // set and old = 0 the compiler folds the move into the add that consumes it
  switch (ctrl_id) {
    case 0: r = __builtin_amdgcn_update_dpp(0, iv, 0x140, 0xF, 0xF, true); break;  // row_mirror: lane ^ 15
    case 1: r = __builtin_amdgcn_update_dpp(0, iv, 0x141, 0xF, 0xF, true); break;  // row_half_mirror: lane ^ 7
    case 2: r = __builtin_amdgcn_update_dpp(0, iv, 0xB1, 0xF, 0xF, true); break;   // quad_perm [1,0,3,2]: lane ^ 1
    default: r = __builtin_amdgcn_update_dpp(0, iv, 0x4E, 0xF, 0xF, true); break;  // quad_perm [2,3,0,1]: lane ^ 2
  }
  return __builtin_bit_cast(float, r);
}

// Register order of the x pencils.  Register i of lane l holds x = i ^ XL0(l), with
// XL0(l) = 4 b2 ^ 12 b3 (b_k the lane's bits), a multiple of 4, so that the reductions over
// lanes are recursive halvings with no lane-dependent selects.  Lanes l, l ^ 16 and l ^ 32
// hold their x in the same registers, so the first halving of each reduction pairs register
// i with register i + 16 of the partner row (or half-wave) through one v_permlane swap: after
// swap(u[i], u[i + 16]) the sum u[i] + u[i + 16] is, in the even row, the pair's sum at
// x = i ^ XL0 and, in the odd row, at x = (i + 16) ^ XL0, so the order of the 16 halved
// registers becomes XL0 ^ 16 b4 (over z, rows l ^ 16) or XL0 ^ 16 b5 (over y, half-waves
// l ^ 32).  The rest of the halving over z: l ^ 15 (h = 8, DPP row_mirror), l ^ 7 (h = 4, DPP
// row_half_mirror), then a sum over the quad (l ^ 1, l ^ 2) of the four remaining registers.
// Blocks of four registers stay four consecutive x, so the pencils' P values load as 128-bit
// vectors.
__device__ __forceinline__ int nyq_xl0(int lane) {
  return (((lane >> 2) & 1) * 4) ^ (((lane >> 3) & 1) * 12);
}

// Rows z >= 16 of the P2 plane and of the R2 partials store element x at x ^ 16: lanes z and
// z + 16 read the same x block, and their rows, 16 x 36 floats apart, fall on the same banks.
__device__ __forceinline__ int nyq_rsw(int z) { return ((z >> 4) & 1) * 16; }

// v_permlane16_swap (odd rows of a <-> even rows of b) and v_permlane32_swap (upper half-wave
// of a <-> lower half-wave of b) with both registers written back; the compiler's builtins for
// these return the first register for both results.  s_nop 1: the two wait states a VALU
// write of an operand needs before the swap reads it.
__device__ __forceinline__ void swap_rows16(float& a, float& b) {
  asm volatile("s_nop 1\n\tv_permlane16_swap_b32 %0, %1" : "+v"(a), "+v"(b));
}
__device__ __forceinline__ void swap_halves32(float& a, float& b) {
  asm volatile("s_nop 1\n\tv_permlane32_swap_b32 %0, %1" : "+v"(a), "+v"(b));
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

// Per-angle weight tables of the plane points: the weight w_sigma Im(prod_{S(k)} g(16))
// prod_{not S(k)} g(k_d) is separable on each plane apart from the kz = 16 column:
//   plane kx = 16:      A0[ky] gz[kz] (kz != 16), B0[ky] (kz = 16)
//   plane ky = 16:      A1[kx] gz[kz] (kz != 16), B1[kx] (kz = 16)
//   plane kz = 16:      A2[kx] gy[ky]
enum NyqTab { W_A0 = 0, W_B0 = 1, W_A1 = 2, W_B1 = 3, W_A2 = 4 };

// Angles are processed in groups of G: the 2D transforms of a group's angles run at
// once (wave q of waves 0 .. G-1 owns angle q: its 32 rows, then its 32 columns, of
// both transforms), the N^3 reductions run angle by angle on all waves, and the plane
// weights of the whole group are applied in one pass.
__global__ __launch_bounds__(512, 2) void k_f3_nyq(F3NArgs A) {
  using namespace f3n;
  constexpr int RPS = N + 4;  // padded row of the R2 partials and of P2 (conflict-free 128-bit access)
  // E factors: read in phases A and B only, so the next group's are written during D in place
  __shared__ cx<float> tE[G][3][N];
  __shared__ cx<float> tG[2][G][3][N];
  __shared__ cx<float> tW[2][G][5][N];
  // per angle: complex transform workspaces Za (P1 + i P2, then R1 + i R2) and Zb (P3, R3);
  // between the transforms the same bytes hold the real planes P1T[z][y], P2T[z][x],
  // P3T[y][x] (rows of RPS floats: 128-bit accesses of consecutive rows hit distinct banks)
  // 128-bit LDS accesses below: 16-byte alignment of both arrays
  __shared__ __align__(16) cx<float> Z[G][2][N][PAD];
  // R2 partials per wave, two buffers by angle parity: angle q + 1's phase B writes the other
  // buffer while slower waves still read angle q's in phase C, so one barrier per angle suffices
  __shared__ __align__(16) float RP[2][8][N][RPS];
  const int tid = threadIdx.x;
  const long long id = blockIdx.x;
  const int nrs = A.nr * A.ns;
  long long cell;
  int jj;  // partial index: radial node j, angle range asp
  cell_of_block(id, nrs, A.xg, cell, jj);
  const int j = jj / A.ns, asp = jj % A.ns;
  if (cell >= A.batch) return;  // uniform over the workgroup
  const int m0 = (asp * A.m) / A.ns, mend = ((asp + 1) * A.m) / A.ns;  // angles m0 .. mend - 1
  const int ps = A.pidx ? A.pidx[cell] : 0;
  const double aj = A.a[ps * A.nr + j] * A.rj[j];
  const double rpl_pi = A.rj[j] / A.L;
  const double aj_pi = aj / M_PI;
  const cx<float>* X1 = A.xp + cell * SLOTS;
  const cx<float>* X2 = X1 + PLANE;
  const cx<float>* X3 = X2 + PLANE;
  const float* fc = A.f32 + cell * (long long)PLANE * N;
  const float fppp = X1[H * N + H].re;  // fhat(16, 16, 16) / n (real)

  // phase tables of the group's angles: E and G factors per axis, from t in [0, 6 G N)
  auto make_tables = [&](int g0, int bb, int t) {
    const int q = t / (6 * N), r = t % (6 * N);
    const int which = r / (3 * N), ax = (r % (3 * N)) / N, k = r % N;
    const int mm = g0 + q;
    if (mm < mend) {
      const double sg = A.sigma[mm * 3 + ax];
      const double kf = (double)kfreq(k, N);
      double tt = which == 0 ? -(rpl_pi * sg * kf) : (aj_pi * sg * kf);
      tt -= 2.0 * rint(0.5 * tt);  // half turns reduced to [-1, 1] in FP64
      float sn, cs;
      sincospif((float)tt, &sn, &cs);
      if (which)
        tG[bb][q][ax][k] = cx<float>{cs, sn};
      else
        tE[q][ax][k] = cx<float>{cs, sn};
    }
  };
  // plane weight tables of the group's angles from the G factors, t in [0, 5 G N)
  auto make_weights = [&](int g0, int bb, int t) {
    const int q = t / (5 * N), r = t % (5 * N), tab = r / N, k = r % N;
    const int mm = g0 + q;
    if (mm < mend) {
      const float ws = (float)A.wsig[mm];
      const cx<float>* g = &tG[bb][q][0][0];
      const cx<float> gx = g[H], gy = g[N + H], gz = g[2 * N + H];
      cx<float> v;
      if (tab == 0) v = k == H ? cx<float>{ws * (gx * gy).im, 0.f} : (ws * gx.im) * g[N + k];
      else if (tab == 1) v = k == H ? cx<float>{ws * ((gx * gy) * gz).im, 0.f} : (ws * (gx * gz).im) * g[N + k];
      else if (tab == 2) v = (ws * gy.im) * g[k];
      else if (tab == 3) v = (ws * (gy * gz).im) * g[k];
      else v = (ws * gz.im) * g[k];
      tW[bb][q][tab][k] = v;
    }
  };

  const int lane = tid & 63, wv = tid >> 6;
  const int z = tid & 31, yq = tid >> 5;  // N^3 roles: x pencils (y, z) and (y + 16, z)
  const int xl0 = nyq_xl0(lane);                // order of the pencils' registers
  const int xl1 = xl0 ^ (((lane >> 4) & 1) * 16);  // order after the halving over z
  const float sgz = (z & 1) ? -1.f : 1.f, sgy = (yq & 1) ? -1.f : 1.f;

  // f of the two pencils in the register order above, for the whole kernel
  float f0[N], f1[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const int x = i ^ xl0;
    f0[i] = fc[x * PLANE + yq * N + z];
    f1[i] = fc[x * PLANE + (yq + 16) * N + z];
  }

  cx<float> acc[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) acc[i] = cx<float>{0.f, 0.f};

  for (int t = tid; t < 6 * G * N; t += THREADS) make_tables(m0, 0, t);
  __syncthreads();
  for (int t = tid; t < 5 * G * N; t += THREADS) make_weights(m0, 0, t);
  for (int g0 = m0, bb = 0; g0 < mend; g0 += G, bb ^= 1) {
    const int ng = mend - g0 < G ? mend - g0 : G;
    // A (wave q < ng): P1 + i P2 = ifft2(h_z (h_y X1 + i h_x X2)), P3 = ifft2(h_x h_y X3),
    // left as the real planes P1T[z][y], P2T[z][x], P3T[y][x]
    if (wv < ng) {
      const int q = wv, a = lane & 31, t = lane >> 5;
      cx<float>(*Zt)[PAD] = Z[q][t];
      const cx<float>* tEq = &tE[q][0][0];
      cx<float> v[N];
      // rows in blocks of 8 points, so that at most one block's loads are in flight
      if (t == 0) {
        const cx<float> hy = hfac(tEq + N, a), hx = hfac(tEq, a);
#pragma unroll
        for (int c0 = 0; c0 < N; c0 += 8) {
#pragma unroll
          for (int c = c0; c < c0 + 8; ++c) {
            const cx<float> t1 = hy * X1[a * N + c], t2 = hx * X2[a * N + c];
            v[c] = hfac(tEq + 2 * N, c) * cx<float>{t1.re - t2.im, t1.im + t2.re};
          }
          __builtin_amdgcn_sched_barrier(0);
        }
      } else {
        const cx<float> hx = hfac(tEq, a);
#pragma unroll
        for (int c0 = 0; c0 < N; c0 += 8) {
#pragma unroll
          for (int c = c0; c < c0 + 8; ++c) v[c] = (hx * hfac(tEq + N, c)) * X3[a * N + c];
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      fft_reg<N, +1>(v);
#pragma unroll
      for (int c = 0; c < N; ++c) Zt[a][c] = v[c];
      nyq_wave_sync();
#pragma unroll
      for (int r = 0; r < N; ++r) v[r] = Zt[r][a];
      fft_reg<N, +1>(v);
      nyq_wave_sync();  // every lane of the wave has read its column before the planes overwrite them
      float* P1T = (float*)&Z[q][0][0][0];
      float* P2T = P1T + N * RPS;
      float* P3T = P2T + N * RPS;
      if (t == 0) {  // column a = z of P1 + i P2
#pragma unroll
        for (int r = 0; r < N; r += 4) {
          *(float4*)&P1T[a * RPS + r] = make_float4(v[r].re, v[r + 1].re, v[r + 2].re, v[r + 3].re);
          *(float4*)&P2T[a * RPS + (r ^ nyq_rsw(a))] = make_float4(v[r].im, v[r + 1].im, v[r + 2].im, v[r + 3].im);
        }
      } else {  // column a = y of P3
#pragma unroll
        for (int r = 0; r < N; r += 4)
          *(float4*)&P3T[a * RPS + r] = make_float4(v[r].re, v[r + 1].re, v[r + 2].re, v[r + 3].re);
      }
    }
    __syncthreads();
    for (int q = 0; q < ng; ++q) {
      const float* P1T = (const float*)&Z[q][0][0][0];
      const float* P2T = P1T + N * RPS;
      const float* P3T = P2T + N * RPS;
      // B (all): u = sigma_y sigma_z f beta on the pencils (y, z), (y + 16, z); reductions
      float r1[2], r3[2][4];
      {
        const float sx = tE[q][0][H].im, sy = tE[q][1][H].im, sz = tE[q][2][H].im;
        const float C = sx * sy * sz * fppp;
        const float K2 = sy * sgz, K3 = sz * sgy;
        // roles from an opaque copy of tid, so that the pencils' LDS addresses are formed
        // next to their loads each angle instead of occupying registers across the loop
        int tq = tid;
        asm volatile("" : "+v"(tq));
        const int zz = tq & 31, y0 = tq >> 5, y1 = y0 + 16, xq = nyq_xl0(tq & 63);
        const int xq2 = xq ^ (((tq >> 5) & 1) * 16);  // order after the halving over y
        const int rsw = nyq_rsw(zz);
        // sigma_y sigma_z (s_x P1 - C sigma_y sigma_z) of the two pencils
        const float K1a = sgy * sgz * sx * P1T[zz * RPS + y0] - C;
        const float K1b = sgy * sgz * sx * P1T[zz * RPS + y1] - C;
        float u0[N], u1[N];
        float s1a = 0.f, s1b = 0.f;
#pragma unroll
        for (int k = 0; k < N / 4; ++k) {
          const int xb = (4 * k) ^ xq;  // first x of the register block 4k .. 4k + 3
          const float4 p2 = *(const float4*)&P2T[zz * RPS + (xb ^ rsw)];
          const float4 pa = *(const float4*)&P3T[y0 * RPS + xb];
          const float4 pb = *(const float4*)&P3T[y1 * RPS + xb];
          const float p2e[4] = {p2.x, p2.y, p2.z, p2.w};
          const float pae[4] = {pa.x, pa.y, pa.z, pa.w};
          const float pbe[4] = {pb.x, pb.y, pb.z, pb.w};
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int i = 4 * k + e;  // (-1)^x = (-1)^i: xq is even
            const float ta = fmaf(K2, p2e[e], (e & 1) ? -K1a : K1a);
            const float tb = fmaf(K2, p2e[e], (e & 1) ? -K1b : K1b);
            u0[i] = f0[i] * fmaf(K3, pae[e], ta);
            u1[i] = f1[i] * fmaf(K3, pbe[e], tb);
            s1a = (e & 1) ? s1a - u0[i] : s1a + u0[i];
            s1b = (e & 1) ? s1b - u1[i] : s1b + u1[i];
          }
        }
        r1[0] = sgy * sgz * s1a;
        r1[1] = sgy * sgz * s1b;
        // R2 over y: the two pencils in registers, the half-waves (l ^ 32), waves via LDS
        float t2[H];
#pragma unroll
        for (int i = 0; i < H; ++i) {
          float a = u0[i] + u1[i], b = u0[i + H] + u1[i + H];
          swap_halves32(a, b);
          t2[i] = a + b;
        }
#pragma unroll
        for (int k = 0; k < H / 4; ++k)
          *(float4*)&RP[q & 1][wv][zz][(4 * k) ^ xq2 ^ rsw] = make_float4(t2[4 * k], t2[4 * k + 1], t2[4 * k + 2], t2[4 * k + 3]);
        // R3 over z: recursive halving across the 32 lanes of the half-wave
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          float* u = p ? u1 : u0;
#pragma unroll
          for (int i = 0; i < 16; ++i) {
            float a = u[i], b = u[i + 16];
            swap_rows16(a, b);
            u[i] = a + b;
          }
#pragma unroll
          for (int i = 0; i < 8; ++i) u[i] += dpp_f(u[i + 8], 0);
#pragma unroll
          for (int i = 0; i < 4; ++i) u[i] += dpp_f(u[i + 4], 1);
#pragma unroll
          for (int i = 0; i < 4; ++i) u[i] += dpp_f(u[i], 2);
#pragma unroll
          for (int i = 0; i < 4; ++i) u[i] += dpp_f(u[i], 3);
#pragma unroll
          for (int i = 0; i < 4; ++i) r3[p][i] = sgy * u[i];  // R3(x = xl1 + i, y)
        }
      }
      __syncthreads();
      // C (all): Za <- R1 + i R2 ((y, z) and (x, z) share the index), Zb <- R3 (x, y)
      {
        cx<float>(*Za)[PAD] = Z[q][0];
        cx<float>(*Zb)[PAD] = Z[q][1];
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          const int y = yq + 16 * p;
          float s2 = 0.f;
#pragma unroll
          for (int w = 0; w < 8; ++w) s2 += RP[q & 1][w][z][y ^ nyq_rsw(z)];
          Za[y][z] = cx<float>{r1[p], sgz * s2};
          // lane e of the quad writes x = xl1 + e (every lane of the quad holds all four)
          const int e = lane & 3;
          const float r3e = e == 0 ? r3[p][0] : e == 1 ? r3[p][1] : e == 2 ? r3[p][2] : r3[p][3];
          Zb[xl1 + e][y] = cx<float>{r3e, 0.f};
        }
      }
    }
    __syncthreads();  // every angle's R planes are written before D transforms them
    // D (wave q < ng): forward 2D transforms of angle q; waves G and up: next group's tables
    if (wv < ng) {
      const int q = wv, a = lane & 31, t = lane >> 5;
      cx<float>(*Zt)[PAD] = Z[q][t];
      cx<float> v[N];
#pragma unroll
      for (int c = 0; c < N; ++c) v[c] = Zt[a][c];
      fft_reg<N, -1>(v);
#pragma unroll
      for (int c = 0; c < N; ++c) Zt[a][c] = v[c];
      nyq_wave_sync();
#pragma unroll
      for (int r = 0; r < N; ++r) v[r] = Zt[r][a];
      fft_reg<N, -1>(v);
#pragma unroll
      for (int r = 0; r < N; ++r) Zt[r][a] = v[r];
    } else if (wv >= G && g0 + G < mend) {
      for (int t = tid - G * 64; t < 6 * G * N; t += THREADS - G * 64) make_tables(g0 + G, bb ^ 1, t);
    }
    __syncthreads();
    if (g0 + G < mend)
      for (int t = tid; t < 5 * G * N; t += THREADS) make_weights(g0 + G, bb ^ 1, t);
    // E (all): V on the plane points times the separable weights, summed over angles.  Slot i of
    // a thread is plane i / 2, row a = tid / 32 + 16 (i % 2), column c = tid % 32 (one column for
    // all six), so the angle loop runs outside the slots and each angle's loads issue together;
    // on the kz = 16 column the row table B0 / B1 is taken with a unit column factor
    {
      static_assert(THREADS == 512 && PER == 6, "slot i: plane i / 2, row tid / 32 + 16 (i % 2)");
      const int c = tid & 31, a0 = tid >> 5, cm = (N - c) % N;
      const bool ch = c == H;
      const int r0 = ch ? W_B0 : W_A0, r1 = ch ? W_B1 : W_A1;
#pragma unroll
      for (int q = 0; q < G; ++q) {
        if (q < ng) {
          const cx<float>* tw = &tW[bb][q][0][0];
          const cx<float> gy = tG[bb][q][1][c];
          const cx<float> gz = ch ? cx<float>{1.f, 0.f} : tG[bb][q][2][c];
#pragma unroll
          for (int i = 0; i < PER; ++i) {
            const int plane = i >> 1, a = a0 + 16 * (i & 1);
            cx<float> V, Wt;
            if (plane == 2) {
              V = Z[q][1][a][c];
              Wt = tw[W_A2 * N + a] * gy;
            } else {
              const cx<float> W = Z[q][0][a][c], Wm = Z[q][0][(N - a) % N][cm];
              if (plane == 0) {
                V = cx<float>{0.5f * (W.re + Wm.re), 0.5f * (W.im - Wm.im)};
                Wt = tw[r0 * N + a] * gz;
              } else {
                V = cx<float>{0.5f * (W.im + Wm.im), 0.5f * (Wm.re - W.re)};
                Wt = tw[r1 * N + a] * gz;
              }
            }
            acc[i] = acc[i] + Wt * V;
          }
        }
      }
    }
    __syncthreads();
  }
  const float cj = (float)A.crad[ps * A.nr + j];
  const float* sp = A.spg + ((long long)ps * A.nr + j) * A.nq;
  cx<float>* out = A.out + (cell * nrs + jj) * SLOTS;
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

size_t nyq_scratch_bytes(int nparts) {
  // out planes per (j, angle range) + packed fhat planes
  return (size_t)(nparts + 1) * f3n::SLOTS * sizeof(cx<float>);
}

// Launch the Nyquist term for nb cells whose FP32 f (f32) and fhat/n planes (xp) are packed;
// out holds nr * splits * SLOTS complex FP32 per cell; acc64 receives the FP64 sum.
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
  A.ns = P->splits();
  A.xg = nb >= 8;
  A.L = P->L;
  hipLaunchKernelGGL(k_f3_nyq, dim3((unsigned)cell_grid(nb, P->nr * A.ns, A.xg)), dim3(THREADS), 0, st, A);
  IB_HIP(hipGetLastError());
  const long long tot = nb * SLOTS;
  hipLaunchKernelGGL(k_f3n_reduce, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, out, acc64,
                     nb, P->nr * A.ns);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

}  // namespace ib
