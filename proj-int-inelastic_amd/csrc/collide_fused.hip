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
#include "group.h"
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

template <bool KAHAN>
__global__ __launch_bounds__(512, 1) void k_fused3(F3Args A) {
  using namespace f3;
  __shared__ cx<float> W[NPL][N][PAD];
  __shared__ cx<float> tE[4][3][N];
  __shared__ cx<float> tG[4][3][N];
  __shared__ float sp[3 * (N / 2) * (N / 2) + 1];
  __shared__ cx<float> S16[N][PAD];
  __shared__ cx<float> SP[N][PAD];
  long long cell;
  int jj;
  cell_of_block(blockIdx.x, A.nr * A.ns, A.xg, cell, jj);
  if (cell >= A.batch) return;  // uniform over the workgroup
  fused3_body<KAHAN>(A, cell, jj, threadIdx.x, W, tE, tG, sp, S16, SP);
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
// partial planes per (radial node, angle range) (twice with Kahan compensation), FP32 f-hat
// and f, the Nyquist planes and the Nyquist term's partials
static size_t f3_cell_bytes(const Plan* P, size_t* part_bytes, size_t* fh32_bytes) {
  using namespace f3;
  const size_t b64 = (size_t)NPTS * sizeof(cx<double>);
  const int kp = P->kahan ? 2 : 1;
  const int parts = P->nr * P->splits();  // partial planes per cell: one per (j, angle range)
  *part_bytes = (size_t)parts * kp * NPL * PLANE * sizeof(cx<float>);
  *fh32_bytes = (size_t)17 * PLANE * sizeof(cx<float>) + (size_t)NPTS * sizeof(float);
  return 4 * b64 + *part_bytes + *fh32_bytes + nyq_scratch_bytes(parts);
}

static void f3_layout(const Plan* P, long long sub, ScratchSizes* z) {
  const size_t b64 = (size_t)f3::NPTS * sizeof(cx<double>);
  size_t pb, fb;
  const size_t cell = f3_cell_bytes(P, &pb, &fb);
  z->fhat = sub * b64;
  z->acc = sub * b64;
  z->work = 2 * sub * b64;
  z->aux = sub * (cell - 4 * b64);
}

static long long f3_sizes(const Plan* P, long long batch, ScratchSizes* z) {
  size_t pb, fb;
  const long long sub = fused_sub(P, batch, 256, f3_cell_bytes(P, &pb, &fb));
  f3_layout(P, sub, z);
  return sub;
}

// the aux pieces of a sub-batch of `sub` cells: partials, FP32 f-hat, FP32 f, Nyquist planes
// (sub x 3 planes) and the Nyquist term's partials (sub x nr x splits x 3 planes)
struct F3Ptrs {
  cx<float>* part;
  cx<float>* fh32;
  float* f32;
  cx<float>* xp;
  cx<float>* nyq_out;
};

static F3Ptrs f3_ptrs(const Plan* P, long long sub) {
  using namespace f3;
  size_t pb, fb;
  f3_cell_bytes(P, &pb, &fb);
  F3Ptrs r;
  r.part = (cx<float>*)P->aux.ptr;
  r.fh32 = (cx<float>*)((char*)P->aux.ptr + sub * pb);
  r.f32 = (float*)((char*)r.fh32 + sub * 17 * PLANE * sizeof(cx<float>));
  r.xp = (cx<float>*)(r.f32 + sub * NPTS);
  r.nyq_out = r.xp + sub * 3 * PLANE;
  return r;
}

static F3Args f3_args(const Plan* P, const F3Ptrs& b, long long nb, const int* pidx) {
  F3Args A{};
  A.fh = b.fh32;
  A.f32 = b.f32;
  A.part = b.part;
  A.pidx = pidx;  // indexed from the sub-batch start
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
  return A;
}

// prologue fhat = fft(f) in FP64 and the FP32 packing of a sub-batch
static int f3_front(Plan* P, const double* fc, long long nb, const F3Ptrs& b, hipStream_t st) {
  using namespace f3;
  int rc;
  if ((rc = staged_prologue_fp64(P, fc, nb, P->fhat.ptr, P->work.ptr))) return rc;
  const long long tot = nb * NPTS;
  hipLaunchKernelGGL(k_f3_pack, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                     (const cx<double>*)P->fhat.ptr, fc, b.fh32, b.f32, b.xp, nb);
  IB_HIP(hipGetLastError());
  return IB_OK;
}

// after the pair loop: FP64 sum of the partials, the Nyquist-plane term, the FP64 epilogue
static int f3_back(Plan* P, const double* fc, double* qc, long long nb, const F3Ptrs& b, const int* pc,
                   int with_loss, double hb, hipStream_t st) {
  using namespace f3;
  const long long half = nb * NPL * PLANE;
  hipLaunchKernelGGL(k_f3_reduce, dim3((unsigned)((half + 255) / 256)), dim3(256), 0, st, b.part,
                     (cx<double>*)P->acc.ptr, nb, P->nr * P->splits(), P->kahan ? 2 : 1);
  IB_HIP(hipGetLastError());
  int rc;
  if ((rc = launch_f3_nyq(P, b.xp, b.f32, b.nyq_out, (cx<double>*)P->acc.ptr, nb, pc, st))) return rc;
  return staged_epilogue_fp64(P, fc, qc, nb, P->acc.ptr, P->fhat.ptr, P->work.ptr, with_loss, hb, pc);
}

int fused_sizes(const Plan* P, long long batch, ScratchSizes* z) {
  const int k = fused_kind(P);
  if (k == 2) return fused2_sizes(P, batch, z);
  if (k == 0) return 0;
  f3_sizes(P, batch, z);
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
  const long long sub = f3_sizes(P, batch, &z);
  int rc;
  if ((rc = P->reserve(z))) return rc;
  const F3Ptrs b = f3_ptrs(P, sub);
  for (long long c0 = 0; c0 < batch; c0 += sub) {
    const long long nb = std::min(sub, batch - c0);
    const double* fc = f + c0 * NPTS;
    const int* pc = pidx ? pidx + c0 : nullptr;
    if ((rc = f3_front(P, fc, nb, b, st))) return rc;
    const F3Args A = f3_args(P, b, nb, pc);
    if ((rc = P->timed_begin(st))) return rc;
    const unsigned grid = (unsigned)cell_grid(nb, P->nr * A.ns, A.xg);
    if (P->kahan)
      hipLaunchKernelGGL(k_fused3<true>, dim3(grid), dim3(THREADS), 0, st, A);
    else
      hipLaunchKernelGGL(k_fused3<false>, dim3(grid), dim3(THREADS), 0, st, A);
    IB_HIP(hipGetLastError());
    P->timed_end(st, nb);
    if ((rc = f3_back(P, fc, q + c0 * NPTS, nb, b, pc, with_loss, hb, st))) return rc;
  }
  return IB_OK;
}

// Grouped launch (group.h), 3D 32^3 form: the whole batch as one segment when its scratch
// fits the plan's budget (*ok = 0 otherwise: the plan then runs alone).
int f3_group_front(Plan* P, const double* f, long long batch, const int* pidx, hipStream_t st,
                   GroupSeg* seg, int* ok) {
  size_t pb, fb;
  *ok = 0;
  if ((size_t)batch * f3_cell_bytes(P, &pb, &fb) > P->scratch_budget) return IB_OK;
  ScratchSizes z;
  f3_layout(P, batch, &z);
  int rc;
  if ((rc = P->reserve(z))) return rc;
  const F3Ptrs b = f3_ptrs(P, batch);
  if ((rc = f3_front(P, f, batch, b, st))) return rc;
  seg->kind = P->kahan ? GK_F3K : GK_F3;
  seg->f3 = f3_args(P, b, batch, pidx);
  seg->units = batch * P->nr;
  seg->blocks = group_blocks(seg->kind, seg->units, seg->f3.ns, cell_grid(batch, P->nr * seg->f3.ns, seg->f3.xg));
  *ok = 1;
  return IB_OK;
}

int f3_group_back(Plan* P, const double* f, double* q, long long batch, const int* pidx, int with_loss,
                  double hb, hipStream_t st) {
  return f3_back(P, f, q, batch, f3_ptrs(P, batch), pidx, with_loss, hb, st);
}

int f2_group_front(Plan* P, const double* f, long long batch, const int* pidx, hipStream_t st,
                   GroupSeg* seg, int* ok);
int f2_group_back(Plan* P, const double* f, double* q, long long batch, const int* pidx, int with_loss,
                  double hb, hipStream_t st);

int f2_group_fits(const Plan* P, long long batch);

// Whether a plan's whole batch can be one segment of a grouped launch (an FP32 fused form whose
// scratch for the whole batch fits the plan's budget); no work is queued.
int fused_group_fits(const Plan* P, long long batch) {
  if (P->precision == IB_PREC_FP64 || batch <= 0) return 0;
  const int k = fused_kind(P);
  if (k == 1) {
    size_t pb, fb;
    return (size_t)batch * f3_cell_bytes(P, &pb, &fb) <= P->scratch_budget;
  }
  return k == 2 ? f2_group_fits(P, batch) : 0;
}

// A plan's part of a grouped launch: *ok = 1 when its whole batch became one segment (the
// front work queued), with *which naming the engine for fused_group_back.
int fused_group_front(Plan* P, const double* f, long long batch, const int* pidx, hipStream_t st,
                      GroupSeg* seg, int* ok, int* which) {
  *ok = 0;
  *which = fused_kind(P);
  if (P->precision == IB_PREC_FP64) return IB_OK;
  if (*which == 1) return f3_group_front(P, f, batch, pidx, st, seg, ok);
  if (*which == 2) return f2_group_front(P, f, batch, pidx, st, seg, ok);
  return IB_OK;
}

int fused_group_back(Plan* P, int which, const double* f, double* q, long long batch, const int* pidx,
                     int with_loss, double hb, hipStream_t st) {
  return which == 1 ? f3_group_back(P, f, q, batch, pidx, with_loss, hb, st)
                    : f2_group_back(P, f, q, batch, pidx, with_loss, hb, st);
}

}  // namespace ib
