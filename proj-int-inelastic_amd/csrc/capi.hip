// C ABI (include/inelastic.h): argument checks, plan lifetime, dispatch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "common.h"
#include "group.h"
#include "plan.h"

namespace ib {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

thread_local Arena* g_arena = nullptr;

size_t arena_left() {
  if (!g_arena) return SIZE_MAX;
  const size_t at = (g_arena->used + 255) & ~(size_t)255;
  return at >= g_arena->size ? 0 : g_arena->size - at;
}

long long fused_sub(const Plan* P, long long batch, long long dflt, size_t per_cell) {
  long long cap = P->fused_cells > 0 ? P->fused_cells : dflt;
  if (!g_arena && per_cell > 0)
    cap = std::min<long long>(cap, std::max<long long>(1, (long long)(P->scratch_budget / per_cell)));
  if (batch <= cap) return std::max<long long>(batch, 1);
  const long long launches = (batch + cap - 1) / cap;
  return (batch + launches - 1) / launches;
}

int DevBuf::ensure(size_t want) {
  if (want <= bytes) return IB_OK;
  if (g_arena) {  // static-scratch call: carve, never allocate
    const size_t at = (g_arena->used + 255) & ~(size_t)255;
    if (at > g_arena->size || want > g_arena->size - at) {
      set_error("ib_collide_static: this batch needs more scratch than the " +
                std::to_string(g_arena->size >> 20) +
                " MiB static region holds; pass fewer cells per call (ib_collide lets the "
                "plan allocate)");
      return IB_ERR_ARG;
    }
    ptr = g_arena->base + at;
    bytes = want;
    g_arena->used = at + want;
    return IB_OK;
  }
  // growing: work queued earlier on any stream may still read the old buffer
  if (ptr) {
    IB_HIP(hipDeviceSynchronize());
    (void)hipFree(ptr);
  }
  ptr = nullptr;
  bytes = 0;
  IB_HIP(hipMalloc(&ptr, want));
  bytes = want;
  return IB_OK;
}

int Plan::timed_begin(hipStream_t st) {
  if (!timing || ev_start.size() >= 8192) return IB_OK;
  hipEvent_t a, b;
  IB_HIP(hipEventCreate(&a));
  IB_HIP(hipEventCreate(&b));
  IB_HIP(hipEventRecord(a, st));
  ev_start.push_back(a);
  ev_stop.push_back(b);
  ev_cells.push_back(0);
  return IB_OK;
}

void Plan::timed_end(hipStream_t st, long long cells) {
  if (!timing || ev_stop.empty() || ev_cells.back() != 0) return;
  (void)hipEventRecord(ev_stop.back(), st);
  ev_cells.back() = cells;
}

static void clear_timing(Plan* p) {
  for (auto e : p->ev_start) (void)hipEventDestroy(e);
  for (auto e : p->ev_stop) (void)hipEventDestroy(e);
  p->ev_start.clear();
  p->ev_stop.clear();
  p->ev_cells.clear();
}

void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

int collide_plain(Plan* P, const double* f, double* q, long long batch, const int* pidx,
                  int with_loss, double hb, int no_gain, hipStream_t st);
bool staged_runs_alone(const Plan* P, long long batch);
int collide_fused(Plan* P, const double* f, double* q, long long batch, const int* pidx,
                  int with_loss, double hb, hipStream_t st, int* handled);
int launch_tables(Plan* P, const double* d_wr, hipStream_t st);
int launch_phase_tables2(Plan* P, hipStream_t st);
int collide_group(Plan** P, int np, const double* const* f, double* const* q,
                  const long long* batch, const int* const* pidx, int with_loss, double hb,
                  hipStream_t st);
// grouped launch of the fused pair loops (group.h, collide_fused.hip, collide_group.hip)
int fused_group_fits(const Plan* P, long long batch);
int fused_group_front(Plan* P, const double* f, long long batch, const int* pidx, hipStream_t st,
                      GroupSeg* seg, int* ok, int* which);
int fused_group_back(Plan* P, int which, const double* f, double* q, long long batch, const int* pidx,
                     int with_loss, double hb, hipStream_t st);
// 1: ib_collide_group runs the fused pair loops of its FP32 plans as one grouped launch
static int g_group_fused = 1;
int launch_tvd1d(const double* f, double* df, long long outer, int nx, long long nv,
                 const double* v1, long long vmod, double dtdx, int order, int limiter, int bc_lo,
                 int bc_hi, int accumulate, hipStream_t st);
int launch_axpy(double* y, const double* x, double a, long long n, hipStream_t st);
int launch_tvd_limit(const double* w, double* out, long long n, long long inner, const double* s,
                     long long smod, int limiter, hipStream_t st);
int launch_embed(const double* a, double* b, long long n0, long long n1, long long inner, int g,
                 int pad1, int extract, hipStream_t st);
int launch_ghost_axis(double* a, long long outer, long long n, long long inner, int g, int bc_lo,
                      int bc_hi, hipStream_t st);
int launch_axpy_chain(double* y, const double* x, double a, double b, double c, long long n,
                      hipStream_t st);
int launch_df_combine(const double* adv, const double* q, double* out, double dt, double kn,
                      long long n, hipStream_t st);
int launch_moments(const double* f, double* out, long long batch, int dim, int n, double L,
                   hipStream_t st);
int launch_project(double* q, long long batch, int dim, int n, double L, hipStream_t st);
int launch_upwind(const double* f, const double* v, double* out, long long n, long long inner,
                  double dx, double dt, int mode, hipStream_t st);
int launch_fast_sweep(double* f, const double* gain, const double* rho, const double* vx, int nx,
                      int nvr, long long nvc, double dx, double nu, double* err, hipStream_t st);

template <typename X>
static int upload(X** dst, const X* src, size_t count) {
  IB_HIP(hipMalloc((void**)dst, count * sizeof(X)));
  IB_HIP(hipMemcpy(*dst, src, count * sizeof(X), hipMemcpyHostToDevice));
  return IB_OK;
}

static void free_param_tables(Plan* p) {
  if (p->d_a) (void)hipFree(p->d_a);
  if (p->d_spg) (void)hipFree(p->d_spg);
  if (p->d_crad) (void)hipFree(p->d_crad);
  if (p->d_lam) (void)hipFree(p->d_lam);
  if (p->d_lam64) (void)hipFree(p->d_lam64);
  if (p->d_gphase) (void)hipFree(p->d_gphase);
  p->d_gphase = nullptr;
  p->d_a = nullptr;
  p->d_spg = nullptr;
  p->d_crad = nullptr;
  p->d_lam = nullptr;
  p->d_lam64 = nullptr;
}

// Per parameter set p = (restitution profile e_j, VHS exponent gamma_p):
// a_j = pi (1 + e_j) / (4 L) (inelastic_collisions.py:80), the radial gain weight
// r_j^(gamma_p + d - 1) w_j (:97-103 with |S| divided out, :225-227), and on the
// device the sp(a_j r_j |k|) and loss-symbol tables.
static int set_params(Plan* p, int n_param, const double* e_r, const double* gamma, hipStream_t st,
                      const double* d_wr) {
  free_param_tables(p);
  p->n_param = n_param;
  p->e_r.assign(e_r, e_r + (size_t)n_param * p->nr);
  p->gamma_p.assign((size_t)n_param, p->gamma);
  if (gamma)
    for (int i = 0; i < n_param; ++i) p->gamma_p[i] = gamma[i];
  std::vector<double> a((size_t)n_param * p->nr), crad((size_t)n_param * p->nr);
  for (int q = 0; q < n_param; ++q)
    for (int j = 0; j < p->nr; ++j) {
      const size_t i = (size_t)q * p->nr + j;
      a[i] = 0.25 * M_PI * (1 + p->e_r[i]) / p->L;
      crad[i] = std::pow(p->r[j], p->gamma_p[q] + p->dim - 1) * p->wr[j];
    }
  int rc = upload(&p->d_a, a.data(), a.size());
  if (!rc) rc = upload(&p->d_crad, crad.data(), crad.size());
  if (rc) return rc;
  const size_t tsz = p->precision == IB_PREC_FP64 ? sizeof(double) : sizeof(float);
  IB_HIP(hipMalloc(&p->d_spg, (size_t)n_param * p->nr * p->nq * tsz));
  IB_HIP(hipMalloc(&p->d_lam, (size_t)n_param * p->nq * tsz));
  IB_HIP(hipMalloc((void**)&p->d_lam64, (size_t)n_param * p->nq * sizeof(double)));
  rc = launch_tables(p, d_wr, st);
  if (rc) return rc;
  const bool fp32_2d = p->precision == IB_PREC_FP32 && (p->n == 16 || p->n == 32 || p->n == 64);
  const bool fp64_2d = p->precision == IB_PREC_FP64 && p->n == 32;
  if (p->dim == 2 && (fp32_2d || fp64_2d)) {  // the 2D fused kernels' phase tables
    const size_t per = (size_t)p->nr * p->m * 2 * p->n * 2 * (fp64_2d ? sizeof(double) : sizeof(float));
    if (!p->d_ephase) IB_HIP(hipMalloc(&p->d_ephase, per));
    if (n_param <= Plan::kMaxPhaseSets) IB_HIP(hipMalloc(&p->d_gphase, (size_t)n_param * per));
    rc = launch_phase_tables2(p, st);
    if (rc) return rc;
  }
  return IB_OK;
}

// Static scratch of ib_collide_static: one copy per device, held until the process ends.
// Host code uses the per-device address from hipGetSymbolAddress, never the array's name.
constexpr size_t kStaticScratchBytes = (size_t)128 << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];
static std::mutex g_scratch_mu;
static void* g_scratch_dev[256];

static int static_scratch(void** out) {
  int dev = 0;
  IB_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 256) {
    set_error("device index out of range");
    return IB_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {  // once per device, with that device current
    void* p = nullptr;
    IB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return IB_OK;
}

}  // namespace ib

using namespace ib;

struct ib_plan : public Plan {
  double* d_wr = nullptr;
};

extern "C" {

const char* ib_last_error(void) { return g_err.c_str(); }

const char* ib_version(void) { return "inelastic-mi355x 0.1 (gfx950)"; }

double ib_host_sp(int dim, double x) { return sp_func(dim, x); }

int ib_plan_create(const ib_plan_desc* d, ib_plan** out) {
  if (!d || !out) {
    set_error("null argument");
    return IB_ERR_ARG;
  }
  *out = nullptr;
  if (d->dim != 2 && d->dim != 3) {
    set_error("Dimension must be 2 or 3.");
    return IB_ERR_ARG;
  }
  if (!(d->n == 8 || d->n == 16 || d->n == 32 || (d->n == 64 && d->dim == 2))) {
    set_error("n must be 8, 16 or 32 (3D) or 8 .. 64 (2D)");
    return IB_ERR_UNSUPPORTED;
  }
  if (d->nr < 1 || d->m < 1 || d->n_param < 1 || !(d->L > 0)) {
    set_error("nr, m, n_param must be >= 1 and L > 0");
    return IB_ERR_ARG;
  }
  if (d->precision != IB_PREC_FP64 && d->precision != IB_PREC_FP32 &&
      d->precision != IB_PREC_FP32_KAHAN) {
    set_error("precision must be IB_PREC_FP64, IB_PREC_FP32 or IB_PREC_FP32_KAHAN");
    return IB_ERR_ARG;
  }
  if (!d->r || !d->wr || !d->sigma || !d->wsigma || !d->e_r) {
    set_error("null quadrature array");
    return IB_ERR_ARG;
  }
  ib_plan* p = new ib_plan();
  p->dim = d->dim;
  p->n = d->n;
  p->nr = d->nr;
  p->m = d->m;
  p->L = d->L;
  p->gamma = d->gamma;
  p->precision = d->precision == IB_PREC_FP64 ? IB_PREC_FP64 : IB_PREC_FP32;
  p->kahan = d->precision == IB_PREC_FP32_KAHAN;
  p->npts = 1;
  for (int a = 0; a < p->dim; ++a) p->npts *= p->n;
  p->nq = p->dim * (p->n / 2) * (p->n / 2) + 1;
  p->r.assign(d->r, d->r + d->nr);
  p->wr.assign(d->wr, d->wr + d->nr);
  p->sigma.assign(d->sigma, d->sigma + (size_t)d->m * d->dim);
  p->wsigma.assign(d->wsigma, d->wsigma + d->m);
  p->sphere_area = 2 * std::pow(M_PI, 0.5 * p->dim) / std::tgamma(0.5 * p->dim);
  int rc = upload(&p->d_sigma, p->sigma.data(), p->sigma.size());
  if (!rc) rc = upload(&p->d_rj, p->r.data(), p->r.size());
  if (!rc) rc = upload(&p->d_wr, p->wr.data(), p->wr.size());
  if (!rc) rc = upload(&p->d_wsig, p->wsigma.data(), p->wsigma.size());
  if (!rc) rc = set_params(p, d->n_param, d->e_r, nullptr, nullptr, p->d_wr);
  if (!rc) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      set_error(std::string("table build failed: ") + hipGetErrorString(e));
      rc = IB_ERR_HIP;
    }
  }
  if (rc) {
    ib_plan_destroy(p);
    return rc;
  }
  *out = p;
  return IB_OK;
}

int ib_plan_set_params(ib_plan* p, int n_param, const double* e_r, const double* gamma,
                       void* stream) {
  if (!p || !e_r || n_param < 1) {
    set_error("invalid parameter sets");
    return IB_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  IB_HIP(hipStreamSynchronize(st));  // old tables may still be in use
  int rc = set_params(p, n_param, e_r, gamma, st, p->d_wr);
  if (rc) return rc;
  IB_HIP(hipStreamSynchronize(st));
  return IB_OK;
}

int ib_plan_scratch_need(ib_plan* p, long long cells, int grouped, long long* need, long long* held) {
  if (!p || cells < 0 || (!need && !held)) {
    set_error("invalid scratch_need arguments");
    return IB_ERR_ARG;
  }
  const DevBuf* bufs[4] = {&p->fhat, &p->acc, &p->work, &p->aux};
  if (held) {
    long long h = 0;
    for (const DevBuf* b : bufs) h += (long long)b->bytes;
    *held = h;
  }
  if (need) {
    ScratchSizes z;
    if (cells > 0 && !fused_sizes(p, cells, &z)) {
      const int rc = staged_sizes(p, cells, grouped, &z);
      if (rc) return rc;
    }
    const size_t want[4] = {z.fhat, z.acc, z.work, z.aux};
    long long n = 0;
    for (int i = 0; i < 4; ++i) n += (long long)std::max(want[i], bufs[i]->bytes);
    *need = n;
  }
  return IB_OK;
}

int ib_plan_set_tuning(ib_plan* p, long long chunk_pairs, long long fused_cells,
                       long long scratch_budget) {
  if (!p || chunk_pairs < 0 || fused_cells < 0 || scratch_budget < 0) {
    set_error("invalid tuning values");
    return IB_ERR_ARG;
  }
  p->chunk_pairs = chunk_pairs;
  p->fused_cells = fused_cells;
  if (scratch_budget > 0) p->scratch_budget = (size_t)scratch_budget;
  return IB_OK;
}

int ib_plan_set_tiles(ib_plan* p, int angle_split, int pair_rounds) {
  if (!p || angle_split < 1 || angle_split > 64 || pair_rounds < 0 || pair_rounds > 1024) {
    set_error("invalid tile shape: angle_split in [1, 64], pair_rounds in [0, 1024]");
    return IB_ERR_ARG;
  }
  p->angle_split = angle_split;
  p->pair_rounds = pair_rounds;
  return IB_OK;
}

int ib_plan_get_tiles(ib_plan* p, int* angle_split, int* pair_rounds) {
  if (!p || !angle_split || !pair_rounds) {
    set_error("null argument");
    return IB_ERR_ARG;
  }
  *angle_split = p->splits();
  *pair_rounds = p->pair_rounds;
  return IB_OK;
}

int ib_plan_set_timing(ib_plan* p, int enable) {
  if (!p) {
    set_error("null plan");
    return IB_ERR_ARG;
  }
  IB_HIP(hipDeviceSynchronize());
  clear_timing(p);
  p->timing = enable ? 1 : 0;
  return IB_OK;
}

int ib_plan_get_timing(ib_plan* p, double* ms_total, long long* launches, long long* cells) {
  if (!p || !ms_total || !launches || !cells) {
    set_error("null argument");
    return IB_ERR_ARG;
  }
  double ms = 0;
  long long nl = 0, nc = 0;
  for (size_t i = 0; i < p->ev_start.size(); ++i) {
    if (p->ev_cells[i] == 0) continue;
    IB_HIP(hipEventSynchronize(p->ev_stop[i]));
    float t = 0.f;
    IB_HIP(hipEventElapsedTime(&t, p->ev_start[i], p->ev_stop[i]));
    ms += t;
    nl += 1;
    nc += p->ev_cells[i];
  }
  *ms_total = ms;
  *launches = nl;
  *cells = nc;
  return IB_OK;
}

int ib_plan_destroy(ib_plan* p) {
  if (!p) return IB_OK;
  (void)hipDeviceSynchronize();
  clear_timing(p);
  free_param_tables(p);
  if (p->d_sigma) (void)hipFree(p->d_sigma);
  if (p->d_rj) (void)hipFree(p->d_rj);
  if (p->d_wr) (void)hipFree(p->d_wr);
  if (p->d_wsig) (void)hipFree(p->d_wsig);
  if (p->d_ephase) (void)hipFree(p->d_ephase);
  p->fhat.release();
  p->acc.release();
  p->work.release();
  p->aux.release();
  delete p;
  return IB_OK;
}

int ib_collide(ib_plan* p, const double* f, double* q, long long batch, const int* param_idx,
               int with_loss, double heat_bath, void* stream) {
  if (!p || (!f && batch > 0) || (!q && batch > 0) || batch < 0) {
    set_error("invalid collide arguments");
    return IB_ERR_ARG;
  }
  if (batch == 0) return IB_OK;
  hipStream_t st = (hipStream_t)stream;
  p->stream = st;
  int handled = 0;  // shapes with a fused kernel run it
  int rc = collide_fused(p, f, q, batch, param_idx, with_loss, heat_bath, st, &handled);
  if (rc || handled) return rc;
  return collide_plain(p, f, q, batch, param_idx, with_loss, heat_bath, 0, st);
}

int ib_collide_static(ib_plan* p, const double* f, double* q, long long batch, const int* param_idx,
                      int with_loss, double heat_bath, void* stream) {
  if (!p || (!f && batch > 0) || (!q && batch > 0) || batch < 0) {
    set_error("invalid collide arguments");
    return IB_ERR_ARG;
  }
  if (batch == 0) return IB_OK;
  void* base = nullptr;
  int rc = static_scratch(&base);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  p->stream = st;
  // the call's scratch comes from the static region: the plan's own buffers are set
  // aside for its duration and restored after, untouched
  const DevBuf saved[4] = {p->fhat, p->acc, p->work, p->aux};
  p->fhat = p->acc = p->work = p->aux = DevBuf{};
  const int timing = p->timing;
  const size_t budget = p->scratch_budget;
  p->timing = 0;  // the timing hooks create events
  p->scratch_budget = std::min(budget, kStaticScratchBytes / 4);
  Arena ar;
  ar.base = (char*)base;
  ar.size = kStaticScratchBytes;
  g_arena = &ar;
  int handled = 0;
  rc = collide_fused(p, f, q, batch, param_idx, with_loss, heat_bath, st, &handled);
  if (!rc && !handled) rc = collide_plain(p, f, q, batch, param_idx, with_loss, heat_bath, 0, st);
  g_arena = nullptr;
  p->fhat = saved[0];
  p->acc = saved[1];
  p->work = saved[2];
  p->aux = saved[3];
  p->timing = timing;
  p->scratch_budget = budget;
  return rc;
}

int ib_collide_group(int nplans, ib_plan* const* plans, const double* const* f, double* const* q,
                     const long long* batch, const int* const* param_idx, int with_loss,
                     double heat_bath, void* stream) {
  if (nplans < 1 || !plans || !f || !q || !batch) {
    set_error("invalid grouped collide arguments");
    return IB_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  const int prec = plans[0]->precision;
  std::vector<Plan*> gp;
  std::vector<const double*> gf;
  std::vector<double*> gq;
  std::vector<long long> gb;
  std::vector<const int*> gi;
  for (int i = 0; i < nplans; ++i) {
    ib_plan* p = plans[i];
    if (!p || batch[i] < 0 || (batch[i] > 0 && (!f[i] || !q[i]))) {
      set_error("invalid grouped collide arguments");
      return IB_ERR_ARG;
    }
    if (p->precision != prec) {
      set_error("all plans of a grouped launch must have the same precision");
      return IB_ERR_ARG;
    }
    // a plan's scratch and parameter-set tables serve one segment: the same plan twice
    // would share them inside one launch
    for (int j = 0; j < i; ++j)
      if (plans[j] == p) {
        set_error("a plan appears twice in one grouped launch; concatenate its cells instead");
        return IB_ERR_ARG;
      }
  }
  // FP32 plans with a fused form whose whole batch fits their scratch budget: their pair
  // loops run as segments of one grouped launch (group.h) between their front and back work,
  // when there are at least two of them (one alone runs its own kernel: the same results,
  // without the combined kernel's costs)
  int n_fit = 0;
  if (g_group_fused && prec != IB_PREC_FP64)
    for (int i = 0; i < nplans; ++i) n_fit += batch[i] > 0 && fused_group_fits(plans[i], batch[i]);
  const bool group_fused = n_fit >= 2;
  std::vector<GroupSeg> segs;
  struct Grouped {
    Plan* p;
    int i, which;
  };
  std::vector<Grouped> grouped;
  for (int i = 0; i < nplans; ++i) {
    ib_plan* p = plans[i];
    if (batch[i] == 0) continue;
    p->stream = st;
    const int* pi = param_idx ? param_idx[i] : nullptr;
    if (group_fused) {
      GroupSeg seg{};
      int ok = 0, which = 0;
      int rc = fused_group_front(p, f[i], batch[i], pi, st, &seg, &ok, &which);
      if (rc) return rc;
      if (ok) {
        segs.push_back(seg);
        grouped.push_back(Grouped{p, i, which});
        continue;
      }
    }
    {  // shapes with a fused kernel run it
      int handled = 0;
      int rc = collide_fused(p, f[i], q[i], batch[i], pi, with_loss, heat_bath, st, &handled);
      if (rc) return rc;
      if (handled) continue;
    }
    if (staged_runs_alone(p, batch[i])) {  // large batch: the plan's own sub-batches
      int rc = collide_plain(p, f[i], q[i], batch[i], pi, with_loss, heat_bath, 0, st);
      if (rc) return rc;
      continue;
    }
    gp.push_back(p);
    gf.push_back(f[i]);
    gq.push_back(q[i]);
    gb.push_back(batch[i]);
    gi.push_back(pi);
  }
  for (size_t s0 = 0; s0 < gp.size(); s0 += 8) {
    const int n = (int)std::min<size_t>(8, gp.size() - s0);
    int rc = collide_group(gp.data() + s0, n, gf.data() + s0, gq.data() + s0, gb.data() + s0,
                           gi.data() + s0, with_loss, heat_bath, st);
    if (rc) return rc;
  }
  // the 3D 32^3 segments first: their workgroups run longest, and the small shapes' short
  // workgroups then fill the launch's tail
  std::stable_sort(segs.begin(), segs.end(), [](const GroupSeg& a, const GroupSeg& b) {
    return (a.kind == GK_F3 || a.kind == GK_F3K) > (b.kind == GK_F3 || b.kind == GK_F3K);
  });
  for (size_t s0 = 0; s0 < segs.size(); s0 += kMaxGroupSegs) {
    const int n = (int)std::min<size_t>(kMaxGroupSegs, segs.size() - s0);
    int rc = launch_fused_group(segs.data() + s0, n, st);
    if (rc) return rc;
  }
  for (const Grouped& g : grouped) {
    const int* pi = param_idx ? param_idx[g.i] : nullptr;
    int rc = fused_group_back(g.p, g.which, f[g.i], q[g.i], batch[g.i], pi, with_loss, heat_bath, st);
    if (rc) return rc;
  }
  return IB_OK;
}

int ib_set_group_mode(int fused_grouped) {
  if (fused_grouped != 0 && fused_grouped != 1) {
    set_error("group mode must be 0 or 1");
    return IB_ERR_ARG;
  }
  g_group_fused = fused_grouped;
  return IB_OK;
}

int ib_laplacian(ib_plan* p, const double* f, double* out, long long batch, void* stream) {
  if (!p || batch < 0 || (batch > 0 && (!f || !out))) {
    set_error("invalid laplacian arguments");
    return IB_ERR_ARG;
  }
  if (batch == 0) return IB_OK;
  return collide_plain(p, f, out, batch, nullptr, 0, 1.0, 1, (hipStream_t)stream);
}

int ib_moments(const double* f, double* out, long long batch, int dim, int n, double L,
               void* stream) {
  if (batch < 0 || (dim != 2 && dim != 3) || n < 1 || n > 256 || (batch > 0 && (!f || !out))) {
    set_error("invalid moments arguments");
    return IB_ERR_ARG;
  }
  if (batch == 0) return IB_OK;
  return launch_moments(f, out, batch, dim, n, L, (hipStream_t)stream);
}

int ib_project_conservative(double* q, long long batch, int dim, int n, double L, void* stream) {
  if (batch < 0 || (dim != 2 && dim != 3) || n < 1 || n > 256 || (batch > 0 && !q)) {
    set_error("invalid projection arguments");
    return IB_ERR_ARG;
  }
  if (batch == 0) return IB_OK;
  return launch_project(q, batch, dim, n, L, (hipStream_t)stream);
}

int ib_tvd1d(const double* f, double* df, int nx, long long nv, const double* v1, double dtdx,
             int order, int limiter, int bc_lower, int bc_upper, int vx_rows, void* stream) {
  if (bc_lower == 0 || bc_upper == 0) {
    set_error(bc_lower == 0 ? "Lower boundary condition not specified."
                            : "Upper boundary condition not specified.");
    return IB_ERR_BC;
  }
  if ((bc_lower != 1 && bc_lower != 2) || (bc_upper != 1 && bc_upper != 2)) {
    set_error("Boundary condition not implemented");
    return IB_ERR_UNSUPPORTED;
  }
  if (nx < 1 || nv < 1 || !f || !df || !v1 || (order != 1 && order != 2) ||
      (limiter != 0 && limiter != 1)) {
    set_error("invalid tvd1d arguments");
    return IB_ERR_ARG;
  }
  if ((bc_lower == 2 || bc_upper == 2) && nx < 2) {
    set_error("periodic boundaries need nx >= 2");
    return IB_ERR_ARG;
  }
  if (order == 2 && nx + 2 < vx_rows) {
    // the reference slices s[LL-1:UL-1] along v_x and fails to broadcast
    set_error("operands could not be broadcast together (nx + 2 < N)");
    return IB_ERR_ARG;
  }
  return launch_tvd1d(f, df, 1, nx, nv, v1, nv, dtdx, order, limiter, bc_lower, bc_upper, 0,
                      (hipStream_t)stream);
}

int ib_tvd_axis(const double* f, double* df, long long outer, int n, long long inner,
                const double* v, long long vmod, double dtdx, int order, int limiter,
                int bc_lower, int bc_upper, int accumulate, void* stream) {
  if (bc_lower == 0 || bc_upper == 0) {
    set_error(bc_lower == 0 ? "Lower boundary condition not specified."
                            : "Upper boundary condition not specified.");
    return IB_ERR_BC;
  }
  if ((bc_lower != 1 && bc_lower != 2) || (bc_upper != 1 && bc_upper != 2)) {
    set_error("Boundary condition not implemented");
    return IB_ERR_UNSUPPORTED;
  }
  if (outer < 1 || n < 1 || inner < 1 || vmod < 1 || !f || !df || !v ||
      (order != 1 && order != 2) || (limiter != 0 && limiter != 1)) {
    set_error("invalid tvd_axis arguments");
    return IB_ERR_ARG;
  }
  if ((bc_lower == 2 || bc_upper == 2) && n < 2) {
    set_error("periodic boundaries need n >= 2");
    return IB_ERR_ARG;
  }
  return launch_tvd1d(f, df, outer, n, inner, v, vmod, dtdx, order, limiter, bc_lower, bc_upper,
                      accumulate ? 1 : 0, (hipStream_t)stream);
}

int ib_upwind_flux(const double* f, const double* v, double* out, long long n, long long inner,
                   double dx, double dt, int mode, void* stream) {
  if (mode < 0 || mode > 3 || inner < 0 || !f || !v || !out) {
    set_error("invalid upwind_flux arguments");
    return IB_ERR_ARG;
  }
  if (n < (mode == 0 ? 2 : 4)) {
    set_error(mode == 0 ? "grad_vf1 needs at least 2 rows" : "the second-order flux needs at least 4 rows");
    return IB_ERR_ARG;
  }
  if (inner == 0) return IB_OK;
  return launch_upwind(f, v, out, n, inner, dx, dt, mode, (hipStream_t)stream);
}

int ib_fast_sweep(double* f, const double* gain, const double* rho, const double* vx, int nx,
                  int nvr, long long nvc, double dx, double nu, double* err, void* stream) {
  if (nx < 1 || nvr < 1 || nvc < 1 || !f || !gain || !rho || !vx || !err) {
    set_error("invalid fast_sweep arguments");
    return IB_ERR_ARG;
  }
  return launch_fast_sweep(f, gain, rho, vx, nx, nvr, nvc, dx, nu, err, (hipStream_t)stream);
}

int ib_tvd_limit(const double* wave, double* out, long long n, long long inner, const double* s,
                 long long smod, int limiter, void* stream) {
  if (n < 0 || inner < 0 || smod < 1 || (limiter != 0 && limiter != 1) ||
      (n * inner > 0 && (!wave || !out || !s))) {
    set_error("invalid tvd_limit arguments");
    return IB_ERR_ARG;
  }
  if (n > 2 && smod > 1 && smod != inner && smod != (n - 2) * inner) {
    set_error("tvd_limit: speeds must be one value, one per column or one per limited wave");
    return IB_ERR_ARG;
  }
  if (n * inner > 0 && out == wave) {
    set_error("tvd_limit: out must not alias wave");
    return IB_ERR_ARG;
  }
  if (n * inner == 0) return IB_OK;
  return launch_tvd_limit(wave, out, n, inner, s, smod, limiter, (hipStream_t)stream);
}

int ib_ghost_embed(const double* f, double* fbc, int num_dim_x, long long n0, long long n1,
                   long long inner, int num_ghost, int extract, void* stream) {
  if ((num_dim_x != 1 && num_dim_x != 2) || n0 < 0 || n1 < 0 || inner < 0 || num_ghost < 0 ||
      (num_dim_x == 1 && n1 != 1) || (n0 * n1 * inner > 0 && (!f || !fbc))) {
    set_error(num_dim_x == 1 || num_dim_x == 2 ? "invalid ghost_embed arguments"
                                               : "Assumption (1 <= num_dim <= 2) violated.");
    return IB_ERR_ARG;
  }
  if (n0 * n1 * inner == 0) return IB_OK;
  return extract ? launch_embed(fbc, (double*)f, n0, n1, inner, num_ghost, num_dim_x == 2, 1,
                                (hipStream_t)stream)
                 : launch_embed(f, fbc, n0, n1, inner, num_ghost, num_dim_x == 2, 0, (hipStream_t)stream);
}

int ib_ghost_fill(double* fbc, long long outer, long long n, long long inner, int num_ghost, int bc_lower,
                  int bc_upper, void* stream) {
  if (outer < 0 || n < 0 || inner < 0 || num_ghost < 0 || (outer * n * inner > 0 && !fbc) ||
      bc_lower < 0 || bc_lower > 2 || bc_upper < 0 || bc_upper > 2) {
    set_error("invalid ghost_fill arguments");
    return IB_ERR_ARG;
  }
  if ((bc_lower == 2 || bc_upper == 2) && n < num_ghost) {
    set_error("periodic ghost cells need at least num_ghost interior cells");
    return IB_ERR_ARG;
  }
  if (outer * inner * num_ghost == 0 || n == 0) return IB_OK;
  return launch_ghost_axis(fbc, outer, n, inner, num_ghost, bc_lower, bc_upper, (hipStream_t)stream);
}

int ib_axpy(double* y, const double* x, double a, long long n, void* stream) {
  if (n < 0 || (n > 0 && (!y || !x))) {
    set_error("invalid axpy arguments");
    return IB_ERR_ARG;
  }
  if (n == 0) return IB_OK;
  return launch_axpy(y, x, a, n, (hipStream_t)stream);
}

int ib_axpy_chain(double* y, const double* x, double a, double b, double c, long long n,
                  void* stream) {
  if (n < 0 || (n > 0 && (!y || !x))) {
    set_error("invalid axpy_chain arguments");
    return IB_ERR_ARG;
  }
  if (n == 0) return IB_OK;
  return launch_axpy_chain(y, x, a, b, c, n, (hipStream_t)stream);
}

int ib_df_combine(const double* adv, const double* q, double* out, double dt, double kn,
                  long long n, void* stream) {
  if (n < 0 || (n > 0 && (!q || !out))) {
    set_error("invalid df_combine arguments");
    return IB_ERR_ARG;
  }
  if (n == 0) return IB_OK;
  return launch_df_combine(adv, q, out, dt, kn, n, (hipStream_t)stream);
}

}  // extern "C"
