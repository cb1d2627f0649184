// Host-side plan of the spectral collision operator: quadrature, per-parameter
// set tables, device scratch.  Shared by the plain (staged) engine and the
// fused FP32 engine.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/inelastic.h"

namespace ib {

void set_error(const std::string& msg);

#define IB_HIP(call)                                                                \
  do {                                                                              \
    hipError_t _e = (call);                                                         \
    if (_e != hipSuccess) {                                                         \
      ::ib::set_error(std::string(#call " failed: ") + hipGetErrorString(_e));      \
      return IB_ERR_HIP;                                                            \
    }                                                                               \
  } while (0)

struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  int ensure(size_t want);  // grow-only; returns IB_OK or IB_ERR_HIP (IB_ERR_ARG in an arena)
  void release();
};

// Caller-provided scratch region of one call (ib_collide_static): while g_arena is set on
// the calling thread, DevBuf::ensure carves 256-byte-aligned pieces from it instead of
// allocating, and fails with IB_ERR_ARG when the region is exhausted.
struct Arena {
  char* base = nullptr;
  size_t size = 0, used = 0;
};
extern thread_local Arena* g_arena;
size_t arena_left();  // bytes left in the active arena; SIZE_MAX when none is active

// Bytes a call asks of a plan's four scratch buffers.  Each engine computes its sizes with
// the function it then allocates by, so ib_plan_scratch_need reports what a call will hold.
struct ScratchSizes {
  size_t fhat = 0, acc = 0, work = 0, aux = 0;
};

struct Plan {
  // geometry and quadrature (host copies)
  int dim = 0, n = 0, nr = 0, m = 0, precision = 0;
  bool kahan = false;  // FP32 plan created as IB_PREC_FP32_KAHAN
  double L = 0, gamma = 0;
  long long npts = 0;  // n^dim
  int nq = 0;          // number of |k|^2 values: dim (n/2)^2 + 1
  int n_param = 0;
  std::vector<double> r, wr, sigma, wsigma, e_r;
  double sphere_area = 0;
  // device tables
  double* d_sigma = nullptr;  // [m][dim]
  double* d_rj = nullptr;     // [nr]
  double* d_wsig = nullptr;   // [m]  sphere weights w_sigma
  double* d_crad = nullptr;   // [n_param][nr]  r_j^(gamma_p + d - 1) w_j (gain weight / w_sigma)
  double* d_a = nullptr;      // [n_param][nr]  pi (1 + e_j) / (4 L)
  void* d_spg = nullptr;      // [n_param][nr][nq]  sp(a_j r_j |k|), T
  void* d_lam = nullptr;      // [n_param][nq]  sum_j |S| r^(g+d-1) w_j sp(pi r_j |k| / L), T
  double* d_lam64 = nullptr;  // [n_param][nq]  the same in FP64 (FP64 epilogue of the fused form)
  // 1D phase factors per quadrature pair of the 2D fused FP32 form (float2):
  // E [nr m][2][n] = exp(-i pi r_j sigma_a k / L); G [n_param][nr m][2][n] =
  // exp(i a_j r_j sigma_a k), built only for n_param <= kMaxPhaseSets (else null:
  // the kernel evaluates G itself)
  void* d_ephase = nullptr;
  void* d_gphase = nullptr;
  static constexpr int kMaxPhaseSets = 64;
  std::vector<double> gamma_p;  // per parameter set
  // device scratch (grow-only)
  DevBuf fhat, acc, work, aux;
  long long chunk_pairs = 0;  // pairs per staged chunk (0 = automatic)
  long long fused_cells = 0;  // cells per fused launch (0 = automatic)
  // tile shape of the fused engines: workgroups per (cell, radial node) of the forms with one
  // workgroup per (cell, j), each taking a contiguous range of the angles (1 = the whole loop);
  // rounds of pairs per workgroup of the 2D 16^2 / 32^2 form (0 = 16)
  int angle_split = 1;
  int pair_rounds = 0;
  // splits actually used: at most one per angle
  int splits() const { return angle_split < 1 ? 1 : (angle_split > m ? m : angle_split); }
  hipStream_t stream = nullptr;  // stream of the current call
  // optional HIP-event timing of the dominant kernel (k_fused3 / staged slab pass)
  int timing = 0;
  std::vector<hipEvent_t> ev_start, ev_stop;
  std::vector<long long> ev_cells;
  int timed_begin(hipStream_t st);
  void timed_end(hipStream_t st, long long cells);
  size_t scratch_budget = (size_t)1 << 30;
  // ensure the four buffers hold z
  int reserve(const ScratchSizes& z) {
    int rc;
    if ((rc = fhat.ensure(z.fhat)) || (rc = acc.ensure(z.acc)) || (rc = work.ensure(z.work)) ||
        (rc = aux.ensure(z.aux)))
      return rc;
    return IB_OK;
  }
};

int build_tables(Plan* p, hipStream_t s);

// Cells per launch of a fused engine: the plan's fused_cells (0: the engine's default),
// cut so that a sub-batch's scratch (per_cell bytes per cell) stays within the scratch
// budget, and balanced over the launches the batch then needs.  A static-scratch call keeps
// the uncut count: its region is the bound there.
long long fused_sub(const Plan* P, long long batch, long long dflt, size_t per_cell);
// Scratch sizes of the engine that runs a plan's batch of `batch` cells: the fused engines
// (return 1), or 0 when the plan has no fused kernel.
int fused_sizes(const Plan* P, long long batch, ScratchSizes* z);
int fused2_sizes(const Plan* P, long long batch, ScratchSizes* z);
// Upper bound of the staged engine's sizes, alone (grouped = 0) or as a segment of a grouped
// launch (grouped = 1).
int staged_sizes(const Plan* P, long long batch, int grouped, ScratchSizes* z);
bool staged_runs_alone(const Plan* P, long long batch);

}  // namespace ib
