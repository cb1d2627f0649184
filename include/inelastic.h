/* C ABI of the MI355X (gfx950) fast spectral inelastic collision library.
 *
 * Every entry point takes plain pointers and sizes.  Device pointers are HIP
 * device memory owned by the caller (e.g. a torch tensor's data_ptr()); the
 * `stream` argument is a hipStream_t (NULL = the null stream).  Every call
 * returns IB_OK (0) or a negative status; ib_last_error() gives the message
 * of the last failure on the calling thread.
 *
 * A plan owns grow-only device scratch shared by its calls: calls on one plan
 * must not overlap (use one stream per plan, or synchronize between streams);
 * different plans are independent.  A call that needs more scratch than the
 * plan holds synchronizes the device before it grows the buffer (so a plan's
 * first call at a new maximum batch size does not belong in a graph capture).
 *
 * Citations are to the reference interface each entry point replaces
 * (paths relative to the reference repository).
 */
#ifndef INELASTIC_H_
#define INELASTIC_H_

#ifdef __cplusplus
extern "C" {
#endif

#define IB_OK 0
#define IB_ERR_ARG -1       /* ValueError in the reference          */
#define IB_ERR_HIP -2       /* HIP runtime failure                  */
#define IB_ERR_UNSUPPORTED -3 /* NotImplementedError in the reference */
#define IB_ERR_BC -4        /* missing boundary condition (Exception) */

#define IB_PREC_FP64 0      /* plain FP64 form: C2C, staged batched FFTs (2D 32^2:
                               a fused FP64 kernel), FP64 throughout            */
#define IB_PREC_FP32 1      /* FP32 fast form: FP32 transforms and angular sums,
                               FP64 radial sums, final transform and loss term;
                               fused kernels for 3D 32^3 and 16^3 and 2D 16^2 / 32^2 / 64^2
                               (3D 32^3 and 2D 32^2 on the half spectrum of real f, with the
                               exact Nyquist-plane / -line term: the reference's C2C result) */
#define IB_PREC_FP32_KAHAN 2 /* FP32 fast form whose angular sums in the fused
                               3D 32^3 kernel are Kahan-compensated, the
                               compensation kept between angles as bf16 (other
                               shapes: as IB_PREC_FP32) */

typedef struct ib_plan ib_plan;

/* Operator description.  Replaces the state built by
 * src/collision/spectral_mesh.py:8-66 (SpectralMesh) and
 * src/collision/inelastic_collisions.py:30-52 (load_parameters).
 * r/wr: radial Gauss-Legendre nodes on [0, R]; sigma: (m, dim) unit vectors;
 * wsigma: (m,) weights; e_r: (n_param, nr) restitution profile
 * e_j = (e0 - e)/2 tanh(r_j - R/2) + (e0 + e)/2 per parameter set. */
typedef struct {
  int dim;        /* 2 or 3 */
  int n;          /* points per velocity axis: 8, 16, 32, and 64 in 2D */
  int nr;         /* radial nodes */
  int m;          /* circle / sphere nodes */
  double L;       /* velocity box half width */
  double gamma;   /* VHS exponent */
  int precision;  /* IB_PREC_FP64 or IB_PREC_FP32 */
  int n_param;    /* number of restitution parameter sets (>= 1) */
  const double* r;
  const double* wr;
  const double* sigma;
  const double* wsigma;
  const double* e_r;
} ib_plan_desc;

/* Precompute the kernel tables on the device.
 * Replaces src/collision/inelastic_collisions.py:57-137 (perform_precomputation). */
int ib_plan_create(const ib_plan_desc* desc, ib_plan** out);

/* Replace the parameter sets of a plan: (n_param x nr) restitution profiles and,
 * optionally, one VHS exponent per set (gamma = NULL keeps the plan's gamma).
 * Cells select a set through ib_collide's param_idx.
 * Replaces re-running inelastic_collisions.py:30-103 for a new e / gamma. */
int ib_plan_set_params(ib_plan* plan, int n_param, const double* e_r, const double* gamma,
                       void* stream);

int ib_plan_destroy(ib_plan* plan);

/* Tuning knobs of a plan (0 keeps the automatic choice): quadrature pairs per
 * staged chunk, cells per fused launch, scratch budget in bytes (the staged engine's
 * sub-batches and chunks, and the cells per fused launch, are cut to fit it).  New (the reference's FFTW plans have no such knobs); used by the
 * autotuner. */
int ib_plan_set_tuning(ib_plan* plan, long long chunk_pairs, long long fused_cells,
                       long long scratch_budget);

/* Tile shape of the fused engines (new, like set_tuning; used by the autotuner).
 * angle_split in [1, 64]: workgroups per (cell, radial node) of the forms that give one
 * workgroup to a (cell, radial node j) and loop it over the M angles (FP32 3D 32^3 with its
 * Nyquist-plane term, FP32 3D 16^3, FP32 2D 64^2, FP64 2D 32^2): split s takes the angles
 * [s M / S, (s + 1) M / S) and writes its own partial spectrum, so a small batch fills more
 * of the GPU (values above M act as M).  pair_rounds in [0, 1024]: rounds of quadrature
 * pairs per workgroup of the FP32 2D 16^2 / 32^2 form, which sets its workgroups per cell
 * (0 = the default: 16 at 16^2, 8 at 32^2, where a round holds twice the pairs).  Results
 * change only by the order of the FP32 sums (the FP64 2D 32^2 form: of
 * its FP64 sums); for a fixed tile shape a cell's result is the same in any batch.
 * get_tiles returns the values in use. */
int ib_plan_set_tiles(ib_plan* plan, int angle_split, int pair_rounds);
int ib_plan_get_tiles(ib_plan* plan, int* angle_split, int* pair_rounds);

/* Device scratch of a plan: *held = the bytes its buffers hold now; *need = the bytes they
 * will hold after a call of `cells` cells (ib_collide for grouped = 0, a segment of
 * ib_collide_group for grouped = 1) under the current tuning: exact for the fused engines,
 * an upper bound for the staged engine (its pair-sum split is sized by the whole group).
 * Buffers only grow, so need >= held.  Either pointer may be NULL.  Fused launches are cut so
 * that a sub-batch's scratch stays within the plan's scratch budget (at least one cell).
 * New: admission control of a server that shares one GPU among many plans (the reference
 * runs one operator per process). */
int ib_plan_scratch_need(ib_plan* plan, long long cells, int grouped, long long* need, long long* held);

/* HIP-event timing of the dominant kernel of ib_collide: the fused FP32 kernel
 * (cells counted per launch) or, on the staged path, the slab pass (cell x
 * quadrature-pair items counted per launch).  set_timing clears the record;
 * get_timing waits for the recorded events and returns the summed time.
 * New (measurement hooks; the reference has no counterpart). */
int ib_plan_set_timing(ib_plan* plan, int enable);
int ib_plan_get_timing(ib_plan* plan, double* ms_total, long long* launches, long long* units);

/* Q(f, f) for `batch` cells: f, q are (batch, n^dim) float64 device arrays.
 * param_idx: device int32 (batch,) parameter-set index per cell, or NULL for 0;
 * every entry must lie in [0, n_param) (device-resident, so not checked here).
 * with_loss = 0 returns the gain term only (the reference default);
 * heat_bath != 0 adds heat_bath * Laplacian(f).
 * Replaces src/collision/base.py:32-58 (__call__) and
 * src/collision/inelastic_collisions.py:208-227 (collide), :229-230 (laplacian). */
int ib_collide(ib_plan* plan, const double* f, double* q, long long batch,
               const int* param_idx, int with_loss, double heat_bath, void* stream);

/* ib_collide that allocates nothing: the same kernels, with their scratch taken from a
 * static 128 MiB device array (one per GPU, held until the process ends) instead of the
 * plan's buffers, which are left untouched.  A batch whose scratch exceeds 128 MiB is
 * rejected with IB_ERR_ARG before any launch (3D 32^3 FP32: about 7 MiB per cell;
 * 2D 32^2 FP64: about 0.6 MiB per cell).  Calls on one GPU share the array: calls that
 * overlap in time on one GPU (other streams, other threads) race.  The kernels assume no
 * zeroed scratch, so the array's contents from an earlier call do not matter.
 * No counterpart in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). */
int ib_collide_static(ib_plan* plan, const double* f, double* q, long long batch,
                      const int* param_idx, int with_loss, double heat_bath, void* stream);

/* Grouped form of ib_collide for an uneven batch: nplans plans (any grid shapes,
 * one precision), each with its own cells f[i], q[i], batch[i], param_idx[i]
 * (param_idx may be NULL).  FP32 plans with a fused kernel (3D 32^3 and 16^3, 2D 16^2,
 * 32^2 and 64^2) run their quadrature-pair loops as ONE grouped kernel launch (up to 8
 * plans per launch; see ib_set_group_mode); every stage of the staged engine runs as ONE
 * persistent kernel launch over the work of the staged plans (up to 8 per launch); the
 * FP64 2D 32^2 fused kernel runs per plan, and a plan whose batch is too large for its
 * scratch budget (fused) or for chunks of 32 quadrature pairs (staged) runs alone in
 * sub-batches.  A plan may
 * appear only once (its scratch and parameter-set tables serve one segment): a
 * repeated plan is IB_ERR_ARG before any work is queued.  New: the reference
 * evaluates one operator per call (src/collision/base.py:32-58). */
int ib_collide_group(int nplans, ib_plan* const* plans, const double* const* f, double* const* q,
                     const long long* batch, const int* const* param_idx, int with_loss,
                     double heat_bath, void* stream);

/* Mode of ib_collide_group for FP32 plans with a fused form (3D 32^3 and 16^3, 2D 16^2, 32^2
 * and 64^2): 1 (default) runs the quadrature-pair loops of all of them as ONE grouped kernel
 * launch per call (plans whose whole batch fits their scratch budget, when there are at least
 * two such plans; the others, or a single such plan, run their own kernels),
 * with each plan's FP64 prologue, Nyquist-plane term, FP64 reduce and epilogue around it; 0
 * runs one fused launch per plan.  Both modes give bit-identical results.  Process-wide; set it
 * between calls, not while another thread is inside ib_collide_group.  New: the grouped launch
 * of an uneven batch (the reference evaluates one operator per call,
 * src/collision/base.py:32-58). */
int ib_set_group_mode(int fused_grouped);

/* Laplacian of f (spectral), out = Re ifft(-pi^2 |k|^2 / L^2 fft f).
 * Replaces src/collision/inelastic_collisions.py:229-230. */
int ib_laplacian(ib_plan* plan, const double* f, double* out, long long batch, void* stream);

/* FP64 moments per cell: out (batch, dim + 2) = [rho, rho u_1..u_d (momentum), energy],
 * each times dv^dim, v_i = -L + (i + 1/2) dv, n <= 256.  Replaces src/state.py:31-52 (2D only
 * there). */
int ib_moments(const double* f, double* out, long long batch, int dim, int n, double L,
               void* stream);

/* In-place FP64 L2 projection of q (batch, n^dim) onto {sum q = 0, sum q v_a = 0}, n <= 256.
 * New (the reference has no conservative correction). */
int ib_project_conservative(double* q, long long batch, int dim, int n, double L, void* stream);

/* Second-order TVD wave-propagation increment of v_x d/dx f on a 1D grid.
 * f, df: (nx, nv) float64 with nv velocity points per cell; v1: (nv,) advection speed
 * per velocity point; dtdx = dt / dx.  order: 1 or 2; limiter: 0 minmod, 1 MC;
 * bc: 1 extrapolation, 2 periodic.  vx_rows: the leading velocity extent (N), used to
 * reproduce the reference's s[LL-1:UL-1] slice check (nx + 2 >= vx_rows).
 * Replaces src/boltzmann.py:177-205 (df_advection) with src/solver.py:51-105
 * (ghost cells), src/riemann/advection_1D.py:4-21 and src/limiters/tvd.py:6-40. */
int ib_tvd1d(const double* f, double* df, int nx, long long nv, const double* v1,
             double dtdx, int order, int limiter, int bc_lower, int bc_upper, int vx_rows,
             void* stream);

/* Flux-form second-order upwind scheme with the van Leer limiter along axis 0
 * of (n, inner) float64 arrays (the stencil's ghost rows included, as the
 * reference's callers pass them).  mode 0: grad_vf1 (out (n-2, inner));
 * 1: grad_vf2 (out (n-4, inner)); v is (n, inner) for modes 0 and 1.
 * mode 2: F_p_2, mode 3: F_m_2 (out (n-4, inner)) with v the speed aligned with
 * f[1:-1], (n-2, inner).  Same evaluation order as numpy.
 * Replaces src/riemann/upwind.py:40-50 (grad_vf2, grad_vf1) and :20-37 (F_p_2, F_m_2). */
int ib_upwind_flux(const double* f, const double* v, double* out, long long n, long long inner,
                   double dx, double dt, int mode, void* stream);

/* One fast-sweeping iteration, in place on f (nx, nvr, nvc): for velocity rows
 * with vx > 0, f[i] = (nu gain[i] + (vx/dx) f[i-1]) / (nu rho[i] + vx/dx) for
 * i = 1..nx-2 in order; rows with vx < 0 sweep i = nx-2..1 from f[i+1]; rows with
 * vx = 0 and the end cells are kept.  gain (nx, nvr, nvc) is the gain term Q+(f)
 * of the iterate, rho (nx,) its density, vx (nvr,) the row velocities.  *err
 * (device) receives max |f_new - f_old|.
 * Replaces src/iter_solver.py:38-62 (BoltzmannFastSweep.step after its Q+ call). */
int ib_fast_sweep(double* f, const double* gain, const double* rho, const double* vx, int nx,
                  int nvr, long long nvc, double dx, double nu, double* err, void* stream);

/* The TVD transport step of ib_tvd1d along the middle axis of (outer, n, inner)
 * float64 arrays: the speed of column j is v[j % vmod]; ghost cells follow the
 * BC codes of that axis; accumulate != 0 adds the increment to df instead of
 * storing it.  Two calls (x with outer = 1, then y with accumulate = 1) give the
 * dimension-split increment df_x + df_y of a 2D-space grid f (nx, ny, N, N).
 * New (the reference keeps ghost cells for 2D space, src/solver.py:51-105 and
 * src/state.py:95-122, but has no 2D df_advection); per axis it is the step of
 * src/boltzmann.py:177-205. */
int ib_tvd_axis(const double* f, double* df, long long outer, int n, long long inner,
                const double* v, long long vmod, double dtdx, int order, int limiter,
                int bc_lower, int bc_upper, int accumulate, void* stream);

/* Wave limiting alone: out[i, j] = the limited wave of interface i, column j, of the (n, inner)
 * float64 waves, for rows 1 .. n - 2 (rows 0 and n - 1 copied): theta = (s > 0 ? W[i-1] W[i] :
 * W[i+1] W[i]) / W[i]^2 with the masked division (0 where numpy's masked true_divide masks),
 * W~ = phi(theta) W where W != 0, limiter 0 = minmod, 1 = MC.  The speed of (i, j) is
 * s[((i - 1) inner + j) % smod] with smod 1, inner or (n - 2) inner; out must not alias wave.
 * Replaces src/limiters/tvd.py:6-27 (limit; its dtdx feeds only the cfl argument, which
 * neither limiter reads). */
int ib_tvd_limit(const double* wave, double* out, long long n, long long inner, const double* s,
                 long long smod, int limiter, void* stream);

/* Ghost arrays.  ib_ghost_embed copies f (n0, inner) [num_dim_x = 1; n1 = 1] or (n0, n1, inner)
 * [num_dim_x = 2] into the interior of fbc (n0 + 2g, inner) or (n0 + 2g, n1 + 2g, inner), g =
 * num_ghost; extract != 0 copies the interior of fbc back into f.  Replaces src/state.py:109-122
 * (get_fbc_from_f) and :95-107 (set_f_from_fbc). */
int ib_ghost_embed(const double* f, double* fbc, int num_dim_x, long long n0, long long n1,
                   long long inner, int num_ghost, int extract, void* stream);

/* Fill the num_ghost ghost rows on each side of the middle axis of fbc (outer, n + 2 g, inner) in
 * place from that axis's interior rows: bc 1 (BC.extrap) repeats the first / last interior row,
 * 2 (BC.periodic) wraps, 0 leaves that side unchanged.  Replaces src/solver.py:67-105 (_bc_lower,
 * _bc_upper); one call per x dimension gives _apply_bcs (:57-65). */
int ib_ghost_fill(double* fbc, long long outer, long long n, long long inner, int num_ghost, int bc_lower,
                  int bc_upper, void* stream);

/* y <- y + a x over n float64 values.  Replaces the `state.f += M * df` updates of
 * src/boltzmann.py:84,126-128,137-140,153-158. */
int ib_axpy(double* y, const double* x, double a, long long n, void* stream);

/* y <- y + a (b (c x)) over n float64 values, the products taken in that order.  Replaces the
 * chained increments of the telescopic steps, src/boltzmann.py:138-140 (df1 = M0 * df0;
 * state.f += df1; state.f += M1 * df1) and :153-158 (df2 = M1 * df1; state.f += M2 * df2):
 * the caller keeps df0 and passes the multipliers, so df1 and df2 are never stored, and the
 * result rounds as those numpy products do. */
int ib_axpy_chain(double* y, const double* x, double a, double b, double c, long long n,
                  void* stream);

/* out <- adv + (dt * q) / kn (adv may be NULL for 0).  Replaces src/boltzmann.py:160-164 (df). */
int ib_df_combine(const double* adv, const double* q, double* out, double dt, double kn,
                  long long n, void* stream);

/* Host-side evaluation of the device special functions (no GPU needed), for tests:
 * dim = 2 -> J0(x), dim = 3 -> sin(x)/x. */
double ib_host_sp(int dim, double x);

/* Message of the last failed call on this thread ("" if none). */
const char* ib_last_error(void);

/* Library version string. */
const char* ib_version(void);

#ifdef __cplusplus
}
#endif
#endif /* INELASTIC_H_ */
