/* mmh.h -- C ABI of the MI355X moving-mesh reconstruction library (libmmh_amd.so).
 *
 * Drop-in boundary for the reconstruction solve of Pen's MMH code
 * (pandaPQY/Moving-Mesh-reconstruction).  The reference exposes Fortran externals with implicit
 * interfaces, every argument by reference; this header gives the same operations with explicit
 * C types.  Each entry point names the reference routine it replaces (file:line in the reference).
 *
 * Conventions
 *  - Fields are N^3 arrays in Fortran order (x fastest), periodic; u is the reference's AoS
 *    u(nu, N, N, N) whose component 1 is the density.
 *  - _d entry points compute in FP64, _f in FP32.  Pointers of the device entry points are HIP
 *    device pointers; work is queued on the context stream and the call returns after it has
 *    completed.  _host entry points take host pointers and copy through context buffers.
 *  - Every entry point returns MMH_OK or an error code; nothing exits the process (the reference
 *    calls stop / pause, mmh/multigrid.f90:28-35,83-86, mmh/gauss1.f90:243-254).
 *  - N must be a power of two, 4 <= N <= 1024 for the solver (16 <= N for relaxing and the
 *    limiter, which the reference also requires, mmh/relaxing.f90:18).
 */
#ifndef MMH_H
#define MMH_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mmh_ctx mmh_ctx;

#define MMH_OK 0
#define MMH_EINVAL 1   /* invalid size or argument */
#define MMH_EDEVICE 2  /* HIP runtime error */
#define MMH_ENOMEM 3   /* device allocation failed */
#define MMH_ENOCONV 4  /* mmh_solve did not reach the tolerance in max_cycles */

#define MMH_MAXLEV 16

/* Per-level V-cycle record (the reference prints these: mmh/mg64.f90:43-48). */
typedef struct {
  int nlev;                     /* levels visited, finest first */
  int n[MMH_MAXLEV];            /* grid size of each level */
  double r1[MMH_MAXLEV];        /* max|r| after pre-smoothing */
  double relaxerr[MMH_MAXLEV];  /* max|r| after post-smoothing */
  int guard[MMH_MAXLEV];        /* 1 if the divergence guard zeroed the level */
  int entry_zerosum;            /* 1 if the initial guess was re-centred (mmh/multigrid.f90:42) */
} mmh_mg_stats;

/* calcdefp diagnostics (mmh/limiter.f90:657-667). */
typedef struct {
  double cmax, detmin, detmax, fluxmin, fluxmax;
  int nexpand, ncompress;
  mmh_mg_stats mg;
} mmh_lim_stats;

/* Solve to a relative residual (builder-defined; the reference does one V-cycle per outer
 * iteration, mmh/limiter.f90:408). */
typedef struct {
  double tol;        /* stop when max|rhs - L u| / max|rhs| <= tol */
  int max_cycles;
  int schedule;      /* 0: reference sweeps nu = 4*ng/N per level, 32+32 on 4^3; 1: fixed */
  int nu_pre, nu_post, nu_coarse;  /* schedule 1: sweeps per level and on 4^3 */
  int ng;            /* schedule 0: the reference's NG (top size if 0) */
  int mixed;         /* 0: FP64; 1: FP32 V-cycles inside FP64 defect correction, or (krylov) as the
                        preconditioner of the FP64 PCG; 2 (krylov): FP32 PCG solves of the
                        correction inside FP64 iterative refinement */
  int krylov;        /* 1: flexible preconditioned CG, one symmetric V-cycle per iteration as the
                        preconditioner (schedule 1 with nu_pre == nu_post; no divergence guard) */
} mmh_solve_opts;

typedef struct {
  int cycles;
  double relres[256];  /* relative residual after each cycle */
  double rhs_norm;
  /* Krylov solve: iterations where p.Lp >= 0 or z.r >= 0.  L is negative definite on mean-free
   * fields and so is the V-cycle preconditioner, so both must stay negative; a count > 0 means the
   * discrete operator (or the preconditioner) lost definiteness, where CG cannot converge. */
  int nonneg_pLp, nonneg_zr;
  /* Krylov solve: 1 if its preconditioner used the Galerkin coarse operators (option "galerkin"),
   * 0 if it fell back to the rebuilt ones because they did not fit in device memory. */
  int galerkin;
} mmh_solve_stats;

/* Per outer iteration: the diagnostics the reference prints in its iteration loop
 * (mmh/main.f90:94-108; sums in FP64 as there) and calcdefp's limiter record. */
typedef struct {
  int iter;
  double defp_sigma;     /* sqrt(sum defp^2 / N^3) */
  double def_sigma;      /* sqrt(sum def^2 / N^3) after def += defp */
  double u_sigma;        /* sqrt(sum (u1 - 1)^2 / N^3) after relaxing */
  double u_min, u_max;   /* extremes of u1 after relaxing */
  long long u_negative;  /* count(u1 < 0) */
  mmh_lim_stats lim;
} mmh_iter_stats;

/* Record of a loss/gradient evaluation (mmh_recon_power_grad_d, mmh_recon_power_grad_mesh_d). */
typedef struct {
  double loss;
  int cycles_forward, cycles_adjoint;      /* PCG iterations of the two solves */
  double relres_forward, relres_adjoint;   /* their final relative residuals */
  int ncompress;                           /* compression-limiter cells at the forward point */
} mmh_adjoint_stats;

/* ---- context ------------------------------------------------------------------------------ */
int mmh_create(mmh_ctx** ctx, int device);
int mmh_destroy(mmh_ctx* ctx);
const char* mmh_error_string(mmh_ctx* ctx);
int mmh_set_stream(mmh_ctx* ctx, void* hip_stream); /* NULL = the null stream (default) */
int mmh_release_workspace(mmh_ctx* ctx);            /* free cached device buffers */
/* Options.  All are bit-identical except "mixed_delta", "galerkin" and "nu_top", which change the
 * preconditioner of the Krylov solve (its iterates and iteration count, not the tolerance reached):
 *   "sweep_impl"  mmh_time_sweeps: 0 = the solver's smoother (sub-lattice layout; the wavefront
 *                 kernel where "wave" is on), 1 = colour kernels on column-major arrays,
 *                 2 = colour kernels on sub-lattice arrays
 *   "wave"        0 (default) / 1: whole-sweep wavefront kernel on levels with n >= "wave_min_n"
 *                 (it measures slower than the colour kernels on MI355X; kept as an option)
 *   "wave_min_n"  smallest level size for the wavefront kernel (default 128)
 *   "wave_p"      1 (default), 2 or 4 points per thread in flight per colour phase
 *   "mixed_delta" 0 (default) / 1: the FP32 preconditioner's stencil sums neighbour differences
 *                 (changes the mixed path's FP32 rounding, not its converged result)
 *   "galerkin"    1 (default) / 0: Galerkin coarse operators (mmh_galerkin_d) on the levels of
 *                 the Krylov solve's preconditioner; 0 rebuilds them from restricted fields as the
 *                 reference's V-cycle does (changes the preconditioner, not the converged result)
 *   "nu_top"      0 (default) or the sweeps each way on the finest level of the Krylov solve's
 *                 preconditioner (nu_pre/nu_post then apply below it); with Galerkin coarse levels
 *                 1 reaches the iteration count of 2 at 512^3
 *   "coarse_fused" 16 (default), 4, 8, 32 or 0: the preconditioner's levels up to this size run in
 *                 one workgroup (bit-identical)
 *   "coarse_fused_ref" 16 (default), 4, 8, 32 or 0: the same for the reference V-cycle (multigrid,
 *                 schedule 0), its residual norms and divergence guard included (bit-identical)
 *   "dots32"      1 (default) / 0: the mixed Krylov solve (mixed = 1) forms z.r and z.r_old from
 *                 r and r_old rounded to FP32 (as its FP32 preconditioner receives r) and keeps
 *                 r_old in FP32 only, and the same pass forms the sum of z that removes its mean;
 *                 the residual recurrence and the stopping test stay FP64
 *   "coarse_lds"  1 (default) / 0: the one-workgroup cycle holds its smallest levels (u, rhs, G)
 *                 in LDS while it runs (bit-identical)
 *   "sweep_slabs" 4 (default), 2, 8, 16, 32 or 0: the mixed solve's finest-level sweeps (n >= 512)
 *                 run all colours on one k-slab before the next (this many slabs; the reverse
 *                 sweep in exactly reverse order).  Another Gauss-Seidel order: the preconditioner
 *                 changes, the FP64 tolerance reached does not
 *   "colour_col"  8 (default), 1, 2, 4 or 16: Gauss-Seidel colour launches on levels
 *                 n >= "colour_col_min_n" (256 by default) update this many points per thread along
 *                 k, carrying the k+1 plane neighbours of one point to the next (1 = one point per
 *                 thread; same values, less memory traffic)
 *   "zero_start"  1 (default) / 0: a V-cycle level's first Gauss-Seidel sweep from u = 0 (every
 *                 level below the finest; the finest too in a preconditioner) loads only the colour
 *                 classes updated earlier in that sweep (same values; less memory traffic)
 *   "pcg_graph"   1 (default) / 0: the Krylov solve (mmh_solve_d, krylov = 1) captures its
 *                 preconditioning V-cycle as a HIP graph on its second call and replays it after;
 *                 on the null stream the solve then runs on a stream of its own, ordered after
 *                 the work already queued
 *   "precond_bf16" 1 (default) / 0: in the mixed Krylov solve (mixed = 1, krylov = 1) the FP32
 *                 preconditioner reads the finest level's metric and the Galerkin operators of levels
 *                 n >= 32 from 16-bit (bfloat16) copies; its arithmetic stays FP32.  Changes the
 *                 preconditioner, not the FP64 tolerance reached
 *   "mesh_solve"  0 (default) / 1 / 2: the mesh solve inside calcdefp and the reconstruction
 *                 iteration (FP64 entry points).  0 is the reference's one V-cycle per outer
 *                 iteration (mmh/limiter.f90:408); 1 repeats V(2,2) cycles of the reference V-cycle
 *                 (divergence guard, 32 + 32 sweeps on 4^3) and 2 runs the Krylov solve, both until
 *                 max|rhs - L u| / max|rhs| <= 10^-"mesh_tol" (MMH_ENOCONV if not reached)
 *   "mesh_tol"    8 (default): the tolerance exponent of "mesh_solve" 1 and 2
 *   "batch_graph", "batch_autotune", "batch_streams", "batch_coarse_fused", "batch_persistent":
 *                 see mmh_solve_batch_d */
int mmh_set_option(mmh_ctx* ctx, const char* key, int value);
/* Current value of an option above, or the read-only "mesh_cycles": the cycles (mesh_solve 1) or
 * iterations (mesh_solve 2) of the last mesh solve inside calcdefp.  No reference counterpart. */
int mmh_get_option(mmh_ctx* ctx, const char* key, int* value);

/* ---- solver (device pointers) --------------------------------------------------------------- */
/* relax (mmh/gauss1.f90:64): nrelax colour sweeps of arr; mod(iopt,8) selects residual output:
 * 1 restricted residual into resid ((N/2)^3), 2 max-norm into *rinf, 3 both. */
int mmh_relax_d(mmh_ctx*, double* arr, double* resid, const double* rhs, const double* deformpot, const double* u,
                int nu, double* rinf, double dx, int n, int nrelax, int iopt);
int mmh_relax_f(mmh_ctx*, float* arr, float* resid, const float* rhs, const float* deformpot, const float* u, int nu,
                float* rinf, float dx, int n, int nrelax, int iopt);

/* multigrid (mmh/multigrid.f90:16): one V-cycle on u with rhs projected in place (rgzerosum);
 * ng is the reference's compile-time NG that sets the sweep schedule (0 = n). */
int mmh_multigrid_d(mmh_ctx*, double* u, double* rhs, const double* defpot, const double* ubig, double dx, int n, int nu,
                    int iopt, int ng, mmh_mg_stats* st);
int mmh_multigrid_f(mmh_ctx*, float* u, float* rhs, const float* defpot, const float* ubig, float dx, int n, int nu,
                    int iopt, int ng, mmh_mg_stats* st);

/* Allocation-free multigrid: the same V-cycle as mmh_multigrid_d (identical results), for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Every workspace is carved from a
 * static 32 MiB scratch array that each GPU holds from library load to process end; the call makes
 * no hipMalloc/hipFree.  n <= 64 (MMH_EINVAL beyond; a 64^3 FP64 cycle needs about 27 MiB).  All
 * calls on one GPU share that GPU's array, which keeps the last call's contents: calls that overlap
 * on one GPU (other host threads, other contexts) race. */
int mmh_multigrid_noalloc_d(mmh_ctx*, double* u, double* rhs, const double* defpot, const double* ubig, double dx,
                            int n, int nu, int iopt, int ng, mmh_mg_stats* st);

/* solve to tolerance with repeated V-cycles (FP64 or mixed precision); rhs projected in place. */
int mmh_solve_d(mmh_ctx*, double* u, double* rhs, const double* defpot, const double* ubig, int n, int nu,
                const mmh_solve_opts* opts, mmh_solve_stats* st);

/* Full residual r = rhs - L phi in FP64 (the formula of calcresid, mmh/gauss1.f90:683-709, at every
 * point; the reference only returns its norm and restriction), with the metric from def and u(1). */
int mmh_residual_d(mmh_ctx*, double* r, const double* phi, const double* rhs, const double* def, const double* u,
                   int nu, int n, double dx);

/* Galerkin coarse operator A_level = R A_(level-1) P of the Krylov solve's preconditioner (A_0 the
 * fine 19-point operator of def and u(1), P the V-cycle's trilinear prolongation, R = P^T / 8 its
 * full weighting): 27 fields of (n >> level)^3 values, column-major, field (ex+1) + 3(ey+1) +
 * 9(ez+1) holding the coefficient of offset (ex, ey, ez).  No counterpart in the reference, whose
 * coarse levels rebuild the metric from restricted fields (mgtemplate.f90); exposed to check the
 * operators the preconditioner uses. */
int mmh_galerkin_d(mmh_ctx*, double* A_out, const double* defpot, const double* ubig, int nu, int n, double dx,
                   int level);

/* Per-realisation record of mmh_solve_batch_d. */
typedef struct {
  int cycles;     /* PCG iterations */
  double relres;  /* final max|rhs - L u| / max|rhs| */
  int converged;
} mmh_batch_stats;

/* Many realisations solved together (no reference counterpart: the reference loops one process
 * per realisation, mmh/run100multigrid.sh).  Item i is the system of mmh_solve_d on device arrays
 * u[i], rhs[i] (projected in place), defpot[i], ubig[i] of size n[i]^3; sizes may differ and are
 * grouped.  Each group advances with one launch per stage for all its grids, one PCG iteration
 * replayed as a HIP graph, and colour-kernel launch shapes tuned on first use.  Only the FP64
 * Krylov form (opts: krylov = 1, schedule = 1, nu_pre = nu_post, mixed = 0); every item's result
 * is bit-identical to its mmh_solve_d.  The groups iterate side by side, each on a stream of its
 * own.  Options "batch_graph", "batch_autotune" and "batch_streams" (mmh_set_option) switch graph
 * replay, tuning and the group streams off (results unchanged).  "batch_coarse_fused" (16 by
 * default, 4, 8, 32 or 0) runs the preconditioner's levels up to that size with one workgroup per
 * grid, one launch for all of them (results unchanged).  "batch_persistent" (0 by default, 8 to 64)
 * solves groups of n up to that size whole in one workgroup per grid, one launch for the entire
 * PCG; its sums run in another order, so results agree with the default to rounding, not bit for
 * bit, and it is slower than the default on MI355X (one CU then carries each grid's fine levels). */
int mmh_solve_batch_d(mmh_ctx*, int count, const int* n, double* const* u, double* const* rhs,
                      double* const* defpot, double* const* ubig, int nu, const mmh_solve_opts* opts,
                      mmh_batch_stats* st);

/* calcdefp (mmh/limiter.f90:17): mesh update defp from def and the density u(1,...); tmp (N^3)
 * and tmp2 ((N+2) N^2) are scratch outputs as in the reference. */
int mmh_calcdefp_d(mmh_ctx*, double* defp, double* tmp, double* tmp2, const double* def, const double* u, double dtold,
                   double dtaumesh, int nfluid, int n, double compressmax, int ng, mmh_lim_stats* st);
int mmh_calcdefp_f(mmh_ctx*, float* defp, float* tmp, float* tmp2, const float* def, const float* u, float dtold,
                   float dtaumesh, int nfluid, int n, float compressmax, int ng, mmh_lim_stats* st);

/* Steps 1-3 of calcdefp (mmh/limiter.f90:134-392): the right-hand side calcdefp passes to
 * multigrid at limiter.f90:408 (clipped mismatch, smoothing, eflux with the expansion limiter and
 * compression gate, smoothing), without the solve. */
int mmh_limiter_rhs_d(mmh_ctx*, double* tmp, const double* def, const double* u, double dtold, double dtaumesh,
                      int nfluid, int n, double compressmax);

/* The reconstruction iteration (mmh/main.f90:89-110): for iter = 1..niter,
 * calcdefp(defp, tmp, tmp2, def, u, dtold, dtaumesh, nfluid); def += defp; relaxing(u, def, defp, dt).
 * u (nfluid = 5 AoS) and def are updated in place, defp holds the last update; st (optional) has
 * niter entries.  The reference uses dtold = dtaumesh = dt = 1 and compressmax = 10. */
int mmh_reconstruct_d(mmh_ctx*, double* u, double* def, double* defp, int n, int nfluid, int niter, double dtold,
                      double dtaumesh, double dt, double compressmax, mmh_iter_stats* st);
int mmh_reconstruct_f(mmh_ctx*, float* u, float* def, float* defp, int n, int nfluid, int niter, float dtold,
                      float dtaumesh, float dt, float compressmax, mmh_iter_stats* st);

/* Density of ng^3 cells from np particle positions xp[3 p + d] (device, in [0, nf) per axis after
 * a periodic wrap), as haloden.f90:49-195 assigns it: each particle's unit mass is shared equally
 * between its nearest-grid-point cell and the empty cells whose nearest particle (searched in
 * growing cubes, haloden.f90:119-167) it is; den (device, ng^3, Fortran order) is then scaled to
 * mean 1 (haloden.f90:187).  wp_out (device, np, may be NULL): the cell count of every particle. */
int mmh_particle_density_f(mmh_ctx*, const float* xp, long long np, int ng, float nf, float* den, int* wp_out);

/* Downsampling (from1024to512.f90:17-23, coarsen 54-66): `levels` rounds of 2x2x2 block sums of
 * den (n^3) into out ((n >> levels)^3), then scaled to mean 1 with an FP64 total (the reference
 * divides by an FP32 SUM).  Device arrays, Fortran order. */
int mmh_coarsen_f(mmh_ctx*, const float* den, int n, int levels, float* out);

/* Column density through the deformed mesh (column_density.f90:14-61): den is scaled to mean 1
 * (FP64 total), and proj (ng^2) is den integrated along z over the deformed column of the cells
 * ng/2 .. ng/2 + nplot - 1, the column's ends moved by the mesh's z displacement from def.
 * Device arrays, Fortran order. */
int mmh_column_density_f(mmh_ctx*, const float* den, const float* def, int ng, int nplot, float* proj);

/* Mesh picture of gridfigure.f90:41-131 (device pointers, FP32 as the reference).  def: the
 * deformation potential of a coarse reconstruction (nf^3); den: density on a finer grid (ng^3,
 * ng a multiple of nf).  pos (2 nf^2): the new (x, y) positions of the coarse nodes in the fine
 * mid-plane, newposition(1:2, ratio::ratio, ratio::ratio, ng/2) with newposition = index + 0.5 +
 * trilinearly interpolated cell-corner displacement; proj (ng^2): the column density of cells
 * ng/2 .. ng/2 + nplot - 1 through the deformed mesh, den scaled to mean 1 first (FP64 total; the
 * reference divides by an FP32 sum); dsp_slab (2 (nf+1)^2 4, or NULL): the reference's
 * dsp_proj.dat, the x/y corner displacements of planes nf/2 .. nf/2+3 before its ghost fill.
 * The ghost layer is filled as the reference does (a copy shifted along all three axes at once).
 * nplot <= ng/2 - ng/nf keeps every corner read inside the displacement array. */
int mmh_gridfigure_f(mmh_ctx*, const float* def, int nf, const float* den, int ng, int nplot, float* pos,
                     float* proj, float* dsp_slab);

/* Binned power spectrum of a real N^3 field (no reference counterpart: the reference reads P(k)
 * from text files, mmh/calcfisherinfo.f90:26-36).  P_b = sum over modes k in bin b of |F(k)|^2 /
 * (N_b N^3), F the unnormalised DFT, N_b the mode count; nbins bins of |k| (grid units, k = 2 pi m / N)
 * with log-spaced edges from 2 pi / N to pi, the last bin closed.  pk, nmodes: host arrays. */
int mmh_power_d(mmh_ctx*, const double* field, int n, int nbins, double* pk, double* nmodes);

/* Loss and gradient of binned reconstructed power for the first outer iteration (no reference
 * counterpart).  Forward map at def = 0: the calcdefp right-hand side from rho
 * (mmh/limiter.f90:134-392), the mesh equation L(rho) phi = rhs solved to tol (FP64 PCG, symmetric
 * V-cycle preconditioner), defp = nsmooth^2(phi) (mmh/limiter.f90:421-423), delta = defp2delta(defp)
 * (mmh/reconstruct.f90:1); loss = sum_b w[b] P_b(delta) with P_b as in mmh_power_d.  grad (device,
 * N^3, may be NULL) receives dloss/drho by the implicit-function adjoint: one adjoint solve of the
 * same symmetric operator and the analytic sensitivity of the metric G = rhobar I to rho.  Refuses
 * (MMH_EINVAL) an input where the compression limiter fires.  rho, grad: device; w, pk: host. */
int mmh_recon_power_grad_d(mmh_ctx*, const double* rho, int n, int nbins, const double* w, double tol, double* pk,
                           double* grad, mmh_adjoint_stats* st);

/* The same loss and gradient at a curved mesh (no reference counterpart): the outer iteration
 * that starts from the mesh potential def (device, N^3; the earlier iterations' result, held
 * fixed) with density rho.  Forward map: the calcdefp right-hand side at (def, rho), whose
 * expansion limiter, compression gate and +-1/dtaumesh clamp are taken on the branches the
 * forward pass takes (mmh/limiter.f90:335-375); rgzerosum with det(def) (mmh/gauss1.f90:1259-1407);
 * L(def, rho) phi = rhs to tol with the curvilinear metric G = rhobar adj(A)/det, A = I + H(def)
 * (mmh/gauss1.f90:1128-1204, the tangle clamp det <= 1e-5 frozen); defp = nsmooth^2(phi);
 * delta = defp2delta(def + defp) (the reconstructed density after the iteration, mmh/main.f90:98,
 * 164).  grad_rho and grad_def (device, N^3, either may be NULL) receive dloss/drho and
 * dloss/d(def): the implicit-function adjoint (one solve of the symmetric operator) with the
 * transposed curvilinear operator's metric sensitivity, dG/d(def) through adj(A) and det, the
 * sensitivity of rgzerosum's det, and the direct dependence of delta on def.  Refuses
 * (MMH_EINVAL) where the compression limiter fires. */
int mmh_recon_power_grad_mesh_d(mmh_ctx*, const double* rho, const double* def, int n, int nbins, const double* w,
                                double tol, double* pk, double* grad_rho, double* grad_def, mmh_adjoint_stats* st);

/* Fisher information as mmh/calcfisherinfo.f90 computes it from nsim realisations of nk binned
 * spectra pk[s*nk + b]: cov (nk x nk, column-major) and covnorm = cov / (mean_i mean_j)
 * (calccov, 280-292), the error bar of each bin's mean (calerror, 269-277), and for every kmax bin
 * m the cumulative information fisher[m] = c^T inv(covnorm[0..m, 0..m]) c over the first m+1 bins
 * (120-132), with the inverse by the same Gauss-Jordan elimination (inverse/Upper/Lower,
 * 294-365).  c (nk) is the derivative vector; host arrays; no device work. */
int mmh_fisher_d(const double* pk, int nsim, int nk, const double* c, double* cov, double* covnorm, double* fisher,
                 double* error);

/* relaxing (mmh/relaxing.f90:21): relaxing-TVD transport of u(5,N,N,N) onto the new mesh. */
int mmh_relaxing_d(mmh_ctx*, double* u, const double* def, const double* defp, double dt, int nfluid, int n);
int mmh_relaxing_f(mmh_ctx*, float* u, const float* def, const float* defp, float dt, int nfluid, int n);

/* scalar2gradient (mmh/gradient.f90:1): psi (N,N,N,3) = spectral gradient with the rlft3 sign. */
int mmh_gradient_d(mmh_ctx*, const double* phi, double* psi, int n);
int mmh_gradient_f(mmh_ctx*, const float* phi, float* psi, int n);
/* defp2delta (mmh/reconstruct.f90:1): delta = spectral -laplacian(phi). */
int mmh_delta_d(mmh_ctx*, const double* phi, double* delta, int n);
int mmh_delta_f(mmh_ctx*, const float* phi, float* delta, int n);

/* nsmooth (mmh/limiter.f90:677), restrict (mmh/mgutil.f90:74), inject (mmh/mgutil.f90:38). */
int mmh_nsmooth_d(mmh_ctx*, double* a, int n);
int mmh_nsmooth_f(mmh_ctx*, float* a, int n);
int mmh_restrict_d(mmh_ctx*, double* half, const double* full, int n, int stride);
int mmh_restrict_f(mmh_ctx*, float* half, const float* full, int n, int stride);
int mmh_inject_d(mmh_ctx*, double* u, const double* uh, int n);
int mmh_inject_f(mmh_ctx*, float* u, const float* uh, int n);
/* rlft3 (mmh/recipes12.f90:77): in-place real 3-D FFT in the Numerical Recipes packing. */
int mmh_rlft3_d(mmh_ctx*, double* data, double* speq, int n, int isign);

/* Synthetic input: lognormal density rho = exp(G - sigma^2/2), mean 1, with G Gaussian of
 * spectrum |k|^slope exp(-(|k|/(ks_frac*pi))^2) and variance sigma^2 (no reference counterpart:
 * the reference reads density files, mmh/main.f90:61-65). */
int mmh_lognormal_d(mmh_ctx*, double* rho, int n, double sigma, double slope, double ks_frac, uint64_t seed);

/* Smoother timing: nsweeps colour sweeps of phi with a precomputed metric, timed with HIP events
 * on the context stream; *ms = average per sweep.  G holds 6 N^3 coefficients (mmh_metric_d). */
int mmh_metric_d(mmh_ctx*, double* G, const double* def, const double* u, int nu, int n, double dx, int lsquare);
int mmh_metric_f(mmh_ctx*, float* G, const float* def, const float* u, int nu, int n, float dx, int lsquare);
int mmh_time_sweeps_d(mmh_ctx*, double* phi, const double* rhs, const double* G, int n, double dx, int nsweeps, float* ms);
int mmh_time_sweeps_f(mmh_ctx*, float* phi, const float* rhs, const float* G, int n, float dx, int nsweeps, float* ms);
/* The same for FP32 phi with the coefficients rounded to bfloat16 (the finest-level smoother of
 * the mixed Krylov solve, option "precond_bf16"); G is the FP64 metric (6 N^3). */
int mmh_time_sweeps_fb(mmh_ctx*, float* phi, const float* rhs, const double* G, int n, float dx, int nsweeps, float* ms);
/* (mmh_time_sweeps_fb sweeps in the slab order of option "sweep_slabs", as the mixed solve does.) */

/* ---- host-pointer entry points (used by the Fortran drop-ins) ------------------------------- */
int mmh_relax_host_f(mmh_ctx*, float* arr, float* resid, const float* rhs, const float* deformpot, const float* u,
                     int nu, float* rinf, float dx, int n, int nrelax, int iopt);
int mmh_relax_host_d(mmh_ctx*, double* arr, double* resid, const double* rhs, const double* deformpot, const double* u,
                     int nu, double* rinf, double dx, int n, int nrelax, int iopt);
int mmh_multigrid_host_f(mmh_ctx*, float* u, float* rhs, const float* defpot, const float* ubig, float dx, int n, int nu,
                         int iopt, int ng);
int mmh_multigrid_host_d(mmh_ctx*, double* u, double* rhs, const double* defpot, const double* ubig, double dx, int n,
                         int nu, int iopt, int ng);
int mmh_calcdefp_host_f(mmh_ctx*, float* defp, float* tmp, float* tmp2, const float* def, const float* u, float dtold,
                        float dtaumesh, int nfluid, int n, float compressmax);
int mmh_calcdefp_host_d(mmh_ctx*, double* defp, double* tmp, double* tmp2, const double* def, const double* u,
                        double dtold, double dtaumesh, int nfluid, int n, double compressmax);
int mmh_relaxing_host_f(mmh_ctx*, float* u, const float* def, const float* defp, float dt, int nfluid, int n);
int mmh_relaxing_host_d(mmh_ctx*, double* u, const double* def, const double* defp, double dt, int nfluid, int n);
int mmh_gradient_host_f(mmh_ctx*, const float* phi, float* psi, int n);
int mmh_delta_host_f(mmh_ctx*, const float* phi, float* delta, int n);
int mmh_nsmooth_host_f(mmh_ctx*, float* a, int n);
int mmh_lognormal_host_d(mmh_ctx*, double* rho, int n, double sigma, double slope, double ks_frac, uint64_t seed);
/* host form of mmh_reconstruct_f for the Fortran driver: diag (7 x niter, may be NULL) receives per
 * iteration defp_sigma, def_sigma, u_sigma, u_min, u_max, u_negative, ncompress. */
int mmh_reconstruct_host_f(mmh_ctx*, float* u, float* def, float* defp, int n, int nfluid, int niter, float dtold,
                           float dtaumesh, float dt, float compressmax, double* diag);

/* The process-wide context used by the Fortran drop-ins (created on first use on device 0). */
mmh_ctx* mmh_default_ctx(void);

#ifdef __cplusplus
}
#endif
#endif
