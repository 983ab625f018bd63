/*
 * mppd.h -- thin C ABI over micropp_amd's gfx950 HIP kernels.
 *
 * This is the layer the host micropp<3> class (and the test harness, through
 * ctypes) calls for each step of the RVE hot path.  Each entry point names
 * the reference function it replaces (paths relative to the Micropp tree).
 *
 * Conventions
 *  - Every function returns an int status: MPPD_OK (0) or a negative code;
 *    mppd_last_error() gives the message of the calling thread's last failure.
 *  - Arrays named *_d are device pointers (allocate with mppd_malloc); arrays
 *    named *_h or fixed-size arrays (eps[6], sig[6]) are host memory.
 *  - Vectors indexed by DOF have nrow = 3 * nx * ny * nz entries in the
 *    reference numbering (DOF = node * 3 + d, node = (k * ny + j) * nx + i).
 *  - Internal variables have nvars = nelem * 8 * 7 entries, index
 *    (e * 8 + gp) * 7 + v (reference include/params.hpp:41).
 *  - Each solver owns one HIP stream; calls on one solver are serialised,
 *    calls on different solvers may run concurrently.
 */
#ifndef MICROPP_AMD_MPPD_H
#define MICROPP_AMD_MPPD_H

#include <stddef.h>

#include "material_base.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MPPD_OK = 0,
  MPPD_ERR_ARG = -1,     /* invalid argument */
  MPPD_ERR_HIP = -2,     /* HIP runtime failure */
  MPPD_ERR_NOMEM = -3,   /* device allocation failed */
  MPPD_ERR_NODEV = -4,   /* no usable gfx950 device */
  MPPD_ERR_NOTCONV = -5  /* iteration limit reached (informational) */
};

typedef struct mppd_rve mppd_rve;
typedef struct mppd_solver mppd_solver;

const char *mppd_last_error(void);
int mppd_device_count(int *n);

/* ---- plain device memory (plumbing for the callers of this ABI) ---- */
int mppd_malloc(void **ptr_d, size_t bytes);
int mppd_free(void *ptr_d);
int mppd_memcpy_h2d(void *dst_d, const void *src_h, size_t bytes);
int mppd_memcpy_d2h(void *dst_h, const void *src_d, size_t bytes);
int mppd_memcpy_d2d(void *dst_d, const void *src_d, size_t bytes);
int mppd_memset(void *ptr_d, int value, size_t bytes);
int mppd_synchronize(void);
/* free and total device memory of the current device, bytes (hipMemGetInfo) */
int mppd_mem_info(size_t *free_bytes, size_t *total_bytes);

/* ---- RVE geometry ----
 * Replaces the geometric part of the micropp<3> constructor: calc_bmat
 * (src/micro3D.cpp:80-120), get_elem_type for every element
 * (src/micropp.cpp:117-124, 338-544), calc_volume_fractions (:609-627) and
 * make_material (src/material.cpp:26-46).  size[] counts nodes per side, as
 * micropp_params_t::size does.  Fails with MPPD_ERR_ARG for an unknown
 * micro_type or material type. */
int mppd_rve_create(mppd_rve **out, const int size[3], int micro_type, const double geo[4],
                    const struct material_base mats[3]);
int mppd_rve_destroy(mppd_rve *r);
/* nrow = 3 nx ny nz, nelem = (nx-1)(ny-1)(nz-1), nvars = 56 nelem */
int mppd_rve_info(const mppd_rve *r, int *nrow, int *nelem, int *nvars);
int mppd_rve_bmat(const mppd_rve *r, double *bmat_h /* [8][6][24] */);
int mppd_rve_elem_types(const mppd_rve *r, int *et_h /* [nelem] */);
int mppd_rve_volume_fractions(const mppd_rve *r, double vf_h[2] /* Vm, Vf */);

/* ---- constitutive laws on n points ----
 * material_t::get_stress / get_ctan / evolute (src/material.cpp:76-302).
 * hist_d may be NULL (no history); sig_d, ctan_d (36 per point, ctan[j*6+i]
 * = dsig_j/deps_i), hist_new_d (7 per point) and nl_d may each be NULL to
 * skip that output. */
int mppd_material_eval(const struct material_base *m, int n, const double *eps_d, const double *hist_d,
                       double *sig_d, double *ctan_d, double *hist_new_d, int *nl_d);

/* ---- one solver: ELL matrix + Jacobi-PCG workspace for one RVE ----
 * ell_init (src/ell-common.cpp:34-139): 27-point stencil, 81 slots per row,
 * out-of-grid neighbours point at column fj with a zero value. */
int mppd_solver_create(mppd_solver **out, mppd_rve *r);
int mppd_solver_destroy(mppd_solver *s);
/* Device-to-device copy and byte fill ordered on the solver's own stream,
 * complete on return.  Unlike mppd_memcpy_d2d / mppd_memset they do not wait
 * for other solvers' work, so solvers driven from other host threads keep
 * running (no counterpart in the reference, which keeps the GP state in host
 * memory: memcpy in src/homogenize.cpp:119-187). */
int mppd_solver_copy(mppd_solver *s, void *dst_d, const void *src_d, size_t bytes);
int mppd_solver_fill(mppd_solver *s, void *ptr_d, int value, size_t bytes);
/* The matrix in the reference layout: cols_h/vals_h are [nrow][81]
 * row-major (either may be NULL). */
int mppd_ell_download(mppd_solver *s, int *cols_h, double *vals_h);
int mppd_ell_upload_vals(mppd_solver *s, const double *vals_h);

/* set_displ_bc (src/micro3D.cpp:26-78): u = eps_t x on the boundary nodes */
int mppd_set_displ_bc(mppd_solver *s, const double eps[6], double *u_d);
/* assembly_rhs (src/assembly.cpp:27-103): b = -(sum_e B^T sigma wg), zero on
 * Dirichlet DOFs; *norm = ||b||_2 (may be NULL) */
int mppd_assembly_rhs(mppd_solver *s, const double *u_d, const double *vars_d, double *b_d, double *norm);
/* assembly_mat (src/assembly.cpp:105-121) with ell_add_3D / ell_set_bc_3D
 * (src/ell-common.cpp:166-297): Jacobian into the solver's own matrix */
int mppd_assembly_mat(mppd_solver *s, const double *u_d, const double *vars_d);
/* ell_mvp (src/ell.cpp:35-44): y = A x */
int mppd_ell_mvp(mppd_solver *s, const double *x_d, double *y_d);
/* ell_solve_cgpd (src/ell.cpp:66-122): Jacobi-PCG from x = 0 on the solver's
 * matrix (or on A0's matrix when A0 != NULL).  *its = iterations, *err = r.z
 * at exit, as in the reference.  Convergence is decided on the device. */
int mppd_cg(mppd_solver *s, const mppd_solver *A0, const double *b_d, double *x_d, double min_err, double rel_err,
            int max_its, int *its, double *err);
/* mppd_cg without device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The same Jacobi-PCG as
 * mppd_cg, with the same arguments and results, but its working set (x, r,
 * z, p, Ap, D^-1, the partial sums and the CG state) lives in a fixed device
 * array of the library (about 19 MB on each GPU the library runs on, held
 * until the process ends) instead of the solver's buffers.  RVEs of more
 * than MPPD_NOALLOC_MAX_ROWS DOF rows (after padding to a multiple of 256;
 * 51^3 nodes fit) fail with MPPD_ERR_ARG.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (several host threads, or several
 * solvers' streams) race and must be serialised by the caller.  Runs on the
 * solver's stream and is complete on return. */
#define MPPD_NOALLOC_MAX_ROWS 398080
int mppd_cg_noalloc(mppd_solver *s, const mppd_solver *A0, const double *b_d, double *x_d, double min_err,
                    double rel_err, int max_its, int *its, double *err);
/* calc_ave_stress (src/average.cpp:57-82) */
int mppd_ave_stress(mppd_solver *s, const double *u_d, const double *vars_d, double sig[6]);
/* calc_vars_new (src/update.cpp:32-56); *nonlinear = OR of the evolute flags */
int mppd_vars_new(mppd_solver *s, const double *u_d, const double *vars_old_d, double *vars_new_d, int *nonlinear);
/* calc_fields (src/average.cpp:84-112): per element, the mean strain and
 * stress over its Gauss points, nelem x 6 each, element-major (host arrays);
 * the cell data of write_vtu (src/output.cpp:67-214). */
int mppd_elem_fields(mppd_solver *s, const double *u_d, const double *vars_d, double *strain_h, double *stress_h);

/* newton_raphson (src/solve.cpp:28-82) */
typedef struct {
  int nr_max_its;
  double nr_max_tol, nr_rel_tol;
  int cg_max_its;
  double cg_abs_tol, cg_rel_tol;
  int its_with_A0; /* used when A0 != NULL */
} mppd_newton_params;
/* out = {its, solver_its, converged}.  u_d is in/out (warm start kept on the
 * interior, boundary overwritten by eps). */
int mppd_newton(mppd_solver *s, const mppd_solver *A0, const double eps[6], const double *vars_old_d, double *u_d,
                const mppd_newton_params *p, int out[3]);

/* ---- mixed-precision multi-right-hand-side solves (no reference
 * counterpart; the reference runs these solves one by one: the six of
 * calc_ctan_lin_fe_models, src/micropp.cpp:255-284, and one newton_raphson
 * per Gauss point, src/homogenize.cpp:111-186) ----
 * One msolver holds an FP32 copy of the stiffness of an RVE whose phases are
 * all linear elastic, and K = 4, 8, 12 or 16 displacement fields.  A solve
 * takes n <= K macro strains, sets u = eps x on the boundary and 0 inside,
 * and iterates FP64 defect correction with FP32 Jacobi-PCG inner solves of
 * all fields at once until, for every field, ||D^-1 (b - A u)|| (FP64)
 * falls below max(abs_tol, rel_tol * its initial value) -- the reference
 * CG's stopping metric, evaluated on the true FP64 residual.  Outputs the
 * FP64 volume-averaged stress of each field (calc_ave_stress,
 * src/average.cpp:57-82). */
typedef struct mppd_msolver mppd_msolver;
typedef struct {
  double rel_tol;    /* outer relative tolerance (e.g. 1e-8) */
  double abs_tol;    /* outer absolute tolerance (e.g. 1e-50) */
  double inner_tol;  /* smallest inner FP32 reduction asked for (e.g. 1e-5) */
  int max_inner_its; /* per inner solve */
  int max_outer;     /* defect-correction steps */
} mppd_mixed_params;
int mppd_msolver_create(mppd_msolver **out, mppd_rve *r, int nrhs);
int mppd_msolver_destroy(mppd_msolver *m);
/* eps_h: n x 6; sig_h: n x 6; stats (may be NULL): {outer steps, inner
 * iterations}; rel_h (may be NULL): n final residual ratios.  Returns
 * MPPD_ERR_NOTCONV when a field did not reach the tolerance. */
int mppd_msolver_solve(mppd_msolver *m, int n, const double *eps_h, const mppd_mixed_params *p, double *sig_h,
                       int stats[2], double *rel_h);
/* displacement field k (nrow doubles, reference DOF numbering) of the last solve */
int mppd_msolver_get_u(mppd_msolver *m, int k, double *u_h);
/* newton_raphson (src/solve.cpp:28-82) of n <= K Gauss points of the
 * msolver's RVE at once (the reference runs one per GP, src/homogenize.cpp:
 * 111-186, OpenMP over the GPs).  GP k starts from u_in[k] (nrow doubles on
 * the device, reference DOF numbering; u_in or u_in[k] NULL: zero) with the
 * boundary values of eps_h[k * 6 ..]; the Newton loop of each GP is the
 * reference's (stop at nr_max_its unconverged, or when ||b||_2 < max(
 * nr_max_tol, nr_rel_tol ||b_0||), tested at the top of an iteration), each
 * correction solved to cg_rel_tol of the current ||D^-1 b|| by FP64 defect
 * correction around the FP32 inner CG of all live GPs together.  Outputs:
 * out[k * 3 ..] = {its, cost, converged} with cost = the inner iterations of
 * the passes the GP took part in; sig_h (may be NULL) n x 6 FP64 volume
 * averages (calc_ave_stress); u_out[k] (may be NULL) the final displacement
 * (may equal u_in[k]); stats (may be NULL) {outer steps, inner iterations}.
 * mp: inner_tol, max_inner_its, max_outer (NULL: 1e-5, cg_max_its, 10). */
int mppd_msolver_newton(mppd_msolver *m, int n, const double *eps_h, const double *const *u_in,
                        double *const *u_out, const mppd_newton_params *p, const mppd_mixed_params *mp,
                        double *sig_h, int *out, int stats[2]);
/* Adjoint derivatives of the volume-averaged stress (no reference
 * counterpart: the reference differentiates by six perturbed solves,
 * calc_ctan_lin_fe_models src/micropp.cpp:255-284 and the FE_FULL loop
 * src/homogenize.cpp:256-275).  One multi-RHS solve carries the six adjoint
 * fields K_II lambda_v = d(sigma_v)/du (one transposed solve per stress
 * component; K is symmetric on the interior DOFs) and the n <= K - 6 forward
 * fields of eps_h.  Outputs: ctan_h[36] = d sigma_v / d eps_i (v * 6 + i);
 * sig_h[n * 6]; dmat_h[n * 36] = d sigma_v / d (E, nu) of phase p of strain s
 * at (s * 6 + v) * 6 + p * 2 + {0: E, 1: nu}; rel_h (may be NULL): the 6 + n
 * final residual ratios.  n = 0 gives ctan alone. */
int mppd_msolver_adjoint(mppd_msolver *m, int n, const double *eps_h, const mppd_mixed_params *p, double *sig_h,
                         double *ctan_h, double *dmat_h, int stats[2], double *rel_h);
/* The adjoint derivatives of a batch of RVEs (no reference counterpart):
 * mppd_msolver_adjoint of each of the nm distinct solvers ms[0 .. nm-1], all
 * with the same n and parameters, run concurrently: up to `workers` host
 * threads (0: 4) take solvers in turn, each solver on its own stream, all on
 * the calling thread's device.  Per-solver arrays are concatenated in solver
 * order: eps_h and sig_h n * 6 each, ctan_h 36, dmat_h n * 36, stats 2, rel_h
 * 6 + n (stats and rel_h may be NULL).  Each solver's results equal
 * mppd_msolver_adjoint's bit for bit.  On failure every solver still runs and
 * the error of the lowest-numbered failing one is returned. */
int mppd_msolver_adjoint_many(int nm, mppd_msolver *const *ms, int n, const double *eps_h, const mppd_mixed_params *p,
                              int workers, double *sig_h, double *ctan_h, double *dmat_h, int *stats, double *rel_h);
/* timing aid: average ms of the FP32 multi-RHS SpMV kernel and of a whole
 * inner iteration, over iters launches, events on the solver's stream */
int mppd_msolver_profile(mppd_msolver *m, int iters, double *spmv_ms, double *iter_ms);
/* The coarse level of the solver's two-level inner preconditioner (no
 * reference counterpart: the reference's CG is Jacobi-preconditioned, src/
 * ell.cpp:80-117).  dims = {nc, hx, hy, hz, mx, my, mz}: coarse DOFs, fine
 * elements per coarse element along x, y, z, interior coarse nodes along x,
 * y, z (all 0 without a coarse level).  ainv_h (may be NULL): nc x nc floats,
 * the inverse of the Galerkin matrix P^T A P, P the trilinear hats of the
 * interior coarse nodes; coarse DOF q * 3 + d of node q = ((cz - 1) my +
 * (cy - 1)) mx + (cx - 1), hat of (cx, cy, cz) at fine node (i, j, k) = the
 * product over the sides of max(0, 1 - |i - cx hx| / hx). */
int mppd_msolver_coarse(const mppd_msolver *m, int dims[7], float *ainv_h);
/* newton_raphson (src/solve.cpp:28-82) with mixed-precision linear solves,
 * for an RVE of any material law (a plastic or damaged RVE's Newton tangent
 * included).  The Newton control flow, the residual (assembly_rhs with the
 * return mapping / damage update), u and the norms are FP64 as in
 * mppd_newton; each linear solve A du = b (the FP64 tangent of assembly_mat)
 * runs as FP64 defect correction around FP32 Jacobi-PCG on an FP32 copy of
 * the matrix's interior block, and stops on the reference CG's metric
 * ||D^-1 (b - A du)|| < max(cg_abs_tol, cg_rel_tol ||D^-1 b||) evaluated in
 * FP64.  mp supplies inner_tol, max_inner_its and max_outer (rel_tol /
 * abs_tol unused; NULL: 1e-5, cg_max_its, 10).  A linear solve whose FP32
 * inner CG breaks down (non-finite: the tangent of a softening law can be
 * indefinite), or that is still short of the tolerance after max_outer
 * outer steps (the FP32 CG can stall on such a tangent), is redone by the
 * FP64 Jacobi-PCG of mppd_newton from 0, so every linear solve meets the
 * reference CG's tolerance.  out as mppd_newton, with solver_its = inner
 * (and any FP64 fallback) iterations; stats (may be NULL) = {outer steps,
 * iterations, FP64 fallbacks (either cause), fallbacks caused by the
 * max_outer cap}.  A0 (may be NULL) as
 * in mppd_newton: the first p->its_with_A0 Newton steps solve with A0's
 * matrix instead of assembling (micropp<3> with use_A0 and mixed precision,
 * for GPs that do not run as a batch);
 * A0 must be assembled (mppd_assembly_mat), not uploaded. */
int mppd_newton_mixed(mppd_solver *s, const mppd_solver *A0, const double eps[6], const double *vars_old_d,
                      double *u_d, const mppd_newton_params *p, const mppd_mixed_params *mp, int out[3],
                      int stats[4]);

/* Timing aid (no reference counterpart; for the benchmark and rocprof
 * cross-checks): runs `iters` CG iterations on the solver's matrix with b_d
 * as right-hand side and the stopping test disabled, timing every SpMV kernel
 * of the iteration (SpMV fused with p.Ap) and every whole iteration with HIP
 * events on the solver's stream.  Averages in milliseconds. */
int mppd_cg_profile(mppd_solver *s, const double *b_d, int iters, double *spmv_ms, double *iter_ms);

/* ---- batches of RVEs (replaces the per-GP loop of homogenize,
 * src/homogenize.cpp:68-99 with homogenize_fe_one_way / homogenize_fe_full,
 * :111-282, for many GPs at once; the reference runs one geometry per
 * micropp<3> instance and its GPs one after the other) ----
 * A batch holds nrve RVEs, each with its own size, geometry, materials,
 * coupling, strain, displacements and history: one per macro Gauss point.
 * Every step of every RVE runs in one launch per finite-element operation
 * over the whole batch; each RVE follows the reference's per-GP algorithm
 * (Newton with the exit quirk, subiterations, sigma from calc_ave_stress or
 * ctan.eps, calc_vars_new with allocation on the first nonlinearity, the
 * FE_FULL perturbation loop reusing u, the stale cost addition), so its
 * results equal a micropp<3> instance of that one RVE. */
typedef struct mppd_batch mppd_batch;
enum { MPPD_FE_LINEAR = 0, MPPD_FE_ONE_WAY = 1, MPPD_FE_FULL = 2 }; /* = FE_LINEAR, FE_ONE_WAY, FE_FULL */
typedef struct {
  int size[3];         /* nodes per side (micropp_params_t::size) */
  int micro_type;      /* MIC_* (include/micropp.hpp) */
  double geo_params[4];
  struct material_base materials[3];
  int coupling;        /* MPPD_FE_LINEAR, MPPD_FE_ONE_WAY or MPPD_FE_FULL */
} mppd_batch_rve;
typedef struct {
  mppd_newton_params newton; /* its_with_A0: used with use_A0 */
  int lin_stress;            /* sigma = ctan . eps (micropp_params_t::lin_stress) */
  int calc_ctan_lin;         /* initial ctan from the six linear solves */
  int subiterations;         /* subiterate a non-converged Newton */
  int nsubiterations;
  int graph_chunk;           /* CG iterations per captured graph; 0 = choose by timing */
  int spmv_chunk;            /* ELL slots per load round of the SpMV: 9, 27 or 81; 0 = choose by timing */
  int block_order;           /* SpMV blocks: 0 in order, 1 grouped per XCD; -1 = choose by timing */
  int use_A0;                /* Newton on the initial matrix (micropp_params_t::use_A0): each distinct
                                RVE's assembly_mat at u = 0 is made at mppd_batch_create, and the first
                                newton.its_with_A0 Newton steps of every RVE solve with it */
  int mixed_precision;       /* every linear solve as FP64 defect correction around an FP32 Jacobi-PCG
                                on an FP32 copy of the matrix, to the reference CG's FP64 metric (the
                                rules of mppd_newton_mixed with mp = NULL); cost = inner iterations */
} mppd_batch_params;
/* the micropp_params_t defaults (include/types.hpp:43-83); graph_chunk,
 * spmv_chunk and block_order chosen by timing at mppd_batch_create */
void mppd_batch_params_default(mppd_batch_params *p);
int mppd_batch_create(mppd_batch **out, int nrve, const mppd_batch_rve *rves, const mppd_batch_params *p);
int mppd_batch_destroy(mppd_batch *b);
/* eps: nrve x 6 (set_strain of every RVE) */
int mppd_batch_set_strains(mppd_batch *b, const double *eps);
int mppd_batch_homogenize(mppd_batch *b);
/* sig: nrve x 6; ctan: nrve x 36 (row-major, ctan[v * 6 + i]) */
int mppd_batch_get_stress(const mppd_batch *b, double *sig);
int mppd_batch_get_ctan(const mppd_batch *b, double *ctan);
/* info: nrve x 4 = {is_non_linear (history allocated), cost, converged, subiterated} */
int mppd_batch_get_info(const mppd_batch *b, int *info);
int mppd_batch_update_vars(mppd_batch *b);
/* stats: {CG iterations queued, graph chunk length, graph launches, ms of the last homogenize, host
 * synchronisations of the last homogenize, batched Newton iterations of the last homogenize (each one
 * launch sequence over all RVEs still iterating)} */
int mppd_batch_stats(const mppd_batch *b, double stats[6]);
/* mixed_precision totals since mppd_batch_create: {outer steps, linear solves redone by the FP64 CG
 * (after an FP32 breakdown, or short of the tolerance at the outer-step cap), those of them caused by
 * the outer-step cap} (no reference counterpart) */
int mppd_batch_mixed_stats(const mppd_batch *b, long out[3]);
/* Timing aid (no reference counterpart): the batch's CG SpMV (fp32 = 0: FP64 on vals; fp32 = 1: the
 * mixed-precision inner FP32 SpMV) in its tuned shape over every RVE, iters launches timed with HIP
 * events on the batch's stream with every CG live: *ms per launch, *bytes (may be NULL) its algorithmic
 * bytes per launch (81 values + p in + Ap out per padded row, at the matrix precision). */
int mppd_batch_spmv_profile(mppd_batch *b, int fp32, int iters, double *ms, double *bytes);
/* the tuned shape: {graph chunk length, SpMV slots per load round, SpMV block order (1 = per XCD)} */
int mppd_batch_tuning(const mppd_batch *b, int out[3]);

/* ---- a batch as working slots for GPs whose state the caller keeps ----
 * micropp<3>::homogenize (src/homogenize.cpp:68-99, OpenMP over the GPs of
 * one instance) runs its FE GPs through a batch of identical RVEs: the GP
 * state (strains, ctan, flags on the host; u_n, u_k, vars_n, vars_k on the
 * device) stays with the instance, and each slot is loaded before and stored
 * after mppd_batch_homogenize.  A slot loaded with coupling MPPD_FE_LINEAR
 * does no FE work (an idle slot). */
typedef struct {
  double strain[6], strain_old[6], stress[6], ctan[36];
  int coupling;   /* MPPD_FE_* */
  int allocated;  /* the GP has history (gp_t::allocated) */
  int cost, converged, subiterated;
} mppd_batch_gp;
/* device bytes one slot of this geometry takes in a batch (columns shared
 * by RVEs of one size not included) */
int mppd_batch_slot_bytes(const mppd_rve *r, size_t *bytes);
/* slot r <- gp; u_n_d (nrow doubles) is copied in when not NULL (else the
 * slot's u_n is zeroed); vars_n_d (nvars doubles) is required when
 * gp->allocated and only then read. */
int mppd_batch_load(mppd_batch *b, int r, const mppd_batch_gp *gp, const double *u_n_d, const double *vars_n_d);
/* gp <- slot r; u_k_d (nrow) is written when not NULL; vars_k_d (nvars) is
 * required when the slot holds history (mppd_batch_get_info tells which). */
int mppd_batch_store(mppd_batch *b, int r, mppd_batch_gp *gp, double *u_k_d, double *vars_k_d);

#ifdef __cplusplus
}
#endif
#endif
