// batch.hpp -- device tables of a batch of RVEs and the batched launchers
// (batch.hip); the host side is batch_api.hip.
#pragma once

#include "internal.hpp"

namespace mpp {

constexpr int kSpmvShapes = 6, kSpmvDefault = 4;  // default: 27 slots, XCD-grouped blocks

// One RVE of a batch: its geometry and every device array the FP64 path
// touches.  Row-space arrays hold nrow_pad entries, element arrays nelem
// entries; the history arrays are null for an all-elastic RVE.
struct BRve {
  Grid g;
  const double *bmat;          // 8 x 144
  const material_base *mats;   // 3
  const int *et;               // nelem phase ids
  const int *cols;             // SELL-64 columns (shared by RVEs of one size)
  double *vals;                // SELL-64 values
  const double *a0;            // the initial matrix (use_A0; shared by equal RVEs), else nullptr
  double *ke;                  // 576 per element, or per phase in phase mode
  double *be;                  // 24 per element
  double *uw, *u_n, *u_k;      // working / converged / current displacements
  double *x, *r, *z, *p, *Ap, *kinv, *b;
  double *vars_n, *vars_k, *vars_aux;
  double *part;                // partials: max(3 nb_row, 6 nb_elem)
  CgState *cg;                 // ping-pong pair
  // mixed precision (nullptr without): the FP32 copy of the matrix being
  // solved (vals' layout), the inner CG's FP32 vectors and Jacobi weights;
  // x holds the FP64 correction du, r the FP64 residual, kinv the FP64 D^-1
  float *vals32, *x32, *r32, *z32, *p32, *Ap32, *kinv32;
  // mixed precision, two-level preconditioner of the inner CG (cgrid.nc = 0:
  // Jacobi alone): the inverse of the Galerkin coarse matrix of the RVE's
  // initial (elastic) tangent, the x-restricted residual (3 nz ny mx), the
  // restricted residual and the coarse correction (nc each)
  CGrid cgrid;
  const float *cainv;
  float *ct1, *crc, *cec;
  int csplit;    // column splits of the coarse solve
  float *cecp;   // their partial products (csplit > 1), csplit x nc
  int phase_mode;              // all phases elastic: one Ke per phase
  int use_vars;                // the RVE has history (vars_n / vars_k in use)
  int nb_row, nb_elem;
};

// One launch shape: nb blocks in all; map[block] = RVE; blk0[rve] = the
// RVE's first block.
struct BShape {
  int nb;
  const int *map;
  const int *blk0;
};

// Outer (defect-correction) test of one RVE: pn = ||D^-1 (b - A du)||, pn0
// its value at du = 0, target = max(abs, rel pn0), conv = pn < target.
struct BOuter {
  double pn, pn0, target;
  int conv, pad;
};

// The Newton loop of one RVE (newton_raphson, src/solve.cpp:28-82) kept on
// the device, so that a batched Newton iteration needs no host decision
// between its kernels: its = Newton iterations, sits = CG iterations, run =
// still iterating, conv = converged, applied = this iteration's correction
// is in uw, vals_ready = phase mode: vals assembled once.
struct BNewton {
  double norm0, norm;
  int its, sits, run, conv, applied, vals_ready, pad[2];
};

// Newton parameters the device-side loop tests (mppd_newton_params)
struct BNewtonTest {
  int max_its;
  double max_tol, rel_tol;
};

struct BLaunch {
  int nrve;
  const BRve *rv;
  const int *act;  // per-RVE activity mask of this launch
  const int *a0on;  // per RVE: the CG solves with a0 instead of vals
  BShape row, elem, ke;
  // the two-level preconditioner's shapes: x-restriction (threads per
  // (direction, interior x-line, coarse x node)), y/z-restriction (a wave per
  // coarse DOF), coarse solve (64 coarse rows x one column split per block),
  // the splits' sum (RVEs with more than one split); coarse = some RVE has a
  // coarse level
  BShape crx, cryz, cgemv, cgsum;
  int coarse = 0;
  int spmv_shape = kSpmvDefault;    // see launch_b_spmv
  int spmv_shape32 = kSpmvDefault;  // the same choice for the inner FP32 SpMV (launch_bm_spmv)
};

hipError_t launch_b_set_bc(const BLaunch &L, const double *eps, hipStream_t st);
// which: 0 uw <- u_n, 1 u_k <- uw, 2 uw <- u_k, 3 uw <- 0
hipError_t launch_b_move(const BLaunch &L, int which, hipStream_t st);
// what 0 (first nonlinearity): vars_n <- 0, vars_k <- vars_aux; what 1: vars_aux <- 0
hipError_t launch_b_vars_move(const BLaunch &L, int what, hipStream_t st);
// element stiffness (RVEs in act_ke), then vals (RVEs in act_gather)
hipError_t launch_b_mat(const BLaunch &L, const int *act_ke, const int *act_gather, hipStream_t st);
// b = rhs(uw), norms[rve] = ||b||
hipError_t launch_b_rhs(const BLaunch &L, double *norms, hipStream_t st);
// sig[rve * 6 ..] = calc_ave_stress of uw (which 0) or u_k (which 1)
hipError_t launch_b_ave(const BLaunch &L, int which, double *sig, hipStream_t st);
// calc_vars_new of u_k; flags[rve] = nonlinear
hipError_t launch_b_vars_new(const BLaunch &L, int *flags, hipStream_t st);
hipError_t launch_b_cg_init(const BLaunch &L, double min_err, double rel_err, int max_its, hipStream_t st);
// one CG iteration of every RVE (inactive / converged RVEs return at once)
hipError_t launch_b_cg_iter(const BLaunch &L, int parity, hipStream_t st);
// the CG SpMV alone in one of kSpmvShapes shapes: shape = 3 * xcd + c, with
// c = 0, 1, 2 for 9, 27, 81 slots in flight and xcd = 1 for the XCD-grouped
// block order (timing of the autotuner; same results in every shape)
hipError_t launch_b_spmv(const BLaunch &L, int shape, int parity, hipStream_t st);
// uw += x
hipError_t launch_b_axpy(const BLaunch &L, hipStream_t st);

// Device-side Newton control (one thread per RVE unless said otherwise).
// start: after the first residual (norms) of the RVEs in `on`, its = sits =
// 0, norm0 = norm, and the test of the loop's top: run = its < max_its and
// the norm above max(max_tol, rel_tol norm0), conv when the norm passes.
hipError_t launch_b_nt_start(const BLaunch &L, const int *on, const double *norms, BNewtonTest t, BNewton *nst,
                             hipStream_t st);
// prepare one Newton iteration: act = run; the tangent masks of the
// iteration (ke: element stiffness, gather: vals, a0on: solve with the
// initial matrix while its < its_with_A0 when has_a0); applied = pend = 0
hipError_t launch_b_nt_prepare(const BLaunch &L, BNewton *nst, int has_a0, int its_with_A0, int *act, int *ke,
                               int *gather, int *a0on, int *pend, hipStream_t st);
// finish the RVEs whose CG has stopped (state cg[0]: the host queues an even
// number of iterations) and whose correction is not yet applied: uw += x
// (rows), then applied = pend = 1 and sits += CG iterations, then (with
// L.act = pend) the residual b and its norm, then its += 1 and the test of
// the loop's top for the next iteration.  Idempotent for RVEs already
// applied, so it can follow every further chunk of CG iterations.
hipError_t launch_b_nt_finish(const BLaunch &L, BNewtonTest t, BNewton *nst, int *pend, double *norms,
                              hipStream_t st);

// Mixed precision: FP32 inner Jacobi-PCG with FP64 defect correction.
// prep: vals32 = the matrix being solved (vals, or a0 where a0on), kinv /
// kinv32 = 1 / diag, x (du) = 0
hipError_t launch_bm_prep(const BLaunch &L, hipStream_t st);
// r = b - A du (FP64) and the outer test into out[rve]; the inner CG starts
// from r32 = r (x32 = 0) with rel_err = max(tau_min, 0.3 target / pn) and
// max_inner iterations; first: du = 0, so pn0 and target are set, with rel
// times tight_factor for the RVEs whose entry in tight (may be null) is set
hipError_t launch_bm_resid(const BLaunch &L, int first, double rel, double abs_tol, double tau_min, int max_inner,
                           const int *tight, double tight_factor, BOuter *out, hipStream_t st);
// one inner CG iteration of every RVE (done RVEs return at once); with a
// coarse level, z = D^-1 r + P_c A_c^-1 P_c^T r
hipError_t launch_bm_iter(const BLaunch &L, int parity, hipStream_t st);
// the inner FP32 SpMV alone in one of kSpmvShapes shapes, as launch_b_spmv
hipError_t launch_bm_spmv(const BLaunch &L, int shape, int parity, hipStream_t st);
// du += x32
hipError_t launch_bm_accum(const BLaunch &L, hipStream_t st);

}  // namespace mpp
