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

struct BLaunch {
  int nrve;
  const BRve *rv;
  const int *act;  // per-RVE activity mask of this launch
  const int *a0on;  // per RVE: the CG solves with a0 instead of vals
  BShape row, elem, ke;
  int spmv_shape = kSpmvDefault;  // see launch_b_spmv
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

}  // namespace mpp
