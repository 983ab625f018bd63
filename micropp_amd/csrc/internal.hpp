// internal.hpp -- device-side data layout of micropp_amd and the launchers
// shared by the .hip translation units.
//
// ELL on the device (SELL-64): rows are grouped into slices of 64 (one
// wavefront); inside a slice the 81 slots are slot-major, so slot s of the 64
// rows of slice q sits at [(q * 81 + s) * 64 + lane].  One wavefront streams
// a slice with 81 fully coalesced 512-byte loads of values and 256-byte loads
// of columns.  Row and slot numbering are the reference's (row = node*3 + fi,
// slot = nb*3 + fj, nb = (dz+1)*9 + (dy+1)*3 + (dx+1)); rows are padded to a
// multiple of 256 with zero values and column 0.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <string>

#include "material_base.h"
#include "mppd.h"

namespace mpp {

constexpr int kSlots = 81;
constexpr int kBlock = 256;     // threads per block of the row-space kernels
constexpr int kRowPad = 256;    // row padding (a block of rows)
constexpr int kMaxSpheres = 64; // MIC3D_RANDOM_SPHERES list length
constexpr int kNumMicroTypes = 14;     // MIC_HOMOGENEOUS .. MIC3D_RANDOM_SPHERES
constexpr int kMicRandomSpheres = 13;  // MIC3D_RANDOM_SPHERES

// Geometry passed by value to kernels.
struct Grid {
  int nx, ny, nz;
  int nex, ney, nez;
  int nn, nrow, nrow_pad, nelem;
  double dx, dy, dz, wg, vol_tot, evol;
};

// Jacobi-PCG scalars of one system (device resident).
struct CgState {
  double rz, pn0, pn, alpha, beta, min_err, rel_err, pAp;
  int its, max_its, done, pad;
  unsigned cnt;
  unsigned pad2[3];
};

// Status of an averaging / norm reduction.
struct RedState {
  double v[8];
  unsigned cnt;
  int flag;
  unsigned pad[2];
};

size_t round_up(size_t v, size_t m);
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define MPP_HIP(call)                                  \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) return mpp::hip_fail(_e, #call); \
  } while (0)

// Device allocation of n elements, zero-filled unless asked otherwise.
template <typename T>
inline int dalloc(T **p, size_t n, bool zero = true) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  if (zero) {
    e = hipMemset(*p, 0, n * sizeof(T));
    // hipMemset runs on the null stream, which the solvers' non-blocking
    // streams do not order against: finish it before the buffer is used
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_fail(e, "hipMemset");
  }
  return MPPD_OK;
}

// Device allocation of n elements, zero-filled on stream st: the fill is
// ordered before st's later work, and the caller synchronizes st before the
// buffer is used from another stream or by a blocking copy.
template <typename T>
inline int dalloc(T **p, size_t n, hipStream_t st) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  e = hipMemsetAsync(*p, 0, n * sizeof(T), st);
  return e == hipSuccess ? MPPD_OK : hip_fail(e, "hipMemsetAsync");
}

inline void dfree(void *p) {
  if (p) (void)hipFree(p);
}

// Test hook: with MPPD_FAIL_ALLOC=1 in the environment, the buffers an entry
// point makes on its first call (mppd_newton_mixed, mppd_msolver_adjoint) are
// freed again after they are allocated and the call fails with
// MPPD_ERR_NOMEM, as an out-of-memory failure would; a test then checks that
// the next call allocates afresh and gives the normal result.
inline bool alloc_fault_injected() {
  const char *v = std::getenv("MPPD_FAIL_ALLOC");
  return v && std::atoi(v) != 0;
}

// Two-level preconditioner of the multi-RHS inner CG (mixed.hip k_c_*): the
// Jacobi weights plus a Galerkin coarse space of trilinear hats on a coarse
// grid aligned with the fine one (h fine elements per coarse element and
// side).  Only interior coarse nodes carry unknowns (the boundary is
// Dirichlet); coarse DOF q * 3 + d of interior coarse node q = ((cz - 1) my +
// (cy - 1)) mx + (cx - 1).
struct CGrid {
  int hx, hy, hz;  // fine elements per coarse element
  int mx, my, mz;  // interior coarse nodes per side
  int nq, nc;      // interior coarse nodes, coarse DOFs (3 nq)
};
// device buffers of the coarse level of one K-field system
struct CoarseDev {
  CGrid cg;
  const float *ainv;  // nc x nc: the inverse of the Galerkin coarse matrix (symmetric)
  float *t1;          // x-restricted residual: 3 K nz ny mx floats
  float *rc;          // restricted residual, nc x K
  float *ecp;         // split partials of the coarse solve, kCoarseSplit x nc x K
  float *ec;          // coarse correction, nc x K
  float *z;           // the preconditioned residual z, in R's layout
  double *prz;        // partials of (r, z) at the inner CG's start: 3 K x (vector blocks)
  void *arena;        // the one device allocation holding all of the above
};
constexpr int kCoarseSplit = 16;    // column splits of the coarse solve (one partial each)
constexpr int kCoarseMaxLine = 256;  // nodes per side a coarse level allows
constexpr int kCoarseMaxDofs = 2200;  // coarse DOFs of a multi-RHS solver (3 x 9^3 = 2187: 10^3 coarse elements)
constexpr int kCoarseMaxPos = 4096;   // fine elements per coarse element (the set-up's table size)
// MPPD_COARSE=0 in the environment turns the coarse levels off
bool coarse_enabled();
// the largest coarse grid aligned with g with at most max_dofs coarse DOFs
// (false: none)
bool coarse_grid(const Grid &g, int max_dofs, CGrid *out);
// ainv (nc x nc floats, device) = the FP32 inverse of the Galerkin coarse
// matrix of the RVE (phase map et_d, per-phase element stiffness ke_d) on cg;
// *ok = false when that matrix is not positive definite
int coarse_inverse(const Grid &g, const CGrid &cg, const int *et_d, const double *ke_d, float *ainv, hipStream_t st,
                   bool *ok);
// coarse.hip: the coarse level of a K-field system on the RVE grid g (per-phase
// element stiffness ke_d, phase map et_d); *out = nullptr when the grid has
// none (or MPPD_COARSE=0 in the environment)
int coarse_setup(const Grid &g, const int *et_d, const double *ke_d, int K, hipStream_t st, CoarseDev **out);
void coarse_free(CoarseDev *c);

// Mixed-precision Newton (mppd_newton_mixed, batches with mixed_precision):
// the linear solve of the last Newton iteration nr_max_its permits stops at
// cg_rel_tol times this factor.  When the Newton ends there unconverged (the
// reference's exit at the iteration cap), that correction is the last term of
// the returned state and no later iteration corrects its error, so that
// solve is taken closer to the exact Newton step.
constexpr double kMixedLastStepTighten = 1e-2;

// Mixed-precision defect correction: each FP32 inner solve stops at
// max(inner_tol, kInnerAim * target / ||D^-1 r||), target being the outer
// test's on the true FP64 residual.  The outer test alone decides
// convergence, so this factor only trades inner iterations against the risk
// of one more outer pass.  A second pass reduces its residual by about 1e-3,
// where the FP32 recursive residual stays within a fraction of a percent of
// the true one; on the 50^3 random-sphere RVE with 12 fields the final
// ||D^-1 r|| came to 0.70 of the target at most (0.30 with a factor of 0.3),
// with 73 instead of 76-77 inner iterations and still two outer passes.
constexpr double kInnerAim = 0.7;

// material_base with the derived constants of the reference's constructors
// (mppd_api.hip); MPPD_ERR_ARG for an unknown law.
int make_material(const material_base &in, material_base &m);

// ---------------- launchers (geometry.hip) ----------------
hipError_t launch_bmat(const Grid &g, double *bmat, hipStream_t st);
hipError_t launch_random_spheres(const double *geo, double *spheres, hipStream_t st);
hipError_t launch_phase_map(const Grid &g, int type, const double *geo, const double *spheres, int nsph, int *et,
                            int *bad, hipStream_t st);
// elements of phase 0 and of phases >= 1 (counts[2], zero on entry)
hipError_t launch_phase_counts(const Grid &g, const int *et, unsigned long long *counts, hipStream_t st);
hipError_t launch_ell_cols(const Grid &g, int *cols, hipStream_t st);
hipError_t launch_sell_to_rows(const Grid &g, const int *cols, const double *vals, int *cols_rm, double *vals_rm,
                               hipStream_t st);
hipError_t launch_rows_to_sell(const Grid &g, const double *vals_rm, double *vals, hipStream_t st);

// ---------------- launchers (assembly.hip) ----------------
hipError_t launch_material_eval(const material_base &m, int n, const double *eps, const double *hist, double *sig,
                                double *ctan, double *hist_new, int *nl, hipStream_t st);
hipError_t launch_set_displ_bc(const Grid &g, const double *eps6, double *u, hipStream_t st);
// Element stiffness: per phase (phase_mode: one 24x24 block per material,
// u/vars ignored) or per element.
hipError_t launch_elem_ke(const Grid &g, const double *bmat, const material_base *mats_d, const int *et, const double *u,
                          const double *vars, int phase_mode, double *ke, hipStream_t st);
hipError_t launch_gather_vals(const Grid &g, const int *et, const double *ke, int phase_mode, double *vals,
                              hipStream_t st);
hipError_t launch_elem_be(const Grid &g, const double *bmat, const material_base *mats, const int *et, const double *u,
                          const double *vars, double *be, hipStream_t st);
hipError_t launch_gather_rhs(const Grid &g, const double *be, double *b, double *partials, RedState *red,
                             hipStream_t st);
hipError_t launch_ave_stress(const Grid &g, const double *bmat, const material_base *mats, const int *et,
                             const double *u, const double *vars, double *partials, RedState *red, hipStream_t st);
hipError_t launch_vars_new(const Grid &g, const double *bmat, const material_base *mats, const int *et, const double *u,
                           const double *vars_old, double *vars_new, RedState *red, hipStream_t st);
// calc_fields: strain / stress nelem x 6 each
hipError_t launch_elem_fields(const Grid &g, const double *bmat, const material_base *mats, const int *et,
                              const double *u, const double *vars, double *strain, double *stress, hipStream_t st);

// ---------------- launchers (solver.hip) ----------------
// CG state is a pair: iteration k reads state[k & 1] and writes
// state[(k + 1) & 1] (parity = k).  part needs 3 * (nrow_pad / 256) doubles.
hipError_t launch_spmv(const Grid &g, const int *cols, const double *vals, const double *x, double *y, hipStream_t st);
hipError_t launch_cg_init(const Grid &g, const double *vals, const double *b, double *x, double *r, double *z,
                          double *p, double *kinv, double *part, double min_err, double rel_err, int max_its,
                          CgState *state, hipStream_t st);
hipError_t launch_cg_iter(const Grid &g, const int *cols, const double *vals, double *x, double *r, double *z,
                          double *p, double *Ap, const double *kinv, double *part, CgState *state, int parity,
                          hipStream_t st);
hipError_t launch_cg_step1(const Grid &g, const int *cols, const double *vals, const double *p, double *Ap,
                           double *part, CgState *state, int parity, hipStream_t st);
hipError_t launch_cg_step23(const Grid &g, double *x, double *r, double *z, double *p, const double *Ap,
                            const double *kinv, double *part, CgState *state, int parity, hipStream_t st);
hipError_t launch_axpy(int n, double a, const double *x, double *y, hipStream_t st);
hipError_t launch_copy(int n, const double *x, double *y, hipStream_t st);

// ---------------- launchers (mixed.hip): K right-hand sides, K in {4,8,12,16} ----------------
size_t mixed_matrix_floats(const Grid &g);
size_t mixed_vector_len(const Grid &g, int K);
size_t mixed_partials(const Grid &g, int K);
size_t mixed_state_bytes();
// FP32 matrix A D^-1 and the Jacobi weights (scratch64: 3 NP doubles)
hipError_t launch_m_asm32(const Grid &g, const int *et, const double *ke, int phase_mode, double *scratch64, float *A,
                          float *kinv, hipStream_t st);
hipError_t launch_m_setbc(const Grid &g, int K, const double *eps, double *U, hipStream_t st);
// F: adjoint sources (FP64, the layout of U) or null
hipError_t launch_m_resid(const Grid &g, int K, const int *et, const double *ke, const double *U, const double *F,
                          const float *kinv, float *R, float *X, float *P, double *part, int first, double rel,
                          double abs_tol, double tau_min, int max_inner, void *state, hipStream_t st,
                          const int *active = nullptr);
// field f of U from / into a displacement vector in the reference numbering (node * 3 + d)
hipError_t launch_m_load_u(const Grid &g, int K, int f, const double *u, double *U, hipStream_t st);
hipError_t launch_m_store_u(const Grid &g, int K, int f, const double *U, double *u, hipStream_t st);
// adjoint: G (3 x 24 x 6 doubles) and the sources F of fields 0 .. nadj-1
hipError_t launch_m_adj_src(const Grid &g, int K, int nadj, const double *bmat, const material_base *mats_d,
                            const int *et, double *G, double *F, hipStream_t st);
// adjoint derivatives from U (fields 0..5 = adjoint fields, s0 .. s0+ns-1 =
// forward fields): ctan[36] = d sigma_v / d eps_i (v * 6 + i), dmat[ns * 36] =
// d sigma_v / d (E, nu) of phase p ((s * 6 + v) * 6 + p * 2 + {0, 1})
size_t mixed_adj_partials(const Grid &g, int ns);
// (ke: the per-phase FP64 element stiffness, 3 x 576)
hipError_t launch_m_adj(const Grid &g, int K, const double *bmat, const material_base *mats_d, const int *et,
                        const double *ke, const double *U, int s0, int ns, double *part, double *ctan, double *dmat,
                        hipStream_t st);
// one inner CG iteration; R X P AP in the 4-field-group layout
// Two-level preconditioner: coarse (may be null: Jacobi alone).  start:
// after launch_m_resid, the inner CG's start direction P = z = D^-1 R +
// P_c A_c^-1 P_c^T R and the state's (r, z); launch_m_iter with coarse applies
// it in every iteration.
hipError_t launch_c_start(const Grid &g, int K, const CoarseDev &c, const float *kinv, const float *R, float *P,
                          void *state, hipStream_t st);
// the full symmetric FP32 copy of the lower triangle (column-major) of the
// FP64 inverse a64 (n x n): ainv[i * n + j]
hipError_t launch_m_iter(const Grid &g, int K, const float *A, const float *kinv, float *R, float *X, float *P,
                         float *AP, double *part, void *state, int parity, hipStream_t st,
                         const CoarseDev *coarse = nullptr);
// blocks per component of the vector kernels (the partial tables' stride)
size_t coarse_vec_blocks(const Grid &g);
hipError_t launch_m_spmv_only(const Grid &g, int K, const float *A, const float *P, float *AP, double *part,
                              const void *state, hipStream_t st);
hipError_t launch_m_accum(const Grid &g, int K, const float *X, double *U, hipStream_t st);
// mixed-precision solves of an assembled FP64 ELL system (field 0 of K = 4):
// FP32 node-block copy of the interior block + Jacobi weights; FP64 residual
// b - A du with the outer test and inner start state; du += X
hipError_t launch_m_from_ell(const Grid &g, const double *vals, float *A, double *kinv64, float *kinv,
                             hipStream_t st);
hipError_t launch_m_resid_ell(const Grid &g, int K, const int *cols, const double *vals, const double *b,
                              const double *du, const double *kinv64, const float *kinv, float *R, float *X,
                              float *P, double *part, int first, double rel, double abs_tol, double tau_min,
                              int max_inner, void *state, hipStream_t st);
hipError_t launch_m_accum_ell(const Grid &g, int K, const float *X, double *du, hipStream_t st);
hipError_t launch_m_ave(const Grid &g, int K, const double *bmat, const material_base *mats_d, const int *et,
                        const double *U, double *part, double *out, hipStream_t st);
}  // namespace mpp

// One RVE: geometry, materials, B, phase map (the opaque mppd_rve of mppd.h).
struct mppd_rve {
  mpp::Grid g;
  int micro_type = 0;
  double geo[4];
  material_base mats[3];
  material_base *mats_d = nullptr;  // the same three on the device
  double *bmat = nullptr;
  int *et = nullptr;
  double *spheres = nullptr;
  int nsph = 0;
  double vf[2] = {0, 0};
  int all_elastic = 1;
};

namespace mpp {

void mixed_state_view(const void *state_h, int parity, int *its, int *all_done);
// field k of the inner state: returns its done flag; *pn its preconditioned
// residual norm, *target the norm at which it stops
int mixed_state_field(const void *state_h, int parity, int k, double *pn, double *target);
void mixed_outer_view(const void *state_h, int K, int *all_conv, double *pn, double *pn0);
// per field: converged flag of the last outer test, ||r||_2 now and at the last `first` test
void mixed_outer_norms(const void *state_h, int K, int *conv, double *rn, double *rn0);

}  // namespace mpp
