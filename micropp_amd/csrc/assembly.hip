// assembly.hip -- finite-element kernels of the RVE hot path (gfx950), one
// RVE per launch: boundary displacements, element stiffness and forces, the
// atomic-free row-owner gathers that build the ELL matrix and the residual,
// the volume average of the stress and the history update.  The bodies live
// in kernels.hpp (shared with the batched kernels).
//
// The gathers add each entry's element contributions in the reference's
// element order (assembly_mat: ex, ey, ez with ez fastest; assembly_rhs: ez,
// ey, ex with ex fastest) starting from 0.0, with contraction off, so the
// assembled values and residuals are bit-identical to the reference's
// element-by-element scatter.
#include "kernels.hpp"

namespace mpp {

// ---------------------------------------------------------------------------
// material evaluation on n points
// ---------------------------------------------------------------------------
__global__ void k_material_eval(material_base m, int n, const double *eps, const double *hist, double *sig,
                                double *ctan, double *hist_new, int *nl) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double e[6], s[6];
  for (int i = 0; i < 6; ++i) e[i] = eps[p * 6 + i];
  const double *h = hist ? hist + (size_t)p * NVAR : nullptr;
  if (sig) {
    mat_stress(m, e, h, s);
    for (int i = 0; i < 6; ++i) sig[p * 6 + i] = s[i];
  }
  if (ctan) {
    double C[36];
    mat_ctan(m, e, h, C);
    for (int i = 0; i < 36; ++i) ctan[(size_t)p * 36 + i] = C[i];
  }
  if (hist_new || nl) {
    double hn[NVAR];
    for (int i = 0; i < NVAR; ++i) hn[i] = hist_new ? hist_new[(size_t)p * NVAR + i] : 0.0;
    const int f = mat_evolute(m, e, h, hn);
    if (hist_new)
      for (int i = 0; i < NVAR; ++i) hist_new[(size_t)p * NVAR + i] = hn[i];
    if (nl) nl[p] = f;
  }
}

struct Eps6 {
  double v[6];
};

__global__ void __launch_bounds__(256) k_set_displ_bc(Grid g, Eps6 eps, double *u) {
  set_bc_node(g, blockIdx.x * blockDim.x + threadIdx.x, eps.v, u);
}

__global__ void __launch_bounds__(256) k_elem_ke(Grid g, const double *bmat, const material_base *mats, const int *et,
                                                 const double *u, const double *vars, int phase_mode, double *ke) {
  elem_ke_body(g, blockIdx.x, bmat, mats, et, u, vars, phase_mode, ke);
}

__global__ void __launch_bounds__(256) k_gather_vals(Grid g, const int *et, const double *ke, int phase_mode,
                                                     double *vals) {
  gather_vals_row(g, blockIdx.x * blockDim.x + threadIdx.x, et, ke, phase_mode, vals);
}

__global__ void __launch_bounds__(256) k_elem_be(Grid g, const double *bmat, const material_base *mats, const int *et,
                                                 const double *u, const double *vars, double *be) {
  elem_be_body(g, blockIdx.x, bmat, mats, et, u, vars, be);
}

__global__ void __launch_bounds__(256) k_gather_rhs(Grid g, const double *be, double *bvec, double *part) {
  gather_rhs_body(g, blockIdx.x, be, bvec, part);
}

// one block: ||b|| from the partials of ||b||^2
__global__ void __launch_bounds__(256) k_norm_finish(const double *part, int nparts, RedState *red) {
  double t[1];
  seg_sum<1>(part, nparts, t);
  if (threadIdx.x == 0) red->v[0] = sqrt(t[0]);
}

__global__ void __launch_bounds__(256) k_ave_stress(Grid g, const double *bmat, const material_base *mats,
                                                    const int *et, const double *u, const double *vars, double *part) {
  ave_stress_body(g, blockIdx.x, bmat, mats, et, u, vars, part);
}

// one block: the 6 averages from the per-block partials
__global__ void __launch_bounds__(256) k_ave_finish(const double *part, int nparts, double vol, RedState *red) {
  double t[6];
  seg_sum<6>(part, nparts, t);
  if (threadIdx.x == 0)
    for (int v = 0; v < 6; ++v) red->v[v] = t[v] / vol;
}

__global__ void __launch_bounds__(256) k_vars_new(Grid g, const double *bmat, const material_base *mats, const int *et,
                                                  const double *u, const double *vars_old, double *vars_new,
                                                  RedState *red) {
  vars_new_body(g, blockIdx.x, bmat, mats, et, u, vars_old, vars_new, &red->flag);
}

__global__ void __launch_bounds__(256) k_elem_fields(Grid g, const double *bmat, const material_base *mats,
                                                     const int *et, const double *u, const double *vars,
                                                     double *strain, double *stress) {
  elem_fields_body(g, blockIdx.x, bmat, mats, et, u, vars, strain, stress);
}

// ---------------------------------------------------------------------------
static unsigned nblk(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t launch_material_eval(const material_base &m, int n, const double *eps, const double *hist, double *sig,
                                double *ctan, double *hist_new, int *nl, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_material_eval, dim3(nblk(n, 128)), dim3(128), 0, st, m, n, eps, hist, sig, ctan, hist_new, nl);
  return hipGetLastError();
}

hipError_t launch_set_displ_bc(const Grid &g, const double *eps6, double *u, hipStream_t st) {
  Eps6 e;
  for (int i = 0; i < 6; ++i) e.v[i] = eps6[i];
  hipLaunchKernelGGL(k_set_displ_bc, dim3(nblk(g.nn, 256)), dim3(256), 0, st, g, e, u);
  return hipGetLastError();
}

hipError_t launch_elem_ke(const Grid &g, const double *bmat, const material_base *mats_d, const int *et,
                          const double *u, const double *vars, int phase_mode, double *ke, hipStream_t st) {
  const int items = phase_mode ? 3 : g.nelem;
  hipLaunchKernelGGL(k_elem_ke, dim3(nblk(items, kKeWaves)), dim3(64 * kKeWaves), 0, st, g, bmat, mats_d, et, u,
                     vars, phase_mode, ke);
  return hipGetLastError();
}

hipError_t launch_gather_vals(const Grid &g, const int *et, const double *ke, int phase_mode, double *vals,
                              hipStream_t st) {
  hipLaunchKernelGGL(k_gather_vals, dim3(nblk(g.nrow_pad, 256)), dim3(256), 0, st, g, et, ke, phase_mode, vals);
  return hipGetLastError();
}

hipError_t launch_elem_be(const Grid &g, const double *bmat, const material_base *mats_d, const int *et,
                          const double *u, const double *vars, double *be, hipStream_t st) {
  hipLaunchKernelGGL(k_elem_be, dim3(nblk(g.nelem, 256)), dim3(256), 0, st, g, bmat, mats_d, et, u, vars, be);
  return hipGetLastError();
}

hipError_t launch_gather_rhs(const Grid &g, const double *be, double *b, double *partials, RedState *red,
                             hipStream_t st) {
  const unsigned nb = nblk(g.nrow_pad, 256);
  hipLaunchKernelGGL(k_gather_rhs, dim3(nb), dim3(256), 0, st, g, be, b, partials);
  hipLaunchKernelGGL(k_norm_finish, dim3(1), dim3(256), 0, st, partials, (int)nb, red);
  return hipGetLastError();
}

hipError_t launch_ave_stress(const Grid &g, const double *bmat, const material_base *mats_d, const int *et,
                             const double *u, const double *vars, double *partials, RedState *red, hipStream_t st) {
  const unsigned nb = nblk(g.nelem, 256);
  hipLaunchKernelGGL(k_ave_stress, dim3(nb), dim3(256), 0, st, g, bmat, mats_d, et, u, vars, partials);
  hipLaunchKernelGGL(k_ave_finish, dim3(1), dim3(256), 0, st, partials, (int)nb, g.vol_tot, red);
  return hipGetLastError();
}

hipError_t launch_elem_fields(const Grid &g, const double *bmat, const material_base *mats_d, const int *et,
                              const double *u, const double *vars, double *strain, double *stress, hipStream_t st) {
  hipLaunchKernelGGL(k_elem_fields, dim3(nblk(g.nelem, 256)), dim3(256), 0, st, g, bmat, mats_d, et, u, vars, strain,
                     stress);
  return hipGetLastError();
}

hipError_t launch_vars_new(const Grid &g, const double *bmat, const material_base *mats_d, const int *et,
                           const double *u, const double *vars_old, double *vars_new, RedState *red, hipStream_t st) {
  hipLaunchKernelGGL(k_vars_new, dim3(nblk(g.nelem, 256)), dim3(256), 0, st, g, bmat, mats_d, et, u, vars_old,
                     vars_new, red);
  return hipGetLastError();
}

}  // namespace mpp
