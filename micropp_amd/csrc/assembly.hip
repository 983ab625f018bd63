// assembly.hip -- finite-element kernels of the RVE hot path (gfx950):
// boundary displacements, element stiffness and forces, the atomic-free
// row-owner gathers that build the ELL matrix and the residual, the volume
// average of the stress and the history update.
//
// The gathers add each entry's element contributions in the reference's
// element order (assembly_mat: ex, ey, ez with ez fastest; assembly_rhs: ez,
// ey, ex with ex fastest) starting from 0.0, with contraction off, so the
// assembled values and residuals are bit-identical to the reference's
// element-by-element scatter.
#include "internal.hpp"
#include "physics.hpp"
#include "reduce.hpp"

namespace mpp {

struct Mats3 {
  material_base m[3];
};

__device__ __forceinline__ void elem_coords(const Grid &g, int e, int &ex, int &ey, int &ez) {
  ex = e % g.nex;
  ey = (e / g.nex) % g.ney;
  ez = e / (g.nex * g.ney);
}

// Global node id of local node a of element (ex, ey, ez) (src/common.cpp:28-41)
__device__ __forceinline__ int elem_node(const Grid &g, int ex, int ey, int ez, int a) {
  return ((ez + kHexPos[a][2]) * g.ny + (ey + kHexPos[a][1])) * g.nx + ex + kHexPos[a][0];
}

__device__ __forceinline__ bool on_boundary(const Grid &g, int i, int j, int k) {
  return i == 0 || j == 0 || k == 0 || i == g.nx - 1 || j == g.ny - 1 || k == g.nz - 1;
}

// get_strain (src/common.cpp:56-72) from the gathered element displacements
__device__ __forceinline__ void strain_at(const double *b /* [6][24] */, const double *ue, double *eps) {
#pragma clang fp contract(off)
  for (int v = 0; v < 6; ++v) {
    double s = 0;
    for (int i = 0; i < 24; ++i) s += b[v * 24 + i] * ue[i];
    eps[v] = s;
  }
}

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

// ---------------------------------------------------------------------------
// set_displ_bc (src/micro3D.cpp:26-78).  The reference visits the faces
// z=0, z=lz, y=0, y=ly, x=0, x=lx and later faces overwrite shared edges, so
// a node takes the coordinates of the last face containing it.
// ---------------------------------------------------------------------------
struct Eps6 {
  double v[6];
};

__global__ void k_set_displ_bc(Grid g, Eps6 eps, double *u) {
#pragma clang fp contract(off)
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= g.nn) return;
  const int i = node % g.nx, j = (node / g.nx) % g.ny, k = node / (g.nx * g.ny);
  if (!on_boundary(g, i, j, k)) return;
  const double L = 1.0;
  double x[3];
  if (i == g.nx - 1) {
    x[0] = L, x[1] = j * g.dy, x[2] = k * g.dz;
  } else if (i == 0) {
    x[0] = 0, x[1] = j * g.dy, x[2] = k * g.dz;
  } else if (j == g.ny - 1) {
    x[0] = i * g.dx, x[1] = L, x[2] = k * g.dz;
  } else if (j == 0) {
    x[0] = i * g.dx, x[1] = 0, x[2] = k * g.dz;
  } else if (k == g.nz - 1) {
    x[0] = i * g.dx, x[1] = j * g.dy, x[2] = L;
  } else {
    x[0] = i * g.dx, x[1] = j * g.dy, x[2] = 0;
  }
  const double *e = eps.v;
  const double et[3][3] = {{e[0], 0.5 * e[3], 0.5 * e[4]}, {0.5 * e[3], e[1], 0.5 * e[5]}, {0.5 * e[4], 0.5 * e[5], e[2]}};
  for (int d = 0; d < 3; ++d) {
    double s = 0.0;
    for (int c = 0; c < 3; ++c) s += et[d][c] * x[c];
    u[(size_t)node * 3 + d] = s;
  }
}

// ---------------------------------------------------------------------------
// Element stiffness, get_elem_mat (src/assembly.cpp:140-178).  One wavefront
// per element (or per material in phase mode); the 8 tangents and C*B*wg are
// staged in LDS; each lane produces 9 of the 576 entries, summing over
// (gp, m) in the reference order.
// ---------------------------------------------------------------------------
constexpr int kKeWaves = 4;

__global__ void __launch_bounds__(256) k_elem_ke(Grid g, const double *bmat_g, Mats3 mats, const int *et,
                                                 const double *u, const double *vars, int phase_mode, double *ke) {
#pragma clang fp contract(off)
  __shared__ double bm[8 * 144];
  __shared__ double ue_s[kKeWaves][24];
  __shared__ double C_s[kKeWaves][8 * 36];
  __shared__ double cb_s[kKeWaves][8 * 144];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  for (int i = tid; i < 8 * 144; i += blockDim.x) bm[i] = bmat_g[i];
  const int item = blockIdx.x * kKeWaves + w;
  const int nitems = phase_mode ? 3 : g.nelem;
  const bool active = item < nitems;
  int ex = 0, ey = 0, ez = 0, ph = 0;
  if (active) {
    if (phase_mode) {
      ph = item;
    } else {
      elem_coords(g, item, ex, ey, ez);
      ph = et[item];
    }
  }
  if (lane < 24) {
    double v = 0.0;
    if (active && !phase_mode && u) v = u[(size_t)elem_node(g, ex, ey, ez, lane / 3) * 3 + lane % 3];
    ue_s[w][lane] = v;
  }
  __syncthreads();
  if (active && lane < 8) {
    double eps[6];
    strain_at(&bm[lane * 144], ue_s[w], eps);
    const double *h = (vars && !phase_mode) ? vars + ((size_t)item * 8 + lane) * NVAR : nullptr;
    mat_ctan(mats.m[ph], eps, h, &C_s[w][lane * 36]);
  }
  __syncthreads();
  if (active) {
    for (int q = lane; q < 8 * 144; q += 64) {
      const int gp = q / 144, m = (q / 24) % 6, jj = q % 24;
      const double *C = &C_s[w][gp * 36 + m * 6];
      const double *b = &bm[gp * 144];
      double s = 0.0;
      for (int k = 0; k < 6; ++k) s += C[k] * b[k * 24 + jj];
      cb_s[w][q] = s * g.wg;
    }
  }
  __syncthreads();
  if (active) {
    double *out = ke + (size_t)item * 576;
    for (int q = lane; q < 576; q += 64) {
      const int ii = q / 24, jj = q % 24;
      double acc = 0.0;
      for (int gp = 0; gp < 8; ++gp)
        for (int m = 0; m < 6; ++m) acc += bm[gp * 144 + m * 24 + ii] * cb_s[w][gp * 144 + m * 24 + jj];
      out[q] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// Row-owner gather of the element blocks into SELL vals: assembly_mat's
// element loop + ell_add_3D (src/assembly.cpp:105-121,
// src/ell-common.cpp:166-198) and ell_set_bc_3D (:238-297).  For row
// (node N, fi) and slot (neighbour M, fj) the contributing elements are those
// holding both N and M, visited in reference order ex, ey, ez.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_gather_vals(Grid g, const int *et, const double *ke, int phase_mode,
                                                     double *vals) {
#pragma clang fp contract(off)
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= g.nrow_pad) return;
  double *out = vals + (size_t)(row >> 6) * kSlots * 64 + (row & 63);
  if (row >= g.nrow) {
    for (int s = 0; s < kSlots; ++s) out[s * 64] = 0.0;
    return;
  }
  const int node = row / 3, fi = row % 3;
  const int i = node % g.nx, j = (node / g.nx) % g.ny, k = node / (g.nx * g.ny);
  if (on_boundary(g, i, j, k)) {
    for (int s = 0; s < kSlots; ++s) out[s * 64] = (s == 13 * 3 + fi) ? 1.0 : 0.0;
    return;
  }
  // element blocks around the node, reference order (a, b, c) = (x, y, z)
  const double *blk[8];
  int la[8];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) {
        const int q = (a * 2 + b) * 2 + c;
        const int ex = i - 1 + a, ey = j - 1 + b, ez = k - 1 + c;  // interior node: always inside
        const int e = (ez * g.ney + ey) * g.nex + ex;
        blk[q] = ke + (size_t)(phase_mode ? et[e] : e) * 576;
        la[q] = local_node(1 - a, 1 - b, 1 - c);
      }
  for (int n = 0; n < 27; ++n) {
    const int dz = n / 9 - 1, dy = (n / 3) % 3 - 1, dx = n % 3 - 1;
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    for (int a = 0; a < 2; ++a) {
      const int px = 1 - a + dx;
      if (px < 0 || px > 1) continue;
      for (int b = 0; b < 2; ++b) {
        const int py = 1 - b + dy;
        if (py < 0 || py > 1) continue;
        for (int c = 0; c < 2; ++c) {
          const int pz = 1 - c + dz;
          if (pz < 0 || pz > 1) continue;
          const int q = (a * 2 + b) * 2 + c;
          const double *r = blk[q] + (la[q] * 3 + fi) * 24 + local_node(px, py, pz) * 3;
          acc0 += r[0];
          acc1 += r[1];
          acc2 += r[2];
        }
      }
    }
    out[(n * 3 + 0) * 64] = acc0;
    out[(n * 3 + 1) * 64] = acc1;
    out[(n * 3 + 2) * 64] = acc2;
  }
}

// ---------------------------------------------------------------------------
// Element internal forces, get_elem_rhs (src/assembly.cpp:123-138): one
// thread per element; B staged in LDS.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_elem_be(Grid g, const double *bmat_g, Mats3 mats, const int *et,
                                                 const double *u, const double *vars, double *be) {
#pragma clang fp contract(off)
  __shared__ double bm[8 * 144];
  for (int i = threadIdx.x; i < 8 * 144; i += blockDim.x) bm[i] = bmat_g[i];
  __syncthreads();
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.nelem) return;
  int ex, ey, ez;
  elem_coords(g, e, ex, ey, ez);
  double ue[24];
  for (int a = 0; a < 8; ++a) {
    const size_t base = (size_t)elem_node(g, ex, ey, ez, a) * 3;
    ue[a * 3 + 0] = u[base + 0];
    ue[a * 3 + 1] = u[base + 1];
    ue[a * 3 + 2] = u[base + 2];
  }
  const material_base &m = mats.m[et[e]];
  double acc[24];
  for (int i = 0; i < 24; ++i) acc[i] = 0;
  for (int gp = 0; gp < 8; ++gp) {
    double eps[6], sig[6];
    strain_at(&bm[gp * 144], ue, eps);
    mat_stress(m, eps, vars ? vars + ((size_t)e * 8 + gp) * NVAR : nullptr, sig);
    for (int i = 0; i < 24; ++i)
      for (int v = 0; v < 6; ++v) acc[i] += bm[gp * 144 + v * 24 + i] * sig[v] * g.wg;
  }
  double *out = be + (size_t)e * 24;
  for (int i = 0; i < 24; ++i) out[i] = acc[i];
}

// ---------------------------------------------------------------------------
// Row-owner gather of the element forces: assembly_rhs (src/assembly.cpp:
// 27-103): element order ez, ey, ex; Dirichlet DOFs zeroed; negated; ||b||.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_gather_rhs(Grid g, const double *be, double *bvec, double *partials,
                                                    RedState *red) {
#pragma clang fp contract(off)
  __shared__ double lds[(256 / 64) * 1];
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  double sq = 0.0;
  if (row < g.nrow) {
    const int node = row / 3, d = row % 3;
    const int i = node % g.nx, j = (node / g.nx) % g.ny, k = node / (g.nx * g.ny);
    double acc = 0.0;
    if (!on_boundary(g, i, j, k)) {
      for (int c = 0; c < 2; ++c)
        for (int b = 0; b < 2; ++b)
          for (int a = 0; a < 2; ++a) {
            const int e = ((k - 1 + c) * g.ney + (j - 1 + b)) * g.nex + (i - 1 + a);
            acc += be[(size_t)e * 24 + local_node(1 - a, 1 - b, 1 - c) * 3 + d];
          }
    }
    const double v = -acc;
    bvec[row] = v;
    sq = v * v;
  }
  double val[1] = {sq}, tot[1];
  if (grid_sum<1, 256>(val, partials, &red->cnt, tot, lds)) {
    if (threadIdx.x == 0) red->v[0] = sqrt(tot[0]);
  }
}

// ---------------------------------------------------------------------------
// calc_ave_stress (src/average.cpp:57-82): per-element partial sums over the
// Gauss points, then a grid sum (tree order).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_ave_stress(Grid g, const double *bmat_g, Mats3 mats, const int *et,
                                                    const double *u, const double *vars, double *partials,
                                                    RedState *red) {
#pragma clang fp contract(off)
  __shared__ double bm[8 * 144];
  __shared__ double lds[(256 / 64) * 6];
  for (int i = threadIdx.x; i < 8 * 144; i += blockDim.x) bm[i] = bmat_g[i];
  __syncthreads();
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double part[6] = {0, 0, 0, 0, 0, 0};
  if (e < g.nelem) {
    int ex, ey, ez;
    elem_coords(g, e, ex, ey, ez);
    double ue[24];
    for (int a = 0; a < 8; ++a) {
      const size_t base = (size_t)elem_node(g, ex, ey, ez, a) * 3;
      ue[a * 3 + 0] = u[base + 0];
      ue[a * 3 + 1] = u[base + 1];
      ue[a * 3 + 2] = u[base + 2];
    }
    const material_base &m = mats.m[et[e]];
    for (int gp = 0; gp < 8; ++gp) {
      double eps[6], sig[6];
      strain_at(&bm[gp * 144], ue, eps);
      mat_stress(m, eps, vars ? vars + ((size_t)e * 8 + gp) * NVAR : nullptr, sig);
      for (int v = 0; v < 6; ++v) part[v] += sig[v] * g.wg;
    }
  }
  double tot[6];
  if (grid_sum<6, 256>(part, partials, &red->cnt, tot, lds)) {
    if (threadIdx.x == 0)
      for (int v = 0; v < 6; ++v) red->v[v] = tot[v] / g.vol_tot;
  }
}

// ---------------------------------------------------------------------------
// calc_vars_new (src/update.cpp:32-56): evolute per Gauss point, OR of the
// nonlinear flags.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_vars_new(Grid g, const double *bmat_g, Mats3 mats, const int *et,
                                                  const double *u, const double *vars_old, double *vars_new,
                                                  RedState *red) {
#pragma clang fp contract(off)
  __shared__ double bm[8 * 144];
  for (int i = threadIdx.x; i < 8 * 144; i += blockDim.x) bm[i] = bmat_g[i];
  __syncthreads();
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.nelem) return;
  int ex, ey, ez;
  elem_coords(g, e, ex, ey, ez);
  double ue[24];
  for (int a = 0; a < 8; ++a) {
    const size_t base = (size_t)elem_node(g, ex, ey, ez, a) * 3;
    ue[a * 3 + 0] = u[base + 0];
    ue[a * 3 + 1] = u[base + 1];
    ue[a * 3 + 2] = u[base + 2];
  }
  const material_base &m = mats.m[et[e]];
  int nl = 0;
  for (int gp = 0; gp < 8; ++gp) {
    double eps[6];
    strain_at(&bm[gp * 144], ue, eps);
    const size_t off = ((size_t)e * 8 + gp) * NVAR;
    nl |= mat_evolute(m, eps, vars_old ? vars_old + off : nullptr, vars_new + off);
  }
  if (nl) atomicOr(&red->flag, 1);
}

// ---------------------------------------------------------------------------
static unsigned nblk(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }
static Mats3 pack(const material_base *m) {
  Mats3 M;
  for (int i = 0; i < 3; ++i) M.m[i] = m[i];
  return M;
}

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

hipError_t launch_elem_ke(const Grid &g, const double *bmat, const material_base *mats, const int *et, const double *u,
                          const double *vars, int phase_mode, double *ke, hipStream_t st) {
  const int items = phase_mode ? 3 : g.nelem;
  hipLaunchKernelGGL(k_elem_ke, dim3(nblk(items, kKeWaves)), dim3(64 * kKeWaves), 0, st, g, bmat, pack(mats), et, u,
                     vars, phase_mode, ke);
  return hipGetLastError();
}

hipError_t launch_gather_vals(const Grid &g, const int *et, const double *ke, int phase_mode, double *vals,
                              hipStream_t st) {
  hipLaunchKernelGGL(k_gather_vals, dim3(nblk(g.nrow_pad, 256)), dim3(256), 0, st, g, et, ke, phase_mode, vals);
  return hipGetLastError();
}

hipError_t launch_elem_be(const Grid &g, const double *bmat, const material_base *mats, const int *et, const double *u,
                          const double *vars, double *be, hipStream_t st) {
  hipLaunchKernelGGL(k_elem_be, dim3(nblk(g.nelem, 256)), dim3(256), 0, st, g, bmat, pack(mats), et, u, vars, be);
  return hipGetLastError();
}

hipError_t launch_gather_rhs(const Grid &g, const double *be, double *b, double *partials, RedState *red,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_gather_rhs, dim3(nblk(g.nrow, 256)), dim3(256), 0, st, g, be, b, partials, red);
  return hipGetLastError();
}

hipError_t launch_ave_stress(const Grid &g, const double *bmat, const material_base *mats, const int *et,
                             const double *u, const double *vars, double *partials, RedState *red, hipStream_t st) {
  hipLaunchKernelGGL(k_ave_stress, dim3(nblk(g.nelem, 256)), dim3(256), 0, st, g, bmat, pack(mats), et, u, vars,
                     partials, red);
  return hipGetLastError();
}

hipError_t launch_vars_new(const Grid &g, const double *bmat, const material_base *mats, const int *et, const double *u,
                           const double *vars_old, double *vars_new, RedState *red, hipStream_t st) {
  hipLaunchKernelGGL(k_vars_new, dim3(nblk(g.nelem, 256)), dim3(256), 0, st, g, bmat, pack(mats), et, u, vars_old,
                     vars_new, red);
  return hipGetLastError();
}

}  // namespace mpp
