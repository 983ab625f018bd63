// kernels.hpp -- device bodies of the FP64 hot path, shared by the
// single-RVE kernels (assembly.hip, solver.hip) and the batched kernels
// (batch.hip).
//
// A body works on one block's share of one RVE: `lb` is the block's index
// inside that RVE and `nb` the RVE's block count, so a single-RVE kernel
// passes blockIdx.x / gridDim.x and a batched kernel passes its block's
// position inside its RVE.  Both therefore do the same arithmetic in the same
// order, and a batched RVE's results are bit-identical to a single solve.
//
// Reductions are fence-free: every block writes its partial, and a later
// kernel re-sums the partials of an RVE in a fixed order (seg_sum).
#pragma once

#include "internal.hpp"
#include "physics.hpp"
#include "reduce.hpp"

namespace mpp {

constexpr int kKeWaves = 4;  // element-stiffness items (waves) per block

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

__device__ __forceinline__ void gather_ue(const Grid &g, int ex, int ey, int ez, const double *u, double *ue) {
  for (int a = 0; a < 8; ++a) {
    const size_t base = (size_t)elem_node(g, ex, ey, ez, a) * 3;
    ue[a * 3 + 0] = u[base + 0];
    ue[a * 3 + 1] = u[base + 1];
    ue[a * 3 + 2] = u[base + 2];
  }
}

__device__ __forceinline__ void stage_bmat(const double *bmat_g, double *bm) {
  for (int i = threadIdx.x; i < 8 * 144; i += blockDim.x) bm[i] = bmat_g[i];
}

// Fixed-order sum of NV interleaved partials [nparts][NV] by one block of 256
// threads; every thread gets the totals.
template <int NV>
__device__ __forceinline__ void seg_sum(const double *part, int nparts, double (&out)[NV]) {
  __shared__ double lds[4 * NV];
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (int b = threadIdx.x; b < nparts; b += 256)
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] += part[(size_t)b * NV + k];
  block_sum<NV, 256>(acc, lds);
#pragma unroll
  for (int k = 0; k < NV; ++k) out[k] = acc[k];
}

// ---------------------------------------------------------------------------
// set_displ_bc (src/micro3D.cpp:26-78).  The reference visits the faces
// z=0, z=lz, y=0, y=ly, x=0, x=lx and later faces overwrite shared edges, so
// a node takes the coordinates of the last face containing it.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void set_bc_node(const Grid &g, int node, const double *e, double *u) {
#pragma clang fp contract(off)
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
// (gp, m) in the reference order.  Block lb holds items lb*4 .. lb*4+3.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void elem_ke_body(const Grid &g, int lb, const double *bmat_g, const material_base *mats,
                                             const int *et, const double *u, const double *vars, int phase_mode,
                                             double *ke) {
#pragma clang fp contract(off)
  __shared__ double bm[8 * 144];
  __shared__ double ue_s[kKeWaves][24];
  __shared__ double C_s[kKeWaves][8 * 36];
  __shared__ double cb_s[kKeWaves][8 * 144];
  __shared__ double eps_s[kKeWaves][8 * 6];
  __shared__ double sig_s[kKeWaves][8 * 7 * 6];  // FD tangent: base and 6 perturbed stresses per GP
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  stage_bmat(bmat_g, bm);
  const int item = lb * kKeWaves + w;
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
  // mat_ctan per GP.  Elastic: the closed form, one lane per GP.  Otherwise
  // the forward-difference tangent's 7 stress evaluations per GP (base and
  // 6 perturbed strains) run on 56 lanes at once, then each tangent entry is
  // (s1 - s0) / D_EPS_CTAN as in mat_ctan: the same values bit for bit.
  const bool fd = active && mats[ph].type != MATERIAL_ELASTIC;
  if (active && lane < 8) {
    double eps[6];
    strain_at(&bm[lane * 144], ue_s[w], eps);
    if (fd) {
      for (int i = 0; i < 6; ++i) eps_s[w][lane * 6 + i] = eps[i];
    } else {
      mat_ctan(mats[ph], eps, nullptr, &C_s[w][lane * 36]);
    }
  }
  __syncthreads();
  if (fd && lane < 56) {
    const int gp = lane / 7, c = lane % 7;  // c = 0: base strain; c > 0: component c - 1 perturbed
    double e1[6];
    for (int i = 0; i < 6; ++i) e1[i] = eps_s[w][gp * 6 + i];
    if (c > 0) e1[c - 1] += D_EPS_CTAN;
    const double *h = (vars && !phase_mode) ? vars + ((size_t)item * 8 + gp) * NVAR : nullptr;
    mat_stress(mats[ph], e1, h, &sig_s[w][(gp * 7 + c) * 6]);
  }
  __syncthreads();
  if (fd) {
    for (int q = lane; q < 8 * 36; q += 64) {
      const int gp = q / 36, r = (q / 6) % 6, c = q % 6;
      const double *sg = &sig_s[w][gp * 7 * 6];
      C_s[w][gp * 36 + r * 6 + c] = (sg[(c + 1) * 6 + r] - sg[r]) / D_EPS_CTAN;
    }
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
  // Ke = sum_gp B^T (C B wg): lane (bi, bj) owns the 3x3 tile of rows 3 bi ..
  // 3 bi + 2 and columns 3 bj .. 3 bj + 2 (6 LDS reads per 9 products); each
  // entry still sums gp, then m, in order
  if (active) {
    double *out = ke + (size_t)item * 576;
    const int bi = lane >> 3, bj = lane & 7;
    double acc[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    for (int gp = 0; gp < 8; ++gp)
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        const double *bq = &bm[gp * 144 + m * 24 + 3 * bi];
        const double *cq = &cb_s[w][gp * 144 + m * 24 + 3 * bj];
        const double b3[3] = {bq[0], bq[1], bq[2]}, c3[3] = {cq[0], cq[1], cq[2]};
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int c = 0; c < 3; ++c) acc[a][c] += b3[a] * c3[c];
      }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int c = 0; c < 3; ++c) out[(3 * bi + a) * 24 + 3 * bj + c] = acc[a][c];
  }
}

// ---------------------------------------------------------------------------
// Row-owner gather of the element blocks into SELL vals: assembly_mat's
// element loop + ell_add_3D (src/assembly.cpp:105-121,
// src/ell-common.cpp:166-198) and ell_set_bc_3D (:238-297).  For row
// (node N, fi) and slot (neighbour M, fj) the contributing elements are those
// holding both N and M, visited in reference order ex, ey, ez.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void gather_vals_row(const Grid &g, int row, const int *et, const double *ke,
                                                int phase_mode, double *vals) {
#pragma clang fp contract(off)
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
__device__ __forceinline__ void elem_be_body(const Grid &g, int lb, const double *bmat_g, const material_base *mats,
                                             const int *et, const double *u, const double *vars, double *be) {
#pragma clang fp contract(off)
  __shared__ double bm[8 * 144];
  stage_bmat(bmat_g, bm);
  __syncthreads();
  const int e = lb * 256 + threadIdx.x;
  if (e >= g.nelem) return;
  int ex, ey, ez;
  elem_coords(g, e, ex, ey, ez);
  double ue[24];
  gather_ue(g, ex, ey, ez, u, ue);
  const material_base &m = mats[et[e]];
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
// 27-103): element order ez, ey, ex; Dirichlet DOFs zeroed; negated.  The
// block's partial of ||b||^2 goes to part[lb].
// ---------------------------------------------------------------------------
__device__ __forceinline__ void gather_rhs_body(const Grid &g, int lb, const double *be, double *bvec, double *part) {
#pragma clang fp contract(off)
  __shared__ double lds[4];
  const int row = lb * 256 + threadIdx.x;
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
  double val[1] = {sq};
  block_sum<1, 256>(val, lds);
  if (threadIdx.x == 0) part[lb] = val[0];
}

// ---------------------------------------------------------------------------
// calc_ave_stress (src/average.cpp:57-82): per-element sums over the Gauss
// points; the block's partial of the 6 components goes to part[lb * 6 ..].
// ---------------------------------------------------------------------------
__device__ __forceinline__ void ave_stress_body(const Grid &g, int lb, const double *bmat_g, const material_base *mats,
                                                const int *et, const double *u, const double *vars, double *part) {
#pragma clang fp contract(off)
  __shared__ double bm[8 * 144];
  __shared__ double lds[4 * 6];
  stage_bmat(bmat_g, bm);
  __syncthreads();
  const int e = lb * 256 + threadIdx.x;
  double s6[6] = {0, 0, 0, 0, 0, 0};
  if (e < g.nelem) {
    int ex, ey, ez;
    elem_coords(g, e, ex, ey, ez);
    double ue[24];
    gather_ue(g, ex, ey, ez, u, ue);
    const material_base &m = mats[et[e]];
    for (int gp = 0; gp < 8; ++gp) {
      double eps[6], sig[6];
      strain_at(&bm[gp * 144], ue, eps);
      mat_stress(m, eps, vars ? vars + ((size_t)e * 8 + gp) * NVAR : nullptr, sig);
      for (int v = 0; v < 6; ++v) s6[v] += sig[v] * g.wg;
    }
  }
  block_sum<6, 256>(s6, lds);
  if (threadIdx.x < 6) part[(size_t)lb * 6 + threadIdx.x] = s6[threadIdx.x];
}

// ---------------------------------------------------------------------------
// calc_fields (src/average.cpp:84-112): per element, the Gauss-point strain
// and stress summed with weight wg, times ivol = 1 / (8 wg)
// (src/micropp.cpp:59); element-major [e * 6 + v].
// ---------------------------------------------------------------------------
__device__ __forceinline__ void elem_fields_body(const Grid &g, int lb, const double *bmat_g,
                                                 const material_base *mats, const int *et, const double *u,
                                                 const double *vars, double *strain, double *stress) {
#pragma clang fp contract(off)
  __shared__ double bm[8 * 144];
  stage_bmat(bmat_g, bm);
  __syncthreads();
  const int e = lb * 256 + threadIdx.x;
  if (e >= g.nelem) return;
  int ex, ey, ez;
  elem_coords(g, e, ex, ey, ez);
  double ue[24];
  gather_ue(g, ex, ey, ez, u, ue);
  const material_base &m = mats[et[e]];
  double ea[6] = {0, 0, 0, 0, 0, 0}, sa[6] = {0, 0, 0, 0, 0, 0};
  for (int gp = 0; gp < 8; ++gp) {
    double eps[6], sig[6];
    strain_at(&bm[gp * 144], ue, eps);
    mat_stress(m, eps, vars ? vars + ((size_t)e * 8 + gp) * NVAR : nullptr, sig);
    for (int v = 0; v < 6; ++v) {
      ea[v] += eps[v] * g.wg;
      sa[v] += sig[v] * g.wg;
    }
  }
  const double ivol = 1.0 / (g.wg * 8);
  for (int v = 0; v < 6; ++v) {
    strain[(size_t)e * 6 + v] = ea[v] * ivol;
    stress[(size_t)e * 6 + v] = sa[v] * ivol;
  }
}

// ---------------------------------------------------------------------------
// calc_vars_new (src/update.cpp:32-56): evolute per Gauss point; any
// nonlinear point sets *flag.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void vars_new_body(const Grid &g, int lb, const double *bmat_g, const material_base *mats,
                                              const int *et, const double *u, const double *vars_old,
                                              double *vars_new, int *flag) {
#pragma clang fp contract(off)
  __shared__ double bm[8 * 144];
  stage_bmat(bmat_g, bm);
  __syncthreads();
  const int e = lb * 256 + threadIdx.x;
  if (e >= g.nelem) return;
  int ex, ey, ez;
  elem_coords(g, e, ex, ey, ez);
  double ue[24];
  gather_ue(g, ex, ey, ez, u, ue);
  const material_base &m = mats[et[e]];
  int nl = 0;
  for (int gp = 0; gp < 8; ++gp) {
    double eps[6];
    strain_at(&bm[gp * 144], ue, eps);
    const size_t off = ((size_t)e * 8 + gp) * NVAR;
    nl |= mat_evolute(m, eps, vars_old ? vars_old + off : nullptr, vars_new + off);
  }
  if (nl) atomicOr(flag, 1);
}

// ---------------------------------------------------------------------------
// ell_mvp (src/ell.cpp:35-44) row: the 81 slots in chunks of CH (9, 27 or
// 81 slots in flight): the chunk's columns and values first (non-temporal,
// streamed once), then its x gathers, then the products added in slot order
// without contraction, so y is bit-identical to the reference SpMV whatever
// the chunk width.
// ---------------------------------------------------------------------------
constexpr int kChunk = 27;

template <int CH = kChunk>
__device__ __forceinline__ double row_dot(const int *cols, const double *vals, const double *x, int row) {
#pragma clang fp contract(off)
  static_assert(kSlots % CH == 0, "chunks tile the row");
  const size_t base = (size_t)(row >> 6) * kSlots * 64 + (row & 63);
  const int *c = cols + base;
  const double *v = vals + base;
  double acc = 0;
#pragma unroll
  for (int ch = 0; ch < kSlots / CH; ++ch) {
    int ci[CH];
    double vi[CH], xv[CH];
#pragma unroll
    for (int k = 0; k < CH; ++k) ci[k] = __builtin_nontemporal_load(&c[(ch * CH + k) * 64]);
#pragma unroll
    for (int k = 0; k < CH; ++k) vi[k] = __builtin_nontemporal_load(&v[(ch * CH + k) * 64]);
#pragma unroll
    for (int k = 0; k < CH; ++k) xv[k] = x[ci[k]];
#pragma unroll
    for (int k = 0; k < CH; ++k) acc += vi[k] * xv[k];
  }
  return acc;
}

// row_dot on the structured grid: the column of slot nb*3 + fj of row 3n + fi
// is 3 m + fj for the stencil neighbour m = n + offset(nb) (ell_init,
// src/ell-common.cpp:102-130), computed instead of loaded, which leaves 8 of
// the 12 bytes per slot.  m is clamped into the grid.  For a matrix assembled
// here only the rows of boundary nodes (Dirichlet rows, zero off the
// diagonal) and padding rows (zero) have stencil slots outside the grid,
// where the reference's column 0 + fj and the clamped column both give a zero
// product for a finite x, and a sum that starts at +0 is unchanged by adding
// either signed zero.  So the result equals row_dot's bit for bit.  (A matrix
// uploaded from the host may hold values in those slots: use row_dot.)
// T = float: the FP32 inner CG of the mixed-precision batch (FP32 sums).
template <int CH = kChunk, typename T>
__device__ __forceinline__ T row_dot_grid(const Grid &g, const T *vals, const T *x, int row) {
#pragma clang fp contract(off)
  static_assert(kSlots % CH == 0 && CH % 3 == 0, "chunks tile the row in whole neighbours");
  const size_t base = (size_t)(row >> 6) * kSlots * 64 + (row & 63);
  const T *v = vals + base;
  const int n = row / 3, nxny = g.nx * g.ny, hi = g.nn - 1;
  T acc = 0;
#pragma unroll
  for (int ch = 0; ch < kSlots / CH; ++ch) {
    T vi[CH], xv[CH];
#pragma unroll
    for (int k = 0; k < CH; ++k) vi[k] = __builtin_nontemporal_load(&v[(ch * CH + k) * 64]);
#pragma unroll
    for (int k = 0; k < CH; k += 3) {
      const int nb = (ch * CH + k) / 3;
      const int off = (nb / 9 - 1) * nxny + ((nb / 3) % 3 - 1) * g.nx + (nb % 3 - 1);
      const int m3 = 3 * min(max(n + off, 0), hi);
#pragma unroll
      for (int fj = 0; fj < 3; ++fj) xv[k + fj] = x[m3 + fj];
    }
#pragma unroll
    for (int k = 0; k < CH; ++k) acc += vi[k] * xv[k];
  }
  return acc;
}

// The gathers of row_dot_grid for a block of 256 consecutive rows, staged in
// LDS: the block's rows belong to nodes n_lo .. n_hi (at most 86), whose
// stencil neighbours lie on 9 runs of consecutive nodes (one per (dz, dy)),
// each n_hi - n_lo + 3 <= kSegNodes long.  The block copies those runs
// (clamped node indices, as row_dot_grid clamps each neighbour) into seg once,
// coalesced, and every row reads its 81 values there instead of gathering
// them from L1/L2: the same values in the same slot order, so the result is
// bit-identical to row_dot_grid.  Every thread of the block must call it.
constexpr int kSegNodes = 88;
template <typename T>
__device__ __forceinline__ void stage_grid_runs(const Grid &g, int row0, const T *x, T (*seg)[kSegNodes * 3]) {
  const int n_lo = row0 / 3, n_hi = (row0 + 255) / 3, len = (n_hi - n_lo + 3) * 3;
  const int nxny = g.nx * g.ny, hi = g.nn - 1;
  for (int idx = threadIdx.x; idx < 9 * len; idx += 256) {
    const int sgi = idx / len, e3 = idx - sgi * len;
    const int base = n_lo - 1 + (sgi / 3 - 1) * nxny + (sgi % 3 - 1) * g.nx;
    const int m = min(max(base + e3 / 3, 0), hi);
    seg[sgi][e3] = x[3 * m + e3 % 3];
  }
  __syncthreads();
}

template <int CH = kChunk, typename T>
__device__ __forceinline__ T row_dot_grid_seg(const T *vals, const T (*seg)[kSegNodes * 3], int row0, int row) {
#pragma clang fp contract(off)
  static_assert(kSlots % CH == 0 && CH % 3 == 0, "chunks tile the row in whole neighbours");
  const size_t base = (size_t)(row >> 6) * kSlots * 64 + (row & 63);
  const T *v = vals + base;
  const int e0 = row / 3 - row0 / 3;  // the row's node within the runs, less 1
  T acc = 0;
#pragma unroll
  for (int ch = 0; ch < kSlots / CH; ++ch) {
    T vi[CH], xv[CH];
#pragma unroll
    for (int k = 0; k < CH; ++k) vi[k] = __builtin_nontemporal_load(&v[(ch * CH + k) * 64]);
#pragma unroll
    for (int k = 0; k < CH; k += 3) {
      const int nb = (ch * CH + k) / 3;
      const T *sv = &seg[nb / 3][(e0 + nb % 3) * 3];
#pragma unroll
      for (int fj = 0; fj < 3; ++fj) xv[k + fj] = sv[fj];
    }
#pragma unroll
    for (int k = 0; k < CH; ++k) acc += vi[k] * xv[k];
  }
  return acc;
}

// ---------------------------------------------------------------------------
// ell_solve_cgpd (src/ell.cpp:66-122), the reference recurrence in three
// kernels per iteration:
//   step1  Ap = A p and the per-block partials of p.Ap
//   step2  alpha = r.z / sum(partials); x += alpha p; r -= alpha Ap; z = k r;
//          per-block partials of z.z and r.z
//   step3  beta = sum(r.z) / r.z_old; p = z + beta p; the new scalars
// Every block of step2/step3 re-sums all partials of its RVE in the same
// fixed order, so every block holds identical scalars.  The scalars of
// iteration k are read from state[k & 1] and written (by block 0 of step3)
// to state[(k+1) & 1].  Once the stopping test holds, later iterations
// return at once and step3 only carries the final state forward.
// Partials: part1 = part[0 .. nb), part2 = part[nb .. 3 nb).  The steps take
// vectors (and, in step 1, the matrix) of type T: double for the FP64 CG,
// float for the inner CG of the mixed-precision batch, whose products and
// partials are still formed in FP64.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int cg_done(const CgState &c) {
  const bool conv = (c.pn < c.min_err) || (c.pn < c.pn0 * c.rel_err);
  return (c.its >= c.max_its || conv) ? 1 : 0;
}

// k = 1/diag; x = 0; r = b - A*0 = b; z = k r; p = z; partials of r.z, z.z
__device__ __forceinline__ void cg_init_body(const Grid &g, int lb, const double *vals, const double *b, double *x,
                                             double *r, double *z, double *p, double *kinv, double *part) {
#pragma clang fp contract(off)
  __shared__ double lds[4 * 2];
  const int row = lb * 256 + threadIdx.x;
  double k = 0.0, rv = 0.0;
  if (row < g.nrow) {
    const size_t d = ((size_t)(row >> 6) * kSlots + 13 * 3 + row % 3) * 64 + (row & 63);
    k = 1 / vals[d];
    rv = b[row] - 0.0;
  }
  const double zv = k * rv;
  kinv[row] = k;
  x[row] = 0.0;
  r[row] = rv;
  z[row] = zv;
  p[row] = zv;
  double v[2] = {rv * zv, zv * zv};
  block_sum<2, 256>(v, lds);
  if (threadIdx.x == 0) {
    part[(size_t)lb * 2 + 0] = v[0];
    part[(size_t)lb * 2 + 1] = v[1];
  }
}

// one block: the initial scalars into state[0]
__device__ __forceinline__ void cg_init_state_body(const double *part, int nparts, double min_err, double rel_err,
                                                   int max_its, CgState *state) {
  double t[2];
  seg_sum<2>(part, nparts, t);
  if (threadIdx.x == 0) {
    CgState c{};
    c.min_err = min_err;
    c.rel_err = rel_err;
    c.max_its = max_its;
    c.rz = t[0];
    c.pn0 = sqrt(t[1]);
    c.pn = c.pn0;
    c.its = 0;
    c.done = cg_done(c);
    state[0] = c;
  }
}

// STENCIL: the matrix has nonzeros only on in-grid stencil slots (assembled
// here), so the columns are computed (row_dot_grid); otherwise they are read.
template <int CH = kChunk, bool STENCIL = true, typename T>
__device__ __forceinline__ void cg_step1_body(const Grid &g, int lb, const int *cols, const T *vals, const T *p,
                                              T *Ap, double *part1, const CgState *cur) {
#pragma clang fp contract(off)
  __shared__ double lds[4];
  if (cur->done) return;
  const int row = lb * 256 + threadIdx.x;
  T y;
  if constexpr (STENCIL) {
    __shared__ T seg[9][kSegNodes * 3];
    stage_grid_runs(g, lb * 256, p, seg);
    y = row_dot_grid_seg<CH>(vals, seg, lb * 256, row);
  } else {
    y = row_dot<CH>(cols, vals, p, row);
  }
  Ap[row] = y;
  double v[1] = {(double)p[row] * (double)y};
  block_sum<1, 256>(v, lds);
  if (threadIdx.x == 0) part1[lb] = v[0];
}

template <typename T>
__device__ __forceinline__ void cg_step2_body(int lb, int nb, T *x, T *r, T *z, const T *p, const T *Ap,
                                              const T *kinv, const double *part1, double *part2,
                                              const CgState *cur) {
#pragma clang fp contract(off)
  __shared__ double lds[4 * 2];
  if (cur->done) return;
  double pap[1];
  seg_sum<1>(part1, nb, pap);
  const double alpha = cur->rz / pap[0];
  const int row = lb * 256 + threadIdx.x;
  x[row] += alpha * p[row];
  const T rv = r[row] - alpha * Ap[row];
  r[row] = rv;
  const T zv = kinv[row] * rv;
  z[row] = zv;
  double v[2] = {(double)zv * zv, (double)rv * zv};
  block_sum<2, 256>(v, lds);
  if (threadIdx.x == 0) {
    part2[(size_t)lb * 2 + 0] = v[0];
    part2[(size_t)lb * 2 + 1] = v[1];
  }
}

template <typename T>
__device__ __forceinline__ void cg_step3_body(int lb, int nb, const T *z, T *p, const double *part1,
                                              const double *part2, const CgState *cur, CgState *nxt) {
#pragma clang fp contract(off)
  if (cur->done) {
    if (lb == 0 && threadIdx.x == 0) *nxt = *cur;
    return;
  }
  double t[2];
  seg_sum<2>(part2, nb, t);
  const double rz_old = cur->rz;
  const double beta = t[1] / rz_old;
  const int row = lb * 256 + threadIdx.x;
  p[row] = z[row] + beta * p[row];
  if (lb == 0) {
    double pap[1];
    seg_sum<1>(part1, nb, pap);
    if (threadIdx.x == 0) {
      CgState c = *cur;
      c.pAp = pap[0];
      c.alpha = rz_old / pap[0];
      c.pn = sqrt(t[0]);
      c.beta = beta;
      c.rz = t[1];
      c.its = cur->its + 1;
      c.done = cg_done(c);
      *nxt = c;
    }
  }
}

// ---------------------------------------------------------------------------
// Coarse-grid hats of the two-level preconditioners (mixed.hip k_c_*,
// batch.hip kb_c_*): trilinear hats of the interior nodes of a coarse grid
// with h fine elements per coarse element and side.
// ---------------------------------------------------------------------------

// hat of interior coarse node q (1 .. m) at fine index i along one side
__device__ __forceinline__ float hat(int i, int q, int h) {
  const int dd = abs(i - q * h);
  return dd < h ? 1.0f - (float)dd / (float)h : 0.0f;
}

// weights and interior coarse indices (0-based) of the two coarse nodes at
// and above fine index i along one side; a node outside 1 .. m gets weight 0
// and index 0 (a valid address)
__device__ __forceinline__ void hat_pair(int i, int h, int m, float (&w)[2], int (&q)[2]) {
  const int q0 = i / h;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int c = q0 + a;
    const bool in = c >= 1 && c <= m;
    w[a] = in ? hat(i, c, h) : 0.0f;
    q[a] = in ? c - 1 : 0;
  }
}

}  // namespace mpp
