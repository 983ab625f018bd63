// coarse.hip -- set-up of the two-level preconditioner of the multi-RHS inner
// CG (the kernels that apply it are mixed.hip k_c_*).
//
// The coarse space: trilinear hats of the interior nodes of a coarse grid
// aligned with the RVE's grid (h fine elements per coarse element and side),
// the Dirichlet boundary excluded.  A_c = P^T A P (Galerkin) is assembled in
// FP64 on the device, element by element: a fine element at local position
// (px, py, pz) inside its coarse element interpolates its 8 nodes from that
// coarse element's 8 corners, so its contribution is C = P_e^T Ke P_e with P_e
// the 24 x 24 map of that position, the same for every element of one phase
// and position (k_cs_table); the coarse element matrices sum those
// (k_cs_elem) and the dense A_c gathers them (k_cs_dense).  A_c is factored
// and inverted on the device (rocSOLVER potrf, then two rocBLAS triangular
// solves with the identity) and kept as a dense
// FP32 inverse, so one application of the coarse solve is one matrix-vector
// product per field.
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>

#include "internal.hpp"

namespace mpp {

namespace {


// coarse elements along a side of ne fine elements: the largest divisor
// d <= dmax of ne with ne / d >= 2 (0: none with d >= 2)
int pick_side(int ne, int dmax) {
  for (int d = std::min(dmax, ne / 2); d >= 2; --d)
    if (ne % d == 0) return d;
  return 0;
}

// corner A of a hexahedron (the element's local node order: (0,0,0), (1,0,0),
// (1,1,0), (0,1,0), then the same at z + 1) and its inverse
__device__ __forceinline__ int hex_ox(int A) { return ((A + 1) >> 1) & 1; }
__device__ __forceinline__ int hex_oy(int A) { return (A >> 1) & 1; }
__device__ __forceinline__ int hex_oz(int A) { return A >> 2; }
__device__ __forceinline__ int hex_corner(int ox, int oy, int oz) { return oz * 4 + (oy ? 3 - ox : ox); }

// C[(pos * 3 + phase) * 576 + r * 24 + c] = (P_e^T Ke_phase P_e)[r][c] for the
// fine element at position pos = (pz hy + py) hx + px: block (pos, phase),
// one thread per entry
__global__ void __launch_bounds__(576) k_cs_table(CGrid cg, const double *ke, double *C) {
  __shared__ double N[8][8];  // N[a][A]: coarse corner A's hat at the fine element's node a
  __shared__ double K[576];
  const int pos = blockIdx.x, ph = blockIdx.y, t = threadIdx.x;
  const int px = pos % cg.hx, py = (pos / cg.hx) % cg.hy, pz = pos / (cg.hx * cg.hy);
  K[t] = ke[(size_t)ph * 576 + t];
  if (t < 64) {
    const int a = t >> 3, A = t & 7;
    const double xi = (double)(px + hex_ox(a)) / cg.hx, yi = (double)(py + hex_oy(a)) / cg.hy,
                 zi = (double)(pz + hex_oz(a)) / cg.hz;
    N[a][A] = (hex_ox(A) ? xi : 1.0 - xi) * (hex_oy(A) ? yi : 1.0 - yi) * (hex_oz(A) ? zi : 1.0 - zi);
  }
  __syncthreads();
  const int A = t / 72, d = (t / 24) % 3, B = (t % 24) / 3, e = t % 3;
  double s = 0.0;
  for (int a = 0; a < 8; ++a) {
    double ta = 0.0;
    for (int b = 0; b < 8; ++b) ta += K[(a * 3 + d) * 24 + b * 3 + e] * N[b][B];
    s += N[a][A] * ta;
  }
  C[((size_t)pos * 3 + ph) * 576 + t] = s;
}

// KE[E * 576 + t]: the stiffness of coarse element E, the sum of its fine
// elements' C entries in position order; block E, one thread per entry
__global__ void __launch_bounds__(576) k_cs_elem(Grid g, CGrid cg, int dx, int dy, const int *et, const double *C,
                                                 double *KE) {
  const int E = blockIdx.x, t = threadIdx.x;
  const int Ex = E % dx, Ey = (E / dx) % dy, Ez = E / (dx * dy);
  double s = 0.0;
  for (int pz = 0; pz < cg.hz; ++pz)
    for (int py = 0; py < cg.hy; ++py)
      for (int px = 0; px < cg.hx; ++px) {
        const int e = ((Ez * cg.hz + pz) * g.ney + (Ey * cg.hy + py)) * g.nex + Ex * cg.hx + px;
        const int pos = (pz * cg.hy + py) * cg.hx + px;
        s += C[((size_t)pos * 3 + et[e]) * 576 + t];
      }
  KE[(size_t)E * 576 + t] = s;
}

// The dense A_c (column-major, nc x nc) over the interior coarse nodes: entry
// (q_i 3 + d, q_j 3 + e) sums the coarse elements holding both nodes, in
// element order (zero when the nodes share none)
__global__ void __launch_bounds__(256) k_cs_dense(CGrid cg, int dx, int dy, const double *KE, double *Ac) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  const int nc = cg.nc;
  if (t >= (size_t)nc * nc) return;
  const int row = (int)(t % nc), col = (int)(t / nc);
  const int qi = row / 3, d = row % 3, qj = col / 3, e = col % 3;
  const int xi = qi % cg.mx + 1, yi = (qi / cg.mx) % cg.my + 1, zi = qi / (cg.mx * cg.my) + 1;
  const int xj = qj % cg.mx + 1, yj = (qj / cg.mx) % cg.my + 1, zj = qj / (cg.mx * cg.my) + 1;
  double s = 0.0;
  if (abs(xi - xj) <= 1 && abs(yi - yj) <= 1 && abs(zi - zj) <= 1)
    for (int ez = max(zi, zj) - 1; ez <= min(zi, zj); ++ez)
      for (int ey = max(yi, yj) - 1; ey <= min(yi, yj); ++ey)
        for (int ex = max(xi, xj) - 1; ex <= min(xi, xj); ++ex) {
          const int A = hex_corner(xi - ex, yi - ey, zi - ez), B = hex_corner(xj - ex, yj - ey, zj - ez);
          s += KE[((size_t)(ez * dy + ey) * dx + ex) * 576 + (A * 3 + d) * 24 + B * 3 + e];
        }
  Ac[t] = s;
}

// X = I (n x n)
__global__ void __launch_bounds__(256) k_cs_eye(int n, double *X) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (size_t)n * n) return;
  X[t] = (t % n == t / n) ? 1.0 : 0.0;
}

// one rocBLAS handle per device for the process's lifetime (the factorisation
// runs under the mutex, with the handle set to the caller's stream)
std::mutex g_solver_mu;
std::map<int, rocblas_handle> g_handles;

}  // namespace

void coarse_free(CoarseDev *c) {
  if (!c) return;
  for (void *p : {(void *)c->ainv, (void *)c->t1, (void *)c->rc, (void *)c->ecp, (void *)c->ec, (void *)c->z, (void *)c->prz})
    dfree(p);
  delete c;
}

bool coarse_enabled() {
  const char *env = std::getenv("MPPD_COARSE");
  return !(env && std::atoi(env) == 0);
}

// The coarsest-first search: the largest coarse grid aligned with g (a
// divisor of each side's element count, at least 2 coarse elements a side)
// with at most max_dofs coarse DOFs and at most kCoarseMaxPos fine elements
// per coarse element.
bool coarse_grid(const Grid &g, int max_dofs, CGrid *out) {
  if (g.nx > kCoarseMaxLine || g.ny > kCoarseMaxLine || g.nz > kCoarseMaxLine) return false;
  int dx = 0, dy = 0, dz = 0;
  for (int dmax = 32; dmax >= 2; --dmax) {
    dx = pick_side(g.nex, dmax);
    dy = pick_side(g.ney, dmax);
    dz = pick_side(g.nez, dmax);
    if (dx && dy && dz && 3 * (dx - 1) * (dy - 1) * (dz - 1) <= max_dofs) break;
    dx = dy = dz = 0;
  }
  if (!dx) return false;
  CGrid cg;
  cg.hx = g.nex / dx;
  cg.hy = g.ney / dy;
  cg.hz = g.nez / dz;
  cg.mx = dx - 1;
  cg.my = dy - 1;
  cg.mz = dz - 1;
  cg.nq = cg.mx * cg.my * cg.mz;
  cg.nc = 3 * cg.nq;
  if (cg.hx * cg.hy * cg.hz > kCoarseMaxPos) return false;
  *out = cg;
  return true;
}

int coarse_inverse(const Grid &g, const CGrid &cg, const int *et_d, const double *ke_d, float *ainv, hipStream_t st,
                   bool *ok) {
  *ok = false;
  const int nc = cg.nc, npos = cg.hx * cg.hy * cg.hz;
  const int dx = cg.mx + 1, dy = cg.my + 1, ncel = dx * dy * (cg.mz + 1);
  double *a64 = nullptr, *C = nullptr, *KE = nullptr, *X = nullptr;
  int *info_d = nullptr;
  int rc = MPPD_OK;
  do {
    if ((rc = dalloc(&a64, (size_t)nc * nc, false))) break;
    if ((rc = dalloc(&X, (size_t)nc * nc, false))) break;
    if ((rc = dalloc(&C, (size_t)npos * 3 * 576, false))) break;
    if ((rc = dalloc(&KE, (size_t)ncel * 576, false))) break;
    if ((rc = dalloc(&info_d, 1))) break;
    hipLaunchKernelGGL(k_cs_table, dim3(npos, 3), dim3(576), 0, st, cg, ke_d, C);
    hipLaunchKernelGGL(k_cs_elem, dim3(ncel), dim3(576), 0, st, g, cg, dx, dy, et_d, C, KE);
    hipLaunchKernelGGL(k_cs_dense, dim3((unsigned)(((size_t)nc * nc + 255) / 256)), dim3(256), 0, st, cg, dx, dy, KE,
                       a64);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      rc = hip_fail(e, "coarse matrix assembly");
      break;
    }
    int dev = 0, info = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) {
      rc = hip_fail(e, "hipGetDevice");
      break;
    }
    std::lock_guard<std::mutex> lock(g_solver_mu);
    rocblas_handle &h = g_handles[dev];
    if (!h && rocblas_create_handle(&h) != rocblas_status_success) {
      h = nullptr;
      rc = fail(MPPD_ERR_HIP, "rocblas handle for the coarse solve");
      break;
    }
    if (rocblas_set_stream(h, st) != rocblas_status_success) {
      rc = fail(MPPD_ERR_HIP, "rocblas stream for the coarse solve");
      break;
    }
    if (rocsolver_dpotrf(h, rocblas_fill_lower, nc, a64, nc, info_d) != rocblas_status_success) {
      rc = fail(MPPD_ERR_HIP, "rocsolver_dpotrf (coarse matrix)");
      break;
    }
    e = hipMemcpyAsync(&info, info_d, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      rc = hip_fail(e, "coarse factor status");
      break;
    }
    if (info != 0) break;  // not positive definite: no coarse level
    // A_c^-1 = L^-T L^-1 by two triangular solves with the identity (level-3
    // BLAS; 0.4 ms less than rocSOLVER's potri at 2 187 coarse DOFs)
    hipLaunchKernelGGL(k_cs_eye, dim3((unsigned)(((size_t)nc * nc + 255) / 256)), dim3(256), 0, st, nc, X);
    const double one = 1.0;
    if (rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_non_unit,
                      nc, nc, &one, a64, nc, X, nc) != rocblas_status_success ||
        rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_transpose,
                      rocblas_diagonal_non_unit, nc, nc, &one, a64, nc, X, nc) != rocblas_status_success) {
      rc = fail(MPPD_ERR_HIP, "rocblas_dtrsm (coarse inverse)");
      break;
    }
    e = launch_c_sym32(nc, X, ainv, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      rc = hip_fail(e, "coarse inverse copy");
      break;
    }
    *ok = true;
  } while (0);
  (void)hipStreamSynchronize(st);  // the temporaries below may still be in use after a failed step
  dfree(a64);
  dfree(X);
  dfree(C);
  dfree(KE);
  dfree(info_d);
  return rc;
}

// Builds the coarse level of a K-field system on the RVE grid g whose phase
// per element is et_d and whose per-phase element stiffness is ke_d (3 x 576,
// FP64, on the device).  *out = nullptr (Jacobi alone) when the grid has no
// usable coarse grid (a side with fewer than 4 elements or prime, a side
// beyond kCoarseMaxLine nodes), when A_c is not positive definite, or when
// MPPD_COARSE=0.
int coarse_setup(const Grid &g, const int *et_d, const double *ke_d, int K, hipStream_t st, CoarseDev **out) {
  *out = nullptr;
  CGrid cg;
  if (!coarse_enabled() || !coarse_grid(g, kCoarseMaxDofs, &cg)) return MPPD_OK;
  const int nc = cg.nc;
  CoarseDev *c = new CoarseDev();
  c->cg = cg;
  int rc = MPPD_OK;
  bool ok = false;
  do {
    float *ainv = nullptr;
    if ((rc = dalloc(&ainv, (size_t)nc * nc, false))) break;
    c->ainv = ainv;
    if ((rc = dalloc(&c->t1, (size_t)3 * K * g.nz * g.ny * cg.mx))) break;
    if ((rc = dalloc(&c->rc, (size_t)nc * K))) break;
    if ((rc = dalloc(&c->ecp, (size_t)kCoarseSplit * nc * K))) break;
    if ((rc = dalloc(&c->ec, (size_t)nc * K))) break;
    if ((rc = dalloc(&c->z, mixed_vector_len(g, K)))) break;
    if ((rc = dalloc(&c->prz, (size_t)3 * K * coarse_vec_blocks(g)))) break;
    rc = coarse_inverse(g, cg, et_d, ke_d, ainv, st, &ok);
  } while (0);
  if (rc || !ok) {
    coarse_free(c);
    return rc;
  }
  *out = c;
  return MPPD_OK;
}

}  // namespace mpp
