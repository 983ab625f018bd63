// solver.hip -- FP64 ELL SpMV and Jacobi-preconditioned CG (gfx950).
//
// ell_mvp (src/ell.cpp:35-44): one thread per row of a SELL-64 slice; a
// wavefront streams the 81 slot columns of its 64 rows with coalesced loads
// and sums each row's slots in order 0..80 without contraction, so y is
// bit-identical to the reference SpMV.  Bound: HBM bandwidth, 12 bytes per
// slot plus 16 bytes per row.
//
// ell_solve_cgpd (src/ell.cpp:66-122): the reference recurrence, one
// iteration = three kernels (SpMV + p.Ap; x, r, z update + z.z and r.z;
// p = z + beta p).  The scalars, the stopping test and the iteration count
// live on the device (CgState); a kernel whose system is done returns at
// once, so the host can queue iterations ahead and look at the state only
// now and then.
#include "internal.hpp"
#include "reduce.hpp"

namespace mpp {

__device__ __forceinline__ double row_dot(const int *cols, const double *vals, const double *x, int row) {
#pragma clang fp contract(off)
  const size_t base = (size_t)(row >> 6) * kSlots * 64 + (row & 63);
  const int *c = cols + base;
  const double *v = vals + base;
  double acc = 0;
#pragma unroll 27
  for (int s = 0; s < kSlots; ++s) acc += v[s * 64] * x[c[s * 64]];
  return acc;
}

__global__ void __launch_bounds__(256) k_spmv(Grid g, const int *cols, const double *vals, const double *x, double *y) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= g.nrow_pad) return;
  const double acc = row_dot(cols, vals, x, row);
  if (row < g.nrow) y[row] = acc;
}

__device__ __forceinline__ void cg_check(CgState *cg) {
  const bool conv = (cg->pn < cg->min_err) || (cg->pn < cg->pn0 * cg->rel_err);
  cg->done = (cg->its >= cg->max_its || conv) ? 1 : 0;
}

// k = 1/diag; x = 0; r = b - A*0 = b; z = k r; p = z; r.z and z.z.
__global__ void __launch_bounds__(256) k_cg_init(Grid g, const int *cols, const double *vals, const double *b,
                                                 double *x, double *r, double *z, double *p, double *kinv,
                                                 double *partials, CgState *cg) {
#pragma clang fp contract(off)
  __shared__ double lds[(256 / 64) * 2];
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  double rz = 0.0, zz = 0.0;
  if (row < g.nrow_pad) {
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
    rz = rv * zv;
    zz = zv * zv;
  }
  double v[2] = {rz, zz}, tot[2];
  if (grid_sum<2, 256>(v, partials, &cg->cnt, tot, lds)) {
    if (threadIdx.x == 0) {
      cg->rz = tot[0];
      cg->pn0 = sqrt(tot[1]);
      cg->pn = cg->pn0;
      cg->its = 0;
      cg_check(cg);
    }
  }
}

// Ap = A p, alpha = r.z / p.Ap
__global__ void __launch_bounds__(256) k_cg_step1(Grid g, const int *cols, const double *vals, const double *p,
                                                  double *Ap, double *partials, CgState *cg) {
#pragma clang fp contract(off)
  __shared__ double lds[(256 / 64) * 1];
  if (cg->done) return;
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  double pap = 0.0;
  if (row < g.nrow_pad) {
    const double y = row_dot(cols, vals, p, row);
    Ap[row] = y;
    pap = p[row] * y;
  }
  double v[1] = {pap}, tot[1];
  if (grid_sum<1, 256>(v, partials, &cg->cnt, tot, lds)) {
    if (threadIdx.x == 0) {
      cg->pAp = tot[0];
      cg->alpha = cg->rz / tot[0];
    }
  }
}

// x += alpha p; r -= alpha Ap; z = k r; ||z||, r.z, beta, stopping test
__global__ void __launch_bounds__(256) k_cg_step2(Grid g, double *x, double *r, double *z, const double *p,
                                                  const double *Ap, const double *kinv, double *partials,
                                                  CgState *cg) {
#pragma clang fp contract(off)
  __shared__ double lds[(256 / 64) * 2];
  if (cg->done) return;
  const double alpha = cg->alpha;
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  double zz = 0.0, rz = 0.0;
  if (row < g.nrow_pad) {
    x[row] += alpha * p[row];
    const double rv = r[row] - alpha * Ap[row];
    r[row] = rv;
    const double zv = kinv[row] * rv;
    z[row] = zv;
    zz = zv * zv;
    rz = rv * zv;
  }
  double v[2] = {zz, rz}, tot[2];
  if (grid_sum<2, 256>(v, partials, &cg->cnt, tot, lds)) {
    if (threadIdx.x == 0) {
      cg->pn = sqrt(tot[0]);
      cg->beta = tot[1] / cg->rz;
      cg->rz = tot[1];
      cg->its += 1;
      cg_check(cg);
    }
  }
}

// p = z + beta p (the reference updates p before re-testing, so the last
// update is harmless; it is skipped once the system is done)
__global__ void __launch_bounds__(256) k_cg_step3(Grid g, const double *z, double *p, const CgState *cg) {
#pragma clang fp contract(off)
  if (cg->done) return;
  const double beta = cg->beta;
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row < g.nrow_pad) p[row] = z[row] + beta * p[row];
}

__global__ void k_axpy(int n, double a, const double *x, double *y) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}

__global__ void k_copy(int n, const double *x, double *y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i];
}

static unsigned nblk(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t launch_spmv(const Grid &g, const int *cols, const double *vals, const double *x, double *y,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_spmv, dim3(nblk(g.nrow_pad, 256)), dim3(256), 0, st, g, cols, vals, x, y);
  return hipGetLastError();
}

hipError_t launch_cg_init(const Grid &g, const int *cols, const double *vals, const double *b, double *x, double *r,
                          double *z, double *p, double *kinv, double *partials, CgState *cg, hipStream_t st) {
  hipLaunchKernelGGL(k_cg_init, dim3(nblk(g.nrow_pad, 256)), dim3(256), 0, st, g, cols, vals, b, x, r, z, p, kinv,
                     partials, cg);
  return hipGetLastError();
}

hipError_t launch_cg_iter(const Grid &g, const int *cols, const double *vals, double *x, double *r, double *z,
                          double *p, double *Ap, const double *kinv, double *partials, CgState *cg, hipStream_t st) {
  const dim3 grid(nblk(g.nrow_pad, 256)), blk(256);
  hipLaunchKernelGGL(k_cg_step1, grid, blk, 0, st, g, cols, vals, p, Ap, partials, cg);
  hipLaunchKernelGGL(k_cg_step2, grid, blk, 0, st, g, x, r, z, p, Ap, kinv, partials, cg);
  hipLaunchKernelGGL(k_cg_step3, grid, blk, 0, st, g, z, p, cg);
  return hipGetLastError();
}

hipError_t launch_axpy(int n, double a, const double *x, double *y, hipStream_t st) {
  hipLaunchKernelGGL(k_axpy, dim3(nblk(n, 256)), dim3(256), 0, st, n, a, x, y);
  return hipGetLastError();
}

hipError_t launch_copy(int n, const double *x, double *y, hipStream_t st) {
  hipLaunchKernelGGL(k_copy, dim3(nblk(n, 256)), dim3(256), 0, st, n, x, y);
  return hipGetLastError();
}

}  // namespace mpp
