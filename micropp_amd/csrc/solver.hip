// solver.hip -- FP64 ELL SpMV and Jacobi-preconditioned CG (gfx950).
//
// ell_mvp (src/ell.cpp:35-44): one thread per row of a SELL-64 slice.  The
// 81 slots are consumed in three chunks of 27: the chunk's columns and values
// are loaded first (non-temporal, they are streamed once), then the 27 x
// gathers, then the 27 products are added in slot order without contraction,
// so y is bit-identical to the reference SpMV while every wavefront keeps 27
// value/column loads in flight.  Bound: HBM bandwidth, 12 bytes per slot plus
// 16 bytes per row.
//
// ell_solve_cgpd (src/ell.cpp:66-122): the reference recurrence, one
// iteration = three kernels:
//   step1  Ap = A p and the per-block partials of p.Ap
//   step2  alpha = r.z / sum(partials); x += alpha p; r -= alpha Ap; z = k r;
//          per-block partials of z.z and r.z
//   step3  beta = sum(r.z) / r.z_old; p = z + beta p; the new scalars
// No kernel uses atomics or fences: each block writes its partial, and every
// block of the next kernel re-sums all partials in the same fixed order, so
// every block holds identical scalars.  The scalars of iteration k are read
// from state[k & 1] and written (by block 0 of step3) to state[(k+1) & 1]; the
// parity is a kernel argument, so a run of iterations can be queued or
// captured in a graph ahead of time.  Once the stopping test holds, the
// kernels of later iterations return at once and step3 only carries the
// final state forward.
#include "internal.hpp"
#include "reduce.hpp"

namespace mpp {

constexpr int kChunk = 27;

__device__ __forceinline__ double row_dot(const int *cols, const double *vals, const double *x, int row) {
#pragma clang fp contract(off)
  const size_t base = (size_t)(row >> 6) * kSlots * 64 + (row & 63);
  const int *c = cols + base;
  const double *v = vals + base;
  double acc = 0;
#pragma unroll
  for (int ch = 0; ch < kSlots / kChunk; ++ch) {
    int ci[kChunk];
    double vi[kChunk], xv[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; ++k) ci[k] = __builtin_nontemporal_load(&c[(ch * kChunk + k) * 64]);
#pragma unroll
    for (int k = 0; k < kChunk; ++k) vi[k] = __builtin_nontemporal_load(&v[(ch * kChunk + k) * 64]);
#pragma unroll
    for (int k = 0; k < kChunk; ++k) xv[k] = x[ci[k]];
#pragma unroll
    for (int k = 0; k < kChunk; ++k) acc += vi[k] * xv[k];
  }
  return acc;
}

__global__ void __launch_bounds__(256) k_spmv(Grid g, const int *cols, const double *vals, const double *x, double *y) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= g.nrow_pad) return;
  const double acc = row_dot(cols, vals, x, row);
  if (row < g.nrow) y[row] = acc;
}

// Sum NV interleaved partials [nparts][NV] in a fixed order; every thread
// gets the totals.
template <int NV>
__device__ __forceinline__ void sum_partials(const double *part, int nparts, double (&out)[NV], double *lds) {
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

__device__ __forceinline__ int cg_done(const CgState &c) {
  const bool conv = (c.pn < c.min_err) || (c.pn < c.pn0 * c.rel_err);
  return (c.its >= c.max_its || conv) ? 1 : 0;
}

// k = 1/diag; x = 0; r = b - A*0 = b; z = k r; p = z; partials of r.z, z.z
__global__ void __launch_bounds__(256) k_cg_init(Grid g, const double *vals, const double *b, double *x, double *r,
                                                 double *z, double *p, double *kinv, double *part) {
#pragma clang fp contract(off)
  __shared__ double lds[4 * 2];
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
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
    part[(size_t)blockIdx.x * 2 + 0] = v[0];
    part[(size_t)blockIdx.x * 2 + 1] = v[1];
  }
}

// one block: the initial scalars into state[0]
__global__ void __launch_bounds__(256) k_cg_init_state(const double *part, int nparts, double min_err,
                                                       double rel_err, int max_its, CgState *state) {
  __shared__ double lds[4 * 2];
  double t[2];
  sum_partials<2>(part, nparts, t, lds);
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

// Ap = A p; partials of p.Ap
__global__ void __launch_bounds__(256) k_cg_step1(Grid g, const int *cols, const double *vals, const double *p,
                                                  double *Ap, double *part1, const CgState *cur) {
#pragma clang fp contract(off)
  __shared__ double lds[4];
  if (cur->done) return;
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  const double y = row_dot(cols, vals, p, row);
  Ap[row] = y;
  double v[1] = {p[row] * y};
  block_sum<1, 256>(v, lds);
  if (threadIdx.x == 0) part1[blockIdx.x] = v[0];
}

// alpha; x += alpha p; r -= alpha Ap; z = k r; partials of z.z, r.z
__global__ void __launch_bounds__(256) k_cg_step2(Grid g, double *x, double *r, double *z, const double *p,
                                                  const double *Ap, const double *kinv, const double *part1,
                                                  double *part2, const CgState *cur) {
#pragma clang fp contract(off)
  __shared__ double lds[4 * 2];
  if (cur->done) return;
  double pap[1];
  sum_partials<1>(part1, gridDim.x, pap, lds);
  const double alpha = cur->rz / pap[0];
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  x[row] += alpha * p[row];
  const double rv = r[row] - alpha * Ap[row];
  r[row] = rv;
  const double zv = kinv[row] * rv;
  z[row] = zv;
  double v[2] = {zv * zv, rv * zv};
  block_sum<2, 256>(v, lds);
  if (threadIdx.x == 0) {
    part2[(size_t)blockIdx.x * 2 + 0] = v[0];
    part2[(size_t)blockIdx.x * 2 + 1] = v[1];
  }
}

// beta; p = z + beta p; block 0 writes the next state
__global__ void __launch_bounds__(256) k_cg_step3(Grid g, const double *z, double *p, const double *part1,
                                                  const double *part2, const CgState *cur, CgState *nxt) {
#pragma clang fp contract(off)
  __shared__ double lds[4 * 2];
  if (cur->done) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *nxt = *cur;
    return;
  }
  double t[2];
  sum_partials<2>(part2, gridDim.x, t, lds);
  const double rz_old = cur->rz;
  const double beta = t[1] / rz_old;
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  p[row] = z[row] + beta * p[row];
  if (blockIdx.x == 0) {
    double pap[1];
    sum_partials<1>(part1, gridDim.x, pap, lds);
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

hipError_t launch_cg_init(const Grid &g, const double *vals, const double *b, double *x, double *r, double *z,
                          double *p, double *kinv, double *part, double min_err, double rel_err, int max_its,
                          CgState *state, hipStream_t st) {
  const unsigned nb = nblk(g.nrow_pad, 256);
  hipLaunchKernelGGL(k_cg_init, dim3(nb), dim3(256), 0, st, g, vals, b, x, r, z, p, kinv, part);
  hipLaunchKernelGGL(k_cg_init_state, dim3(1), dim3(256), 0, st, part, (int)nb, min_err, rel_err, max_its, state);
  return hipGetLastError();
}

hipError_t launch_cg_step1(const Grid &g, const int *cols, const double *vals, const double *p, double *Ap,
                           double *part, CgState *state, int parity, hipStream_t st) {
  hipLaunchKernelGGL(k_cg_step1, dim3(nblk(g.nrow_pad, 256)), dim3(256), 0, st, g, cols, vals, p, Ap, part,
                     state + (parity & 1));
  return hipGetLastError();
}

hipError_t launch_cg_step23(const Grid &g, double *x, double *r, double *z, double *p, const double *Ap,
                            const double *kinv, double *part, CgState *state, int parity, hipStream_t st) {
  const unsigned nb = nblk(g.nrow_pad, 256);
  double *part1 = part, *part2 = part + nb;
  CgState *cur = state + (parity & 1), *nxt = state + ((parity + 1) & 1);
  hipLaunchKernelGGL(k_cg_step2, dim3(nb), dim3(256), 0, st, g, x, r, z, p, Ap, kinv, part1, part2, cur);
  hipLaunchKernelGGL(k_cg_step3, dim3(nb), dim3(256), 0, st, g, z, p, part1, part2, cur, nxt);
  return hipGetLastError();
}

hipError_t launch_cg_iter(const Grid &g, const int *cols, const double *vals, double *x, double *r, double *z,
                          double *p, double *Ap, const double *kinv, double *part, CgState *state, int parity,
                          hipStream_t st) {
  hipError_t e = launch_cg_step1(g, cols, vals, p, Ap, part, state, parity, st);
  if (e == hipSuccess) e = launch_cg_step23(g, x, r, z, p, Ap, kinv, part, state, parity, st);
  return e;
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
