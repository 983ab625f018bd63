// geometry.hip -- RVE set-up kernels: shape-function matrices, phase maps,
// volume fractions and the ELL column structure (gfx950).
#include "internal.hpp"
#include "physics.hpp"

namespace mpp {

// calc_bmat for the 8 Gauss points (src/micro3D.cpp:80-120; the constructor
// loop src/micropp.cpp:83-85).
__global__ void k_bmat(double dx, double dy, double dz, double *bmat) {
  const int gp = threadIdx.x;
  if (gp < 8) bmat_gp(gp, dx, dy, dz, bmat + gp * 144);
}

// The seeded sphere list of MIC3D_RANDOM_SPHERES: geo = {seed, count, r_min,
// r_max}; per sphere cx, cy, cz ~ U[0,1), r = r_min + (r_max - r_min) U[0,1).
__global__ void k_random_spheres(double seed, double count, double rmin, double rmax, double *out) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint64_t s = (uint64_t)seed;
  int n = (int)count;
  n = n < 0 ? 0 : (n > kMaxSpheres ? kMaxSpheres : n);
  for (int i = 0; i < n; ++i) {
    out[i * 4 + 0] = splitmix_unit(&s);
    out[i * 4 + 1] = splitmix_unit(&s);
    out[i * 4 + 2] = splitmix_unit(&s);
    out[i * 4 + 3] = rmin + (rmax - rmin) * splitmix_unit(&s);
  }
}

struct Geo4 {
  double v[4];
};

// get_elem_type for every element (src/micropp.cpp:117-124, 338-544):
// centroid = (e + 1/2) h per axis.
__global__ void k_phase_map(Grid g, int type, Geo4 geo, const double *spheres, int nsph, int *et, int *bad) {
#pragma clang fp contract(off)
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.nelem) return;
  const int ex = e % g.nex, ey = (e / g.nex) % g.ney, ez = e / (g.nex * g.ney);
  const double c[3] = {ex * g.dx + g.dx / 2., ey * g.dy + g.dy / 2., ez * g.dz + g.dz / 2.};
  const int t = phase_of(type, geo.v, spheres, nsph, c);
  et[e] = t;
  if (t < 0) atomicOr(bad, 1);
}

// calc_volume_fractions (src/micropp.cpp:609-627) adds the same evol once per
// element of each phase, so its sums depend only on the two element counts:
// these are counted here (integer sums, exact in any order) and the host
// repeats the reference's additions (mppd_rve_create).  counts must be zero
// on entry.
__global__ void __launch_bounds__(256) k_phase_counts(Grid g, const int *et, unsigned long long *counts) {
  __shared__ unsigned c_s[2];
  if (threadIdx.x < 2) c_s[threadIdx.x] = 0u;
  __syncthreads();
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < g.nelem) {
    const int t = et[e];
    if (t == 0)
      atomicAdd(&c_s[0], 1u);
    else if (t >= 1)
      atomicAdd(&c_s[1], 1u);
  }
  __syncthreads();
  if (threadIdx.x < 2 && c_s[threadIdx.x]) atomicAdd(&counts[threadIdx.x], (unsigned long long)c_s[threadIdx.x]);
}

// ell_init, dim 3, nfield 3 (src/ell-common.cpp:98-131) written straight into
// the SELL-64 layout; padding rows get column 0.
__global__ void k_ell_cols(Grid g, int *cols) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= g.nrow_pad) return;
  const int slice = row >> 6, lane = row & 63;
  int *out = cols + (size_t)slice * kSlots * 64 + lane;
  if (row >= g.nrow) {
    for (int s = 0; s < kSlots; ++s) out[s * 64] = 0;
    return;
  }
  const int node = row / 3;
  const int i = node % g.nx, j = (node / g.nx) % g.ny, k = node / (g.nx * g.ny);
  const int nxny = g.nx * g.ny;
  for (int n = 0; n < 27; ++n) {
    const int dz = n / 9 - 1, dy = (n / 3) % 3 - 1, dx = n % 3 - 1;
    const bool in = (i + dx >= 0 && i + dx < g.nx) && (j + dy >= 0 && j + dy < g.ny) && (k + dz >= 0 && k + dz < g.nz);
    const int nb = in ? node + dz * nxny + dy * g.nx + dx : 0;
    for (int fj = 0; fj < 3; ++fj) out[(n * 3 + fj) * 64] = nb * 3 + fj;
  }
}

// SELL-64 -> reference row-major [nrow][81]
__global__ void k_sell_to_rows(Grid g, const int *cols, const double *vals, int *cols_rm, double *vals_rm) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)g.nrow * kSlots) return;
  const int row = (int)(idx / kSlots), s = (int)(idx % kSlots);
  const size_t src = ((size_t)(row >> 6) * kSlots + s) * 64 + (row & 63);
  if (cols_rm) cols_rm[idx] = cols[src];
  if (vals_rm) vals_rm[idx] = vals[src];
}

// reference row-major [nrow][81] -> SELL-64 (padding rows zero)
__global__ void k_rows_to_sell(Grid g, const double *vals_rm, double *vals) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)g.nrow_pad * kSlots) return;
  const int lane = (int)(idx & 63);
  const size_t rest = idx >> 6;
  const int s = (int)(rest % kSlots);
  const int slice = (int)(rest / kSlots);
  const int row = slice * 64 + lane;
  vals[idx] = (row < g.nrow) ? vals_rm[(size_t)row * kSlots + s] : 0.0;
}

static unsigned nblk(size_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t launch_bmat(const Grid &g, double *bmat, hipStream_t st) {
  hipLaunchKernelGGL(k_bmat, dim3(1), dim3(64), 0, st, g.dx, g.dy, g.dz, bmat);
  return hipGetLastError();
}

hipError_t launch_random_spheres(const double *geo, double *spheres, hipStream_t st) {
  hipLaunchKernelGGL(k_random_spheres, dim3(1), dim3(64), 0, st, geo[0], geo[1], geo[2], geo[3], spheres);
  return hipGetLastError();
}

hipError_t launch_phase_map(const Grid &g, int type, const double *geo, const double *spheres, int nsph, int *et,
                            int *bad, hipStream_t st) {
  Geo4 G;
  for (int i = 0; i < 4; ++i) G.v[i] = geo[i];
  hipLaunchKernelGGL(k_phase_map, dim3(nblk(g.nelem, 256)), dim3(256), 0, st, g, type, G, spheres, nsph, et, bad);
  return hipGetLastError();
}

hipError_t launch_phase_counts(const Grid &g, const int *et, unsigned long long *counts, hipStream_t st) {
  hipLaunchKernelGGL(k_phase_counts, dim3(nblk(g.nelem, 256)), dim3(256), 0, st, g, et, counts);
  return hipGetLastError();
}

hipError_t launch_ell_cols(const Grid &g, int *cols, hipStream_t st) {
  hipLaunchKernelGGL(k_ell_cols, dim3(nblk(g.nrow_pad, 256)), dim3(256), 0, st, g, cols);
  return hipGetLastError();
}

hipError_t launch_sell_to_rows(const Grid &g, const int *cols, const double *vals, int *cols_rm, double *vals_rm,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_sell_to_rows, dim3(nblk((size_t)g.nrow * kSlots, 256)), dim3(256), 0, st, g, cols, vals,
                     cols_rm, vals_rm);
  return hipGetLastError();
}

hipError_t launch_rows_to_sell(const Grid &g, const double *vals_rm, double *vals, hipStream_t st) {
  hipLaunchKernelGGL(k_rows_to_sell, dim3(nblk((size_t)g.nrow_pad * kSlots, 256)), dim3(256), 0, st, g, vals_rm,
                     vals);
  return hipGetLastError();
}

}  // namespace mpp
