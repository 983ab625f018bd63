// coarse.hip -- set-up of the two-level preconditioner of the multi-RHS inner
// CG (host code; the kernels that apply it are mixed.hip k_c_*).
//
// The coarse space: trilinear hats of the interior nodes of a coarse grid
// aligned with the RVE's grid (h fine elements per coarse element and side),
// the Dirichlet boundary excluded.  A_c = P^T A P (Galerkin) is assembled in
// FP64 element by element: a fine element at local position (px, py, pz)
// inside its coarse element interpolates its 8 nodes from that coarse
// element's 8 corners, so its contribution is C = P_e^T Ke P_e with P_e the
// 24 x 24 map of that position, the same for every element of one phase and
// position.  A_c is factored and inverted on the device (rocSOLVER potrf /
// potri) and kept as a dense FP32 inverse, so one application of the coarse
// solve is one matrix-vector product per field.
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "internal.hpp"

namespace mpp {

namespace {

constexpr int kCoarseMaxDofs = 2200;  // 3 x 9^3 = 2187: a 10^3-element coarse grid
constexpr int kHexOff[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

// coarse elements along a side of ne fine elements: the largest divisor
// d <= dmax of ne with ne / d >= 2 (0: none with d >= 2)
int pick_side(int ne, int dmax) {
  for (int d = std::min(dmax, ne / 2); d >= 2; --d)
    if (ne % d == 0) return d;
  return 0;
}

}  // namespace

void coarse_free(CoarseDev *c) {
  if (!c) return;
  for (void *p : {(void *)c->ainv, (void *)c->t1, (void *)c->rc, (void *)c->ecp, (void *)c->ec, (void *)c->z, (void *)c->prz})
    dfree(p);
  delete c;
}

// Builds the coarse level of a K-field system on the RVE grid g whose phase
// per element is et_d and whose per-phase element stiffness is ke_d (3 x 576,
// FP64, on the device).  *out = nullptr (Jacobi alone) when the grid has no
// usable coarse grid (a side with fewer than 4 elements or prime, a side
// beyond kCoarseMaxLine nodes) or when MPPD_COARSE=0.
int coarse_setup(const Grid &g, const int *et_d, const double *ke_d, int K, hipStream_t st, CoarseDev **out) {
  *out = nullptr;
  const char *env = std::getenv("MPPD_COARSE");
  if (env && std::atoi(env) == 0) return MPPD_OK;
  if (g.nx > kCoarseMaxLine || g.ny > kCoarseMaxLine || g.nz > kCoarseMaxLine) return MPPD_OK;  // k_c_zdot's tables
  int dx = 0, dy = 0, dz = 0;
  for (int dmax = 32; dmax >= 2; --dmax) {
    dx = pick_side(g.nex, dmax);
    dy = pick_side(g.ney, dmax);
    dz = pick_side(g.nez, dmax);
    if (dx && dy && dz && 3 * (dx - 1) * (dy - 1) * (dz - 1) <= kCoarseMaxDofs) break;
    dx = dy = dz = 0;
  }
  if (!dx) return MPPD_OK;
  CGrid cg;
  cg.hx = g.nex / dx;
  cg.hy = g.ney / dy;
  cg.hz = g.nez / dz;
  cg.mx = dx - 1;
  cg.my = dy - 1;
  cg.mz = dz - 1;
  cg.nq = cg.mx * cg.my * cg.mz;
  cg.nc = 3 * cg.nq;
  const int nc = cg.nc;

  std::vector<int> et(g.nelem);
  std::vector<double> ke(3 * 576);
  MPP_HIP(hipMemcpyAsync(et.data(), et_d, sizeof(int) * g.nelem, hipMemcpyDeviceToHost, st));
  MPP_HIP(hipMemcpyAsync(ke.data(), ke_d, sizeof(double) * 3 * 576, hipMemcpyDeviceToHost, st));
  MPP_HIP(hipStreamSynchronize(st));

  // C[(pos * 3 + phase)] = P_e^T Ke_phase P_e for every local position of a
  // fine element inside its coarse element
  const int npos = cg.hx * cg.hy * cg.hz;
  std::vector<double> C((size_t)npos * 3 * 576, 0.0);
  for (int pz = 0; pz < cg.hz; ++pz)
    for (int py = 0; py < cg.hy; ++py)
      for (int px = 0; px < cg.hx; ++px) {
        const int pos = (pz * cg.hy + py) * cg.hx + px;
        double N[8][8];  // N[a][A]: corner A's hat at fine node a
        for (int a = 0; a < 8; ++a) {
          const double xi[3] = {(double)(px + kHexOff[a][0]) / cg.hx, (double)(py + kHexOff[a][1]) / cg.hy,
                                (double)(pz + kHexOff[a][2]) / cg.hz};
          for (int A = 0; A < 8; ++A) {
            double w = 1.0;
            for (int t = 0; t < 3; ++t) w *= kHexOff[A][t] ? xi[t] : 1.0 - xi[t];
            N[a][A] = w;
          }
        }
        for (int ph = 0; ph < 3; ++ph) {
          const double *K_ = &ke[(size_t)ph * 576];
          double T[24][8];  // T[a d][B] for one column dof e: Ke P restricted to dof e
          double *Cp = &C[((size_t)pos * 3 + ph) * 576];
          for (int e = 0; e < 3; ++e) {
            for (int r = 0; r < 24; ++r)
              for (int B = 0; B < 8; ++B) {
                double s = 0.0;
                for (int b = 0; b < 8; ++b) s += K_[r * 24 + b * 3 + e] * N[b][B];
                T[r][B] = s;
              }
            for (int A = 0; A < 8; ++A)
              for (int d = 0; d < 3; ++d)
                for (int B = 0; B < 8; ++B) {
                  double s = 0.0;
                  for (int a = 0; a < 8; ++a) s += N[a][A] * T[a * 3 + d][B];
                  Cp[(A * 3 + d) * 24 + B * 3 + e] = s;
                }
          }
        }
      }
  // coarse element stiffness, then the dense A_c over the interior coarse nodes (column-major)
  std::vector<double> KE((size_t)dx * dy * dz * 576, 0.0);
  for (int ez = 0; ez < g.nez; ++ez)
    for (int ey = 0; ey < g.ney; ++ey)
      for (int ex = 0; ex < g.nex; ++ex) {
        const int e = (ez * g.ney + ey) * g.nex + ex;
        const int E = ((ez / cg.hz) * dy + ey / cg.hy) * dx + ex / cg.hx;
        const int pos = ((ez % cg.hz) * cg.hy + ey % cg.hy) * cg.hx + ex % cg.hx;
        const double *Cp = &C[((size_t)pos * 3 + et[e]) * 576];
        double *KEp = &KE[(size_t)E * 576];
        for (int t = 0; t < 576; ++t) KEp[t] += Cp[t];
      }
  std::vector<double> Ac((size_t)nc * nc, 0.0);
  for (int Ez = 0; Ez < dz; ++Ez)
    for (int Ey = 0; Ey < dy; ++Ey)
      for (int Ex = 0; Ex < dx; ++Ex) {
        const double *KEp = &KE[((size_t)(Ez * dy + Ey) * dx + Ex) * 576];
        int q[8];
        for (int A = 0; A < 8; ++A) {
          const int cx = Ex + kHexOff[A][0], cy = Ey + kHexOff[A][1], cz = Ez + kHexOff[A][2];
          const bool in = cx >= 1 && cx <= cg.mx && cy >= 1 && cy <= cg.my && cz >= 1 && cz <= cg.mz;
          q[A] = in ? ((cz - 1) * cg.my + (cy - 1)) * cg.mx + (cx - 1) : -1;
        }
        for (int A = 0; A < 8; ++A) {
          if (q[A] < 0) continue;
          for (int B = 0; B < 8; ++B) {
            if (q[B] < 0) continue;
            for (int d = 0; d < 3; ++d)
              for (int e = 0; e < 3; ++e)
                Ac[(size_t)(q[A] * 3 + d) + (size_t)(q[B] * 3 + e) * nc] += KEp[(A * 3 + d) * 24 + B * 3 + e];
          }
        }
      }

  // device: Cholesky factor and inverse in FP64, then the FP32 symmetric copy
  CoarseDev *c = new CoarseDev();
  c->cg = cg;
  double *a64 = nullptr;
  int *info_d = nullptr;
  int rc = MPPD_OK;
  rocblas_handle h = nullptr;
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
    if ((rc = dalloc(&a64, (size_t)nc * nc, false))) break;
    if ((rc = dalloc(&info_d, 1))) break;
    hipError_t e = hipMemcpyAsync(a64, Ac.data(), sizeof(double) * Ac.size(), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) {
      rc = hip_fail(e, "coarse matrix upload");
      break;
    }
    if (rocblas_create_handle(&h) != rocblas_status_success || rocblas_set_stream(h, st) != rocblas_status_success) {
      rc = fail(MPPD_ERR_HIP, "rocblas handle for the coarse solve");
      break;
    }
    int info = 0;
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
    if (info != 0) break;  // not positive definite: no coarse level (c stays unused)
    if (rocsolver_dpotri(h, rocblas_fill_lower, nc, a64, nc, info_d) != rocblas_status_success) {
      rc = fail(MPPD_ERR_HIP, "rocsolver_dpotri (coarse matrix)");
      break;
    }
    e = hipMemcpyAsync(&info, info_d, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      rc = hip_fail(e, "coarse inverse status");
      break;
    }
    if (info != 0) break;
    e = launch_c_sym32(nc, a64, ainv, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      rc = hip_fail(e, "coarse inverse copy");
      break;
    }
    *out = c;
  } while (0);
  if (h) (void)rocblas_destroy_handle(h);
  dfree(a64);
  dfree(info_d);
  if (*out != c) coarse_free(c);
  return rc;
}

}  // namespace mpp
