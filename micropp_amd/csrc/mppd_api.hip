// mppd_api.hip -- the thin C ABI (include/mppd.h) over the gfx950 kernels.
//
// Host code here only allocates, launches and steers: every floating-point
// operation of the hot path runs in the kernels of geometry.hip,
// assembly.hip and solver.hip.  The host reads back only the scalars that
// steer control flow (residual norms, CG state, nonlinearity flags) and the
// 6-component averages.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "internal.hpp"

namespace mpp {

static thread_local std::string g_err;

size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }
void set_error(const std::string &msg) { g_err = msg; }
int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
int hip_fail(hipError_t e, const char *what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? MPPD_ERR_NOMEM : MPPD_ERR_HIP;
}

// make_material (src/material.cpp:26-46) with the constructors of
// include/material.hpp:66-135: derived constants rebuilt from E and nu, the
// fields a law does not use set to -1.
int make_material(const material_base &in, material_base &m) {
  if (in.type < MATERIAL_ELASTIC || in.type > MATERIAL_DAMAGE) return fail(MPPD_ERR_ARG, "invalid material type");
  m = in;
  m.k = in.E / (3. * (1. - 2. * in.nu));
  m.mu = in.E / (2. * (1. + in.nu));
  m.lambda = in.nu * in.E / ((1. + in.nu) * (1. - 2. * in.nu));
  if (in.type != MATERIAL_PLASTIC) {
    m.Ka = -1.0;
    m.Sy = -1.0;
  }
  if (in.type != MATERIAL_DAMAGE) m.Xt = -1.0;
  return MPPD_OK;
}

}  // namespace mpp

using namespace mpp;

struct mppd_solver {
  mppd_rve *r = nullptr;
  hipStream_t st = nullptr;
  int *cols = nullptr;
  double *vals = nullptr;
  double *ke = nullptr;
  int phase_mode = 0;
  int ke_phase_ready = 0;
  double *be = nullptr;
  double *x = nullptr, *rr = nullptr, *z = nullptr, *p = nullptr, *Ap = nullptr, *kinv = nullptr;
  double *b = nullptr, *du = nullptr;
  double *partials = nullptr;
  CgState *cg = nullptr, *cg_h = nullptr;
  RedState *red = nullptr, *red_h = nullptr;
  struct MixedLin *ml = nullptr;  // mppd_newton_mixed buffers, made on first use
  // vals came from this library's assembly (nonzeros only on in-grid stencil
  // slots), so the CG SpMV computes the columns; false after a host upload
  bool vals_stencil = true;
};

// FP32 inner-solve buffers of mppd_newton_mixed: a 4-field system (field 0
// carries the Newton correction) in the layout of mixed.hip
struct MixedLin {
  static constexpr int K = 4;
  float *A = nullptr, *kinv = nullptr, *R = nullptr, *X = nullptr, *P = nullptr, *AP = nullptr;
  double *kinv64 = nullptr, *part = nullptr;
  void *state = nullptr, *state_h = nullptr;
  // two-level preconditioner of the inner CG from the RVE's initial (elastic)
  // tangent (null: Jacobi alone); the tangent of a Newton iteration changes,
  // the coarse matrix does not, so it stays a symmetric positive definite
  // preconditioner
  CoarseDev *coarse = nullptr;
};

static void mixed_lin_free(MixedLin *m) {
  if (!m) return;
  coarse_free(m->coarse);
  for (void *p : {(void *)m->A, (void *)m->kinv, (void *)m->R, (void *)m->X, (void *)m->P, (void *)m->AP,
                  (void *)m->kinv64, (void *)m->part, m->state})
    dfree(p);
  if (m->state_h) (void)hipHostFree(m->state_h);
  delete m;
}

extern "C" {

const char *mppd_last_error(void) { return g_err.c_str(); }

int mppd_device_count(int *n) {
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    *n = 0;
    return hip_fail(e, "hipGetDeviceCount");
  }
  *n = c;
  return c > 0 ? MPPD_OK : fail(MPPD_ERR_NODEV, "no HIP device");
}

int mppd_malloc(void **ptr_d, size_t bytes) {
  MPP_HIP(hipMalloc(ptr_d, bytes ? bytes : 1));
  return MPPD_OK;
}
int mppd_free(void *ptr_d) {
  if (ptr_d) MPP_HIP(hipFree(ptr_d));
  return MPPD_OK;
}
// The plain-memory helpers complete before they return (the solvers' streams
// are non-blocking, so null-stream work is not ordered against them).
int mppd_memcpy_h2d(void *dst_d, const void *src_h, size_t bytes) {
  MPP_HIP(hipMemcpy(dst_d, src_h, bytes, hipMemcpyHostToDevice));
  MPP_HIP(hipDeviceSynchronize());
  return MPPD_OK;
}
int mppd_memcpy_d2h(void *dst_h, const void *src_d, size_t bytes) {
  MPP_HIP(hipMemcpy(dst_h, src_d, bytes, hipMemcpyDeviceToHost));
  MPP_HIP(hipDeviceSynchronize());
  return MPPD_OK;
}
int mppd_memcpy_d2d(void *dst_d, const void *src_d, size_t bytes) {
  MPP_HIP(hipMemcpy(dst_d, src_d, bytes, hipMemcpyDeviceToDevice));
  MPP_HIP(hipDeviceSynchronize());
  return MPPD_OK;
}
int mppd_memset(void *ptr_d, int value, size_t bytes) {
  MPP_HIP(hipMemset(ptr_d, value, bytes));
  MPP_HIP(hipDeviceSynchronize());
  return MPPD_OK;
}
int mppd_synchronize(void) {
  MPP_HIP(hipDeviceSynchronize());
  return MPPD_OK;
}

int mppd_mem_info(size_t *free_bytes, size_t *total_bytes) {
  if (!free_bytes || !total_bytes) return fail(MPPD_ERR_ARG, "null argument");
  MPP_HIP(hipMemGetInfo(free_bytes, total_bytes));
  return MPPD_OK;
}

// ---------------------------------------------------------------------------
// RVE
// ---------------------------------------------------------------------------
int mppd_rve_create(mppd_rve **out, const int size[3], int micro_type, const double geo[4],
                    const struct material_base mats[3]) {
  *out = nullptr;
  if (!size || !geo || !mats) return fail(MPPD_ERR_ARG, "null argument");
  for (int d = 0; d < 3; ++d)
    if (size[d] < 2) return fail(MPPD_ERR_ARG, "size must be >= 2 nodes per side");
  if (micro_type < 0 || micro_type >= kNumMicroTypes) return fail(MPPD_ERR_ARG, "invalid micro_type");
  const double nrow_d = 3.0 * size[0] * size[1] * size[2];
  if (nrow_d * kSlots > 2147483647.0) return fail(MPPD_ERR_ARG, "RVE too large for 32-bit column indices");
  mppd_rve *r = new mppd_rve();
  for (int i = 0; i < 3; ++i) {
    const int rc = make_material(mats[i], r->mats[i]);
    if (rc) {
      delete r;
      return rc;
    }
    if (mats[i].type != MATERIAL_ELASTIC) r->all_elastic = 0;
  }
  Grid &g = r->g;
  g.nx = size[0];
  g.ny = size[1];
  g.nz = size[2];
  g.nex = g.nx - 1;
  g.ney = g.ny - 1;
  g.nez = g.nz - 1;
  g.nn = g.nx * g.ny * g.nz;
  g.nrow = 3 * g.nn;
  g.nrow_pad = (int)round_up(g.nrow, kRowPad);
  g.nelem = g.nex * g.ney * g.nez;
  // micropp constructor (src/micropp.cpp:29-71): unit cube
  const double lx = 1.0, ly = 1.0, lz = 1.0;
  g.dx = lx / g.nex;
  g.dy = ly / g.ney;
  g.dz = lz / g.nez;
  g.wg = g.dx * g.dy * g.dz / 8;
  g.vol_tot = lx * ly * lz;
  g.evol = g.dx * g.dy * g.dz;
  r->micro_type = micro_type;
  for (int i = 0; i < 4; ++i) r->geo[i] = geo[i];
  int rc = MPPD_OK;
  int *bad = nullptr;
  unsigned long long *cnt_d = nullptr;
  do {
    if ((rc = dalloc(&r->bmat, 8 * 144))) break;
    if ((rc = dalloc(&r->mats_d, 3))) break;
    if ((rc = dalloc(&r->et, (size_t)g.nelem))) break;
    if (hipMemcpy(r->mats_d, r->mats, 3 * sizeof(material_base), hipMemcpyHostToDevice) != hipSuccess) {
      rc = fail(MPPD_ERR_HIP, "materials upload");
      break;
    }
    if ((rc = dalloc(&r->spheres, (size_t)kMaxSpheres * 4))) break;
    if ((rc = dalloc(&bad, 1))) break;
    if ((rc = dalloc(&cnt_d, 2))) break;
    hipError_t e = launch_bmat(g, r->bmat, nullptr);
    if (e == hipSuccess && micro_type == kMicRandomSpheres) {
      int n = (int)geo[1];
      r->nsph = n < 0 ? 0 : (n > kMaxSpheres ? kMaxSpheres : n);
      e = launch_random_spheres(geo, r->spheres, nullptr);
    }
    if (e == hipSuccess) e = launch_phase_map(g, micro_type, geo, r->spheres, r->nsph, r->et, bad, nullptr);
    if (e == hipSuccess) e = launch_phase_counts(g, r->et, cnt_d, nullptr);
    if (e != hipSuccess) {
      rc = hip_fail(e, "rve set-up kernels");
      break;
    }
    int bad_h = 0;
    unsigned long long cnt_h[2] = {0, 0};
    e = hipMemcpy(&bad_h, bad, sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(cnt_h, cnt_d, sizeof(cnt_h), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      rc = hip_fail(e, "rve set-up readback");
      break;
    }
    // calc_volume_fractions (src/micropp.cpp:609-627): evol added once per
    // element of each phase, in the reference's sequence of additions
    double vm = 0.0, vfb = 0.0;
    for (unsigned long long q = 0; q < cnt_h[0]; ++q) vm += g.evol;
    for (unsigned long long q = 0; q < cnt_h[1]; ++q) vfb += g.evol;
    r->vf[0] = vm / g.vol_tot;
    r->vf[1] = vfb / g.vol_tot;
    if (bad_h) rc = fail(MPPD_ERR_ARG, "invalid micro_type");
  } while (0);
  dfree(bad);
  dfree(cnt_d);
  if (rc) {
    mppd_rve_destroy(r);
    return rc;
  }
  *out = r;
  return MPPD_OK;
}

int mppd_rve_destroy(mppd_rve *r) {
  if (!r) return MPPD_OK;
  dfree(r->bmat);
  dfree(r->mats_d);
  dfree(r->et);
  dfree(r->spheres);
  delete r;
  return MPPD_OK;
}

int mppd_rve_info(const mppd_rve *r, int *nrow, int *nelem, int *nvars) {
  if (!r) return fail(MPPD_ERR_ARG, "null rve");
  if (nrow) *nrow = r->g.nrow;
  if (nelem) *nelem = r->g.nelem;
  if (nvars) *nvars = r->g.nelem * 8 * 7;
  return MPPD_OK;
}

int mppd_rve_bmat(const mppd_rve *r, double *bmat_h) {
  MPP_HIP(hipMemcpy(bmat_h, r->bmat, 8 * 144 * sizeof(double), hipMemcpyDeviceToHost));
  return MPPD_OK;
}

int mppd_rve_elem_types(const mppd_rve *r, int *et_h) {
  MPP_HIP(hipMemcpy(et_h, r->et, (size_t)r->g.nelem * sizeof(int), hipMemcpyDeviceToHost));
  return MPPD_OK;
}

int mppd_rve_volume_fractions(const mppd_rve *r, double vf_h[2]) {
  vf_h[0] = r->vf[0];
  vf_h[1] = r->vf[1];
  return MPPD_OK;
}

int mppd_material_eval(const struct material_base *m, int n, const double *eps_d, const double *hist_d,
                       double *sig_d, double *ctan_d, double *hist_new_d, int *nl_d) {
  if (!m || n < 0 || (n > 0 && !eps_d)) return fail(MPPD_ERR_ARG, "bad material_eval arguments");
  material_base mm;
  int rc = make_material(*m, mm);
  if (rc) return rc;
  MPP_HIP(launch_material_eval(mm, n, eps_d, hist_d, sig_d, ctan_d, hist_new_d, nl_d, nullptr));
  MPP_HIP(hipDeviceSynchronize());
  return MPPD_OK;
}

// ---------------------------------------------------------------------------
// solver
// ---------------------------------------------------------------------------
int mppd_solver_create(mppd_solver **out, mppd_rve *r) {
  *out = nullptr;
  if (!r) return fail(MPPD_ERR_ARG, "null rve");
  mppd_solver *s = new mppd_solver();
  s->r = r;
  const Grid &g = r->g;
  const size_t np = (size_t)g.nrow_pad;
  s->phase_mode = r->all_elastic;
  int rc = MPPD_OK;
  do {
    hipError_t e = hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking);
    if (e != hipSuccess) {
      rc = hip_fail(e, "hipStreamCreate");
      break;
    }
    // zero fills on s->st (synchronized below)
    hipStream_t st = s->st;
    if ((rc = dalloc(&s->cols, np * kSlots, false))) break;
    if ((rc = dalloc(&s->vals, np * kSlots, st))) break;
    if ((rc = dalloc(&s->ke, (s->phase_mode ? 3 : (size_t)g.nelem) * 576, false))) break;
    if ((rc = dalloc(&s->be, (size_t)g.nelem * 24, false))) break;
    double **vecs[] = {&s->x, &s->rr, &s->z, &s->p, &s->Ap, &s->kinv, &s->b, &s->du};
    for (double **v : vecs)
      if ((rc = dalloc(v, np, st))) break;
    if (rc) break;
    const size_t nb = std::max(round_up(np, 256) / 256, round_up((size_t)g.nelem, 256) / 256) + 1;
    if ((rc = dalloc(&s->partials, nb * 8, st))) break;
    if ((rc = dalloc(&s->cg, 2, st))) break;  // ping-pong pair
    if ((rc = dalloc(&s->red, 1, st))) break;
    e = hipHostMalloc((void **)&s->cg_h, sizeof(CgState), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&s->red_h, sizeof(RedState), hipHostMallocDefault);
    if (e == hipSuccess) e = launch_ell_cols(g, s->cols, s->st);
    if (e == hipSuccess) e = hipStreamSynchronize(s->st);
    if (e != hipSuccess) rc = hip_fail(e, "solver set-up");
  } while (0);
  if (rc) {
    mppd_solver_destroy(s);
    return rc;
  }
  *out = s;
  return MPPD_OK;
}

int mppd_solver_destroy(mppd_solver *s) {
  if (!s) return MPPD_OK;
  if (s->st) (void)hipStreamSynchronize(s->st);
  for (void *p : {(void *)s->cols, (void *)s->vals, (void *)s->ke, (void *)s->be, (void *)s->x, (void *)s->rr,
                  (void *)s->z, (void *)s->p, (void *)s->Ap, (void *)s->kinv, (void *)s->b, (void *)s->du,
                  (void *)s->partials, (void *)s->cg, (void *)s->red})
    dfree(p);
  if (s->cg_h) (void)hipHostFree(s->cg_h);
  if (s->red_h) (void)hipHostFree(s->red_h);
  mixed_lin_free(s->ml);
  if (s->st) (void)hipStreamDestroy(s->st);
  delete s;
  return MPPD_OK;
}

int mppd_solver_copy(mppd_solver *s, void *dst_d, const void *src_d, size_t bytes) {
  if (!s || ((!dst_d || !src_d) && bytes)) return fail(MPPD_ERR_ARG, "null argument");
  if (!bytes) return MPPD_OK;
  MPP_HIP(hipMemcpyAsync(dst_d, src_d, bytes, hipMemcpyDeviceToDevice, s->st));
  MPP_HIP(hipStreamSynchronize(s->st));
  return MPPD_OK;
}

int mppd_solver_fill(mppd_solver *s, void *ptr_d, int value, size_t bytes) {
  if (!s || (!ptr_d && bytes)) return fail(MPPD_ERR_ARG, "null argument");
  if (!bytes) return MPPD_OK;
  MPP_HIP(hipMemsetAsync(ptr_d, value, bytes, s->st));
  MPP_HIP(hipStreamSynchronize(s->st));
  return MPPD_OK;
}

int mppd_ell_download(mppd_solver *s, int *cols_h, double *vals_h) {
  const Grid &g = s->r->g;
  const size_t n = (size_t)g.nrow * kSlots;
  int *c = nullptr;
  double *v = nullptr;
  int rc = MPPD_OK;
  if (cols_h && (rc = dalloc(&c, n, false))) return rc;
  if (vals_h && (rc = dalloc(&v, n, false))) {
    dfree(c);
    return rc;
  }
  hipError_t e = launch_sell_to_rows(g, s->cols, s->vals, c, v, s->st);
  if (e == hipSuccess && c) e = hipMemcpyAsync(cols_h, c, n * sizeof(int), hipMemcpyDeviceToHost, s->st);
  if (e == hipSuccess && v) e = hipMemcpyAsync(vals_h, v, n * sizeof(double), hipMemcpyDeviceToHost, s->st);
  if (e == hipSuccess) e = hipStreamSynchronize(s->st);
  dfree(c);
  dfree(v);
  return e == hipSuccess ? MPPD_OK : hip_fail(e, "ell_download");
}

int mppd_ell_upload_vals(mppd_solver *s, const double *vals_h) {
  const Grid &g = s->r->g;
  const size_t n = (size_t)g.nrow * kSlots;
  double *v = nullptr;
  int rc = dalloc(&v, n, false);
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(v, vals_h, n * sizeof(double), hipMemcpyHostToDevice, s->st);
  if (e == hipSuccess) e = launch_rows_to_sell(g, v, s->vals, s->st);
  if (e == hipSuccess) e = hipStreamSynchronize(s->st);
  dfree(v);
  s->vals_stencil = false;  // any slot may hold a value: the CG reads the columns
  return e == hipSuccess ? MPPD_OK : hip_fail(e, "ell_upload_vals");
}

int mppd_set_displ_bc(mppd_solver *s, const double eps[6], double *u_d) {
  MPP_HIP(launch_set_displ_bc(s->r->g, eps, u_d, s->st));
  MPP_HIP(hipStreamSynchronize(s->st));
  return MPPD_OK;
}

static hipError_t rhs_async(mppd_solver *s, const double *u_d, const double *vars_d, double *b_d) {
  const mppd_rve *r = s->r;
  hipError_t e = launch_elem_be(r->g, r->bmat, r->mats_d, r->et, u_d, vars_d, s->be, s->st);
  if (e == hipSuccess) e = launch_gather_rhs(r->g, s->be, b_d, s->partials, s->red, s->st);
  return e;
}

static hipError_t read_red(mppd_solver *s) {
  hipError_t e = hipMemcpyAsync(s->red_h, s->red, sizeof(RedState), hipMemcpyDeviceToHost, s->st);
  if (e == hipSuccess) e = hipStreamSynchronize(s->st);
  return e;
}

int mppd_assembly_rhs(mppd_solver *s, const double *u_d, const double *vars_d, double *b_d, double *norm) {
  MPP_HIP(rhs_async(s, u_d, vars_d, b_d));
  MPP_HIP(read_red(s));
  if (norm) *norm = s->red_h->v[0];
  return MPPD_OK;
}

static hipError_t mat_async(mppd_solver *s, const double *u_d, const double *vars_d) {
  const mppd_rve *r = s->r;
  hipError_t e = hipSuccess;
  if (s->phase_mode) {
    if (!s->ke_phase_ready) {
      e = launch_elem_ke(r->g, r->bmat, r->mats_d, r->et, nullptr, nullptr, 1, s->ke, s->st);
      if (e == hipSuccess) s->ke_phase_ready = 1;
    }
  } else {
    e = launch_elem_ke(r->g, r->bmat, r->mats_d, r->et, u_d, vars_d, 0, s->ke, s->st);
  }
  if (e == hipSuccess) e = launch_gather_vals(r->g, r->et, s->ke, s->phase_mode, s->vals, s->st);
  if (e == hipSuccess) s->vals_stencil = true;  // every slot written, zero off the in-grid stencil
  return e;
}

int mppd_assembly_mat(mppd_solver *s, const double *u_d, const double *vars_d) {
  MPP_HIP(mat_async(s, u_d, vars_d));
  MPP_HIP(hipStreamSynchronize(s->st));
  return MPPD_OK;
}

int mppd_ell_mvp(mppd_solver *s, const double *x_d, double *y_d) {
  MPP_HIP(launch_spmv(s->r->g, s->cols, s->vals, x_d, y_d, s->st));
  MPP_HIP(hipStreamSynchronize(s->st));
  return MPPD_OK;
}

// The working set of one Jacobi-PCG: vectors of nrow_pad doubles, the
// partial table (3 * nrow_pad / 256 doubles), the ping-pong state pair on the
// device and one host (pinned) state for the reads between chunks.
struct CgWork {
  double *x, *r, *z, *p, *Ap, *kinv, *part;
  CgState *state, *state_h;
};

// CG on the device; the host queues iterations in chunks sized from the
// observed convergence rate and reads the 64-byte state between chunks.
// After `queued` iterations the current state is state[queued & 1].
static int cg_run_on(const Grid &g, const mppd_solver *A, const CgWork &w, hipStream_t st, const double *b_d,
                     double min_err, double rel_err, int max_its, int *its_out, double *err_out) {
  MPP_HIP(launch_cg_init(g, A->vals, b_d, w.x, w.r, w.z, w.p, w.kinv, w.part, min_err, rel_err, max_its, w.state,
                         st));
  int queued = 0;
  MPP_HIP(hipMemcpyAsync(w.state_h, w.state, sizeof(CgState), hipMemcpyDeviceToHost, st));
  MPP_HIP(hipStreamSynchronize(st));
  double pn_prev = w.state_h->pn;
  int its_prev = 0;
  int chunk = 32;
  while (!w.state_h->done) {
    const int n = std::min(chunk, max_its - queued);
    if (n <= 0) break;
    for (int i = 0; i < n; ++i)
      MPP_HIP(launch_cg_iter(g, A->vals_stencil ? nullptr : A->cols, A->vals, w.x, w.r, w.z, w.p, w.Ap, w.kinv,
                             w.part, w.state, queued + i, st));
    queued += n;
    MPP_HIP(hipMemcpyAsync(w.state_h, w.state + (queued & 1), sizeof(CgState), hipMemcpyDeviceToHost, st));
    MPP_HIP(hipStreamSynchronize(st));
    const CgState &c = *w.state_h;
    if (c.done) break;
    // next chunk: iterations still needed at the observed rate, within [4, 256]
    const double target = std::max(c.min_err, c.pn0 * c.rel_err);
    const int dit = c.its - its_prev;
    int next = chunk * 2;
    if (dit > 0 && c.pn > 0 && pn_prev > 0 && c.pn < pn_prev) {
      const double rate = std::log(c.pn / pn_prev) / dit;
      const double need = std::log(target / c.pn) / rate;
      if (std::isfinite(need)) next = (int)std::ceil(need) + 2;
    }
    chunk = std::max(4, std::min(256, next));
    pn_prev = c.pn;
    its_prev = c.its;
  }
  if (its_out) *its_out = w.state_h->its;
  if (err_out) *err_out = w.state_h->rz;
  return MPPD_OK;
}

static int cg_run(mppd_solver *s, const mppd_solver *A, const double *b_d, double min_err, double rel_err,
                  int max_its, int *its_out, double *err_out) {
  const CgWork w{s->x, s->rr, s->z, s->p, s->Ap, s->kinv, s->partials, s->cg, s->cg_h};
  return cg_run_on(s->r->g, A, w, s->st, b_d, min_err, rel_err, max_its, its_out, err_out);
}

// A0's matrix is read with s's row layout: it must be a solver of an RVE grid
// of the same size (in micropp<3>, the instance's own RVE).
static int check_A0(const mppd_solver *s, const mppd_solver *A0) {
  if (!A0 || A0 == s) return MPPD_OK;
  const Grid &a = s->r->g, &b = A0->r->g;
  if (a.nx != b.nx || a.ny != b.ny || a.nz != b.nz) return fail(MPPD_ERR_ARG, "A0 belongs to an RVE of another size");
  return MPPD_OK;
}

int mppd_cg(mppd_solver *s, const mppd_solver *A0, const double *b_d, double *x_d, double min_err, double rel_err,
            int max_its, int *its, double *err) {
  if (!s || !b_d || !x_d) return fail(MPPD_ERR_ARG, "null argument");
  if (int rc = check_A0(s, A0)) return rc;
  const mppd_solver *A = A0 ? A0 : s;
  int rc = cg_run(s, A, b_d, min_err, rel_err, max_its, its, err);
  if (rc) return rc;
  MPP_HIP(launch_copy(s->r->g.nrow, s->x, x_d, s->st));
  MPP_HIP(hipStreamSynchronize(s->st));
  return MPPD_OK;
}

// mppd_cg_noalloc: the working set of the CG (x, r, z, p, Ap, D^-1, the
// partial table and the state pair) lives in `scratch`, a fixed device array
// of this file, instead of the solver's own buffers.  Layout (doubles): the
// state pair in the first 256 bytes, then 6 vectors of kNoAllocRows, then the
// partial table.  x is a scratch vector too: the CG kernels write all
// nrow_pad rows, the caller's x_d holds nrow.  On each device the array's address comes from
// hipGetSymbolAddress (once per device, with that device current); in host
// code the array's name is not a device address and is never passed on.
constexpr size_t kNoAllocRows = MPPD_NOALLOC_MAX_ROWS;
static_assert(kNoAllocRows % 256 == 0, "whole row blocks");
constexpr size_t kNoAllocHead = 256 / sizeof(double);
constexpr size_t kNoAllocParts = 3 * (kNoAllocRows / 256);
constexpr size_t kNoAllocDoubles = kNoAllocHead + 6 * kNoAllocRows + kNoAllocParts;
static_assert(2 * sizeof(CgState) <= 256, "the state pair fits the head");
static __device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_noalloc_mu;
static double *g_noalloc_ptr[kNoAllocMaxDevices];

// the current device's copy of `scratch` (looked up on the first call there)
static int noalloc_scratch(double **out) {
  int dev = 0;
  MPP_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) return fail(MPPD_ERR_ARG, "mppd_cg_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!ptr) return fail(MPPD_ERR_HIP, "hipGetSymbolAddress(scratch) gave no address");
    g_noalloc_ptr[dev] = static_cast<double *>(ptr);
  }
  *out = g_noalloc_ptr[dev];
  return MPPD_OK;
}

int mppd_cg_noalloc(mppd_solver *s, const mppd_solver *A0, const double *b_d, double *x_d, double min_err,
                    double rel_err, int max_its, int *its, double *err) {
  if (!s || !b_d || !x_d) return fail(MPPD_ERR_ARG, "null argument");
  if (int rc = check_A0(s, A0)) return rc;
  const Grid &g = s->r->g;
  if ((size_t)g.nrow_pad > kNoAllocRows)
    return fail(MPPD_ERR_ARG, "mppd_cg_noalloc: " + std::to_string(g.nrow) + " rows exceed the scratch maximum of " +
                                  std::to_string(kNoAllocRows) + " (MPPD_NOALLOC_MAX_ROWS)");
  double *ptr = nullptr;
  if (int rc = noalloc_scratch(&ptr)) return rc;
  // every entry the CG reads is written by its initialisation (all nrow_pad
  // rows of the vectors, every partial, state[0]) or by an earlier kernel of
  // the same solve, so nothing left by the previous call needs clearing
  double *v = ptr + kNoAllocHead;
  const CgWork w{v,
                 v + kNoAllocRows,
                 v + 2 * kNoAllocRows,
                 v + 3 * kNoAllocRows,
                 v + 4 * kNoAllocRows,
                 v + 5 * kNoAllocRows,
                 v + 6 * kNoAllocRows,
                 reinterpret_cast<CgState *>(ptr),
                 s->cg_h};
  if (int rc = cg_run_on(g, A0 ? A0 : s, w, s->st, b_d, min_err, rel_err, max_its, its, err)) return rc;
  MPP_HIP(launch_copy(g.nrow, w.x, x_d, s->st));
  MPP_HIP(hipStreamSynchronize(s->st));
  return MPPD_OK;
}

int mppd_ave_stress(mppd_solver *s, const double *u_d, const double *vars_d, double sig[6]) {
  const mppd_rve *r = s->r;
  MPP_HIP(launch_ave_stress(r->g, r->bmat, r->mats_d, r->et, u_d, vars_d, s->partials, s->red, s->st));
  MPP_HIP(read_red(s));
  for (int v = 0; v < 6; ++v) sig[v] = s->red_h->v[v];
  return MPPD_OK;
}

int mppd_elem_fields(mppd_solver *s, const double *u_d, const double *vars_d, double *strain_h, double *stress_h) {
  if (!s || !u_d || !strain_h || !stress_h) return fail(MPPD_ERR_ARG, "null argument");
  const mppd_rve *r = s->r;
  const size_t n = (size_t)r->g.nelem * 6;
  double *buf = nullptr;
  int rc = dalloc(&buf, 2 * n, false);
  if (rc) return rc;
  hipError_t e = launch_elem_fields(r->g, r->bmat, r->mats_d, r->et, u_d, vars_d, buf, buf + n, s->st);
  if (e == hipSuccess) e = hipMemcpyAsync(strain_h, buf, sizeof(double) * n, hipMemcpyDeviceToHost, s->st);
  if (e == hipSuccess) e = hipMemcpyAsync(stress_h, buf + n, sizeof(double) * n, hipMemcpyDeviceToHost, s->st);
  if (e == hipSuccess) e = hipStreamSynchronize(s->st);
  dfree(buf);
  return e == hipSuccess ? MPPD_OK : hip_fail(e, "calc_fields");
}

int mppd_vars_new(mppd_solver *s, const double *u_d, const double *vars_old_d, double *vars_new_d, int *nonlinear) {
  const mppd_rve *r = s->r;
  MPP_HIP(hipMemsetAsync(&s->red->flag, 0, sizeof(int), s->st));
  MPP_HIP(launch_vars_new(r->g, r->bmat, r->mats_d, r->et, u_d, vars_old_d, vars_new_d, s->red, s->st));
  MPP_HIP(read_red(s));
  if (nonlinear) *nonlinear = s->red_h->flag ? 1 : 0;
  return MPPD_OK;
}

int mppd_newton(mppd_solver *s, const mppd_solver *A0, const double eps[6], const double *vars_old_d, double *u_d,
                const mppd_newton_params *p, int out[3]) {
  if (!s || !eps || !u_d || !p || !out) return fail(MPPD_ERR_ARG, "null argument");
  if (int rc = check_A0(s, A0)) return rc;
  const Grid &g = s->r->g;
  int its = 0, solver_its = 0, conv = 0;
  MPP_HIP(launch_set_displ_bc(g, eps, u_d, s->st));
  MPP_HIP(rhs_async(s, u_d, vars_old_d, s->b));
  MPP_HIP(read_red(s));
  double norm = s->red_h->v[0];
  const double norm0 = norm;
  while (its < p->nr_max_its) {
    if (norm < p->nr_max_tol || norm < norm0 * p->nr_rel_tol) {
      conv = 1;
      break;
    }
    const mppd_solver *A = s;
    if (!A0 || its > p->its_with_A0 - 1) {
      MPP_HIP(mat_async(s, u_d, vars_old_d));
    } else {
      A = A0;
    }
    int cg_its = 0;
    double err = 0;
    int rc = cg_run(s, A, s->b, p->cg_abs_tol, p->cg_rel_tol, p->cg_max_its, &cg_its, &err);
    if (rc) return rc;
    solver_its += cg_its;
    MPP_HIP(launch_axpy(g.nrow, 1.0, s->x, u_d, s->st));
    MPP_HIP(rhs_async(s, u_d, vars_old_d, s->b));
    MPP_HIP(read_red(s));
    norm = s->red_h->v[0];
    its++;
  }
  out[0] = its;
  out[1] = solver_its;
  out[2] = conv;
  return MPPD_OK;
}

int mppd_cg_profile(mppd_solver *s, const double *b_d, int iters, double *spmv_ms, double *iter_ms) {
  if (!s || !b_d || iters <= 0) return fail(MPPD_ERR_ARG, "bad cg_profile arguments");
  const Grid &g = s->r->g;
  // min_err = rel_err = 0: the stopping test never holds within iters + 1
  MPP_HIP(launch_cg_init(g, s->vals, b_d, s->x, s->rr, s->z, s->p, s->kinv, s->partials, 0.0, 0.0, iters + 1, s->cg,
                         s->st));
  hipEvent_t ev[3];
  for (auto &x : ev) MPP_HIP(hipEventCreate(&x));
  // (1) the SpMV kernel back to back: it reads p and state[0], writes Ap and
  // its partials, so repeating it leaves the CG state unchanged
  hipError_t e = hipEventRecord(ev[0], s->st);
  for (int i = 0; i < iters && e == hipSuccess; ++i)
    e = launch_cg_step1(g, s->vals_stencil ? nullptr : s->cols, s->vals, s->p, s->Ap, s->partials, s->cg, 0, s->st);
  // (2) whole iterations
  if (e == hipSuccess) e = hipEventRecord(ev[1], s->st);
  for (int i = 0; i < iters && e == hipSuccess; ++i)
    e = launch_cg_iter(g, s->vals_stencil ? nullptr : s->cols, s->vals, s->x, s->rr, s->z, s->p, s->Ap, s->kinv,
                       s->partials, s->cg, i, s->st);
  if (e == hipSuccess) e = hipEventRecord(ev[2], s->st);
  if (e == hipSuccess) e = hipStreamSynchronize(s->st);
  float ms1 = 0.f, ms2 = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms1, ev[0], ev[1]);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms2, ev[1], ev[2]);
  for (auto &x : ev) (void)hipEventDestroy(x);
  if (e != hipSuccess) return hip_fail(e, "cg_profile");
  if (spmv_ms) *spmv_ms = (double)ms1 / iters;
  if (iter_ms) *iter_ms = (double)ms2 / iters;
  return MPPD_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// mixed-precision multi-RHS solver
// ---------------------------------------------------------------------------
constexpr int kMaxRhsFields = 16;  // largest K of an msolver

struct mppd_msolver {
  mppd_rve *r = nullptr;
  int K = 0;
  CoarseDev *coarse = nullptr;  // two-level preconditioner of the inner CG (null: Jacobi alone)
  hipStream_t st = nullptr;
  double *ke = nullptr;         // phase table 3 x 576 (FP64)
  material_base *mats_d = nullptr;
  float *A = nullptr, *kinv = nullptr;
  float *X = nullptr, *R = nullptr, *P = nullptr, *AP = nullptr;  // 4-field-group layout
  double *U = nullptr, *eps_d = nullptr, *part = nullptr, *sig_d = nullptr;
  void *state = nullptr, *state_h = nullptr;
  double *eps_h = nullptr, *sig_h = nullptr;  // pinned staging
  int *act_d = nullptr, *act_h = nullptr;     // fields taking part in a Newton iteration (device, pinned)
  // adjoint (allocated on first use): sources, per-phase source table,
  // partials, results (device and pinned)
  double *F = nullptr, *G = nullptr, *apart = nullptr, *adj_d = nullptr, *adj_h = nullptr;
  void *arena = nullptr, *harena = nullptr;  // the device buffers above through act_d, and the pinned staging
};

extern "C" {

int mppd_msolver_create(mppd_msolver **out, mppd_rve *r, int nrhs) {
  *out = nullptr;
  if (!r) return fail(MPPD_ERR_ARG, "null rve");
  if (!(nrhs == 4 || nrhs == 8 || nrhs == 12 || nrhs == 16)) return fail(MPPD_ERR_ARG, "nrhs must be 4, 8, 12 or 16");
  if (!r->all_elastic) return fail(MPPD_ERR_ARG, "the mixed-precision solver needs linear elastic phases");
  mppd_msolver *m = new mppd_msolver();
  m->r = r;
  m->K = nrhs;
  const Grid &g = r->g;
  int rc = MPPD_OK;
  do {
    hipError_t e = hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking);
    if (e != hipSuccess) {
      rc = hip_fail(e, "hipStreamCreate");
      break;
    }
    const size_t nv = mixed_vector_len(g, nrhs);
    // The device buffers in one allocation, 256-byte aligned parts, the FP32
    // matrix last: everything before it is zero-filled on m->st in one fill
    // (ordered before the set-up kernels; mats_d is then written by a
    // blocking copy), the matrix is written whole by the assembly.  The
    // pinned staging in one host allocation.
    hipStream_t st = m->st;
    size_t off = 0;
    auto carve = [&off](size_t bytes) {
      const size_t o = off;
      off += (bytes + 255) & ~(size_t)255;
      return o;
    };
    const size_t o_ke = carve(3 * 576 * sizeof(double)), o_mats = carve(3 * sizeof(material_base));
    const size_t o_kinv = carve(nv / nrhs * sizeof(float)), o_X = carve(nv * sizeof(float)),
                 o_R = carve(nv * sizeof(float)), o_P = carve(nv * sizeof(float)), o_AP = carve(nv * sizeof(float));
    const size_t o_U = carve(nv * sizeof(double)), o_eps = carve(16 * 6 * sizeof(double)),
                 o_sig = carve(16 * 6 * sizeof(double)), o_part = carve(mixed_partials(g, nrhs) * sizeof(double));
    const size_t o_state = carve(mixed_state_bytes()), o_act = carve(kMaxRhsFields * sizeof(int));
    const size_t zeroed = off, o_A = carve(mixed_matrix_floats(g) * sizeof(float));
    e = hipMalloc(&m->arena, off);
    if (e == hipSuccess) e = hipMemsetAsync(m->arena, 0, zeroed, st);
    if (e != hipSuccess) {
      rc = hip_fail(e, "msolver buffers");
      break;
    }
    char *base = static_cast<char *>(m->arena);
    m->ke = reinterpret_cast<double *>(base + o_ke);
    m->mats_d = reinterpret_cast<material_base *>(base + o_mats);
    m->kinv = reinterpret_cast<float *>(base + o_kinv);
    m->X = reinterpret_cast<float *>(base + o_X);
    m->R = reinterpret_cast<float *>(base + o_R);
    m->P = reinterpret_cast<float *>(base + o_P);
    m->AP = reinterpret_cast<float *>(base + o_AP);
    m->U = reinterpret_cast<double *>(base + o_U);
    m->eps_d = reinterpret_cast<double *>(base + o_eps);
    m->sig_d = reinterpret_cast<double *>(base + o_sig);
    m->part = reinterpret_cast<double *>(base + o_part);
    m->state = base + o_state;
    m->act_d = reinterpret_cast<int *>(base + o_act);
    m->A = reinterpret_cast<float *>(base + o_A);
    off = 0;
    const size_t h_state = carve(mixed_state_bytes()), h_act = carve(kMaxRhsFields * sizeof(int)),
                 h_eps = carve(16 * 6 * sizeof(double)), h_sig = carve(16 * 6 * sizeof(double));
    e = hipHostMalloc(&m->harena, off, hipHostMallocDefault);
    if (e == hipSuccess) {
      char *hb = static_cast<char *>(m->harena);
      m->state_h = hb + h_state;
      m->act_h = reinterpret_cast<int *>(hb + h_act);
      m->eps_h = reinterpret_cast<double *>(hb + h_eps);
      m->sig_h = reinterpret_cast<double *>(hb + h_sig);
      e = hipStreamSynchronize(st);  // the fill, before the blocking copy into mats_d
    }
    if (e == hipSuccess) e = hipMemcpy(m->mats_d, r->mats, 3 * sizeof(material_base), hipMemcpyHostToDevice);
    // element stiffness per phase (FP64, the plain path's kernel), then the FP32 node-block matrix
    if (e == hipSuccess) e = launch_elem_ke(g, r->bmat, r->mats_d, r->et, nullptr, nullptr, 1, m->ke, m->st);
    // (U serves as FP64 scratch for the Jacobi weights: 3 NP <= 3 K NP doubles)
    if (e == hipSuccess) e = launch_m_asm32(g, r->et, m->ke, 1, m->U, m->A, m->kinv, m->st);
    if (e == hipSuccess) e = hipStreamSynchronize(m->st);
    if (e != hipSuccess) {
      rc = hip_fail(e, "msolver set-up");
      break;
    }
    rc = coarse_setup(g, r->et, m->ke, nrhs, m->st, &m->coarse);
  } while (0);
  if (rc) {
    mppd_msolver_destroy(m);
    return rc;
  }
  *out = m;
  return MPPD_OK;
}

int mppd_msolver_destroy(mppd_msolver *m) {
  if (!m) return MPPD_OK;
  if (m->st) (void)hipStreamSynchronize(m->st);
  for (void *p : {m->arena, (void *)m->F, (void *)m->G, (void *)m->apart, (void *)m->adj_d}) dfree(p);
  for (void *p : {m->harena, (void *)m->adj_h})
    if (p) (void)hipHostFree(p);
  coarse_free(m->coarse);
  if (m->st) (void)hipStreamDestroy(m->st);
  delete m;
  return MPPD_OK;
}

// The K fields: 0 .. nadj-1 adjoint fields (zero boundary values, source
// a_v / V of field v), then n forward fields with the macro strains eps_h,
// then idle fields.  FP64 defect correction with FP32 inner CG until every
// field converges; the solution is left in m->U.
// Inner FP32 CG of a K-field system, from the start state the outer test has
// written (state_h holds it), queued in chunks; after q iterations the state
// is state[q & 1].  The next chunk is sized from each live field's observed
// convergence rate (iterations still needed, + 2), so few iterations are
// queued past the end.  *its = iterations run; *finite (may be NULL) = every
// field's ||z|| is finite at the end (no breakdown).
static int mixed_inner(const Grid &g, int K, const float *A, const float *kinv, float *R, float *X, float *P,
                       float *AP, double *part, void *state, void *state_h, int max_inner, hipStream_t st,
                       int *its_out, bool *finite = nullptr, bool stop_on_stall = false,
                       const CoarseDev *coarse = nullptr) {
  int queued = 0, its = 0, done = 0, its_prev = 0;
  int chunk = 32;
  std::vector<double> pn_prev(K, 0.0);
  for (int k = 0; k < K; ++k) {
    double target;
    (void)mixed_state_field(state_h, 0, k, &pn_prev[k], &target);
  }
  while (!done) {
    const int nq = std::min(chunk, max_inner - queued);
    if (nq <= 0) break;
    for (int i = 0; i < nq; ++i)
      MPP_HIP(launch_m_iter(g, K, A, kinv, R, X, P, AP, part, state, queued + i, st, coarse));
    queued += nq;
    MPP_HIP(hipMemcpyAsync(state_h, state, mixed_state_bytes(), hipMemcpyDeviceToHost, st));
    MPP_HIP(hipStreamSynchronize(st));
    mixed_state_view(state_h, queued, &its, &done);
    if (done) break;
    int next = 0;
    const int dit = its - its_prev;
    bool stalled = dit >= 32;
    for (int k = 0; k < K; ++k) {
      double pnk, target;
      if (mixed_state_field(state_h, queued, k, &pnk, &target)) continue;
      if (pnk < 0.9 * pn_prev[k]) stalled = false;
      int need = chunk * 2;
      if (dit > 0 && pn_prev[k] > 0 && pnk > 0 && pnk < pn_prev[k]) {
        const double rate = std::log(pnk / pn_prev[k]) / dit;
        const double est = std::log(target / pnk) / rate;
        if (std::isfinite(est)) need = (int)std::ceil(est) + 2;
      }
      next = std::max(next, need);
      pn_prev[k] = pnk;
    }
    its_prev = its;
    chunk = std::max(4, std::min(256, next));
    // FP32 CG that no longer reduces ||z|| by 10% over a chunk of >= 32
    // iterations (rounding, or a nonsymmetric Newton tangent) is stopped: the
    // caller's FP64 defect correction restarts it from the true residual
    if (stalled && stop_on_stall) break;
  }
  *its_out = its;
  if (finite) {
    *finite = true;
    for (int k = 0; k < K; ++k) {
      double pnk, target;
      (void)mixed_state_field(state_h, queued, k, &pnk, &target);
      if (!std::isfinite(pnk)) *finite = false;
    }
  }
  return MPPD_OK;
}

static int msolver_run(mppd_msolver *m, int nadj, int n, const double *eps_h, const mppd_mixed_params *p,
                       int stats[2], std::vector<double> &pn, std::vector<double> &pn0, int *all_conv_out) {
  const Grid &g = m->r->g;
  const int K = m->K;
  for (int i = 0; i < K * 6; ++i) {
    const int f = i / 6;
    m->eps_h[i] = (f >= nadj && f < nadj + n) ? eps_h[(f - nadj) * 6 + i % 6] : 0.0;
  }
  MPP_HIP(hipMemcpyAsync(m->eps_d, m->eps_h, K * 6 * sizeof(double), hipMemcpyHostToDevice, m->st));
  MPP_HIP(hipMemsetAsync(m->U, 0, mixed_vector_len(g, K) * sizeof(double), m->st));
  MPP_HIP(launch_m_setbc(g, K, m->eps_d, m->U, m->st));
  const double *F = nullptr;
  if (nadj > 0) {
    MPP_HIP(launch_m_adj_src(g, K, nadj, m->r->bmat, m->mats_d, m->r->et, m->G, m->F, m->st));
    F = m->F;
  }
  int outer = 0, inner_total = 0, all_conv = 0;
  pn.assign(K, 0.0);
  pn0.assign(K, 0.0);
  for (;;) {
    MPP_HIP(launch_m_resid(g, K, m->r->et, m->ke, m->U, F, m->kinv, m->R, m->X, m->P, m->part, outer == 0,
                           p->rel_tol, p->abs_tol, p->inner_tol, p->max_inner_its, m->state, m->st));
    if (m->coarse) MPP_HIP(launch_c_start(g, K, *m->coarse, m->kinv, m->R, m->P, m->state, m->st));
    MPP_HIP(hipMemcpyAsync(m->state_h, m->state, mixed_state_bytes(), hipMemcpyDeviceToHost, m->st));
    MPP_HIP(hipStreamSynchronize(m->st));
    mixed_outer_view(m->state_h, K, &all_conv, pn.data(), pn0.data());
    if (all_conv || outer >= p->max_outer) break;
    int its = 0, rc = mixed_inner(g, K, m->A, m->kinv, m->R, m->X, m->P, m->AP, m->part, m->state, m->state_h,
                                  p->max_inner_its, m->st, &its, nullptr, false, m->coarse);
    if (rc) return rc;
    inner_total += its;
    MPP_HIP(launch_m_accum(g, K, m->X, m->U, m->st));
    ++outer;
  }
  if (stats) {
    stats[0] = outer;
    stats[1] = inner_total;
  }
  *all_conv_out = all_conv;
  return MPPD_OK;
}

// volume averages of all K fields into m->sig_h
static int msolver_averages(mppd_msolver *m) {
  const Grid &g = m->r->g;
  MPP_HIP(launch_m_ave(g, m->K, m->r->bmat, m->mats_d, m->r->et, m->U, m->part, m->sig_d, m->st));
  MPP_HIP(hipMemcpyAsync(m->sig_h, m->sig_d, m->K * 6 * sizeof(double), hipMemcpyDeviceToHost, m->st));
  MPP_HIP(hipStreamSynchronize(m->st));
  return MPPD_OK;
}

int mppd_msolver_solve(mppd_msolver *m, int n, const double *eps_h, const mppd_mixed_params *p, double *sig_h,
                       int stats[2], double *rel_h) {
  if (!m || !eps_h || !p || !sig_h || n < 1 || n > m->K) return fail(MPPD_ERR_ARG, "bad msolver_solve arguments");
  std::vector<double> pn, pn0;
  int all_conv = 0;
  int rc = msolver_run(m, 0, n, eps_h, p, stats, pn, pn0, &all_conv);
  if (!rc) rc = msolver_averages(m);
  if (rc) return rc;
  for (int i = 0; i < n * 6; ++i) sig_h[i] = m->sig_h[i];
  if (rel_h)
    for (int i = 0; i < n; ++i) rel_h[i] = pn0[i] > 0 ? pn[i] / pn0[i] : 0.0;
  return all_conv ? MPPD_OK : fail(MPPD_ERR_NOTCONV, "mixed solve: tolerance not reached");
}

int mppd_msolver_adjoint(mppd_msolver *m, int n, const double *eps_h, const mppd_mixed_params *p, double *sig_h,
                         double *ctan_h, double *dmat_h, int stats[2], double *rel_h) {
  if (!m || !p || !ctan_h || n < 0 || 6 + n > m->K || (n > 0 && (!eps_h || !sig_h || !dmat_h)))
    return fail(MPPD_ERR_ARG, "bad msolver_adjoint arguments (needs 6 + n <= K)");
  const Grid &g = m->r->g;
  const int K = m->K;
  int rc = MPPD_OK;
  if (!m->F) {
    // all five buffers or none: m->F, tested above, is set last
    double *F = nullptr, *G = nullptr, *apart = nullptr, *adj_d = nullptr, *adj_h = nullptr;
    do {
      if ((rc = dalloc(&F, mixed_vector_len(g, K)))) break;
      if ((rc = dalloc(&G, 3 * 24 * 6))) break;
      if ((rc = dalloc(&apart, mixed_adj_partials(g, K - 6)))) break;
      if ((rc = dalloc(&adj_d, 36 * (size_t)(K - 5)))) break;
      const hipError_t e = hipHostMalloc((void **)&adj_h, 36 * (K - 5) * sizeof(double), hipHostMallocDefault);
      if (e != hipSuccess) rc = hip_fail(e, "hipHostMalloc");
    } while (0);
    if (!rc && alloc_fault_injected()) rc = fail(MPPD_ERR_NOMEM, "injected allocation failure (MPPD_FAIL_ALLOC)");
    if (rc) {
      for (void *q : {(void *)F, (void *)G, (void *)apart, (void *)adj_d}) dfree(q);
      if (adj_h) (void)hipHostFree(adj_h);
      return rc;
    }
    m->G = G;
    m->apart = apart;
    m->adj_d = adj_d;
    m->adj_h = adj_h;
    m->F = F;
  }
  std::vector<double> pn, pn0;
  int all_conv = 0;
  if ((rc = msolver_run(m, 6, n, eps_h, p, stats, pn, pn0, &all_conv))) return rc;
  MPP_HIP(launch_m_adj(g, K, m->r->bmat, m->mats_d, m->r->et, m->ke, m->U, 6, n, m->apart, m->adj_d,
                       m->adj_d + 36, m->st));
  MPP_HIP(hipMemcpyAsync(m->adj_h, m->adj_d, 36 * (n + 1) * sizeof(double), hipMemcpyDeviceToHost, m->st));
  if ((rc = msolver_averages(m))) return rc;  // synchronises the stream
  for (int i = 0; i < 36; ++i) ctan_h[i] = m->adj_h[i];
  for (int i = 0; i < 36 * n; ++i) dmat_h[i] = m->adj_h[36 + i];
  for (int i = 0; i < 6 * n; ++i) sig_h[i] = m->sig_h[36 + i];
  if (rel_h)
    for (int i = 0; i < 6 + n; ++i) rel_h[i] = pn0[i] > 0 ? pn[i] / pn0[i] : 0.0;
  return all_conv ? MPPD_OK : fail(MPPD_ERR_NOTCONV, "adjoint solve: tolerance not reached");
}

int mppd_msolver_get_u(mppd_msolver *m, int k, double *u_h) {
  if (!m || !u_h || k < 0 || k >= m->K) return fail(MPPD_ERR_ARG, "bad msolver_get_u arguments");
  const Grid &g = m->r->g;
  std::vector<double> U(mixed_vector_len(g, m->K));
  MPP_HIP(hipMemcpyAsync(U.data(), m->U, U.size() * sizeof(double), hipMemcpyDeviceToHost, m->st));
  MPP_HIP(hipStreamSynchronize(m->st));
  // component-major: component d * K + k of node n at (d * K + k) * NP + n
  const size_t np = U.size() / (3 * (size_t)m->K);
  for (int n = 0; n < g.nn; ++n)
    for (int d = 0; d < 3; ++d) u_h[(size_t)n * 3 + d] = U[((size_t)d * m->K + k) * np + n];
  return MPPD_OK;
}

// newton_raphson (src/solve.cpp:28-82) of n Gauss points of the msolver's
// (all-elastic) RVE at once.  Field k starts from u_in[k] (zero when the
// pointer is null) with the boundary values of eps_h[k]; the Newton loop of
// every field follows the reference: at the top of an iteration it stops at
// nr_max_its (unconverged) or when ||b||_2 (the FP64 interior residual, the
// norm of assembly_rhs) is below max(nr_max_tol, nr_rel_tol ||b_0||); else it
// solves for the correction to the CG's relative tolerance of the current
// ||D^-1 b|| -- FP64 defect correction around the FP32 inner CG of all live
// fields together -- and adds it.  The averages are FP64 (calc_ave_stress).
int mppd_msolver_newton(mppd_msolver *m, int n, const double *eps_h, const double *const *u_in,
                        double *const *u_out, const mppd_newton_params *p, const mppd_mixed_params *mp,
                        double *sig_h, int *out, int stats[2]) {
  if (!m || !eps_h || !p || !out || n < 1 || n > m->K) return fail(MPPD_ERR_ARG, "bad msolver_newton arguments");
  const Grid &g = m->r->g;
  const int K = m->K;
  const double inner_tol = mp ? mp->inner_tol : 1e-5;
  const int max_inner = mp ? mp->max_inner_its : p->cg_max_its;
  const int max_outer = mp ? mp->max_outer : 10;
  if (max_inner < 1 || max_outer < 1) return fail(MPPD_ERR_ARG, "bad mixed parameters");
  for (int i = 0; i < K * 6; ++i) m->eps_h[i] = (i / 6 < n) ? eps_h[i] : 0.0;
  MPP_HIP(hipMemcpyAsync(m->eps_d, m->eps_h, K * 6 * sizeof(double), hipMemcpyHostToDevice, m->st));
  MPP_HIP(hipMemsetAsync(m->U, 0, mixed_vector_len(g, K) * sizeof(double), m->st));
  for (int k = 0; k < n; ++k)
    if (u_in && u_in[k]) MPP_HIP(launch_m_load_u(g, K, k, u_in[k], m->U, m->st));
  MPP_HIP(launch_m_setbc(g, K, m->eps_d, m->U, m->st));
  std::vector<int> its(K, 0), run(K, 0), conv(K, 0), cost(K, 0), cflag(K, 0);
  std::vector<double> rn(K), rn0(K), norm0(K), pn(K), pn0(K);
  for (int k = 0; k < n; ++k) run[k] = 1;
  int outer_total = 0, inner_total = 0;
  // the residual of the current U for the fields in run, with the outer test
  // and the inner start state (first: targets from this residual)
  auto resid = [&](int first) -> int {
    MPP_HIP(hipStreamSynchronize(m->st));  // act_h may still feed an earlier copy
    for (int k = 0; k < K; ++k) m->act_h[k] = run[k];
    MPP_HIP(hipMemcpyAsync(m->act_d, m->act_h, K * sizeof(int), hipMemcpyHostToDevice, m->st));
    MPP_HIP(launch_m_resid(g, K, m->r->et, m->ke, m->U, nullptr, m->kinv, m->R, m->X, m->P, m->part, first,
                           p->cg_rel_tol, p->cg_abs_tol, inner_tol, max_inner, m->state, m->st, m->act_d));
    if (m->coarse) MPP_HIP(launch_c_start(g, K, *m->coarse, m->kinv, m->R, m->P, m->state, m->st));
    MPP_HIP(hipMemcpyAsync(m->state_h, m->state, mixed_state_bytes(), hipMemcpyDeviceToHost, m->st));
    MPP_HIP(hipStreamSynchronize(m->st));
    mixed_outer_norms(m->state_h, K, cflag.data(), rn.data(), rn0.data());
    return MPPD_OK;
  };
  int rc = resid(1);
  if (rc) return rc;
  for (int k = 0; k < K; ++k) norm0[k] = rn[k];
  bool fresh = true;  // the last residual set the targets of a Newton iteration
  for (;;) {
    // Newton tests of the fields still running (at the top of an iteration)
    bool any = false, changed = false;
    for (int k = 0; k < n; ++k) {
      if (!run[k]) continue;
      if (its[k] >= p->nr_max_its) {
        run[k] = 0;
      } else if (rn[k] < p->nr_max_tol || rn[k] < norm0[k] * p->nr_rel_tol) {
        conv[k] = 1;
        run[k] = 0;
      } else {
        any = true;
        continue;
      }
      changed = true;
    }
    if (!any) break;
    // one Newton iteration of those fields: its linear solve targets the
    // current residual (the last one computed is that of the current U)
    if ((changed || !fresh) && (rc = resid(1))) return rc;
    fresh = false;
    for (int outer = 0;; ++outer) {
      if (outer > 0 && (rc = resid(0))) return rc;
      int all_conv = 0;
      mixed_outer_view(m->state_h, K, &all_conv, pn.data(), pn0.data());
      if (all_conv || outer >= max_outer) break;
      int inner = 0;
      if ((rc = mixed_inner(g, K, m->A, m->kinv, m->R, m->X, m->P, m->AP, m->part, m->state, m->state_h, max_inner,
                            m->st, &inner, nullptr, false, m->coarse)))
        return rc;
      inner_total += inner;
      ++outer_total;
      for (int k = 0; k < n; ++k) cost[k] += run[k] ? inner : 0;
      MPP_HIP(launch_m_accum(g, K, m->X, m->U, m->st));
    }
    // every exit of the loop above follows a residual of the current U, whose
    // norms the next tests read
    for (int k = 0; k < n; ++k) its[k] += run[k];
  }
  if (sig_h) {
    if ((rc = msolver_averages(m))) return rc;
    for (int i = 0; i < n * 6; ++i) sig_h[i] = m->sig_h[i];
  }
  for (int k = 0; k < n; ++k)
    if (u_out && u_out[k]) MPP_HIP(launch_m_store_u(g, K, k, m->U, u_out[k], m->st));
  MPP_HIP(hipStreamSynchronize(m->st));
  for (int k = 0; k < n; ++k) {
    out[3 * k] = its[k];
    out[3 * k + 1] = cost[k];
    out[3 * k + 2] = conv[k];
  }
  if (stats) {
    stats[0] = outer_total;
    stats[1] = inner_total;
  }
  return MPPD_OK;
}

// newton_raphson (src/solve.cpp:28-82) with mixed-precision linear solves.
// Each Newton step assembles the FP64 tangent (assembly_mat: any law), copies
// its interior block to FP32 node blocks, and solves A du = b by FP64 defect
// correction around the FP32 Jacobi-PCG of mixed.hip; the defect-correction
// residual comes from the FP64 ELL matrix and du stays FP64.  The Newton
// residual (assembly_rhs, with the return mapping or damage update of every
// Gauss point), u and the norms stay FP64, and the Newton control flow is the
// reference's.  The linear solve stops on the reference CG's metric,
// ||D^-1 (b - A du)|| < max(cg_abs_tol, cg_rel_tol ||D^-1 b||), in FP64.
// Should the FP32 inner CG break down (a non-finite ||z||: the tangent of a
// softening law can be indefinite), that linear solve is redone by the FP64
// Jacobi-PCG of mppd_newton, so the result is never worse than the FP64 path.
// With A0, the first its_with_A0 Newton steps solve with A0's matrix instead
// of assembling (mppd_newton's rule); its FP32 copy is made once per call.
int mppd_newton_mixed(mppd_solver *s, const mppd_solver *A0, const double eps[6], const double *vars_old_d,
                      double *u_d, const mppd_newton_params *p, const mppd_mixed_params *mp, int out[3],
                      int stats[4]) {
  if (!s || !eps || !u_d || !p || !out) return fail(MPPD_ERR_ARG, "null argument");
  if (int rc = check_A0(s, A0)) return rc;
  // the FP32 copy and the FP64 defect residual take the columns from the grid
  if (A0 && !A0->vals_stencil) return fail(MPPD_ERR_ARG, "A0 holds uploaded values: the mixed solve needs an assembled A0");
  const Grid &g = s->r->g;
  constexpr int K = MixedLin::K;
  const double inner_tol = mp ? mp->inner_tol : 1e-5;
  const int max_inner = mp ? mp->max_inner_its : p->cg_max_its;
  const int max_outer = mp ? mp->max_outer : 10;
  if (max_inner < 1 || max_outer < 1) return fail(MPPD_ERR_ARG, "bad mixed parameters");
  int rc = MPPD_OK;
  if (!s->ml) {
    // built in full before the solver holds it: a failed allocation leaves
    // s->ml null, so the next call allocates again instead of launching on
    // a half-built set of buffers
    MixedLin *ml = new MixedLin();
    const size_t nv = mixed_vector_len(g, K), NP = nv / (3 * K);
    do {
      if ((rc = dalloc(&ml->A, mixed_matrix_floats(g)))) break;
      for (float **v : {&ml->R, &ml->X, &ml->P, &ml->AP})
        if ((rc = dalloc(v, nv))) break;
      if (rc) break;
      if ((rc = dalloc(&ml->kinv, 3 * NP))) break;
      if ((rc = dalloc(&ml->kinv64, 3 * NP))) break;
      if ((rc = dalloc(&ml->part, mixed_partials(g, K)))) break;
      char *st = nullptr;
      if ((rc = dalloc(&st, mixed_state_bytes()))) break;
      ml->state = st;
      hipError_t e = hipHostMalloc(&ml->state_h, mixed_state_bytes(), hipHostMallocDefault);
      if (e != hipSuccess) {
        rc = hip_fail(e, "hipHostMalloc");
        break;
      }
      if (!coarse_enabled()) break;
      // the coarse level from the per-phase element stiffness at u = 0 without history
      double *ke = nullptr;
      if ((rc = dalloc(&ke, 3 * 576, false))) break;
      e = launch_elem_ke(g, s->r->bmat, s->r->mats_d, s->r->et, nullptr, nullptr, 1, ke, s->st);
      if (e != hipSuccess)
        rc = hip_fail(e, "coarse level: element stiffness");
      else
        rc = coarse_setup(g, s->r->et, ke, K, s->st, &ml->coarse);
      (void)hipStreamSynchronize(s->st);
      dfree(ke);
    } while (0);
    if (!rc && alloc_fault_injected()) rc = fail(MPPD_ERR_NOMEM, "injected allocation failure (MPPD_FAIL_ALLOC)");
    if (rc) {
      mixed_lin_free(ml);
      return rc;
    }
    s->ml = ml;
  }
  MixedLin *ml = s->ml;
  int its = 0, solver_its = 0, conv = 0, outers = 0, fallbacks = 0, short_solves = 0;
  double pn[K], pn0[K];
  const mppd_solver *copied = nullptr;  // whose matrix ml->A holds
  MPP_HIP(launch_set_displ_bc(g, eps, u_d, s->st));
  MPP_HIP(rhs_async(s, u_d, vars_old_d, s->b));
  MPP_HIP(read_red(s));
  double norm = s->red_h->v[0];
  const double norm0 = norm;
  while (its < p->nr_max_its) {
    if (norm < p->nr_max_tol || norm < norm0 * p->nr_rel_tol) {
      conv = 1;
      break;
    }
    const mppd_solver *A = s;
    if (!A0 || its > p->its_with_A0 - 1) {
      MPP_HIP(mat_async(s, u_d, vars_old_d));
      MPP_HIP(launch_m_from_ell(g, s->vals, ml->A, ml->kinv64, ml->kinv, s->st));
      copied = nullptr;  // s's matrix changes with every assembly
    } else {
      A = A0;
      if (copied != A0) MPP_HIP(launch_m_from_ell(g, A0->vals, ml->A, ml->kinv64, ml->kinv, s->st));
      copied = A0;
    }
    MPP_HIP(hipMemsetAsync(s->du, 0, sizeof(double) * g.nrow_pad, s->st));
    // test hook: MPPD_FORCE_FP64_FALLBACK=1 counts every FP32 solve as broken
    // down, so each linear solve is redone by the FP64 CG (the fallback path)
    const char *ff = getenv("MPPD_FORCE_FP64_FALLBACK");
    bool broke = ff && atoi(ff) != 0;
    for (int outer = 0; !broke; ++outer) {
      // the last Newton iteration nr_max_its permits solves to a tighter target (kMixedLastStepTighten)
      const double rel = its == p->nr_max_its - 1 ? p->cg_rel_tol * kMixedLastStepTighten : p->cg_rel_tol;
      MPP_HIP(launch_m_resid_ell(g, K, A->cols, A->vals, s->b, s->du, ml->kinv64, ml->kinv, ml->R, ml->X, ml->P,
                                 ml->part, outer == 0, rel, p->cg_abs_tol, inner_tol, max_inner, ml->state, s->st));
      if (ml->coarse) MPP_HIP(launch_c_start(g, K, *ml->coarse, ml->kinv, ml->R, ml->P, ml->state, s->st));
      MPP_HIP(hipMemcpyAsync(ml->state_h, ml->state, mixed_state_bytes(), hipMemcpyDeviceToHost, s->st));
      MPP_HIP(hipStreamSynchronize(s->st));
      int all_conv = 0;
      mixed_outer_view(ml->state_h, K, &all_conv, pn, pn0);
      if (!std::isfinite(pn[0])) {
        broke = true;
        break;
      }
      if (all_conv) break;
      if (outer >= max_outer) {  // short of the tolerance: redone in FP64 below, like a breakdown
        ++short_solves;
        broke = true;
        break;
      }
      int inner = 0;
      bool finite = true;
      if ((rc = mixed_inner(g, K, ml->A, ml->kinv, ml->R, ml->X, ml->P, ml->AP, ml->part, ml->state, ml->state_h,
                            max_inner, s->st, &inner, &finite, true, ml->coarse)))
        return rc;
      solver_its += inner;
      ++outers;
      if (!finite) {
        broke = true;
        break;
      }
      MPP_HIP(launch_m_accum_ell(g, K, ml->X, s->du, s->st));
    }
    // a breakdown, or no convergence within max_outer outer steps: the FP64
    // Jacobi-PCG of mppd_newton on the same system, from 0, so that every
    // linear solve meets the reference CG's tolerance
    if (broke) {
      int cg_its = 0;
      double err = 0;
      if ((rc = cg_run(s, A, s->b, p->cg_abs_tol, p->cg_rel_tol, p->cg_max_its, &cg_its, &err))) return rc;
      solver_its += cg_its;
      ++fallbacks;
      MPP_HIP(launch_axpy(g.nrow, 1.0, s->x, u_d, s->st));
    } else {
      MPP_HIP(launch_axpy(g.nrow, 1.0, s->du, u_d, s->st));
    }
    MPP_HIP(rhs_async(s, u_d, vars_old_d, s->b));
    MPP_HIP(read_red(s));
    norm = s->red_h->v[0];
    its++;
  }
  out[0] = its;
  out[1] = solver_its;
  out[2] = conv;
  if (stats) {
    stats[0] = outers;
    stats[1] = solver_its;
    stats[2] = fallbacks;
    stats[3] = short_solves;
  }
  return MPPD_OK;
}

int mppd_msolver_adjoint_many(int nm, mppd_msolver *const *ms, int n, const double *eps_h, const mppd_mixed_params *p,
                              int workers, double *sig_h, double *ctan_h, double *dmat_h, int *stats, double *rel_h) {
  if (nm < 1 || !ms || !p || !ctan_h || n < 0 || workers < 0 || (n > 0 && (!eps_h || !sig_h || !dmat_h)))
    return fail(MPPD_ERR_ARG, "bad msolver_adjoint_many arguments");
  for (int i = 0; i < nm; ++i) {
    if (!ms[i]) return fail(MPPD_ERR_ARG, "msolver_adjoint_many: null solver " + std::to_string(i));
    if (6 + n > ms[i]->K) return fail(MPPD_ERR_ARG, "msolver_adjoint_many: solver " + std::to_string(i) + " needs 6 + n <= K");
  }
  {  // each solver runs on its own stream from one thread at a time: a solver given twice would race with itself
    std::vector<mppd_msolver *> v(ms, ms + nm);
    std::sort(v.begin(), v.end());
    if (std::adjacent_find(v.begin(), v.end()) != v.end())
      return fail(MPPD_ERR_ARG, "msolver_adjoint_many: a solver appears twice");
  }
  int dev = 0;
  MPP_HIP(hipGetDevice(&dev));
  const int W = std::min(nm, workers ? workers : 4);
  std::atomic<int> next{0};
  std::mutex mu;
  int first_i = nm, first_rc = MPPD_OK;
  std::string first_msg;
  auto work = [&]() {
    if (hipSetDevice(dev) != hipSuccess) {
      std::lock_guard<std::mutex> lk(mu);
      if (first_rc == MPPD_OK) {
        first_rc = MPPD_ERR_HIP;
        first_msg = "hipSetDevice in a worker thread";
        first_i = -1;
      }
      return;
    }
    for (int i; (i = next.fetch_add(1)) < nm;) {
      int st[2] = {0, 0};
      const size_t fo = (size_t)i * n;
      const int rc = mppd_msolver_adjoint(ms[i], n, n ? eps_h + fo * 6 : nullptr, p, n ? sig_h + fo * 6 : nullptr,
                                          ctan_h + (size_t)i * 36, n ? dmat_h + fo * 36 : nullptr, st,
                                          rel_h ? rel_h + (size_t)i * (6 + n) : nullptr);
      if (stats) {
        stats[2 * i] = st[0];
        stats[2 * i + 1] = st[1];
      }
      if (rc) {
        std::lock_guard<std::mutex> lk(mu);
        if (first_rc == MPPD_OK || i < first_i) {
          first_i = i;
          first_rc = rc;
          first_msg = mppd_last_error();  // this thread's message
        }
      }
    }
  };
  std::vector<std::thread> th;
  for (int w = 1; w < W; ++w) th.emplace_back(work);
  work();
  for (auto &t : th) t.join();
  if (first_rc) return fail(first_rc, "msolver_adjoint_many: solver " + std::to_string(first_i) + ": " + first_msg);
  return MPPD_OK;
}

int mppd_msolver_profile(mppd_msolver *m, int iters, double *spmv_ms, double *iter_ms) {
  if (!m || iters <= 0) return fail(MPPD_ERR_ARG, "bad msolver_profile arguments");
  const Grid &g = m->r->g;
  const int K = m->K;
  // a live inner state: unit strains on every field, no inner stopping
  for (int i = 0; i < K * 6; ++i) m->eps_h[i] = (i % 7 == 0) ? 1e-3 : 0.0;
  MPP_HIP(hipMemcpyAsync(m->eps_d, m->eps_h, K * 6 * sizeof(double), hipMemcpyHostToDevice, m->st));
  MPP_HIP(hipMemsetAsync(m->U, 0, mixed_vector_len(g, K) * sizeof(double), m->st));
  MPP_HIP(launch_m_setbc(g, K, m->eps_d, m->U, m->st));
  MPP_HIP(launch_m_resid(g, K, m->r->et, m->ke, m->U, nullptr, m->kinv, m->R, m->X, m->P, m->part, 1, 0.0, 0.0, 0.0,
                         iters + 1, m->state, m->st));
  if (m->coarse) MPP_HIP(launch_c_start(g, K, *m->coarse, m->kinv, m->R, m->P, m->state, m->st));
  hipEvent_t ev[3];
  for (auto &x : ev) MPP_HIP(hipEventCreate(&x));
  hipError_t e = hipEventRecord(ev[0], m->st);
  for (int i = 0; i < iters && e == hipSuccess; ++i)
    e = launch_m_spmv_only(g, K, m->A, m->P, m->AP, m->part, m->state, m->st);
  if (e == hipSuccess) e = hipEventRecord(ev[1], m->st);
  for (int i = 0; i < iters && e == hipSuccess; ++i)
    e = launch_m_iter(g, K, m->A, m->kinv, m->R, m->X, m->P, m->AP, m->part, m->state, i, m->st, m->coarse);
  if (e == hipSuccess) e = hipEventRecord(ev[2], m->st);
  if (e == hipSuccess) e = hipStreamSynchronize(m->st);
  float ms1 = 0.f, ms2 = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms1, ev[0], ev[1]);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms2, ev[1], ev[2]);
  for (auto &x : ev) (void)hipEventDestroy(x);
  if (e != hipSuccess) return hip_fail(e, "msolver_profile");
  if (spmv_ms) *spmv_ms = (double)ms1 / iters;
  if (iter_ms) *iter_ms = (double)ms2 / iters;
  return MPPD_OK;
}

int mppd_msolver_coarse(const mppd_msolver *m, int dims[7], float *ainv_h) {
  if (!m || !dims) return fail(MPPD_ERR_ARG, "bad msolver_coarse arguments");
  const CoarseDev *c = m->coarse;
  const CGrid cg = c ? c->cg : CGrid{};
  const int v[7] = {cg.nc, cg.hx, cg.hy, cg.hz, cg.mx, cg.my, cg.mz};
  for (int i = 0; i < 7; ++i) dims[i] = v[i];
  if (c && ainv_h) {
    MPP_HIP(hipMemcpyAsync(ainv_h, c->ainv, (size_t)cg.nc * cg.nc * sizeof(float), hipMemcpyDeviceToHost, m->st));
    MPP_HIP(hipStreamSynchronize(m->st));
  }
  return MPPD_OK;
}

}  // extern "C"
