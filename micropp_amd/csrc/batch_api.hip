// batch_api.hip -- the mppd_batch_* C ABI: host control flow of a batch of
// RVEs over the batched kernels of batch.hip.
//
// The host keeps, per RVE, the reference's per-GP bookkeeping (strain,
// stress, ctan, cost, flags) and drives the Newton / CG control flow of all
// RVEs together: per Newton step it reads one norm per RVE; per CG chunk it
// reads the CG states.  Everything else stays on the device.  The CG
// iterations of a chunk are replayed from a captured hipGraph.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "batch.hpp"
#include "kernels.hpp"

using namespace mpp;

namespace {

constexpr double kDEpsCtanAve = 1e-8;  // D_EPS_CTAN_AVE (include/params.hpp:32)
// coarse DOFs per RVE of a mixed-precision batch: every inner iteration
// reads the nc^2 FP32 inverse beside the RVE's 324 B/row FP32 matrix, so nc
// is held to max(400, 4 sqrt(rows)), an inverse of at most 64 B per row
// (30^3: 6 coarse elements a side, 375 DOFs; 40^3: 8, 1029; 50^3: 10, 2187)
// columns per split of the coarse solve (kb_c_gemv); more than one split is
// summed by kb_c_gsum
constexpr int kBatchCoarseSplitCols = 512;
int batch_coarse_max_dofs(const Grid &g) { return std::max(400, (int)(4.0 * std::sqrt((double)g.nrow))); }

struct BState {
  double strain[6] = {0}, strain_old[6] = {0}, stress[6] = {0}, ctan[36] = {0};
  int coupling = MPPD_FE_ONE_WAY;
  int nonlinear_mat = 0;  // some phase is not elastic: history arrays exist
  int allocated = 0;      // history in use (gp_t::allocated)
  int converged = 1, subiterated = 0;
  long cost = 0;
  int vals_ready = 0;     // phase mode: vals assembled once
};

enum { M_MAIN = 0, M_KE = 1, M_GATHER = 2, M_A0 = 3, M_TIGHT = 4, M_NUM = 5 };

}  // namespace

struct mppd_batch {
  int n = 0;
  mppd_batch_params prm{};
  hipStream_t st = nullptr;
  std::vector<mppd_rve *> rves;
  std::vector<BRve> h;
  std::vector<BState> s;
  BRve *rv = nullptr;
  // launch shapes
  // launch shapes: rows, elements, element stiffness, then the coarse
  // level's x-restriction, y/z-restriction and coarse solve
  static constexpr int kShapes = 7;
  int *map_d[kShapes] = {}, *blk0_d[kShapes] = {};
  int nb[kShapes] = {};
  int coarse_any = 0;    // mixed precision: some RVE has a coarse level
  float *c32 = nullptr;  // coarse arena: inverses, restriction and correction vectors
  // arenas
  double *vec = nullptr, *vals = nullptr, *ke = nullptr, *be = nullptr, *vars = nullptr, *part = nullptr;
  double *a0 = nullptr;  // initial matrices (use_A0), one per distinct RVE
  float *f32 = nullptr;  // mixed precision: FP32 matrices and inner-CG vectors
  BOuter *outer_d = nullptr, *outer_h = nullptr;
  CgState *cg = nullptr;
  std::map<std::tuple<int, int, int>, int *> cols;
  // per-call device arrays and pinned staging
  int *mask_d[M_NUM] = {};
  int *mask_h[M_NUM] = {};
  int *flags_d = nullptr, *flags_h = nullptr;
  double *eps_d = nullptr, *eps_h = nullptr, *norms_d = nullptr, *norms_h = nullptr, *sig_d = nullptr,
         *sig_h = nullptr;
  CgState *cg_h = nullptr;
  // graph of gchunk CG iterations
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int gchunk = 0;
  // mixed precision: graph of gchunk inner (FP32) CG iterations; counters
  hipGraph_t graph32 = nullptr;
  hipGraphExec_t gexec32 = nullptr;
  long m_outer = 0, m_fallbacks = 0, m_short = 0;
  // device-side Newton control of the FP64 path (newton_batch): per-RVE
  // state (device, pinned), the per-RVE mask of RVEs whose correction was
  // applied in the last finishing pass, and the graphs of a Newton
  // iteration's head (tangent + CG start) and finishing pass
  BNewton *nst_d = nullptr, *nst_h = nullptr;
  int *pend_d = nullptr;
  hipGraph_t graph_head = nullptr, graph_tail = nullptr;
  hipGraphExec_t gexec_head = nullptr, gexec_tail = nullptr;
  int cg_est = 0;  // CG iterations queued before the first read of a Newton iteration
  long syncs = 0, syncs_last = 0, newton_its_last = 0;
  // test hook (MPPD_FORCE_FP64_FALLBACK=1 at mppd_batch_create): every FP32
  // inner solve counts as broken down, so each linear solve is redone by the
  // FP64 CG from du = 0, which exercises the fallback path
  bool force_fallback = false;
  // test hook (MPPD_BATCH_HOST_NEWTON=1 at mppd_batch_create): the FP64 path
  // runs the host-driven Newton loop (newton_batch_host) instead of the
  // device-side one, for comparison
  bool host_newton = false;
  int spmv_shape = kSpmvDefault;    // tuned SpMV shape (launch_b_spmv)
  int spmv_shape32 = kSpmvDefault;  // tuned inner FP32 SpMV shape (launch_bm_spmv)
  double cg_queued = 0, graph_launches = 0, last_ms = 0;

  BLaunch launch(int mask) const {
    BLaunch L;
    L.nrve = n;
    L.rv = rv;
    L.act = mask_d[mask];
    L.a0on = mask_d[M_A0];
    L.row = {nb[0], map_d[0], blk0_d[0]};
    L.elem = {nb[1], map_d[1], blk0_d[1]};
    L.ke = {nb[2], map_d[2], blk0_d[2]};
    L.crx = {nb[3], map_d[3], blk0_d[3]};
    L.cryz = {nb[4], map_d[4], blk0_d[4]};
    L.cgemv = {nb[5], map_d[5], blk0_d[5]};
    L.cgsum = {nb[6], map_d[6], blk0_d[6]};
    L.coarse = coarse_any;
    L.spmv_shape = spmv_shape;
    L.spmv_shape32 = spmv_shape32;
    return L;
  }
};

namespace {

int sync(mppd_batch *B) {
  MPP_HIP(hipStreamSynchronize(B->st));
  ++B->syncs;
  return MPPD_OK;
}

// mask upload: the pinned staging buffer is rewritten only after the stream
// has drained (earlier copies from it have completed)
int set_mask(mppd_batch *B, int k, const std::vector<int> &on) {
  int rc = sync(B);
  if (rc) return rc;
  for (int r = 0; r < B->n; ++r) B->mask_h[k][r] = on[r];
  MPP_HIP(hipMemcpyAsync(B->mask_d[k], B->mask_h[k], sizeof(int) * B->n, hipMemcpyHostToDevice, B->st));
  return MPPD_OK;
}

int upload_table(mppd_batch *B) {
  int rc = sync(B);
  if (rc) return rc;
  MPP_HIP(hipMemcpy(B->rv, B->h.data(), sizeof(BRve) * B->n, hipMemcpyHostToDevice));
  return MPPD_OK;
}

// eps of every RVE: row r of e (6 per RVE)
int upload_eps(mppd_batch *B, const std::vector<double> &e) {
  int rc = sync(B);
  if (rc) return rc;
  std::memcpy(B->eps_h, e.data(), sizeof(double) * 6 * B->n);
  MPP_HIP(hipMemcpyAsync(B->eps_d, B->eps_h, sizeof(double) * 6 * B->n, hipMemcpyHostToDevice, B->st));
  return MPPD_OK;
}

int read_norms(mppd_batch *B) {
  MPP_HIP(hipMemcpyAsync(B->norms_h, B->norms_d, sizeof(double) * B->n, hipMemcpyDeviceToHost, B->st));
  return sync(B);
}

int read_sig(mppd_batch *B) {
  MPP_HIP(hipMemcpyAsync(B->sig_h, B->sig_d, sizeof(double) * 6 * B->n, hipMemcpyDeviceToHost, B->st));
  return sync(B);
}

int read_cg(mppd_batch *B) {
  MPP_HIP(hipMemcpyAsync(B->cg_h, B->cg, sizeof(CgState) * 2 * B->n, hipMemcpyDeviceToHost, B->st));
  return sync(B);
}

// g iterations queued on the stream (parities 0 .. g-1)
hipError_t queue_iters(mppd_batch *B, int g) {
  const BLaunch L = B->launch(M_MAIN);
  for (int i = 0; i < g; ++i) {
    hipError_t e = launch_b_cg_iter(L, i, B->st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// the inner FP32 CG iterations of the mixed-precision solve, gchunk of them
int capture_graph32(mppd_batch *B) {
  int rc = sync(B);
  if (rc) return rc;
  const BLaunch L = B->launch(M_MAIN);
  MPP_HIP(hipStreamBeginCapture(B->st, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  for (int i = 0; i < B->gchunk && e == hipSuccess; ++i) e = launch_bm_iter(L, i, B->st);
  hipGraph_t graph = nullptr;
  const hipError_t e2 = hipStreamEndCapture(B->st, &graph);
  if (e != hipSuccess) return hip_fail(e, "batch graph capture (mixed)");
  if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture");
  B->graph32 = graph;
  MPP_HIP(hipGraphInstantiate(&B->gexec32, graph, nullptr, nullptr, 0));
  return MPPD_OK;
}

int capture_graph(mppd_batch *B, int g) {
  if (B->gexec) (void)hipGraphExecDestroy(B->gexec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  B->gexec = nullptr;
  B->graph = nullptr;
  int rc = sync(B);
  if (rc) return rc;
  MPP_HIP(hipStreamBeginCapture(B->st, hipStreamCaptureModeThreadLocal));
  hipError_t e = queue_iters(B, g);
  hipGraph_t graph = nullptr;
  const hipError_t e2 = hipStreamEndCapture(B->st, &graph);
  if (e != hipSuccess) return hip_fail(e, "batch graph capture");
  if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture");
  B->graph = graph;
  MPP_HIP(hipGraphInstantiate(&B->gexec, graph, nullptr, nullptr, 0));
  B->gchunk = g;
  return MPPD_OK;
}

// ell_solve_cgpd (src/ell.cpp:66-122) of every RVE in `on`, queued in chunks
// of whole graphs sized from the slowest RVE's observed convergence rate.
int cg_batch(mppd_batch *B, const std::vector<int> &on, std::vector<int> &its) {
  const mppd_newton_params &p = B->prm.newton;
  int rc = set_mask(B, M_MAIN, on);
  if (rc) return rc;
  MPP_HIP(launch_b_cg_init(B->launch(M_MAIN), p.cg_abs_tol, p.cg_rel_tol, p.cg_max_its, B->st));
  if ((rc = read_cg(B))) return rc;
  const int n = B->n, G = B->gchunk;
  std::vector<double> pn_prev(n), target(n);
  std::vector<int> its_prev(n, 0);
  bool all = true;
  for (int r = 0; r < n; ++r) {
    const CgState &c = B->cg_h[2 * r];
    pn_prev[r] = c.pn;
    target[r] = std::max(c.min_err, c.pn0 * c.rel_err);
    all &= c.done != 0;
  }
  int queued = 0, chunk = 2 * G;
  while (!all && queued < p.cg_max_its) {
    const int reps = std::max(1, (std::min(chunk, p.cg_max_its - queued) + G - 1) / G);
    for (int i = 0; i < reps; ++i) MPP_HIP(hipGraphLaunch(B->gexec, B->st));
    queued += reps * G;
    B->cg_queued += reps * G;
    B->graph_launches += reps;
    if ((rc = read_cg(B))) return rc;
    all = true;
    int next = 0;
    for (int r = 0; r < n; ++r) {
      const CgState &c = B->cg_h[2 * r + (queued & 1)];
      if (c.done) continue;
      all = false;
      int need = 2 * chunk;
      const int dit = c.its - its_prev[r];
      if (dit > 0 && c.pn > 0 && pn_prev[r] > 0 && c.pn < pn_prev[r]) {
        const double rate = std::log(c.pn / pn_prev[r]) / dit;
        const double est = std::log(target[r] / c.pn) / rate;
        if (std::isfinite(est)) need = (int)std::ceil(est) + 2;
      }
      next = std::max(next, need);
      pn_prev[r] = c.pn;
      its_prev[r] = c.its;
    }
    chunk = std::max(G, std::min(256, next));
  }
  for (int r = 0; r < n; ++r) its[r] = on[r] ? B->cg_h[2 * r + (queued & 1)].its : 0;
  return MPPD_OK;
}

// The linear solves of cg_batch in mixed precision, as mppd_newton_mixed
// does them for one RVE: FP64 defect correction du += e around an FP32
// Jacobi-PCG on an FP32 copy of the matrix, stopping on the reference CG's
// metric ||D^-1 (b - A du)|| < max(cg_abs_tol, cg_rel_tol ||D^-1 b||) in
// FP64.  An inner CG that gains less than 10% over >= 32 iterations is
// stopped (its RVE's next outer step restarts from the true residual); an RVE
// whose inner CG breaks down (non-finite) is solved again by the FP64 CG.
// its[r] = inner iterations (+ FP64 fallback iterations).  RVEs set in
// `tight` solve to cg_rel_tol * kMixedLastStepTighten (the last Newton
// iteration nr_max_its permits).  Leaves the main mask = on and du in x.
int cg_batch_mixed(mppd_batch *B, const std::vector<int> &on, const std::vector<int> &tight, std::vector<int> &its) {
  const mppd_newton_params &p = B->prm.newton;
  constexpr double kInnerTol = 1e-5;
  constexpr int kMaxOuter = 10;
  const int n = B->n, G = B->gchunk, max_inner = p.cg_max_its;
  for (int r = 0; r < n; ++r) its[r] = 0;
  int rc = set_mask(B, M_MAIN, on);
  if (rc) return rc;
  if ((rc = set_mask(B, M_TIGHT, tight))) return rc;
  const BLaunch L = B->launch(M_MAIN);
  MPP_HIP(launch_bm_prep(L, B->st));
  std::vector<int> live(on), broke(n, 0);
  for (int outer = 0;; ++outer) {
    if ((rc = set_mask(B, M_MAIN, live))) return rc;
    MPP_HIP(launch_bm_resid(L, outer == 0, p.cg_rel_tol, p.cg_abs_tol, kInnerTol, max_inner, B->mask_d[M_TIGHT],
                            kMixedLastStepTighten, B->outer_d, B->st));
    MPP_HIP(hipMemcpyAsync(B->outer_h, B->outer_d, sizeof(BOuter) * n, hipMemcpyDeviceToHost, B->st));
    if ((rc = read_cg(B))) return rc;
    bool any = false;
    for (int r = 0; r < n; ++r) {
      if (!live[r]) continue;
      if (!std::isfinite(B->outer_h[r].pn)) {
        live[r] = 0;
        broke[r] = 1;
      } else if (B->outer_h[r].conv) {
        live[r] = 0;
      } else {
        any = true;
      }
    }
    if (B->force_fallback)
      for (int r = 0; r < n; ++r)
        if (live[r]) {
          live[r] = 0;
          broke[r] = 1;
          any = false;
        }
    if (!any) break;
    if (outer >= kMaxOuter) {  // short of the tolerance: redone by the FP64 CG below, like a breakdown
      for (int r = 0; r < n; ++r)
        if (live[r]) {
          ++B->m_short;
          live[r] = 0;
          broke[r] = 1;
        }
      break;
    }
    ++B->m_outer;
    // the inner CGs (the outer kernel started those of the other RVEs done),
    // in graphs of gchunk iterations, as many between reads as the slowest
    // live RVE's observed rate needs to reach its inner target (at most
    // kMaxChunk, which bounds what a stalled RVE runs before it is stopped);
    // each live RVE's ||z|| is checked at every read for breakdown and stall
    constexpr int kMaxChunk = 64;
    std::vector<double> pn_mark(n, 0.0), pn_prev(n, 0.0);
    std::vector<int> its_mark(n, 0), its_prev(n, 0);
    for (int r = 0; r < n; ++r) pn_mark[r] = pn_prev[r] = B->cg_h[2 * r].pn;
    int queued = 0, chunk = 2 * G;
    bool all = false;
    while (!all && queued < max_inner) {
      const int reps = std::max(1, (std::min(chunk, max_inner - queued) + G - 1) / G);
      for (int i = 0; i < reps; ++i) MPP_HIP(hipGraphLaunch(B->gexec32, B->st));
      queued += reps * G;
      B->cg_queued += reps * G;
      B->graph_launches += reps;
      if ((rc = read_cg(B))) return rc;
      all = true;
      int next = 0;
      for (int r = 0; r < n; ++r) {
        CgState &c = B->cg_h[2 * r + (queued & 1)];
        if (!live[r] || c.done) continue;
        bool stop = false;
        if (!std::isfinite(c.pn)) {
          stop = true;
          broke[r] = 1;
        } else if (c.its - its_mark[r] >= 32) {
          if (c.pn > 0.9 * pn_mark[r]) stop = true;  // stalled: restart from the true residual
          its_mark[r] = c.its;
          pn_mark[r] = c.pn;
        }
        if (stop) {
          c.done = 1;
          MPP_HIP(hipMemcpyAsync(B->cg + 2 * r + (queued & 1), &c, sizeof(CgState), hipMemcpyHostToDevice, B->st));
          continue;
        }
        all = false;
        int need = 2 * chunk;
        const int dit = c.its - its_prev[r];
        if (dit > 0 && c.pn > 0 && pn_prev[r] > 0 && c.pn < pn_prev[r]) {
          const double rate = std::log(c.pn / pn_prev[r]) / dit;
          const double est = std::log(c.pn0 * c.rel_err / c.pn) / rate;
          if (std::isfinite(est)) need = (int)std::ceil(est) + 2;
        }
        next = std::max(next, need);
        pn_prev[r] = c.pn;
        its_prev[r] = c.its;
      }
      chunk = std::max(G, std::min(kMaxChunk, next));
    }
    if ((rc = sync(B))) return rc;
    for (int r = 0; r < n; ++r) {
      if (!live[r]) continue;
      its[r] += B->cg_h[2 * r + (queued & 1)].its;
      if (broke[r]) live[r] = 0;
    }
    if ((rc = set_mask(B, M_MAIN, live))) return rc;
    MPP_HIP(launch_bm_accum(L, B->st));
  }
  bool any_broke = false;
  for (int r = 0; r < n; ++r) any_broke |= broke[r] != 0;
  if (any_broke) {  // the FP64 Jacobi-PCG on the same systems, from du = 0
    std::vector<int> fb(n, 0);
    if ((rc = cg_batch(B, broke, fb))) return rc;
    for (int r = 0; r < n; ++r)
      if (broke[r]) {
        its[r] += fb[r];
        ++B->m_fallbacks;
      }
  }
  return set_mask(B, M_MAIN, on);
}

// The tangent of Newton iteration its[r] of the RVEs in `on`: element
// stiffness + vals (phase-mode RVEs: once), or, with use_A0, the initial
// matrix for the first its_with_A0 iterations (src/solve.cpp:50-66).
int mat_batch(mppd_batch *B, const std::vector<int> &on, const std::vector<int> &its) {
  std::vector<int> ke(B->n, 0), ga(B->n, 0), a0(B->n, 0);
  for (int r = 0; r < B->n; ++r) {
    if (!on[r]) continue;
    if (B->a0 && its[r] < B->prm.newton.its_with_A0) {
      a0[r] = 1;
      continue;
    }
    const bool phase = B->h[r].phase_mode != 0;
    ke[r] = !phase;
    ga[r] = !phase || !B->s[r].vals_ready;
    if (phase) B->s[r].vals_ready = 1;
  }
  int rc = set_mask(B, M_KE, ke);
  if (!rc) rc = set_mask(B, M_GATHER, ga);
  if (!rc && B->a0) rc = set_mask(B, M_A0, a0);
  if (rc) return rc;
  MPP_HIP(launch_b_mat(B->launch(M_MAIN), B->mask_d[M_KE], B->mask_d[M_GATHER], B->st));
  return MPPD_OK;
}

// newton_raphson (src/solve.cpp:28-82) of every RVE in `on`, on uw with the
// strains in eps_d, the host deciding between kernels (the mixed-precision
// path, whose linear solves steer from the host): its / solver_its /
// converged per RVE.
int newton_batch_host(mppd_batch *B, const std::vector<int> &on, std::vector<int> &its, std::vector<int> &sits,
                      std::vector<int> &conv) {
  const mppd_newton_params &p = B->prm.newton;
  const int n = B->n;
  int rc = set_mask(B, M_MAIN, on);
  if (rc) return rc;
  const BLaunch L = B->launch(M_MAIN);
  MPP_HIP(launch_b_set_bc(L, B->eps_d, B->st));
  MPP_HIP(launch_b_rhs(L, B->norms_d, B->st));
  if ((rc = read_norms(B))) return rc;
  std::vector<double> norm0(B->norms_h, B->norms_h + n);
  std::vector<int> run(on), cg_its(n, 0);
  for (int r = 0; r < n; ++r) its[r] = sits[r] = conv[r] = 0;
  for (;;) {
    bool any = false;
    for (int r = 0; r < n; ++r) {
      if (!run[r]) continue;
      // the reference tests the residual only at the top of an iteration
      // it is about to run: at the iteration cap it stops unconverged, even
      // when the last residual passes (src/solve.cpp:43-78)
      const double norm = B->norms_h[r];
      if (its[r] >= p.nr_max_its) {
        run[r] = 0;
      } else if (norm < p.nr_max_tol || norm < norm0[r] * p.nr_rel_tol) {
        conv[r] = 1;
        run[r] = 0;
      } else {
        any = true;
      }
    }
    if (!any) break;
    ++B->newton_its_last;
    if ((rc = mat_batch(B, run, its))) return rc;
    if (B->prm.mixed_precision) {
      // the last Newton iteration nr_max_its permits solves to a tighter target: if the Newton stops
      // there unconverged, this correction is the last term of the state it returns
      std::vector<int> tight(n, 0);
      for (int r = 0; r < n; ++r) tight[r] = run[r] && its[r] == p.nr_max_its - 1;
      rc = cg_batch_mixed(B, run, tight, cg_its);
    } else {
      rc = cg_batch(B, run, cg_its);
    }
    if (rc) return rc;  // leaves the main mask = run
    for (int r = 0; r < n; ++r) sits[r] += cg_its[r];
    MPP_HIP(launch_b_axpy(L, B->st));
    MPP_HIP(launch_b_rhs(L, B->norms_d, B->st));
    if ((rc = read_norms(B))) return rc;
    for (int r = 0; r < n; ++r) its[r] += run[r];
  }
  return MPPD_OK;
}

BNewtonTest newton_test(const mppd_batch *B) {
  const mppd_newton_params &p = B->prm.newton;
  return BNewtonTest{p.nr_max_its, p.nr_max_tol, p.nr_rel_tol};
}

// The head of a Newton iteration (masks of the running RVEs, their tangent,
// the start of their CGs) and its finishing pass (correction, residual, norm,
// the loop's test), each captured once as a graph: the device decides per
// RVE, so a batched Newton iteration has no host decision between kernels.
int capture_newton_graphs(mppd_batch *B) {
  int rc = sync(B);
  if (rc) return rc;
  const mppd_newton_params &p = B->prm.newton;
  const BLaunch L = B->launch(M_MAIN);
  for (int k = 0; k < 2; ++k) {
    MPP_HIP(hipStreamBeginCapture(B->st, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    if (k == 0) {
      e = launch_b_nt_prepare(L, B->nst_d, B->a0 != nullptr, p.its_with_A0, B->mask_d[M_MAIN], B->mask_d[M_KE],
                              B->mask_d[M_GATHER], B->mask_d[M_A0], B->pend_d, B->st);
      if (e == hipSuccess) e = launch_b_mat(L, B->mask_d[M_KE], B->mask_d[M_GATHER], B->st);
      if (e == hipSuccess) e = launch_b_cg_init(L, p.cg_abs_tol, p.cg_rel_tol, p.cg_max_its, B->st);
    } else {
      e = launch_b_nt_finish(L, newton_test(B), B->nst_d, B->pend_d, B->norms_d, B->st);
    }
    hipGraph_t graph = nullptr;
    const hipError_t e2 = hipStreamEndCapture(B->st, &graph);
    if (e != hipSuccess) return hip_fail(e, "Newton graph capture");
    if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture");
    (k == 0 ? B->graph_head : B->graph_tail) = graph;
    MPP_HIP(hipGraphInstantiate(k == 0 ? &B->gexec_head : &B->gexec_tail, graph, nullptr, nullptr, 0));
  }
  return MPPD_OK;
}

// the per-RVE Newton states and the CG state pairs in one read
int read_newton(mppd_batch *B) {
  MPP_HIP(hipMemcpyAsync(B->nst_h, B->nst_d, sizeof(BNewton) * B->n, hipMemcpyDeviceToHost, B->st));
  MPP_HIP(hipMemcpyAsync(B->cg_h, B->cg, sizeof(CgState) * 2 * B->n, hipMemcpyDeviceToHost, B->st));
  return sync(B);
}

// newton_raphson (src/solve.cpp:28-82) of every RVE in `on`, on uw with the
// strains in eps_d: its / solver_its / converged per RVE.  FP64: the loop's
// decisions run on the device (BNewton), and each Newton iteration is the
// head graph, chunk graphs of the CG (as many as the last iteration's
// slowest CG needed), the finishing graph and ONE read of the states; only
// RVEs whose CG has not stopped by then get more chunks and another
// finishing pass.  Each RVE runs the same kernels in the same order as
// newton_batch_host, so its results are the same bit for bit.
int newton_batch(mppd_batch *B, const std::vector<int> &on, std::vector<int> &its, std::vector<int> &sits,
                 std::vector<int> &conv) {
  if (B->prm.mixed_precision || B->host_newton) return newton_batch_host(B, on, its, sits, conv);
  const mppd_newton_params &p = B->prm.newton;
  const int n = B->n, G = B->gchunk;
  int rc = set_mask(B, M_MAIN, on);
  if (rc) return rc;
  const BLaunch L = B->launch(M_MAIN);
  MPP_HIP(launch_b_set_bc(L, B->eps_d, B->st));
  MPP_HIP(launch_b_rhs(L, B->norms_d, B->st));
  MPP_HIP(launch_b_nt_start(L, B->mask_d[M_MAIN], B->norms_d, newton_test(B), B->nst_d, B->st));
  if ((rc = read_newton(B))) return rc;
  std::vector<double> pn_prev(n), target(n);
  std::vector<int> its_prev(n);
  for (;;) {
    bool any = false;
    for (int r = 0; r < n; ++r) any |= B->nst_h[r].run != 0;
    if (!any) break;
    MPP_HIP(hipGraphLaunch(B->gexec_head, B->st));
    ++B->newton_its_last;
    int queued = 0, chunk = std::max(G, B->cg_est);
    for (;;) {
      const int reps = std::max(1, (std::min(chunk, p.cg_max_its - queued) + G - 1) / G);
      for (int i = 0; i < reps; ++i) MPP_HIP(hipGraphLaunch(B->gexec, B->st));
      queued += reps * G;
      B->cg_queued += reps * G;
      B->graph_launches += reps;
      MPP_HIP(hipGraphLaunch(B->gexec_tail, B->st));
      if ((rc = read_newton(B))) return rc;
      // RVEs still in their CG (state cg[0]: queued is a multiple of the even G)
      int next = 0;
      bool pending = false;
      for (int r = 0; r < n; ++r) {
        const BNewton &s = B->nst_h[r];
        if (!s.run || s.applied) continue;
        pending = true;
        const CgState &c = B->cg_h[2 * r];
        int need = 2 * chunk;
        const int dit = c.its - its_prev[r];
        if (queued > reps * G && dit > 0 && c.pn > 0 && pn_prev[r] > 0 && c.pn < pn_prev[r]) {
          const double rate = std::log(c.pn / pn_prev[r]) / dit;
          const double est = std::log(std::max(c.min_err, c.pn0 * c.rel_err) / c.pn) / rate;
          if (std::isfinite(est)) need = (int)std::ceil(est) + 2;
        }
        next = std::max(next, need);
        pn_prev[r] = c.pn;
        its_prev[r] = c.its;
      }
      if (!pending || queued >= p.cg_max_its) break;
      chunk = std::max(G, std::min(256, next));
    }
    // the next iteration starts with as many as this one's slowest CG took
    int most = 0;
    for (int r = 0; r < n; ++r)
      if (on[r]) most = std::max(most, B->cg_h[2 * r].its);
    B->cg_est = most + G;
  }
  for (int r = 0; r < n; ++r) {
    const BNewton &s = B->nst_h[r];
    its[r] = on[r] ? s.its : 0;
    sits[r] = on[r] ? s.sits : 0;
    conv[r] = on[r] ? s.conv : 0;
  }
  return MPPD_OK;
}

int move(mppd_batch *B, const std::vector<int> &on, int which) {
  int rc = set_mask(B, M_MAIN, on);
  if (rc) return rc;
  MPP_HIP(launch_b_move(B->launch(M_MAIN), which, B->st));
  return MPPD_OK;
}

int ave_batch(mppd_batch *B, const std::vector<int> &on, int which) {
  int rc = set_mask(B, M_MAIN, on);
  if (rc) return rc;
  MPP_HIP(launch_b_ave(B->launch(M_MAIN), which, B->sig_d, B->st));
  return read_sig(B);
}

std::vector<double> strains_of(const mppd_batch *B) {
  std::vector<double> e(6 * B->n);
  for (int r = 0; r < B->n; ++r)
    for (int v = 0; v < 6; ++v) e[r * 6 + v] = B->s[r].strain[v];
  return e;
}

// calc_ctan_lin_fe_models (src/micropp.cpp:255-284) for every RVE
int ctan_lin_batch(mppd_batch *B) {
  const int n = B->n;
  std::vector<int> all(n, 1), its(n), sits(n), conv(n);
  for (int i = 0; i < 6; ++i) {
    int rc = move(B, all, 3);
    if (rc) return rc;
    std::vector<double> e(6 * n, 0.0);
    for (int r = 0; r < n; ++r) e[r * 6 + i] += kDEpsCtanAve;
    if ((rc = upload_eps(B, e))) return rc;
    if ((rc = newton_batch(B, all, its, sits, conv))) return rc;
    if ((rc = ave_batch(B, all, 0))) return rc;
    for (int r = 0; r < n; ++r)
      for (int v = 0; v < 6; ++v) B->s[r].ctan[v * 6 + i] = B->sig_h[r * 6 + v] / kDEpsCtanAve;
  }
  return MPPD_OK;
}

// CG time per iteration for a graph of g iterations (a converged CG state,
// so every kernel runs its early exit: this measures launch overhead), ms
int time_chunk(mppd_batch *B, int g, double *ms_per_it) {
  int rc = capture_graph(B, g);
  if (rc) return rc;
  hipEvent_t e0, e1;
  MPP_HIP(hipEventCreate(&e0));
  MPP_HIP(hipEventCreate(&e1));
  const int reps = std::max(1, 64 / g);
  hipError_t e = hipGraphLaunch(B->gexec, B->st);  // warm
  if (e == hipSuccess) e = hipEventRecord(e0, B->st);
  for (int i = 0; i < reps && e == hipSuccess; ++i) e = hipGraphLaunch(B->gexec, B->st);
  if (e == hipSuccess) e = hipEventRecord(e1, B->st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return hip_fail(e, "batch graph timing");
  *ms_per_it = ms / (reps * g);
  return MPPD_OK;
}

// SpMV shape: every candidate (slots in flight x block order, restricted by
// the parameters) timed on this batch's own matrices with every RVE's CG
// state live; the fastest is kept.  The shapes differ only in load
// scheduling and block order, so the results do not depend on the choice.
// fp32: the same for the inner FP32 SpMV of the mixed-precision solves.
int tune_spmv(mppd_batch *B, bool fp32 = false) {
  const int chunks[3] = {9, 27, 81};
  std::vector<int> cand;
  for (int x = 0; x < 2; ++x)
    for (int c = 0; c < 3; ++c)
      if ((B->prm.spmv_chunk == 0 || B->prm.spmv_chunk == chunks[c]) &&
          (B->prm.block_order < 0 || B->prm.block_order == x))
        cand.push_back(3 * x + c);
  int &pick = fp32 ? B->spmv_shape32 : B->spmv_shape;
  if (cand.size() == 1) {
    pick = cand[0];
    return MPPD_OK;
  }
  int rc = sync(B);
  if (rc) return rc;
  std::vector<CgState> live(2 * (size_t)B->n);
  for (auto &c : live) {
    c = CgState{};
    c.max_its = 1;
  }
  MPP_HIP(hipMemcpy(B->cg, live.data(), sizeof(CgState) * live.size(), hipMemcpyHostToDevice));
  hipEvent_t e0, e1;
  MPP_HIP(hipEventCreate(&e0));
  MPP_HIP(hipEventCreate(&e1));
  const BLaunch L = B->launch(M_MAIN);
  float best = 0.f;
  hipError_t e = hipSuccess;
  for (int shape : cand) {
    constexpr int kReps = 10;
    auto spmv = [&]() { return fp32 ? launch_bm_spmv(L, shape, 0, B->st) : launch_b_spmv(L, shape, 0, B->st); };
    for (int i = 0; i < 2 && e == hipSuccess; ++i) e = spmv();  // warm
    if (e == hipSuccess) e = hipEventRecord(e0, B->st);
    for (int i = 0; i < kReps && e == hipSuccess; ++i) e = spmv();
    if (e == hipSuccess) e = hipEventRecord(e1, B->st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) break;
    if (shape == cand[0] || ms < best) {
      best = ms;
      pick = shape;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return hip_fail(e, "batch SpMV tuning");
  // every CG state (both parities) back to stopped
  for (auto &c : live) c.done = 1;
  MPP_HIP(hipMemcpyAsync(B->cg, live.data(), sizeof(CgState) * live.size(), hipMemcpyHostToDevice, B->st));
  return sync(B);
}

// Graph chunk length: the shortest whose per-iteration launch cost is within
// 10% of the longest candidate's (short chunks waste fewer iterations past
// convergence, long ones amortise the graph launch).
int autotune(mppd_batch *B) {
  const int cand[] = {2, 4, 8, 16, 32};
  double t[5];
  for (int i = 0; i < 5; ++i) {
    int rc = time_chunk(B, cand[i], &t[i]);
    if (rc) return rc;
  }
  int pick = 32;
  for (int i = 0; i < 5; ++i)
    if (t[i] <= 1.1 * t[4]) {
      pick = cand[i];
      break;
    }
  return capture_graph(B, pick);
}

void destroy(mppd_batch *B) {
  if (B->st) (void)hipStreamSynchronize(B->st);
  if (B->gexec) (void)hipGraphExecDestroy(B->gexec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  if (B->gexec32) (void)hipGraphExecDestroy(B->gexec32);
  if (B->graph32) (void)hipGraphDestroy(B->graph32);
  for (hipGraphExec_t x : {B->gexec_head, B->gexec_tail})
    if (x) (void)hipGraphExecDestroy(x);
  for (hipGraph_t x : {B->graph_head, B->graph_tail})
    if (x) (void)hipGraphDestroy(x);
  for (int k = 0; k < mppd_batch::kShapes; ++k) {
    dfree(B->map_d[k]);
    dfree(B->blk0_d[k]);
  }
  for (void *p : {(void *)B->rv, (void *)B->vec, (void *)B->vals, (void *)B->a0, (void *)B->f32, (void *)B->c32,
                  (void *)B->outer_d, (void *)B->ke, (void *)B->be, (void *)B->vars, (void *)B->part, (void *)B->cg,
                  (void *)B->flags_d, (void *)B->eps_d, (void *)B->norms_d, (void *)B->sig_d, (void *)B->nst_d,
                  (void *)B->pend_d})
    dfree(p);
  for (auto &kv : B->cols) dfree(kv.second);
  for (int k = 0; k < M_NUM; ++k) {
    dfree(B->mask_d[k]);
    if (B->mask_h[k]) (void)hipHostFree(B->mask_h[k]);
  }
  for (void *p : {(void *)B->flags_h, (void *)B->eps_h, (void *)B->norms_h, (void *)B->sig_h, (void *)B->cg_h,
                  (void *)B->outer_h, (void *)B->nst_h})
    if (p) (void)hipHostFree(p);
  for (mppd_rve *r : B->rves) mppd_rve_destroy(r);
  if (B->st) (void)hipStreamDestroy(B->st);
  delete B;
}

template <typename T>
int pinned(T **p, size_t n) {
  MPP_HIP(hipHostMalloc((void **)p, sizeof(T) * std::max<size_t>(n, 1), hipHostMallocDefault));
  std::memset(*p, 0, sizeof(T) * std::max<size_t>(n, 1));
  return MPPD_OK;
}

// RVE descriptors that make the same RVE (the coupling aside)
bool same_rve(const mppd_batch_rve &a, const mppd_batch_rve &b) {
  for (int i = 0; i < 3; ++i)
    if (a.size[i] != b.size[i]) return false;
  if (a.micro_type != b.micro_type) return false;
  for (int i = 0; i < 4; ++i)
    if (a.geo_params[i] != b.geo_params[i]) return false;
  for (int i = 0; i < 3; ++i) {
    const material_base &x = a.materials[i], &y = b.materials[i];
    if (x.E != y.E || x.nu != y.nu || x.Ka != y.Ka || x.Sy != y.Sy || x.k != y.k || x.mu != y.mu ||
        x.lambda != y.lambda || x.Xt != y.Xt || x.type != y.type)
      return false;
  }
  return true;
}

// use_A0: the initial matrix of each distinct RVE (assembly_mat at u = 0
// without history, src/micropp.cpp:128-143), assembled in the slot of the
// first RVE that has it and kept; equal RVEs share it.
int setup_a0(mppd_batch *B, const mppd_batch_rve *in) {
  const int n = B->n;
  std::vector<int> owner(n), first;
  for (int r = 0; r < n; ++r) {
    owner[r] = r;
    for (int q : first)
      if (same_rve(in[q], in[r])) {
        owner[r] = q;
        break;
      }
    if (owner[r] == r) first.push_back(r);
  }
  std::vector<size_t> off(n, 0);
  size_t tot = 0;
  for (int q : first) {
    off[q] = tot;
    tot += (size_t)B->h[q].g.nrow_pad * kSlots;
  }
  int rc = dalloc(&B->a0, tot, false);
  if (rc) return rc;
  std::vector<int> ke(n, 0), ga(n, 0);
  for (int q : first) {
    ke[q] = !B->h[q].phase_mode;  // phase-mode RVEs: per-phase stiffness already made
    ga[q] = 1;
  }
  if ((rc = set_mask(B, M_KE, ke))) return rc;
  if ((rc = set_mask(B, M_GATHER, ga))) return rc;
  MPP_HIP(launch_b_mat(B->launch(M_MAIN), B->mask_d[M_KE], B->mask_d[M_GATHER], B->st));
  for (int q : first)
    MPP_HIP(hipMemcpyAsync(B->a0 + off[q], B->h[q].vals, sizeof(double) * B->h[q].g.nrow_pad * kSlots,
                           hipMemcpyDeviceToDevice, B->st));
  for (int r = 0; r < n; ++r) B->h[r].a0 = B->a0 + off[owner[r]];
  return upload_table(B);
}

// The coarse levels of a mixed-precision batch: per RVE with a coarse grid,
// the FP32 inverse of the Galerkin coarse matrix of its initial tangent
// (per-phase element stiffness at u = 0 without history: the elastic one),
// shared by equal RVEs, and its work vectors.  An RVE whose coarse matrix is
// not positive definite keeps Jacobi alone (cgrid.nc = 0).
int setup_coarse(mppd_batch *B, const mppd_batch_rve *in, size_t c32) {
  const int n = B->n;
  int rc = dalloc(&B->c32, c32, false);
  if (rc) return rc;
  double *ke = nullptr;
  if ((rc = dalloc(&ke, 3 * 576, false))) return rc;
  std::vector<int> owner(n, -1);
  size_t off = 0;
  for (int r = 0; r < n && !rc; ++r) {
    BRve &R = B->h[r];
    const Grid &g = R.g;
    const CGrid cg = R.cgrid;
    if (!cg.nc) continue;
    for (int q = 0; q < r && owner[r] < 0; ++q)
      if (owner[q] == q && same_rve(in[q], in[r])) owner[r] = q;
    if (owner[r] >= 0) {
      R.cainv = B->h[owner[r]].cainv;
    } else {
      float *ainv = B->c32 + off;
      off += (size_t)cg.nc * cg.nc;
      const mppd_rve *v = B->rves[r];
      bool ok = false;
      hipError_t e = launch_elem_ke(g, v->bmat, v->mats_d, v->et, nullptr, nullptr, 1, ke, B->st);
      if (e != hipSuccess) {
        rc = hip_fail(e, "coarse level: element stiffness");
        break;
      }
      if ((rc = coarse_inverse(g, cg, v->et, ke, ainv, B->st, &ok))) break;
      if (!ok) {
        R.cgrid = CGrid{};
        continue;
      }
      owner[r] = r;
      R.cainv = ainv;
    }
    R.ct1 = B->c32 + off;
    off += (size_t)3 * g.nz * g.ny * cg.mx;
    R.crc = B->c32 + off;
    R.cec = B->c32 + off + cg.nc;
    off += 2 * (size_t)cg.nc;
    if (R.csplit > 1) {
      R.cecp = B->c32 + off;
      off += (size_t)R.csplit * cg.nc;
    }
  }
  (void)hipStreamSynchronize(B->st);
  dfree(ke);
  return rc;
}

int setup(mppd_batch *B, const mppd_batch_rve *in) {
  const int n = B->n;
  int rc = MPPD_OK;
  MPP_HIP(hipStreamCreateWithFlags(&B->st, hipStreamNonBlocking));
  B->rves.assign(n, nullptr);
  B->h.assign(n, BRve{});
  B->s.assign(n, BState{});
  for (int r = 0; r < n; ++r) {
    const mppd_batch_rve &d = in[r];
    if (d.coupling != MPPD_FE_LINEAR && d.coupling != MPPD_FE_ONE_WAY && d.coupling != MPPD_FE_FULL)
      return fail(MPPD_ERR_ARG, "batch coupling must be FE_LINEAR, FE_ONE_WAY or FE_FULL");
    if ((rc = mppd_rve_create(&B->rves[r], d.size, d.micro_type, d.geo_params, d.materials))) return rc;
    B->s[r].coupling = d.coupling;
    B->s[r].nonlinear_mat = !B->rves[r]->all_elastic;
  }
  // sizes, block tables and arena offsets; with mixed precision, each RVE's
  // coarse grid (a small one: its inverse is read by every inner iteration)
  constexpr int kS = mppd_batch::kShapes;
  std::vector<int> map[kS], blk0[kS];
  size_t rows = 0, slots = 0, kes = 0, bes = 0, nvars = 0, parts = 0, c32 = 0;
  for (int k = 0; k < kS; ++k) blk0[k].assign(n, 0);
  const bool coarse = B->prm.mixed_precision && coarse_enabled();
  for (int r = 0; r < n; ++r) {
    const mppd_rve *v = B->rves[r];
    const Grid &g = v->g;
    BRve &R = B->h[r];
    R.g = g;
    R.phase_mode = v->all_elastic;
    R.nb_row = g.nrow_pad / 256;
    R.nb_elem = (g.nelem + 255) / 256;
    const int nb_ke = ((R.phase_mode ? 3 : g.nelem) + kKeWaves - 1) / kKeWaves;
    R.cgrid = CGrid{};
    if (coarse && !coarse_grid(g, batch_coarse_max_dofs(g), &R.cgrid)) R.cgrid = CGrid{};
    const CGrid &cg = R.cgrid;
    const int nrx = cg.nc ? (3 * (g.ny - 2) * (g.nz - 2) * cg.mx + 255) / 256 : 0;
    R.csplit = (cg.nc + kBatchCoarseSplitCols - 1) / kBatchCoarseSplitCols;
    const int ngsum = R.csplit > 1 ? (cg.nc + 255) / 256 : 0;
    const int counts[kS] = {R.nb_row, R.nb_elem, nb_ke, nrx, (cg.nc + 3) / 4, (cg.nc + 63) / 64 * R.csplit, ngsum};
    if (cg.nc) {
      B->coarse_any = 1;
      c32 += (size_t)cg.nc * cg.nc + (size_t)3 * g.nz * g.ny * cg.mx + 2 * (size_t)cg.nc;
      if (R.csplit > 1) c32 += (size_t)R.csplit * cg.nc;
    }
    for (int k = 0; k < kS; ++k) {
      blk0[k][r] = (int)map[k].size();
      map[k].insert(map[k].end(), counts[k], r);
    }
    rows += g.nrow_pad;
    slots += (size_t)g.nrow_pad * kSlots;
    kes += (size_t)(R.phase_mode ? 3 : g.nelem) * 576;
    bes += (size_t)g.nelem * 24;
    if (!v->all_elastic) nvars += (size_t)g.nelem * 8 * NVAR;
    parts += std::max<size_t>(3 * (size_t)R.nb_row, 6 * (size_t)R.nb_elem);
  }
  for (int k = 0; k < kS; ++k) {
    B->nb[k] = (int)map[k].size();
    if ((rc = dalloc(&B->map_d[k], map[k].size(), false))) return rc;
    if ((rc = dalloc(&B->blk0_d[k], n, false))) return rc;
    MPP_HIP(hipMemcpy(B->map_d[k], map[k].data(), sizeof(int) * map[k].size(), hipMemcpyHostToDevice));
    MPP_HIP(hipMemcpy(B->blk0_d[k], blk0[k].data(), sizeof(int) * n, hipMemcpyHostToDevice));
  }
  constexpr int kRowVecs = 10;  // uw u_n u_k x r z p Ap kinv b
  if ((rc = dalloc(&B->vec, rows * kRowVecs))) return rc;
  if ((rc = dalloc(&B->vals, slots))) return rc;
  if ((rc = dalloc(&B->ke, kes, false))) return rc;
  if ((rc = dalloc(&B->be, bes, false))) return rc;
  if ((rc = dalloc(&B->vars, nvars * 3))) return rc;
  if ((rc = dalloc(&B->part, parts))) return rc;
  if ((rc = dalloc(&B->cg, 2 * (size_t)n))) return rc;
  if ((rc = dalloc(&B->rv, n))) return rc;
  size_t o_row = 0, o_slot = 0, o_ke = 0, o_be = 0, o_var = 0, o_part = 0;
  for (int r = 0; r < n; ++r) {
    const mppd_rve *v = B->rves[r];
    const Grid &g = v->g;
    BRve &R = B->h[r];
    R.bmat = v->bmat;
    R.mats = v->mats_d;
    R.et = v->et;
    auto key = std::make_tuple(g.nx, g.ny, g.nz);
    if (!B->cols.count(key)) {
      int *c = nullptr;
      if ((rc = dalloc(&c, (size_t)g.nrow_pad * kSlots, false))) return rc;
      B->cols[key] = c;
      MPP_HIP(launch_ell_cols(g, c, B->st));
    }
    R.cols = B->cols[key];
    R.vals = B->vals + o_slot;
    R.ke = B->ke + o_ke;
    R.be = B->be + o_be;
    double **vecs[kRowVecs] = {&R.uw, &R.u_n, &R.u_k, &R.x, &R.r, &R.z, &R.p, &R.Ap, &R.kinv, &R.b};
    for (int q = 0; q < kRowVecs; ++q) *vecs[q] = B->vec + rows * q + o_row;
    if (!v->all_elastic) {
      const size_t nv = (size_t)g.nelem * 8 * NVAR;
      R.vars_n = B->vars + o_var;
      R.vars_k = B->vars + nvars + o_var;
      R.vars_aux = B->vars + 2 * nvars + o_var;
      o_var += nv;
    }
    R.part = B->part + o_part;
    R.cg = B->cg + 2 * r;
    R.use_vars = 0;
    o_row += g.nrow_pad;
    o_slot += (size_t)g.nrow_pad * kSlots;
    o_ke += (size_t)(R.phase_mode ? 3 : g.nelem) * 576;
    o_be += (size_t)g.nelem * 24;
    o_part += std::max<size_t>(3 * (size_t)R.nb_row, 6 * (size_t)R.nb_elem);
  }
  if (B->prm.mixed_precision) {
    constexpr int kVecs32 = 6;  // x32 r32 z32 p32 Ap32 kinv32
    if ((rc = dalloc(&B->f32, slots + rows * kVecs32))) return rc;
    size_t os = 0, orow = 0;
    for (int r = 0; r < n; ++r) {
      BRve &R = B->h[r];
      R.vals32 = B->f32 + os;
      float **v[kVecs32] = {&R.x32, &R.r32, &R.z32, &R.p32, &R.Ap32, &R.kinv32};
      for (int q = 0; q < kVecs32; ++q) *v[q] = B->f32 + slots + rows * q + orow;
      os += (size_t)R.g.nrow_pad * kSlots;
      orow += R.g.nrow_pad;
    }
    if ((rc = dalloc(&B->outer_d, n))) return rc;
    if ((rc = pinned(&B->outer_h, n))) return rc;
  }
  if (B->coarse_any && (rc = setup_coarse(B, in, c32))) return rc;
  for (int k = 0; k < M_NUM; ++k) {
    if ((rc = dalloc(&B->mask_d[k], n))) return rc;
    if ((rc = pinned(&B->mask_h[k], n))) return rc;
  }
  if ((rc = dalloc(&B->flags_d, n))) return rc;
  if ((rc = dalloc(&B->eps_d, 6 * (size_t)n))) return rc;
  if ((rc = dalloc(&B->norms_d, n))) return rc;
  if ((rc = dalloc(&B->sig_d, 6 * (size_t)n))) return rc;
  if ((rc = pinned(&B->flags_h, n))) return rc;
  if ((rc = pinned(&B->eps_h, 6 * (size_t)n))) return rc;
  if ((rc = pinned(&B->norms_h, n))) return rc;
  if ((rc = pinned(&B->sig_h, 6 * (size_t)n))) return rc;
  if ((rc = pinned(&B->cg_h, 2 * (size_t)n))) return rc;
  if ((rc = dalloc(&B->nst_d, n))) return rc;
  if ((rc = pinned(&B->nst_h, n))) return rc;
  if ((rc = dalloc(&B->pend_d, n))) return rc;
  if ((rc = upload_table(B))) return rc;
  // element stiffness per phase of the all-elastic RVEs, once
  std::vector<int> ph(n);
  for (int r = 0; r < n; ++r) ph[r] = B->h[r].phase_mode;
  if ((rc = set_mask(B, M_KE, ph))) return rc;
  if ((rc = set_mask(B, M_GATHER, std::vector<int>(n, 0)))) return rc;
  MPP_HIP(launch_b_mat(B->launch(M_MAIN), B->mask_d[M_KE], B->mask_d[M_GATHER], B->st));
  if (B->prm.use_A0 && (rc = setup_a0(B, in))) return rc;
  // the SpMV shape, then the CG graph (its kernels read the table and masks
  // through fixed pointers)
  if ((rc = tune_spmv(B))) return rc;
  if ((rc = set_mask(B, M_MAIN, std::vector<int>(n, 0)))) return rc;
  MPP_HIP(launch_b_cg_init(B->launch(M_MAIN), 0.0, 0.0, 0, B->st));  // every CG done: timing runs exit early
  // an even chunk length: after whole chunks the CG state is state[0]
  if (B->prm.graph_chunk > 0)
    rc = capture_graph(B, (B->prm.graph_chunk + 1) / 2 * 2);
  else
    rc = autotune(B);
  if (rc) return rc;
  if ((rc = capture_newton_graphs(B))) return rc;
  B->cg_est = 2 * B->gchunk;
  if (B->prm.mixed_precision && ((rc = tune_spmv(B, true)) || (rc = capture_graph32(B)))) return rc;
  if ((rc = sync(B))) return rc;
  // micropp constructor: ctan of every GP from the linear solves
  bool any_fe = false;
  for (int r = 0; r < n; ++r) any_fe |= B->s[r].coupling != MPPD_FE_LINEAR;
  if (B->prm.calc_ctan_lin && any_fe) rc = ctan_lin_batch(B);
  return rc;
}

}  // namespace

extern "C" {

void mppd_batch_params_default(mppd_batch_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  // micropp_params_t defaults (include/types.hpp:43-83)
  p->newton.nr_max_its = 4;
  p->newton.nr_max_tol = 1.0e-10;
  p->newton.nr_rel_tol = 1.0e-3;
  p->newton.cg_max_its = 1000;
  p->newton.cg_abs_tol = 1.0e-50;
  p->newton.cg_rel_tol = 1.0e-5;
  p->newton.its_with_A0 = 1;
  p->lin_stress = 1;
  p->calc_ctan_lin = 1;
  p->subiterations = 0;
  p->nsubiterations = 10;
  p->graph_chunk = 0;
  p->spmv_chunk = 0;
  p->block_order = -1;
  p->use_A0 = 0;
  p->mixed_precision = 0;
}

int mppd_batch_create(mppd_batch **out, int nrve, const mppd_batch_rve *rves, const mppd_batch_params *p) {
  if (!out) return fail(MPPD_ERR_ARG, "null argument");
  *out = nullptr;
  if (nrve < 1 || !rves || !p) return fail(MPPD_ERR_ARG, "bad batch arguments");
  if (p->graph_chunk < 0 || p->graph_chunk > 256 || (p->graph_chunk & 1))
    return fail(MPPD_ERR_ARG, "graph_chunk must be 0 or an even number <= 256");
  if (p->spmv_chunk != 0 && p->spmv_chunk != 9 && p->spmv_chunk != 27 && p->spmv_chunk != 81)
    return fail(MPPD_ERR_ARG, "spmv_chunk must be 0, 9, 27 or 81");
  if (p->block_order < -1 || p->block_order > 1) return fail(MPPD_ERR_ARG, "block_order must be -1, 0 or 1");
  if (p->subiterations && p->nsubiterations < 1) return fail(MPPD_ERR_ARG, "nsubiterations must be >= 1");
  mppd_batch *B = new mppd_batch();
  B->n = nrve;
  B->prm = *p;
  const char *ff = getenv("MPPD_FORCE_FP64_FALLBACK");
  B->force_fallback = ff && atoi(ff) != 0;
  const char *hn = getenv("MPPD_BATCH_HOST_NEWTON");
  B->host_newton = hn && atoi(hn) != 0;
  const int rc = setup(B, rves);
  if (rc) {
    destroy(B);
    return rc;
  }
  *out = B;
  return MPPD_OK;
}

int mppd_batch_destroy(mppd_batch *b) {
  if (b) destroy(b);
  return MPPD_OK;
}

int mppd_batch_set_strains(mppd_batch *b, const double *eps) {
  if (!b || !eps) return fail(MPPD_ERR_ARG, "null argument");
  for (int r = 0; r < b->n; ++r) std::memcpy(b->s[r].strain, eps + 6 * r, 6 * sizeof(double));
  return MPPD_OK;
}

// homogenize (src/homogenize.cpp:68-99): FE_LINEAR -> ctan.eps; FE_ONE_WAY /
// FE_FULL -> homogenize_fe_one_way / homogenize_fe_full (:111-282) for all
// such RVEs together.
int mppd_batch_homogenize(mppd_batch *B) {
  if (!B) return fail(MPPD_ERR_ARG, "null batch");
  const auto t0 = std::chrono::steady_clock::now();
  const long syncs0 = B->syncs;
  B->newton_its_last = 0;
  const int n = B->n;
  int rc = MPPD_OK;
  std::vector<int> fe(n, 0), its(n), sits(n), conv(n);
  bool any = false;
  for (int r = 0; r < n; ++r) {
    BState &q = B->s[r];
    if (q.coupling == MPPD_FE_LINEAR) {  // homogenize_linear
      for (int i = 0; i < 6; ++i) {
        q.stress[i] = 0.0;
        for (int j = 0; j < 6; ++j) q.stress[i] += q.ctan[i * 6 + j] * q.strain[j];
      }
      continue;
    }
    fe[r] = 1;
    any = true;
  }
  if (!any) return MPPD_OK;
  // history in use as of this call
  bool tab = false;
  for (int r = 0; r < n; ++r)
    if (B->h[r].use_vars != B->s[r].allocated) {
      B->h[r].use_vars = B->s[r].allocated;
      tab = true;
    }
  if (tab && (rc = upload_table(B))) return rc;
  // u = u_n; newton(strain, vars_n); u_k = u
  if ((rc = move(B, fe, 0))) return rc;
  if ((rc = upload_eps(B, strains_of(B)))) return rc;
  if ((rc = newton_batch(B, fe, its, sits, conv))) return rc;
  if ((rc = move(B, fe, 1))) return rc;
  std::vector<int> last_sits(sits);
  for (int r = 0; r < n; ++r)
    if (fe[r]) {
      B->s[r].cost = sits[r];
      B->s[r].converged = conv[r];
      B->s[r].subiterated = 0;
    }
  // subiterations of the GPs whose Newton did not converge
  if (B->prm.subiterations) {
    std::vector<int> sub(n, 0);
    bool anysub = false;
    for (int r = 0; r < n; ++r)
      if (fe[r] && !conv[r]) sub[r] = 1, anysub = true;
    if (anysub) {
      if ((rc = move(B, sub, 0))) return rc;
      std::vector<double> e(6 * n, 0.0), de(6 * n, 0.0);
      for (int r = 0; r < n; ++r)
        if (sub[r]) {
          B->s[r].subiterated = 1;
          for (int i = 0; i < 6; ++i) {
            e[r * 6 + i] = B->s[r].strain_old[i];
            de[r * 6 + i] = (B->s[r].strain[i] - B->s[r].strain_old[i]) / B->prm.nsubiterations;
          }
        }
      for (int k = 0; k < B->prm.nsubiterations; ++k) {
        for (int q = 0; q < 6 * n; ++q) e[q] += de[q];
        if ((rc = upload_eps(B, e))) return rc;
        if ((rc = newton_batch(B, sub, its, sits, conv))) return rc;
        for (int r = 0; r < n; ++r)
          if (sub[r]) {
            B->s[r].cost += sits[r];
            last_sits[r] = sits[r];
          }
      }
      for (int r = 0; r < n; ++r)
        if (sub[r]) B->s[r].converged = conv[r];
      if ((rc = move(B, sub, 1))) return rc;
    }
  }
  // sigma
  if (B->prm.lin_stress) {
    for (int r = 0; r < n; ++r) {
      if (!fe[r]) continue;
      BState &q = B->s[r];
      for (int i = 0; i < 6; ++i) {
        q.stress[i] = 0.0;
        for (int j = 0; j < 6; ++j) q.stress[i] += q.ctan[i * 6 + j] * q.strain[j];
      }
    }
  } else {
    if ((rc = ave_batch(B, fe, 1))) return rc;
    for (int r = 0; r < n; ++r)
      if (fe[r]) std::memcpy(B->s[r].stress, B->sig_h + 6 * r, 6 * sizeof(double));
  }
  // calc_vars_new; history allocated on the first nonlinearity
  std::vector<int> nl_on(n, 0);
  bool anynl = false;
  for (int r = 0; r < n; ++r)
    if (fe[r] && B->s[r].nonlinear_mat) nl_on[r] = 1, anynl = true;
  if (anynl) {
    // a GP without history evolutes into a zeroed scratch (vars_aux)
    std::vector<int> fresh(n, 0);
    for (int r = 0; r < n; ++r) fresh[r] = nl_on[r] && !B->s[r].allocated;
    if ((rc = set_mask(B, M_MAIN, fresh))) return rc;
    MPP_HIP(launch_b_vars_move(B->launch(M_MAIN), 1, B->st));
    if ((rc = set_mask(B, M_MAIN, nl_on))) return rc;
    MPP_HIP(launch_b_vars_new(B->launch(M_MAIN), B->flags_d, B->st));
    MPP_HIP(hipMemcpyAsync(B->flags_h, B->flags_d, sizeof(int) * n, hipMemcpyDeviceToHost, B->st));
    if ((rc = sync(B))) return rc;
    std::vector<int> alloc(n, 0);
    bool anyalloc = false;
    for (int r = 0; r < n; ++r)
      if (nl_on[r] && B->flags_h[r] && !B->s[r].allocated) {
        alloc[r] = 1;
        anyalloc = true;
        B->s[r].allocated = 1;
        B->h[r].use_vars = 1;
      }
    if (anyalloc) {
      if ((rc = set_mask(B, M_MAIN, alloc))) return rc;
      MPP_HIP(launch_b_vars_move(B->launch(M_MAIN), 0, B->st));
      if ((rc = upload_table(B))) return rc;
    }
  }
  // FE_FULL: ctan from six perturbed solves reusing u (GPs with history only)
  std::vector<int> full(n, 0);
  bool anyfull = false;
  for (int r = 0; r < n; ++r)
    if (fe[r] && B->s[r].coupling == MPPD_FE_FULL && B->s[r].allocated) full[r] = 1, anyfull = true;
  if (anyfull) {
    if ((rc = move(B, full, 2))) return rc;
    std::vector<double> sig0(6 * n);
    for (int r = 0; r < n; ++r) std::memcpy(&sig0[6 * r], B->s[r].stress, 6 * sizeof(double));
    for (int i = 0; i < 6; ++i) {
      std::vector<double> e = strains_of(B);
      for (int r = 0; r < n; ++r) e[r * 6 + i] += kDEpsCtanAve;
      if ((rc = upload_eps(B, e))) return rc;
      if ((rc = newton_batch(B, full, its, sits, conv))) return rc;
      if ((rc = ave_batch(B, full, 0))) return rc;
      for (int r = 0; r < n; ++r) {
        if (!full[r]) continue;
        B->s[r].cost += last_sits[r];  // the reference adds the sigma solve's count again
        for (int v = 0; v < 6; ++v) B->s[r].ctan[v * 6 + i] = (B->sig_h[r * 6 + v] - sig0[r * 6 + v]) / kDEpsCtanAve;
      }
    }
  }
  if ((rc = sync(B))) return rc;
  B->last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  B->syncs_last = B->syncs - syncs0;
  return MPPD_OK;
}

int mppd_batch_get_stress(const mppd_batch *b, double *sig) {
  if (!b || !sig) return fail(MPPD_ERR_ARG, "null argument");
  for (int r = 0; r < b->n; ++r) std::memcpy(sig + 6 * r, b->s[r].stress, 6 * sizeof(double));
  return MPPD_OK;
}

int mppd_batch_get_ctan(const mppd_batch *b, double *ctan) {
  if (!b || !ctan) return fail(MPPD_ERR_ARG, "null argument");
  for (int r = 0; r < b->n; ++r) std::memcpy(ctan + 36 * r, b->s[r].ctan, 36 * sizeof(double));
  return MPPD_OK;
}

int mppd_batch_get_info(const mppd_batch *b, int *info) {
  if (!b || !info) return fail(MPPD_ERR_ARG, "null argument");
  for (int r = 0; r < b->n; ++r) {
    info[4 * r + 0] = b->s[r].allocated;
    info[4 * r + 1] = (int)b->s[r].cost;
    info[4 * r + 2] = b->s[r].converged;
    info[4 * r + 3] = b->s[r].subiterated;
  }
  return MPPD_OK;
}

// gp_t::update_vars (include/gp.hpp:95-105) of every RVE
int mppd_batch_update_vars(mppd_batch *b) {
  if (!b) return fail(MPPD_ERR_ARG, "null batch");
  for (int r = 0; r < b->n; ++r) {
    std::swap(b->h[r].vars_n, b->h[r].vars_k);
    std::swap(b->h[r].u_n, b->h[r].u_k);
    std::memcpy(b->s[r].strain_old, b->s[r].strain, 6 * sizeof(double));
  }
  return upload_table(b);
}

int mppd_batch_stats(const mppd_batch *b, double stats[6]) {
  if (!b || !stats) return fail(MPPD_ERR_ARG, "null argument");
  stats[0] = b->cg_queued;
  stats[1] = b->gchunk;
  stats[2] = b->graph_launches;
  stats[3] = b->last_ms;
  stats[4] = (double)b->syncs_last;
  stats[5] = (double)b->newton_its_last;
  return MPPD_OK;
}

int mppd_batch_spmv_profile(mppd_batch *b, int fp32, int iters, double *ms, double *bytes) {
  if (!b || iters < 1 || !ms) return fail(MPPD_ERR_ARG, "bad batch SpMV profile arguments");
  if (fp32 && !b->prm.mixed_precision) return fail(MPPD_ERR_ARG, "the batch has no FP32 matrices");
  int rc = sync(b);
  if (rc) return rc;
  // every CG live (one iteration each), as during a solve; stopped again afterwards
  std::vector<CgState> live(2 * (size_t)b->n);
  for (auto &c : live) {
    c = CgState{};
    c.max_its = 1;
  }
  MPP_HIP(hipMemcpy(b->cg, live.data(), sizeof(CgState) * live.size(), hipMemcpyHostToDevice));
  const BLaunch L = b->launch(M_MAIN);
  const int shape = fp32 ? b->spmv_shape32 : b->spmv_shape;
  auto spmv = [&]() { return fp32 ? launch_bm_spmv(L, shape, 0, b->st) : launch_b_spmv(L, shape, 0, b->st); };
  hipEvent_t e0, e1;
  MPP_HIP(hipEventCreate(&e0));
  MPP_HIP(hipEventCreate(&e1));
  hipError_t e = spmv();
  if (e == hipSuccess) e = hipEventRecord(e0, b->st);
  for (int i = 0; i < iters && e == hipSuccess; ++i) e = spmv();
  if (e == hipSuccess) e = hipEventRecord(e1, b->st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return hip_fail(e, "batch SpMV profile");
  for (auto &c : live) c.done = 1;
  MPP_HIP(hipMemcpy(b->cg, live.data(), sizeof(CgState) * live.size(), hipMemcpyHostToDevice));
  *ms = (double)t / iters;
  if (bytes) {
    // per row: 81 values (columns computed from the grid) + p in + Ap out, at the matrix precision
    const double w = fp32 ? 4.0 : 8.0;
    double rows = 0;
    for (int r = 0; r < b->n; ++r) rows += b->h[r].g.nrow_pad;
    *bytes = rows * (81 * w + 2 * w);
  }
  return MPPD_OK;
}

int mppd_batch_mixed_stats(const mppd_batch *b, long out[3]) {
  if (!b || !out) return fail(MPPD_ERR_ARG, "null argument");
  out[0] = b->m_outer;
  out[1] = b->m_fallbacks;
  out[2] = b->m_short;
  return MPPD_OK;
}

int mppd_batch_tuning(const mppd_batch *b, int out[3]) {
  if (!b || !out) return fail(MPPD_ERR_ARG, "null argument");
  const int chunks[3] = {9, 27, 81};
  out[0] = b->gchunk;
  out[1] = chunks[b->spmv_shape % 3];
  out[2] = b->spmv_shape / 3;
  return MPPD_OK;
}

// The arena shares setup() gives one RVE, plus its mppd_rve tables.
int mppd_batch_slot_bytes(const mppd_rve *r, size_t *bytes) {
  if (!r || !bytes) return fail(MPPD_ERR_ARG, "null argument");
  const Grid &g = r->g;
  const size_t nb_row = g.nrow_pad / 256, nb_elem = (g.nelem + 255) / 256;
  size_t b = (size_t)g.nrow_pad * (kSlots + 10) * sizeof(double);
  b += (size_t)(r->all_elastic ? 3 : g.nelem) * 576 * sizeof(double);
  b += (size_t)g.nelem * (24 * sizeof(double) + sizeof(int));
  if (!r->all_elastic) b += (size_t)g.nelem * 8 * NVAR * 3 * sizeof(double);
  b += std::max(3 * nb_row, 6 * nb_elem) * sizeof(double) + 2 * sizeof(CgState) + 8 * 144 * sizeof(double);
  *bytes = b;
  return MPPD_OK;
}

int mppd_batch_load(mppd_batch *B, int r, const mppd_batch_gp *gp, const double *u_n_d, const double *vars_n_d) {
  if (!B || !gp || r < 0 || r >= B->n) return fail(MPPD_ERR_ARG, "bad batch slot");
  if (gp->coupling != MPPD_FE_LINEAR && gp->coupling != MPPD_FE_ONE_WAY && gp->coupling != MPPD_FE_FULL)
    return fail(MPPD_ERR_ARG, "batch coupling must be FE_LINEAR, FE_ONE_WAY or FE_FULL");
  BState &q = B->s[r];
  const BRve &R = B->h[r];
  if (gp->allocated && (!vars_n_d || !q.nonlinear_mat))
    return fail(MPPD_ERR_ARG, "a GP with history needs vars_n and a slot with a nonlinear phase");
  std::memcpy(q.strain, gp->strain, sizeof(q.strain));
  std::memcpy(q.strain_old, gp->strain_old, sizeof(q.strain_old));
  std::memcpy(q.stress, gp->stress, sizeof(q.stress));
  std::memcpy(q.ctan, gp->ctan, sizeof(q.ctan));
  q.coupling = gp->coupling;
  q.allocated = gp->allocated != 0;
  q.cost = gp->cost;
  q.converged = gp->converged;
  q.subiterated = gp->subiterated;
  const size_t nrow = (size_t)R.g.nrow;
  if (u_n_d)
    MPP_HIP(hipMemcpyAsync(R.u_n, u_n_d, sizeof(double) * nrow, hipMemcpyDeviceToDevice, B->st));
  else
    MPP_HIP(hipMemsetAsync(R.u_n, 0, sizeof(double) * nrow, B->st));
  if (q.allocated)
    MPP_HIP(hipMemcpyAsync(R.vars_n, vars_n_d, sizeof(double) * R.g.nelem * 8 * NVAR, hipMemcpyDeviceToDevice,
                           B->st));
  return MPPD_OK;
}

int mppd_batch_store(mppd_batch *B, int r, mppd_batch_gp *gp, double *u_k_d, double *vars_k_d) {
  if (!B || !gp || r < 0 || r >= B->n) return fail(MPPD_ERR_ARG, "bad batch slot");
  const BState &q = B->s[r];
  const BRve &R = B->h[r];
  if (q.allocated && !vars_k_d) return fail(MPPD_ERR_ARG, "the slot holds history: vars_k needed");
  std::memcpy(gp->strain, q.strain, sizeof(q.strain));
  std::memcpy(gp->strain_old, q.strain_old, sizeof(q.strain_old));
  std::memcpy(gp->stress, q.stress, sizeof(q.stress));
  std::memcpy(gp->ctan, q.ctan, sizeof(q.ctan));
  gp->coupling = q.coupling;
  gp->allocated = q.allocated;
  gp->cost = (int)q.cost;
  gp->converged = q.converged;
  gp->subiterated = q.subiterated;
  if (u_k_d)
    MPP_HIP(hipMemcpyAsync(u_k_d, R.u_k, sizeof(double) * R.g.nrow, hipMemcpyDeviceToDevice, B->st));
  if (q.allocated)
    MPP_HIP(hipMemcpyAsync(vars_k_d, R.vars_k, sizeof(double) * R.g.nelem * 8 * NVAR, hipMemcpyDeviceToDevice,
                           B->st));
  return sync(B);
}

}  // extern "C"
