// micropp.cpp -- micropp<3> on the GPU: the reference's public API
// (include/micropp.hpp:176-217) driving the mppd_* kernels.
//
// The per-GP state (u_n, u_k, vars_n, vars_k) lives in device memory; every
// finite-element operation (boundary values, assembly, CG, averages,
// history) runs in the gfx950 kernels.  Host code keeps the reference's
// control flow: the Newton loop, subiterations, the FE_FULL perturbation loop,
// cost bookkeeping, and the 6x6 products of homogenize_linear.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <system_error>
#include <thread>

#include "micropp.hpp"
#include "mppd.h"

namespace {

void check(int rc, const char *what) {
  if (rc != MPPD_OK) throw std::runtime_error(std::string("micropp_amd: ") + what + ": " + mppd_last_error());
}

double *dzeros(size_t n) {
  void *p = nullptr;
  check(mppd_malloc(&p, n * sizeof(double)), "device allocation");
  check(mppd_memset(p, 0, n * sizeof(double)), "device memset");
  return static_cast<double *>(p);
}

// zeroed on the solver's own stream (other solvers keep running)
double *dzeros_on(mppd_solver *s, size_t n) {
  void *p = nullptr;
  check(mppd_malloc(&p, n * sizeof(double)), "device allocation");
  if (mppd_solver_fill(s, p, 0, n * sizeof(double)) != MPPD_OK) {
    mppd_free(p);
    check(MPPD_ERR_HIP, "device fill");
  }
  return static_cast<double *>(p);
}

// One FE solve context: a solver (matrix, CG workspace and its own stream)
// with the working displacement and a scratch history array.
struct worker {
  mppd_solver *sol = nullptr;
  double *u = nullptr;         // working displacement (device)
  double *vars_aux = nullptr;  // vars_new when the GP has no history yet (device)
};

struct gp_state {
  double strain_old[6] = {0, 0, 0, 0, 0, 0};
  double strain[6] = {0, 0, 0, 0, 0, 0};
  double stress[6] = {0, 0, 0, 0, 0, 0};
  double ctan[36] = {0};
  bool allocated = false;
  double *vars_n = nullptr, *vars_k = nullptr;  // device, nvars
  double *u_n = nullptr, *u_k = nullptr;        // device, nndim
  long cost = 0;
  bool converged = true;
  bool subiterated = false;
  int coupling = FE_ONE_WAY;
};

}  // namespace

template <int tdim>
struct micropp<tdim>::impl {
  int ngp, nx, ny, nz, nn, nndim, nelem, nvars;
  int micro_type;
  double geo[4];
  material_base mats[MAX_MATERIALS];
  int nsubiterations, mpi_rank;
  bool subiterations;
  mppd_newton_params np;
  bool calc_ctan_lin_flag, use_A0, lin_stress, write_log_flag, mixed;
  int log_id = 0;
  ofstream log;
  double Vm = 0, Vf = 0;
  double ctan_lin_fe[36] = {0};
  std::vector<gp_state> gp;
  mppd_rve *rve = nullptr;
  mppd_solver *sol = nullptr;
  mppd_solver *A0 = nullptr;
  double *u = nullptr;        // working displacement (device)
  double *vars_aux = nullptr;  // vars_new when the GP has no history yet (device)
  std::vector<int> fe_ids;     // GPs with FE_ONE_WAY / FE_FULL coupling
  mppd_batch *batch = nullptr;  // working slots for those GPs (homogenize)
  int batch_slots = 0;
  bool batch_tried = false;
  std::vector<worker> wk;  // wk[0] = {sol, u, vars_aux}; more for concurrent GPs
  bool wk_tried = false;
  std::mutex err_mu;
  // mixed precision on an all-elastic RVE: the multi-right-hand-side solver
  // with K fields, for the constructor's ctan and the FE GPs of homogenize
  mppd_msolver *ms = nullptr;
  int ms_K = 0;
  // with ms: the constructor's six ctan solves are deferred to the first pass
  // of homogenize, which solves them beside its GPs (one pass of up to K
  // fields instead of two); anything that reads a ctan first (get_ctan,
  // get_ctan_lin, print_info, a linear GP) solves them on their own
  bool ctan_pending = false;

  ~impl() {
    mppd_msolver_destroy(ms);
    mppd_batch_destroy(batch);
    for (size_t k = 1; k < wk.size(); ++k) {
      mppd_free(wk[k].u);
      mppd_free(wk[k].vars_aux);
      mppd_solver_destroy(wk[k].sol);
    }
    for (auto &q : gp)
      for (double *p : {q.vars_n, q.vars_k, q.u_n, q.u_k}) mppd_free(p);
    mppd_free(u);
    mppd_free(vars_aux);
    mppd_solver_destroy(A0);
    mppd_solver_destroy(sol);
    mppd_rve_destroy(rve);
  }

  newton_t newton(worker &w, const double eps[6], const double *vars_old) {
    int out[3] = {0, 0, 0};
    if (mixed)
      check(mppd_newton_mixed(w.sol, use_A0 ? A0 : nullptr, eps, vars_old, w.u, &np, nullptr, out, nullptr),
            "newton_raphson (mixed)");
    else
      check(mppd_newton(w.sol, use_A0 ? A0 : nullptr, eps, vars_old, w.u, &np, out), "newton_raphson");
    newton_t n;
    n.its = out[0];
    n.solver_its = out[1];
    n.converged = out[2] != 0;
    return n;
  }

  void ave_stress(worker &w, const double *ud, const double *vars, double sig[6]) {
    check(mppd_ave_stress(w.sol, ud, vars, sig), "calc_ave_stress");
  }

  void copy(worker &w, double *dst, const double *src) {
    check(mppd_solver_copy(w.sol, dst, src, sizeof(double) * nndim), "device copy");
  }

  // the linear ctan into every GP but the mixing-rule ones (micropp
  // constructor, src/micropp.cpp:113-121)
  void fill_gp_ctan() {
    for (auto &q : gp)
      if (q.coupling != MIX_RULE_CHAMIS) memcpy(q.ctan, ctan_lin_fe, sizeof(q.ctan));
  }

  // the deferred ctan solves, on their own; a device failure is reported on
  // stderr (the accessors that call this cross the C interface) and the
  // solves are tried again at the next read
  void ensure_ctan() {
    if (!ctan_pending) return;
    try {
      ctan_lin_fe_models();
    } catch (const std::exception &ex) {
      cerr << ex.what() << endl;
      return;
    }
    ctan_pending = false;
    fill_gp_ctan();
  }

  // calc_ctan_lin_fe_models (src/micropp.cpp:255-284)
  void ctan_lin_fe_models() {
    if (ms) {  // the six solves from u = 0 as one multi-right-hand-side Newton pass
      double eps[36] = {0}, sig[36];
      int out[18];
      for (int i = 0; i < 6; ++i) eps[i * 6 + i] = D_EPS_CTAN_AVE;
      check(mppd_msolver_newton(ms, 6, eps, nullptr, nullptr, &np, nullptr, sig, out, nullptr),
            "calc_ctan_lin_fe_models (multi-RHS)");
      for (int i = 0; i < 6; ++i)
        for (int v = 0; v < 6; ++v) ctan_lin_fe[v * 6 + i] = sig[i * 6 + v] / D_EPS_CTAN_AVE;
      return;
    }
    worker &w = wk[0];
    for (int i = 0; i < 6; ++i) {
      check(mppd_solver_fill(w.sol, w.u, 0, sizeof(double) * nndim), "device fill");
      double eps[6] = {0, 0, 0, 0, 0, 0}, sig[6];
      eps[i] += D_EPS_CTAN_AVE;
      newton(w, eps, nullptr);
      ave_stress(w, w.u, nullptr, sig);
      for (int v = 0; v < 6; ++v) ctan_lin_fe[v * 6 + i] = sig[v] / D_EPS_CTAN_AVE;
    }
  }

  // homogenize_linear(gp) (src/homogenize.cpp:101-109)
  static void linear(gp_state &q) {
    for (int i = 0; i < 6; ++i) q.stress[i] = 0.0;
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) q.stress[i] += q.ctan[i * 6 + j] * q.strain[j];
  }

  // homogenize_fe_one_way / homogenize_fe_full (src/homogenize.cpp:111-282)
  void fe(worker &w, gp_state &q, bool full) {
    check(mppd_solver_fill(w.sol, w.vars_aux, 0, sizeof(double) * nvars), "device fill");
    double *vars_new = q.allocated ? q.vars_k : w.vars_aux;
    q.cost = 0;
    q.subiterated = false;
    copy(w, w.u, q.u_n);
    newton_t n = newton(w, q.strain, q.vars_n);
    copy(w, q.u_k, w.u);
    q.cost += n.solver_its;
    q.converged = n.converged;
    if (!q.converged && subiterations) {
      double eps_sub[6], deps[6];
      copy(w, w.u, q.u_n);
      for (int i = 0; i < 6; ++i) eps_sub[i] = q.strain_old[i];
      q.subiterated = true;
      for (int i = 0; i < 6; ++i) deps[i] = (q.strain[i] - q.strain_old[i]) / nsubiterations;
      for (int s = 0; s < nsubiterations; ++s) {
        for (int j = 0; j < 6; ++j) eps_sub[j] += deps[j];
        n = newton(w, eps_sub, q.vars_n);
        q.cost += n.solver_its;
      }
      q.converged = n.converged;
      copy(w, q.u_k, w.u);
    }
    if (lin_stress)
      linear(q);
    else
      ave_stress(w, q.u_k, q.vars_n, q.stress);
    int nl = 0;
    check(mppd_vars_new(w.sol, q.u_k, q.vars_n, vars_new, &nl), "calc_vars_new");
    if (nl && !q.allocated) {
      q.vars_n = dzeros_on(w.sol, nvars);
      q.vars_k = dzeros_on(w.sol, nvars);
      q.allocated = true;
      check(mppd_solver_copy(w.sol, q.vars_k, vars_new, sizeof(double) * nvars), "device copy");
    }
    if (full && q.allocated) {
      double eps1[6], sig0[6], sig1[6];
      copy(w, w.u, q.u_k);
      for (int v = 0; v < 6; ++v) sig0[v] = q.stress[v];
      for (int i = 0; i < 6; ++i) {
        for (int v = 0; v < 6; ++v) eps1[v] = q.strain[v];
        eps1[i] += D_EPS_CTAN_AVE;
        newton(w, eps1, q.vars_n);
        q.cost += n.solver_its;  // the reference adds the sigma solve's count again
        ave_stress(w, w.u, q.vars_n, sig1);
        for (int v = 0; v < 6; ++v) q.ctan[v * 6 + i] = (sig1[v] - sig0[v]) / D_EPS_CTAN_AVE;
      }
    }
  }

  // The FE GPs of an all-elastic RVE with mixed precision run through the
  // multi-right-hand-side solver, ms_K GPs per pass, in place of the
  // reference's OpenMP loop over GPs (src/homogenize.cpp:72): each GP's
  // Newton starts from its u_n with its strain's boundary values (as
  // homogenize_fe_one_way does, :111-186) and writes u_k; sigma is ctan.eps
  // (lin_stress) or the pass's FP64 volume average.  The GPs of a pass share
  // one FP32 matrix stream per inner iteration.  An elastic RVE never
  // allocates history, so calc_vars_new and the FE_FULL perturbations have
  // nothing to do.  A GP whose Newton does not converge, with subiterations
  // on, is solved again on the one-GP path, which subiterates.
  // With the constructor's ctan still pending, the first pass carries its
  // six unit-strain solves (fields 0-5, from u = 0, as
  // calc_ctan_lin_fe_models) ahead of its GPs.  Each field's solve does not
  // depend on the others in the pass.
  void fe_multi() {
    const int nfe = (int)fe_ids.size(), K = ms_K;
    std::vector<double> eps(6 * (size_t)K), sig(6 * (size_t)K);
    std::vector<const double *> uin(K);
    std::vector<double *> uout(K);
    std::vector<int> out(3 * (size_t)K);
    for (int c0 = 0; c0 < nfe;) {
      const int lead = ctan_pending ? 6 : 0;
      const int m = std::min(K - lead, nfe - c0);
      try {
        for (int i = 0; i < lead; ++i) {
          for (int v = 0; v < 6; ++v) eps[6 * i + v] = v == i ? D_EPS_CTAN_AVE : 0.0;
          uin[i] = nullptr;
          uout[i] = nullptr;
        }
        for (int r = 0; r < m; ++r) {
          const gp_state &q = gp[fe_ids[c0 + r]];
          memcpy(&eps[6 * (lead + r)], q.strain, sizeof(q.strain));
          uin[lead + r] = q.u_n;
          uout[lead + r] = q.u_k;
        }
        check(mppd_msolver_newton(ms, lead + m, eps.data(), uin.data(), uout.data(), &np, nullptr, sig.data(),
                                  out.data(), nullptr),
              lead ? "calc_ctan_lin_fe_models and newton_raphson (multi-RHS)" : "newton_raphson (multi-RHS)");
        if (lead) {
          for (int i = 0; i < 6; ++i)
            for (int v = 0; v < 6; ++v) ctan_lin_fe[v * 6 + i] = sig[i * 6 + v] / D_EPS_CTAN_AVE;
          ctan_pending = false;
          fill_gp_ctan();
        }
        for (int r = 0; r < m; ++r) {
          gp_state &q = gp[fe_ids[c0 + r]];
          const int f = lead + r;
          q.cost = out[3 * f + 1];
          q.converged = out[3 * f + 2] != 0;
          q.subiterated = false;
          if (!q.converged && subiterations) {
            fe(wk[0], q, q.coupling == FE_FULL);
            continue;
          }
          if (lin_stress)
            linear(q);
          else
            memcpy(q.stress, &sig[6 * f], sizeof(q.stress));
        }
      } catch (const std::exception &ex) {
        for (int r = 0; r < m; ++r) gp[fe_ids[c0 + r]].converged = false;
        cerr << ex.what() << " (gps " << fe_ids[c0] << " .. " << fe_ids[c0 + m - 1] << ")" << endl;
      }
      c0 += m;
    }
  }

  // The FE GPs of homogenize run as one batch of identical RVE slots, in
  // groups of batch_slots GPs: every finite-element step of a group is one
  // launch over all its slots, in place of the reference's OpenMP loop over
  // GPs (src/homogenize.cpp:72).  Slots hold the same RVE and run the same
  // kernel bodies as the one-GP path, so the results are bit-identical.
  // The batch is made at the first homogenize, sized to half the free device
  // memory; MPPD_GP_BATCH=0 in the environment keeps the one-GP-at-a-time
  // path, MPPD_GP_BATCH=n caps the slots at n.  With use_A0 the batch keeps
  // its own copy of the initial matrix; with mixed precision it solves in
  // mixed precision.  With fewer than two FE GPs the one-GP path runs.
  bool batch_ready() {
    if (batch) return true;
    if (batch_tried) return false;
    batch_tried = true;
    if (fe_ids.size() < 2) return false;
    const char *env = getenv("MPPD_GP_BATCH");
    const long cap = env ? strtol(env, nullptr, 10) : -1;
    if (cap == 0) return false;
    size_t per = 0, fr = 0, tot = 0;
    check(mppd_batch_slot_bytes(rve, &per), "batch sizing");
    // mixed precision adds an FP32 copy of the values and six FP32 vectors,
    // less than half of a slot's FP64 bytes
    if (mixed) per += per / 2;
    check(mppd_mem_info(&fr, &tot), "device memory");
    size_t n = std::min(fe_ids.size(), fr / 2 / std::max<size_t>(per, 1));
    if (cap > 0) n = std::min<size_t>(n, (size_t)cap);
    if (n < 2) return false;
    mppd_batch_rve d;
    memset(&d, 0, sizeof(d));
    d.size[0] = nx;
    d.size[1] = ny;
    d.size[2] = nz;
    d.micro_type = micro_type;
    for (int i = 0; i < 4; ++i) d.geo_params[i] = geo[i];
    for (int i = 0; i < MAX_MATERIALS; ++i) d.materials[i] = mats[i];
    d.coupling = MPPD_FE_ONE_WAY;
    std::vector<mppd_batch_rve> ds(n, d);
    mppd_batch_params p;
    mppd_batch_params_default(&p);
    p.newton = np;
    p.lin_stress = lin_stress;
    p.calc_ctan_lin = 0;  // the instance's ctan_lin_fe is loaded into every slot
    p.use_A0 = use_A0;
    p.mixed_precision = mixed;
    p.subiterations = subiterations;
    p.nsubiterations = nsubiterations;
    if (mppd_batch_create(&batch, (int)n, ds.data(), &p) != MPPD_OK) {
      cerr << "micropp_amd: GP batch not available (" << mppd_last_error() << "); GPs run one at a time" << endl;
      batch = nullptr;
      return false;
    }
    batch_slots = (int)n;
    return true;
  }

  void fe_batched() {
    const int S = batch_slots, nfe = (int)fe_ids.size();
    std::vector<int> info(4 * (size_t)S);
    for (int c0 = 0; c0 < nfe; c0 += S) {
      const int m = std::min(S, nfe - c0);
      try {
        for (int r = 0; r < S; ++r) {
          mppd_batch_gp b;
          memset(&b, 0, sizeof(b));
          b.coupling = MPPD_FE_LINEAR;  // idle slot
          b.converged = 1;
          if (r >= m) {
            check(mppd_batch_load(batch, r, &b, nullptr, nullptr), "batch load");
            continue;
          }
          const gp_state &q = gp[fe_ids[c0 + r]];
          memcpy(b.strain, q.strain, sizeof(b.strain));
          memcpy(b.strain_old, q.strain_old, sizeof(b.strain_old));
          memcpy(b.stress, q.stress, sizeof(b.stress));
          memcpy(b.ctan, q.ctan, sizeof(b.ctan));
          b.coupling = q.coupling == FE_FULL ? MPPD_FE_FULL : MPPD_FE_ONE_WAY;
          b.allocated = q.allocated;
          b.cost = (int)q.cost;
          b.converged = q.converged;
          b.subiterated = q.subiterated;
          check(mppd_batch_load(batch, r, &b, q.u_n, q.allocated ? q.vars_n : nullptr), "batch load");
        }
        check(mppd_batch_homogenize(batch), "homogenize");
        check(mppd_batch_get_info(batch, info.data()), "batch info");
        for (int r = 0; r < m; ++r) {
          gp_state &q = gp[fe_ids[c0 + r]];
          if (info[4 * r] && !q.allocated) {  // history allocated on the first nonlinearity
            q.vars_n = dzeros(nvars);
            q.vars_k = dzeros(nvars);
          }
          mppd_batch_gp b;
          check(mppd_batch_store(batch, r, &b, q.u_k, info[4 * r] ? q.vars_k : nullptr), "batch store");
          memcpy(q.stress, b.stress, sizeof(q.stress));
          memcpy(q.ctan, b.ctan, sizeof(q.ctan));
          q.allocated = b.allocated != 0;
          q.cost = b.cost;
          q.converged = b.converged != 0;
          q.subiterated = b.subiterated != 0;
        }
      } catch (const std::exception &ex) {
        for (int r = 0; r < m; ++r) gp[fe_ids[c0 + r]].converged = false;
        cerr << ex.what() << " (gps " << fe_ids[c0] << " .. " << fe_ids[c0 + m - 1] << ")" << endl;
      }
    }
  }

  // FE GPs that do not run as a batch (MPPD_GP_BATCH=0, or no batch could be
  // made) run concurrently, in place of
  // the reference's OpenMP loop over GPs (src/homogenize.cpp:72): up to
  // MPPD_GP_WORKERS (default 4) solvers, each with its own stream, take GPs
  // in turn from host threads.  A GP's kernels and their order do not depend
  // on the solver that runs it, so the results are those of one solver bit
  // for bit.  The extra solvers are made at the first homogenize, while at
  // least a quarter of the device memory stays free; MPPD_GP_WORKERS=1 keeps
  // one solver.
  bool workers_ready() {
    if (wk.size() > 1) return true;
    if (wk_tried) return false;
    wk_tried = true;
    const char *env = getenv("MPPD_GP_WORKERS");
    const long want = std::min<long>(env ? strtol(env, nullptr, 10) : 4, (long)fe_ids.size());
    for (long k = 1; k < want; ++k) {
      size_t fr = 0, tot = 0;
      if (mppd_mem_info(&fr, &tot) != MPPD_OK || fr < tot / 4) break;
      worker w;
      try {
        check(mppd_solver_create(&w.sol, rve), "solver set-up");
        w.u = dzeros_on(w.sol, nndim);
        w.vars_aux = dzeros_on(w.sol, nvars);
      } catch (const std::exception &ex) {
        cerr << ex.what() << "; " << wk.size() << " concurrent GP solver(s)" << endl;
        mppd_free(w.u);
        mppd_free(w.vars_aux);
        mppd_solver_destroy(w.sol);
        break;
      }
      wk.push_back(w);
    }
    return wk.size() > 1;
  }

  void fe_parallel() {
    const int nfe = (int)fe_ids.size();
    std::atomic<int> next{0};
    auto run = [&](worker *w) {
      for (int i = next++; i < nfe; i = next++) {
        gp_state &q = gp[fe_ids[i]];
        try {
          fe(*w, q, q.coupling == FE_FULL);
        } catch (const std::exception &ex) {
          q.converged = false;
          std::lock_guard<std::mutex> lk(err_mu);
          cerr << ex.what() << " (gp " << fe_ids[i] << ")" << endl;
        }
      }
    };
    std::vector<std::thread> th;
    for (size_t k = 1; k < wk.size(); ++k) {
      try {
        th.emplace_back(run, &wk[k]);
      } catch (const std::system_error &) {
        break;  // the threads already started and this one share the GPs
      }
    }
    run(&wk[0]);
    for (auto &t : th) t.join();
  }

  void chamis(double c[36]) const;
  void write_vtu(const gp_state &q, const std::string &name) const;
};

// calc_ctan_lin_mix_rule_Chamis (src/micropp.cpp:286-331) and invert_3x3
// (include/util.hpp:145-163): a closed-form 6x6, no FE work.
template <int tdim>
void micropp<tdim>::impl::chamis(double c[36]) const {
  const double Em = mats[0].E, nu_m = mats[0].nu, Gm = Em / (2 * (1 + nu_m));
  const double Ef = mats[1].E, nu_f = mats[1].nu, Gf = Ef / (2 * (1 + nu_f));
  const double E11 = Vf * Ef + Vm * Em;
  const double E22 = Em / (1 - sqrt(Vf) * (1 - Em / Ef));
  const double nu12 = Vf * nu_f + Vm * nu_m;
  const double G12 = Gm / (1 - sqrt(Vf) * (1 - Gm / Gf));
  const double nu23 = nu12;
  const double m[3][3] = {{1 / E11, -nu12 / E11, -nu12 / E11},
                          {-nu12 / E11, 1 / E22, -nu23 / E22},
                          {-nu12 / E11, -nu23 / E22, 1 / E22}};
  const double det = m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
                     m[0][1] * (m[1][0] * m[2][2] - m[2][0] * m[1][2]) +
                     m[0][2] * (m[1][0] * m[2][1] - m[2][0] * m[1][1]);
  double o[3][3];
  o[0][0] = +(m[1][1] * m[2][2] - m[2][1] * m[1][2]) / det;
  o[0][1] = -(m[1][0] * m[2][2] - m[2][0] * m[1][2]) / det;
  o[0][2] = +(m[1][0] * m[2][1] - m[2][0] * m[1][1]) / det;
  o[1][0] = -(m[0][1] * m[2][2] - m[2][1] * m[0][2]) / det;
  o[1][1] = +(m[0][0] * m[2][2] - m[2][0] * m[0][2]) / det;
  o[1][2] = -(m[0][0] * m[2][1] - m[2][0] * m[0][1]) / det;
  o[2][0] = +(m[0][1] * m[1][2] - m[1][1] * m[0][2]) / det;
  o[2][1] = -(m[0][0] * m[1][2] - m[1][0] * m[0][2]) / det;
  o[2][2] = +(m[0][0] * m[1][1] - m[1][0] * m[0][1]) / det;
  for (int i = 0; i < 36; ++i) c[i] = 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) c[i * 6 + j] = o[i][j];
  for (int i = 3; i < 6; ++i) c[i * 6 + i] = G12;
}

namespace {

// options of an instance made with the reference's constructor
micropp_options_t env_options() {
  micropp_options_t o;
  const char *v = getenv("MICROPP_MIXED_PRECISION");
  o.mixed_precision = v && strtol(v, nullptr, 10) != 0;
  return o;
}

}  // namespace

template <int tdim>
micropp<tdim>::micropp(const micropp_params_t &params) : micropp(params, env_options()) {}

// Constructor (src/micropp.cpp:29-183)
template <int tdim>
micropp<tdim>::micropp(const micropp_params_t &params, const micropp_options_t &options) : p_(new impl()) {
  impl &s = *p_;
  s.ngp = params.ngp;
  s.nx = params.size[0];
  s.ny = params.size[1];
  s.nz = params.size[2];
  s.nn = s.nx * s.ny * s.nz;
  s.nndim = s.nn * 3;
  s.nelem = (s.nx - 1) * (s.ny - 1) * (s.nz - 1);
  s.nvars = s.nelem * 8 * NUM_VAR_GP;
  s.micro_type = params.type;
  for (int i = 0; i < 4; ++i) s.geo[i] = params.geo_params[i];
  for (int i = 0; i < MAX_MATERIALS; ++i) s.mats[i] = params.materials[i];
  s.subiterations = params.subiterations;
  s.nsubiterations = params.nsubiterations;
  s.mpi_rank = params.mpi_rank;
  s.np.nr_max_its = params.nr_max_its;
  s.np.nr_max_tol = params.nr_max_tol;
  s.np.nr_rel_tol = params.nr_rel_tol;
  s.np.cg_max_its = params.cg_max_its;
  s.np.cg_abs_tol = params.cg_abs_tol;
  s.np.cg_rel_tol = params.cg_rel_tol;
  s.np.its_with_A0 = params.its_with_A0;
  s.calc_ctan_lin_flag = params.calc_ctan_lin;
  s.use_A0 = params.use_A0;
  s.lin_stress = params.lin_stress;
  s.write_log_flag = params.write_log;
  s.mixed = options.mixed_precision;
  try {
    if (s.ngp < 0) throw std::invalid_argument("micropp_amd: ngp must be >= 0");
    if (s.micro_type < 0 || s.micro_type > MIC3D_RANDOM_SPHERES)
      throw std::invalid_argument("micropp_amd: invalid micro_type " + std::to_string(s.micro_type));
    check(mppd_rve_create(&s.rve, params.size, s.micro_type, s.geo, s.mats), "RVE set-up");
    check(mppd_solver_create(&s.sol, s.rve), "solver set-up");
    double vf[2];
    check(mppd_rve_volume_fractions(s.rve, vf), "volume fractions");
    s.Vm = vf[0];
    s.Vf = vf[1];
    s.u = dzeros(s.nndim);
    s.vars_aux = dzeros(s.nvars);
    s.wk.push_back(worker{s.sol, s.u, s.vars_aux});
    s.gp.resize(s.ngp);
    int n_fe = 0;
    for (int g = 0; g < s.ngp; ++g) {
      gp_state &q = s.gp[g];
      q.coupling = params.coupling ? params.coupling[g] : FE_ONE_WAY;
      if (q.coupling == FE_LINEAR || q.coupling == FE_ONE_WAY || q.coupling == FE_FULL) n_fe++;
      if (q.coupling == FE_ONE_WAY || q.coupling == FE_FULL) s.fe_ids.push_back(g);
      if (!params.coupling || q.coupling == FE_ONE_WAY || q.coupling == FE_FULL) {
        q.u_n = dzeros(s.nndim);
        q.u_k = dzeros(s.nndim);
      }
    }
    if (s.use_A0) {
      check(mppd_solver_create(&s.A0, s.rve), "A0 set-up");
      check(mppd_assembly_mat(s.A0, s.u, nullptr), "A0 assembly");
    }
    // mixed precision on an all-elastic RVE: the multi-RHS solver, with
    // fields for the six ctan problems or up to 16 FE GPs per pass
    // (MPPD_MSOLVER_ROUTE=0 in the environment keeps the per-GP paths)
    bool elastic = true;
    for (int i = 0; i < MAX_MATERIALS; ++i) elastic &= s.mats[i].type == MATERIAL_ELASTIC;
    const char *route = getenv("MPPD_MSOLVER_ROUTE");
    if (s.mixed && elastic && (!route || strtol(route, nullptr, 10) != 0) &&
        (n_fe > 0 && (s.calc_ctan_lin_flag || !s.fe_ids.empty()))) {
      // fields for the six ctan solves beside the FE GPs of the first homogenize pass
      const int lead = s.calc_ctan_lin_flag ? 6 : 0;
      const int want = std::max<int>(std::max(lead, 1), std::min<int>(16, (int)s.fe_ids.size() + lead));
      s.ms_K = (want + 3) / 4 * 4;
      check(mppd_msolver_create(&s.ms, s.rve, s.ms_K), "multi-RHS solver set-up");
    }
    if (s.calc_ctan_lin_flag && n_fe > 0) {
      if (s.ms)
        s.ctan_pending = true;  // solved with the first homogenize pass, or when first read
      else
        s.ctan_lin_fe_models();
    }
    for (auto &q : s.gp)
      if (q.coupling == MIX_RULE_CHAMIS) s.chamis(q.ctan);
    if (!s.ctan_pending) s.fill_gp_ctan();
    if (s.write_log_flag) {
      std::stringstream fn;
      fn << "micropp-profiling-" << s.mpi_rank << ".log";
      s.log.open(fn.str(), ios::out);
      s.log << "#<gp_id>  <non-linear>  <cost>  <converged>" << endl;
    }
  } catch (...) {
    delete p_;
    p_ = nullptr;
    throw;
  }
}

template <int tdim>
micropp<tdim>::~micropp() {
  delete p_;
}

// set_strain / get_stress / get_ctan (src/homogenize.cpp:30-49)
template <int tdim>
void micropp<tdim>::set_strain(const int gp_id, const double *strain) {
  assert(gp_id >= 0 && gp_id < p_->ngp);
  memcpy(p_->gp[gp_id].strain, strain, 6 * sizeof(double));
}

template <int tdim>
void micropp<tdim>::get_stress(const int gp_id, double *stress) const {
  assert(gp_id >= 0 && gp_id < p_->ngp);
  memcpy(stress, p_->gp[gp_id].stress, 6 * sizeof(double));
}

template <int tdim>
void micropp<tdim>::get_ctan(const int gp_id, double *ctan) const {
  assert(gp_id >= 0 && gp_id < p_->ngp);
  p_->ensure_ctan();
  memcpy(ctan, p_->gp[gp_id].ctan, 36 * sizeof(double));
}

// homogenize (src/homogenize.cpp:68-99).  FE GPs run as a batch (see
// impl::batch_ready), else concurrently on several solvers (see
// impl::workers_ready), else one after the other.  A failed device call marks
// the GP (or the batch group) as not converged and is reported on stderr; the
// other GPs still run.
template <int tdim>
void micropp<tdim>::homogenize() {
  impl &s = *p_;
  bool batched = false;
  try {
    batched = !s.ms && s.batch_ready();
  } catch (const std::exception &ex) {
    cerr << ex.what() << "; GPs run one at a time" << endl;
  }
  bool parallel = false;
  const bool multi = s.ms && !s.fe_ids.empty();
  if (multi) {
    s.fe_multi();
  } else if (batched) {
    s.fe_batched();
  } else if (s.fe_ids.size() > 1) {
    try {
      parallel = s.workers_ready();
    } catch (const std::exception &ex) {
      cerr << ex.what() << "; GPs run one at a time" << endl;
    }
    if (parallel) s.fe_parallel();
  }
  s.ensure_ctan();  // no FE pass carried the pending ctan solves (linear GPs read it below)
  for (int g = 0; g < s.ngp; ++g) {
    gp_state &q = s.gp[g];
    try {
      if (q.coupling == FE_LINEAR || q.coupling == MIX_RULE_CHAMIS)
        impl::linear(q);
      else if (multi || batched || parallel)
        continue;
      else if (q.coupling == FE_ONE_WAY)
        s.fe(s.wk[0], q, false);
      else if (q.coupling == FE_FULL)
        s.fe(s.wk[0], q, true);
    } catch (const std::exception &ex) {
      q.converged = false;
      cerr << ex.what() << " (gp " << g << ")" << endl;
    }
  }
  if (s.write_log_flag) {
    // write_log (src/output.cpp:264-282)
    s.log << "log_id : " << s.log_id << endl;
    for (int g = 0; g < s.ngp; ++g)
      s.log << "\t" << g << "\t" << s.gp[g].allocated << "\t" << s.gp[g].cost << "\t" << s.gp[g].converged << endl;
    s.log_id++;
  }
}

template <int tdim>
void micropp<tdim>::homogenize_linear() {
  p_->ensure_ctan();
  for (auto &q : p_->gp) impl::linear(q);
}

template <int tdim>
int micropp<tdim>::is_non_linear(const int gp_id) const {
  assert(gp_id >= 0 && gp_id < p_->ngp);
  return (int)p_->gp[gp_id].allocated;
}

template <int tdim>
int micropp<tdim>::get_non_linear_gps(void) const {
  int c = 0;
  for (const auto &q : p_->gp) c += q.allocated ? 1 : 0;
  return c;
}

template <int tdim>
int micropp<tdim>::get_cost(int gp_id) const {
  assert(gp_id >= 0 && gp_id < p_->ngp);
  return (int)p_->gp[gp_id].cost;
}

template <int tdim>
bool micropp<tdim>::has_converged(int gp_id) const {
  assert(gp_id >= 0 && gp_id < p_->ngp);
  return p_->gp[gp_id].converged;
}

template <int tdim>
bool micropp<tdim>::has_subiterated(int gp_id) const {
  assert(gp_id >= 0 && gp_id < p_->ngp);
  return p_->gp[gp_id].subiterated;
}

// gp_t::update_vars (include/gp.hpp:95-105)
template <int tdim>
void micropp<tdim>::update_vars() {
  for (auto &q : p_->gp) {
    std::swap(q.vars_n, q.vars_k);
    std::swap(q.u_n, q.u_k);
    memcpy(q.strain_old, q.strain, sizeof(q.strain));
  }
}

// write_restart / read_restart (src/output.cpp:216-262, include/gp.hpp:107-124):
// per GP a bool, then vars_n and u_n if the GP has history.
template <int tdim>
void micropp<tdim>::write_restart(const int restart_id) const {
  const impl &s = *p_;
  std::stringstream fn;
  fn << "micropp-restart-" << s.mpi_rank << "-" << restart_id << ".bin";
  ofstream file(fn.str(), ios::out | ios::binary);
  std::vector<double> buf;
  for (const auto &q : s.gp) {
    const bool a = q.allocated;
    file.write((const char *)&a, sizeof(bool));
    if (a) {
      buf.resize(s.nvars);
      check(mppd_memcpy_d2h(buf.data(), q.vars_n, sizeof(double) * s.nvars), "restart copy");
      file.write((const char *)buf.data(), sizeof(double) * s.nvars);
      buf.resize(s.nndim);
      check(mppd_memcpy_d2h(buf.data(), q.u_n, sizeof(double) * s.nndim), "restart copy");
      file.write((const char *)buf.data(), sizeof(double) * s.nndim);
    }
  }
}

template <int tdim>
void micropp<tdim>::read_restart(const int restart_id) const {
  impl &s = *p_;
  std::stringstream fn;
  fn << "micropp-restart-" << s.mpi_rank << "-" << restart_id << ".bin";
  ifstream file(fn.str(), ios::in | ios::binary);
  std::vector<double> buf;
  for (auto &q : s.gp) {
    bool a = false;
    file.read((char *)&a, sizeof(bool));
    q.allocated = a;
    if (a) {
      if (!q.vars_n) q.vars_n = dzeros(s.nvars);
      if (!q.vars_k) q.vars_k = dzeros(s.nvars);
      buf.resize(s.nvars);
      file.read((char *)buf.data(), sizeof(double) * s.nvars);
      check(mppd_memcpy_h2d(q.vars_n, buf.data(), sizeof(double) * s.nvars), "restart copy");
      buf.resize(s.nndim);
      file.read((char *)buf.data(), sizeof(double) * s.nndim);
      if (q.u_n) check(mppd_memcpy_h2d(q.u_n, buf.data(), sizeof(double) * s.nndim), "restart copy");
    }
  }
}

// write_vtu (src/output.cpp:67-214) of one GP: its u_k and vars_n, with the
// element strain and stress of calc_fields computed on the GPU.  The file is
// the reference's ASCII UnstructuredGrid, reals printed as %e (the
// reference's std::scientific at the default precision of 6).
template <int tdim>
void micropp<tdim>::impl::write_vtu(const gp_state &q, const std::string &name) const {
  if (!q.u_k) throw std::runtime_error("micropp_amd: output: the GP has no displacement field (its coupling is "
                                       "FE_LINEAR or MIX_RULE_CHAMIS)");
  const int ne = nelem;
  std::vector<double> u(nndim), es(6 * (size_t)ne), ss(6 * (size_t)ne), vars;
  std::vector<int> et(ne);
  const double *vars_d = q.allocated ? q.vars_n : nullptr;
  check(mppd_memcpy_d2h(u.data(), q.u_k, sizeof(double) * nndim), "output copy");
  if (vars_d) {
    vars.resize(nvars);
    check(mppd_memcpy_d2h(vars.data(), vars_d, sizeof(double) * nvars), "output copy");
  }
  check(mppd_elem_fields(sol, q.u_k, vars_d, es.data(), ss.data()), "calc_fields");
  check(mppd_rve_elem_types(rve, et.data()), "element types");
  const std::string path = name + ".vtu";
  FILE *f = fopen(path.c_str(), "w");
  if (!f) throw std::runtime_error("micropp_amd: output: cannot open " + path);
  const double dx = 1.0 / (nx - 1), dy = 1.0 / (ny - 1), dz = 1.0 / (nz - 1);
  fprintf(f,
          "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" "
          "byte_order=\"LittleEndian\">\n<UnstructuredGrid>\n<Piece NumberOfPoints=\"%d\" NumberOfCells=\"%d\">\n"
          "<Points>\n<DataArray type=\"Float64\" Name=\"Position\" NumberOfComponents=\"3\" format=\"ascii\">\n",
          nn, ne);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) fprintf(f, "%e %e %e \n", i * dx, j * dy, k * dz);
  fprintf(f, "</DataArray>\n</Points>\n<Cells>\n\n"
             "<DataArray type=\"Int32\" Name=\"connectivity\" NumberOfComponents=\"1\" format=\"ascii\">\n");
  const int nxny = nx * ny;
  for (int ez = 0; ez < nz - 1; ++ez)
    for (int ey = 0; ey < ny - 1; ++ey)
      for (int ex = 0; ex < nx - 1; ++ex) {
        // get_elem_nodes (src/common.cpp:28-41)
        const int n0 = ez * nxny + ey * nx + ex;
        fprintf(f, "%d %d %d %d %d %d %d %d \n", n0, n0 + 1, n0 + nx + 1, n0 + nx, n0 + nxny, n0 + nxny + 1,
                n0 + nxny + nx + 1, n0 + nxny + nx);
      }
  fprintf(f, "</DataArray>\n<DataArray type=\"Int32\" Name=\"offsets\" NumberOfComponents=\"1\" format=\"ascii\">\n");
  for (int e = 0; e < ne; ++e) fprintf(f, "%d ", 8 * (e + 1));
  fprintf(f, "\n</DataArray>\n<DataArray type=\"UInt8\"  Name=\"types\" NumberOfComponents=\"1\" format=\"ascii\">\n");
  for (int e = 0; e < ne; ++e) fputs("12 ", f);
  fprintf(f, "\n</DataArray>\n</Cells>\n<PointData Vectors=\"displ\" >\n"
             "<DataArray type=\"Float64\" Name=\"displ\" NumberOfComponents=\"3\" format=\"ascii\" >\n");
  for (int n = 0; n < nn; ++n) fprintf(f, "%e %e %e \n", u[3 * n], u[3 * n + 1], u[3 * n + 2]);
  fprintf(f, "</DataArray>\n</PointData>\n<CellData>\n");
  for (const auto &arr : {std::make_pair("strain", &es), std::make_pair("stress", &ss)}) {
    fprintf(f, "<DataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"6\" format=\"ascii\">\n", arr.first);
    const double *a = arr.second->data();
    for (int e = 0; e < ne; ++e)
      fprintf(f, "%e %e %e %e %e %e \n", a[6 * e], a[6 * e + 1], a[6 * e + 2], a[6 * e + 3], a[6 * e + 4],
              a[6 * e + 5]);
    fprintf(f, "</DataArray>\n");
  }
  fprintf(f, "<DataArray type=\"Int32\" Name=\"elem_type\" NumberOfComponents=\"1\" format=\"ascii\">\n");
  for (int e = 0; e < ne; ++e) fprintf(f, "%d ", et[e]);
  fprintf(f, "\n</DataArray>\n");
  // history summaries per element, means over the 8 Gauss points of history
  // slot 0..5 squared (plus the square root of that sum, as the reference
  // adds it), slot 0, slot 1 and slot 6; zero without history
  const char *hname[4] = {"plasticity", "damage_e", "damage_D", "hardening"};
  const int hslot[4] = {-1, 0, 1, 6};
  for (int h = 0; h < 4; ++h) {
    fprintf(f, "<DataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"1\" format=\"ascii\">\n", hname[h]);
    for (int e = 0; e < ne; ++e) {
      double acc = 0.0;
      if (!vars.empty())
        for (int gp = 0; gp < 8; ++gp) {
          const double *v = &vars[((size_t)e * 8 + gp) * NUM_VAR_GP];
          if (hslot[h] >= 0) {
            acc += v[hslot[h]];
          } else {
            for (int c = 0; c < 6; ++c) acc += v[c] * v[c];
          }
        }
      if (hslot[h] < 0) acc += sqrt(acc);
      fprintf(f, "%e ", acc / 8);
    }
    fprintf(f, "\n</DataArray>\n");
  }
  fprintf(f, "</CellData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n");
  if (fclose(f) != 0) throw std::runtime_error("micropp_amd: output: write error on " + path);
}

// output (src/output.cpp:29-37): <filename>.vtu of GP gp_id
template <int tdim>
void micropp<tdim>::output(int gp_id, const char *filename) {
  assert(gp_id >= 0 && gp_id < p_->ngp);
  try {
    p_->write_vtu(p_->gp[gp_id], filename);
  } catch (const std::exception &ex) {
    cerr << ex.what() << endl;
  }
}

// output2 (src/output.cpp:39-65): micropp-<elem_global>-<time_step>.vtu
template <int tdim>
void micropp<tdim>::output2(const int gp_id, const int elem_global, const int time_step) {
  assert(gp_id >= 0 && gp_id < p_->ngp);
  std::stringstream fn;
  fn << "micropp-" << elem_global << "-" << time_step;
  try {
    p_->write_vtu(p_->gp[gp_id], fn.str());
  } catch (const std::exception &ex) {
    cerr << ex.what() << endl;
  }
}

// print_info (src/micropp.cpp:546-597)
template <int tdim>
void micropp<tdim>::print_info() const {
  p_->ensure_ctan();
  const impl &s = *p_;
  int cnt[4] = {0, 0, 0, 0};
  for (const auto &q : s.gp)
    if (q.coupling >= 0 && q.coupling < 4) cnt[q.coupling]++;
  cout << "micropp3 (micropp_amd, gfx950)" << endl;
  cout << "Micro-structure   : " << micro_names[s.micro_type] << endl;
  cout << "MATRIX [%]        : " << s.Vm << endl;
  cout << "FIBER  [%]        : " << s.Vf << endl;
  cout << "FE_LINEAR         : " << cnt[FE_LINEAR] << " GPs" << endl;
  cout << "FE_ONE_WAY        : " << cnt[FE_ONE_WAY] << " GPs" << endl;
  cout << "FE_FULL           : " << cnt[FE_FULL] << " GPs" << endl;
  cout << "MIX_RULE_CHAMIS   : " << cnt[MIX_RULE_CHAMIS] << " GPs" << endl;
  cout << "USE A0            : " << s.use_A0 << endl;
  cout << "NUM SUBITS        : " << s.nsubiterations << endl;
  cout << "MPI RANK          : " << s.mpi_rank << endl;
  cout << "ngp :" << s.ngp << " nx :" << s.nx << " ny :" << s.ny << " nz :" << s.nz << " nn :" << s.nn << endl;
  cout << "geo_params:";
  for (int i = 0; i < 4; ++i) cout << " " << s.geo[i];
  cout << endl;
  for (int i = 0; i < MAX_MATERIALS; ++i) material_print(&s.mats[i]);
  cout << endl << "ctan_lin_fe = " << endl;
  for (int i = 0; i < 6; ++i) {
    for (int j = 0; j < 6; ++j) cout << s.ctan_lin_fe[i * 6 + j] << "\t";
    cout << endl;
  }
  cout << endl;
}

template <int tdim>
bool micropp<tdim>::get_state(int gp_id, int which, double *out) const {
  const impl &s = *p_;
  assert(gp_id >= 0 && gp_id < s.ngp);
  const gp_state &q = s.gp[gp_id];
  const double *src = which == 0 ? q.u_n : which == 1 ? q.u_k : which == 2 ? q.vars_n : q.vars_k;
  if (!src) return false;
  const size_t n = which < 2 ? (size_t)s.nndim : (size_t)s.nvars;
  check(mppd_memcpy_d2h(out, src, sizeof(double) * n), "state copy");
  return true;
}

template <int tdim>
void micropp<tdim>::get_ctan_lin(double *ctan) const {
  p_->ensure_ctan();
  memcpy(ctan, p_->ctan_lin_fe, sizeof(p_->ctan_lin_fe));
}

template <int tdim>
void micropp<tdim>::get_volume_fractions(double *vf) const {
  vf[0] = p_->Vm;
  vf[1] = p_->Vf;
}

template class micropp<3>;
