/*
 * ref_shim.cpp -- C entry points into the compiled Micropp reference, used to
 * pin the oracle (TEST INFRASTRUCTURE ONLY).
 *
 * oracle/Makefile compiles this file together with the reference's own
 * sources, read in place from the Micropp tree, into oracle/_ref/.  The
 * reference sources are never copied into this repository.  A subclass of
 * micropp<3> reaches the protected members, the way the reference's own
 * test/test_cg.cpp:38-82 does.  Every ref_* function has the signature of the
 * orc_* function of the same name in oracle.h.
 */
#include <cstdlib>
#include <cstring>

#include "micropp.hpp"
#include "oracle.h"

namespace {

micropp_params_t to_params(const orc_params *p) {
  micropp_params_t q;
  q.ngp = p->ngp;
  for (int i = 0; i < 3; ++i) q.size[i] = p->size[i];
  q.type = p->type;
  for (int i = 0; i < 4; ++i) q.geo_params[i] = p->geo[i];
  static_assert(sizeof(orc_material) == sizeof(material_base), "material layout");
  for (int i = 0; i < 3; ++i) std::memcpy(&q.materials[i], &p->mat[i], sizeof(material_base));
  q.coupling = const_cast<int *>(p->coupling);
  q.subiterations = p->subiterations != 0;
  q.nsubiterations = p->nsubiterations;
  q.nr_max_its = p->nr_max_its;
  q.nr_max_tol = p->nr_max_tol;
  q.nr_rel_tol = p->nr_rel_tol;
  q.calc_ctan_lin = p->calc_ctan_lin != 0;
  q.use_A0 = p->use_A0 != 0;
  q.its_with_A0 = p->its_with_A0;
  q.lin_stress = p->lin_stress != 0;
  return q;
}

class probe : public micropp<3> {
 public:
  explicit probe(const micropp_params_t &p) : micropp<3>(p) {}
  int nrow() const { return nndim; }
  int ne() const { return nelem; }
  void get_bmat(double *o) const { std::memcpy(o, bmat, sizeof(bmat)); }
  void get_et(int *o) const { std::memcpy(o, elem_type, sizeof(int) * nelem); }
  void bc(const double *eps, double *u) { set_displ_bc(eps, u); }
  double rhs(const double *u, const double *v, double *b) { return assembly_rhs(u, v, b); }
  void mat(const double *u, const double *v, double *vals) {
    ell_matrix A;
    const int ns[3] = {nx, ny, nz};
    ell_init(&A, 3, 3, ns);
    assembly_mat(&A, u, v);
    std::memcpy(vals, A.vals, sizeof(double) * A.nrow * A.nnz);
    ell_free(&A);
  }
  void ave(const double *u, const double *v, double *s) const { calc_ave_stress(u, s, v); }
  int vnew(const double *u, const double *vo, double *vn) const { return calc_vars_new(u, vo, vn) ? 1 : 0; }
  void newton(const double *eps, const double *vo, double *u, double abs_tol, double rel_tol, int max_its,
              int out[3]) {
    ell_matrix A;
    const int ns[3] = {nx, ny, nz};
    ell_init(&A, 3, 3, ns, abs_tol, rel_tol, max_its);
    double *b = (double *)calloc(nndim, sizeof(double));
    double *du = (double *)calloc(nndim, sizeof(double));
    newton_t n = newton_raphson(&A, b, u, du, eps, vo);
    out[0] = n.its;
    out[1] = n.solver_its;
    out[2] = n.converged ? 1 : 0;
    free(b);
    free(du);
    ell_free(&A);
  }
  void vtu(const double *u, const double *v, const char *fn) {
    calc_fields(const_cast<double *>(u), const_cast<double *>(v));
    write_vtu(const_cast<double *>(u), const_cast<double *>(v), fn);
  }
  void ctan_lin(double *o) const { std::memcpy(o, ctan_lin_fe, sizeof(ctan_lin_fe)); }
  void vfrac(double *o) const {
    o[0] = Vm;
    o[1] = Vf;
  }
  int state(int g, int which, double *out) const {
    const gp_t<3> &q = gp_list[g];
    const double *src = which == 0 ? q.u_n : which == 1 ? q.u_k : which == 2 ? q.vars_n : q.vars_k;
    const int n = which < 2 ? nndim : nvars;
    if (!src) return 0;
    std::memcpy(out, src, sizeof(double) * n);
    return 1;
  }
};

probe *mk(const orc_params *p) { return new probe(to_params(p)); }

// an instance that writes the profiling log (write_log, src/output.cpp:264-282)
// as MPI rank `rank`; its files go to the current directory
probe *mk_log(const orc_params *p, int rank, int write_log) {
  micropp_params_t q = to_params(p);
  q.mpi_rank = rank;
  q.write_log = write_log != 0;
  return new probe(q);
}

}  // namespace

extern "C" {

void ref_mat_stress(const orc_material *m, const double eps[6], const double *hist, double sig[6]) {
  material_base b;
  std::memcpy(&b, m, sizeof(b));
  material_t *mt = material_t::make_material(b);
  mt->get_stress(eps, sig, hist);
  delete mt;
}

void ref_mat_ctan(const orc_material *m, const double eps[6], const double *hist, double ctan[36]) {
  material_base b;
  std::memcpy(&b, m, sizeof(b));
  material_t *mt = material_t::make_material(b);
  mt->get_ctan(eps, ctan, hist);
  delete mt;
}

int ref_mat_evolute(const orc_material *m, const double eps[6], const double *hist_old, double *hist_new) {
  material_base b;
  std::memcpy(&b, m, sizeof(b));
  material_t *mt = material_t::make_material(b);
  const int r = mt->evolute(eps, hist_old, hist_new) ? 1 : 0;
  delete mt;
  return r;
}

void ref_ell_cols3d(int nx, int ny, int nz, int *cols) {
  ell_matrix A;
  const int ns[3] = {nx, ny, nz};
  ell_init(&A, 3, 3, ns);
  std::memcpy(cols, A.cols, sizeof(int) * A.nrow * A.nnz);
  ell_free(&A);
}

void ref_ell_mvp(int nrow, int nnz, const int *cols, const double *vals, const double *x, double *y) {
  ell_matrix A;
  A.nrow = nrow;
  A.nnz = nnz;
  A.cols = const_cast<int *>(cols);
  A.vals = const_cast<double *>(vals);
  ell_mvp(&A, x, y);
}

int ref_ell_cg(int nn, int nfield, int nnz, int shift, const int *cols, const double *vals, const double *b, double *x,
               double min_err, double rel_err, int max_its, double *err) {
  ell_matrix A;
  A.nn = nn;
  A.nfield = nfield;
  A.nnz = nnz;
  A.shift = shift;
  A.nrow = nn * nfield;
  A.ncol = A.nrow;
  A.cols = const_cast<int *>(cols);
  A.vals = const_cast<double *>(vals);
  A.min_err = min_err;
  A.rel_err = rel_err;
  A.max_its = max_its;
  A.k = (double *)malloc(sizeof(double) * A.nrow);
  A.r = (double *)malloc(sizeof(double) * A.nrow);
  A.z = (double *)malloc(sizeof(double) * A.nrow);
  A.p = (double *)malloc(sizeof(double) * A.nrow);
  A.Ap = (double *)malloc(sizeof(double) * A.nrow);
  const int its = ell_solve_cgpd(&A, b, x, err);
  free(A.k);
  free(A.r);
  free(A.z);
  free(A.p);
  free(A.Ap);
  return its;
}

/* 2-D, nfield 1: ell_init + ell_add_2D over all elements + ell_set_bc_2D,
 * exactly the sequence of test/test_ell_1.cpp:53-79. */
void ref_ell2d_build(int nx, int ny, const double Ae[16], int *cols, double *vals) {
  ell_matrix A;
  const int ns[3] = {nx, ny, 0};
  ell_init(&A, 1, 2, ns, 1.0e-5, 1.0e-5, 50);
  ell_set_zero_mat(&A);
  for (int ex = 0; ex < nx - 1; ++ex)
    for (int ey = 0; ey < ny - 1; ++ey) ell_add_2D(&A, ex, ey, Ae);
  ell_set_bc_2D(&A);
  std::memcpy(cols, A.cols, sizeof(int) * A.nrow * A.nnz);
  std::memcpy(vals, A.vals, sizeof(double) * A.nrow * A.nnz);
  ell_free(&A);
}

void *ref_micro_new(const orc_params *p) { return mk(p); }
void *ref_micro_new_rank(const orc_params *p, int rank, int write_log) { return mk_log(p, rank, write_log); }
/* micropp<3>::write_restart / read_restart (src/output.cpp:216-262): the
 * file micropp-restart-<rank>-<id>.bin in the current directory */
void ref_write_restart(void *m, int id) { static_cast<probe *>(m)->write_restart(id); }
void ref_read_restart(void *m, int id) { static_cast<probe *>(m)->read_restart(id); }
void ref_micro_free(void *m) { delete static_cast<probe *>(m); }
int ref_nrow(void *m) { return static_cast<probe *>(m)->nrow(); }
int ref_nelem(void *m) { return static_cast<probe *>(m)->ne(); }
void ref_bmat(void *m, double *o) { static_cast<probe *>(m)->get_bmat(o); }
void ref_elem_types(void *m, int *o) { static_cast<probe *>(m)->get_et(o); }
void ref_set_displ_bc(void *m, const double eps[6], double *u) { static_cast<probe *>(m)->bc(eps, u); }
double ref_assembly_rhs(void *m, const double *u, const double *v, double *b) {
  return static_cast<probe *>(m)->rhs(u, v, b);
}
void ref_assembly_mat(void *m, const double *u, const double *v, double *vals) {
  static_cast<probe *>(m)->mat(u, v, vals);
}
void ref_ave_stress(void *m, const double *u, const double *v, double sig[6]) {
  static_cast<probe *>(m)->ave(u, v, sig);
}
int ref_write_vtu(void *m, const double *u, const double *v, const char *fn) {
  static_cast<probe *>(m)->vtu(u, v, fn);
  return 0;
}
int ref_vars_new(void *m, const double *u, const double *vo, double *vn) {
  return static_cast<probe *>(m)->vnew(u, vo, vn);
}
void ref_newton(void *m, const double eps[6], const double *vo, double *u, double abs_tol, double rel_tol,
                int max_its, int out[3]) {
  static_cast<probe *>(m)->newton(eps, vo, u, abs_tol, rel_tol, max_its, out);
}
void ref_set_strain(void *m, int gp, const double eps[6]) { static_cast<probe *>(m)->set_strain(gp, eps); }
void ref_get_stress(void *m, int gp, double sig[6]) { static_cast<probe *>(m)->get_stress(gp, sig); }
void ref_get_ctan(void *m, int gp, double c[36]) { static_cast<probe *>(m)->get_ctan(gp, c); }
void ref_homogenize(void *m) { static_cast<probe *>(m)->homogenize(); }
void ref_homogenize_linear(void *m) { static_cast<probe *>(m)->homogenize_linear(); }
void ref_update_vars(void *m) { static_cast<probe *>(m)->update_vars(); }
int ref_is_non_linear(void *m, int gp) { return static_cast<probe *>(m)->is_non_linear(gp); }
int ref_get_non_linear_gps(void *m) { return static_cast<probe *>(m)->get_non_linear_gps(); }
long ref_get_cost(void *m, int gp) { return static_cast<probe *>(m)->get_cost(gp); }
int ref_has_converged(void *m, int gp) { return static_cast<probe *>(m)->has_converged(gp) ? 1 : 0; }
int ref_has_subiterated(void *m, int gp) { return static_cast<probe *>(m)->has_subiterated(gp) ? 1 : 0; }
void ref_get_ctan_lin(void *m, double c[36]) { static_cast<probe *>(m)->ctan_lin(c); }
void ref_get_volume_fractions(void *m, double vf[2]) { static_cast<probe *>(m)->vfrac(vf); }
int ref_get_state(void *m, int gp, int which, double *out) { return static_cast<probe *>(m)->state(gp, which, out); }
int ref_nthreads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
}
