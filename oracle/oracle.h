/*
 * oracle.h -- CPU restatement of Micropp's RVE homogenization hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / the CPU baseline.
 *
 * Every function restates one reference function with the same loop order,
 * the same constants and the same expression order, so that, compiled with
 * -ffp-contract=off and no fast-math, it reproduces the reference bit for bit.
 * Reference citations are relative to the Micropp source tree.
 *
 * Pinning (what checks each function, besides the compiled reference in
 * oracle/_ref): see tests/test_oracle*.py.
 */
#ifndef MPP_ORACLE_H
#define MPP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order as the reference struct material_base
 * (include/material_base.h:40-45). */
typedef struct {
  double E, nu, Ka, Sy;
  double k, mu, lambda;
  double Xt;
  int type;
} orc_material;

/* Subset of micropp_params_t (include/types.hpp:43-83) that changes the
 * arithmetic of the hot path.  cg_* are honoured here; the reference always
 * uses its compile-time CG constants, which are the defaults. */
typedef struct {
  int ngp;
  int size[3];
  int type;
  double geo[4];
  orc_material mat[3];
  const int *coupling; /* NULL: every GP is FE_ONE_WAY (src/micropp.cpp:92-94) */
  int subiterations;
  int nsubiterations;
  int nr_max_its;
  double nr_max_tol, nr_rel_tol;
  int cg_max_its;
  double cg_abs_tol, cg_rel_tol;
  int calc_ctan_lin;
  int use_A0;
  int its_with_A0;
  int lin_stress;
} orc_params;

void orc_params_default(orc_params *p);

/* ---- constitutive models (src/material.cpp, src/material.c) ---- */
void orc_material_set(orc_material *m, int type, double E, double nu, double Ka, double Sy, double Xt);
void orc_mat_stress(const orc_material *m, const double eps[6], const double *hist, double sig[6]);
void orc_mat_ctan(const orc_material *m, const double eps[6], const double *hist, double ctan[36]);
int orc_mat_evolute(const orc_material *m, const double eps[6], const double *hist_old, double *hist_new);

/* ---- geometry ---- */
void orc_bmat(int nx, int ny, int nz, double bmat[8 * 6 * 24]);
void orc_elem_nodes(int nx, int ny, int ex, int ey, int ez, int n[8]);
int orc_elem_types(int micro_type, const double geo[4], int nx, int ny, int nz, int *et);
int orc_random_spheres(const double geo[4], double *out_xyzr, int max_spheres);

/* ---- ELL (src/ell-common.cpp, src/ell.cpp) ---- */
void orc_ell_cols3d(int nx, int ny, int nz, int *cols);
void orc_ell_cols2d(int nx, int ny, int *cols);
void orc_ell_add2d_all(int nx, int ny, const double Ae[16], double *vals);
void orc_ell_bc2d(int nx, int ny, double *vals);
void orc_ell_mvp(int nrow, int nnz, const int *cols, const double *vals, const double *x, double *y);
int orc_ell_cg(int nn, int nfield, int nnz, int shift, const int *cols, const double *vals, const double *b, double *x,
               double min_err, double rel_err, int max_its, double *err);

/* ---- one RVE: assembly, Newton, averages (src/assembly.cpp, src/solve.cpp, ...) ---- */
typedef struct orc_rve orc_rve;
orc_rve *orc_rve_new(int nx, int ny, int nz, int micro_type, const double geo[4], const orc_material mat[3]);
void orc_rve_free(orc_rve *r);
int orc_rve_elem_type(const orc_rve *r, int e);
void orc_set_displ_bc(const orc_rve *r, const double eps[6], double *u);
double orc_assembly_rhs(const orc_rve *r, const double *u, const double *vars, double *b);
void orc_assembly_mat(const orc_rve *r, const double *u, const double *vars, double *vals);
void orc_ave_stress(const orc_rve *r, const double *u, const double *vars, double sig[6]);
int orc_vars_new(const orc_rve *r, const double *u, const double *vars_old, double *vars_new);
/* calc_fields (src/average.cpp:84-112): nelem x 6 element strain and stress */
void orc_calc_fields(const orc_rve *r, const double *u, const double *vars, double *elem_strain, double *elem_stress);
/* calc_fields + write_vtu (src/output.cpp:67-214) into <filename>.vtu; 0 or -1 */
int orc_write_vtu(const orc_rve *r, const double *u, const double *vars, const char *filename);
/* Newton-Raphson (src/solve.cpp:28-82); out = {its, solver_its, converged}.
 * A0 may be NULL; then use_A0 is ignored. */
void orc_newton(const orc_rve *r, const double eps[6], const double *vars_old, double *u, int nr_max_its,
                double nr_max_tol, double nr_rel_tol, int cg_max_its, double cg_abs_tol, double cg_rel_tol,
                const double *A0_vals, int its_with_A0, int out[3]);

/* ---- the micropp<3> instance (src/micropp.cpp, src/homogenize.cpp) ---- */
typedef struct orc_micro orc_micro;
orc_micro *orc_micro_new(const orc_params *p);
void orc_micro_free(orc_micro *m);
void orc_set_strain(orc_micro *m, int gp, const double eps[6]);
void orc_get_stress(const orc_micro *m, int gp, double sig[6]);
void orc_get_ctan(const orc_micro *m, int gp, double ctan[36]);
void orc_homogenize(orc_micro *m);
void orc_homogenize_linear(orc_micro *m);
void orc_update_vars(orc_micro *m);
int orc_is_non_linear(const orc_micro *m, int gp);
int orc_get_non_linear_gps(const orc_micro *m);
long orc_get_cost(const orc_micro *m, int gp);
int orc_has_converged(const orc_micro *m, int gp);
int orc_has_subiterated(const orc_micro *m, int gp);
void orc_get_ctan_lin(const orc_micro *m, double ctan[36]);
void orc_get_volume_fractions(const orc_micro *m, double vf[2]);
/* which: 0 u_n, 1 u_k, 2 vars_n, 3 vars_k; returns 0 if the array is absent */
int orc_get_state(const orc_micro *m, int gp, int which, double *out);
int orc_nthreads(void);

#ifdef __cplusplus
}
#endif
#endif
