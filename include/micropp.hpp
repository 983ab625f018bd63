/*
 * micropp.hpp -- C++ interface of micropp_amd: the micropp<3> RVE
 * homogenization class whose finite-element solves run on an AMD Instinct
 * MI355X (gfx950).
 *
 * Source-compatible drop-in for the reference include/micropp.hpp:54-218,
 * include/types.hpp:29-117, include/params.hpp:26-41 and the CG constants of
 * include/ell.hpp:33-35: the same class template, public member functions,
 * parameter struct (field order and defaults), enums and macros, and the same
 * set of standard headers plus `using namespace std`, which the reference's
 * drivers rely on.  Only dimension 3 exists, as in the reference
 * (src/micropp.cpp:629).
 *
 * Behaviour that differs on purpose:
 *  - cg_max_its / cg_abs_tol / cg_rel_tol are honoured (the reference always
 *    uses CG_MAX_ITS / CG_ABS_TOL / CG_REL_TOL, which are still the defaults);
 *  - an invalid micro_type throws std::invalid_argument instead of indexing
 *    material_list[-1] (reference src/micropp.cpp:542-543);
 *  - micro_type MIC3D_RANDOM_SPHERES (13) is an addition: a seeded list of
 *    spheres, geo_params = {seed, count, r_min, r_max}.
 */
#pragma once

#include <cassert>
#include <cmath>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "material_base.h"

using namespace std;

/* reference include/params.hpp:26-41 */
#define MAX_DIM 3
#define NUM_VAR_GP 7
#define MAX_MATERIALS 3
#define FILTER_REL_TOL 1.0e-5
#define D_EPS_CTAN_AVE 1.0e-8
#define CONSTXG 0.577350269189626
#define NR_MAX_TOL 1.0e-10
#define NR_MAX_ITS 4
#define NR_REL_TOL 1.0e-3

/* reference include/ell.hpp:33-35 */
#define CG_ABS_TOL 1.0e-50
#define CG_MAX_ITS 1000
#define CG_REL_TOL 1.0e-5

/* reference include/types.hpp:29-41 */
typedef struct {
  int its = 0;
  int solver_its = 0;
  bool converged = false;

  void print() {
    cout << "newton.its        : " << its << endl;
    cout << "newton.solver_its : " << solver_its << endl;
    cout << "newton.converged  : " << converged << endl;
  }
} newton_t;

/* reference include/types.hpp:43-83 (same fields, order and defaults) */
typedef struct {
  int ngp = 1;
  int size[3];
  int type = 0;
  double geo_params[4] = {0.1, 0.1, 0.1, 0.1};
  struct material_base materials[4];
  int *coupling = nullptr;
  bool subiterations = false;
  int nsubiterations = 10;
  int mpi_rank = 0;
  int nr_max_its = NR_MAX_ITS;
  double nr_max_tol = NR_MAX_TOL;
  double nr_rel_tol = NR_REL_TOL;
  int cg_max_its = CG_MAX_ITS;
  double cg_abs_tol = CG_ABS_TOL;
  double cg_rel_tol = CG_REL_TOL;
  bool calc_ctan_lin = true;
  bool use_A0 = false;
  int its_with_A0 = 1;
  bool lin_stress = true;
  bool write_log = false;

  void print() {
    cout << "ngp  : " << ngp << endl;
    cout << "size : " << size[0] << endl;
    cout << "type  : " << type << endl;
    cout << "geo_params : " << geo_params[0] << endl;
    cout << "subiterations : " << subiterations << endl;
    cout << "nsubiterations : " << nsubiterations << endl;
    cout << "mpi_rank : " << mpi_rank << endl;
    cout << "nr_max_its : " << nr_max_its << endl;
    cout << "nr_max_tol : " << nr_max_tol << endl;
    cout << "nr_rel_tol : " << nr_rel_tol << endl;
    cout << "calc_ctan_lin : " << calc_ctan_lin << endl;
    cout << "use_A0 : " << use_A0 << endl;
    cout << "its_with_A0 : " << its_with_A0 << endl;
    cout << "lin_stress : " << lin_stress << endl;
    cout << "write_log : " << write_log << endl;
  }
} micropp_params_t;

/* Additive (no reference counterpart): choices of this implementation that
 * micropp_params_t, kept to the reference's layout, does not carry.  An
 * instance made with the reference's constructor takes them from the
 * environment (MICROPP_MIXED_PRECISION=1). */
typedef struct {
  /* every Newton linear solve as FP64 defect correction around FP32
   * Jacobi-PCG, to the reference CG's metric evaluated in FP64; the residual,
   * return mapping, damage state, u and the averages stay FP64.  An RVE whose
   * phases are all linear elastic solves the constructor's six ctan problems
   * in one multi-right-hand-side pass and the FE Gauss points of homogenize
   * in passes of up to 16; other RVEs solve their GPs as a batch (use_A0 is
   * applied in both). */
  bool mixed_precision = false;
} micropp_options_t;

/* reference include/types.hpp:85-99, plus the seeded sphere list */
enum {
  MIC_HOMOGENEOUS,
  MIC_SPHERE,
  MIC_LAYER_Y,
  MIC_CILI_FIB_X,
  MIC_CILI_FIB_Z,
  MIC_CILI_FIB_XZ,
  MIC_QUAD_FIB_XYZ,
  MIC_QUAD_FIB_XZ,
  MIC_QUAD_FIB_XZ_BROKEN_X,
  MIC3D_SPHERES,
  MIC3D_8,
  MIC3D_FIBS_20_ORDER,
  MIC3D_FIBS_20_DISORDER,
  MIC3D_RANDOM_SPHERES
};

/* reference include/types.hpp:101-113 */
static map<int, std::string> micro_names = {{MIC_HOMOGENEOUS, "MIC_HOMOGENEOUS"},
                                            {MIC_SPHERE, "MIC_SPHERE"},
                                            {MIC_LAYER_Y, "MIC_LAYER_Y"},
                                            {MIC_CILI_FIB_X, "MIC_CILI_FIB_X"},
                                            {MIC_CILI_FIB_Z, "MIC_CILI_FIB_Z"},
                                            {MIC_CILI_FIB_XZ, "MIC_CILI_FIB_XZ"},
                                            {MIC_QUAD_FIB_XYZ, "MIC_QUAD_FIB_XYZ"},
                                            {MIC_QUAD_FIB_XZ, "MIC_QUAD_FIB_XZ"},
                                            {MIC_QUAD_FIB_XZ_BROKEN_X, "MIC_QUAD_FIB_XZ_BROKEN_X"},
                                            {MIC3D_SPHERES, "MIC3D_SPHERES"},
                                            {MIC3D_8, "MIC3D_8"},
                                            {MIC3D_FIBS_20_ORDER, "MIC3D_FIBS_20_ORDER"},
                                            {MIC3D_FIBS_20_DISORDER, "MIC3D_FIBS_20_DISORDER"},
                                            {MIC3D_RANDOM_SPHERES, "MIC3D_RANDOM_SPHERES"}};

/* reference include/types.hpp:115 */
enum { FE_LINEAR, FE_ONE_WAY, FE_FULL, MIX_RULE_CHAMIS };

/* reference include/micropp.hpp:54-218: public interface of micropp<3>.  The
 * GP state (displacements, internal variables) lives in GPU memory behind an
 * opaque pointer. */
template <int tdim>
class micropp {
  static_assert(tdim == 3, "micropp_amd implements the 3-D RVE only (as the reference instantiates)");

 public:
  micropp() = delete;
  micropp(const micropp_params_t &params);
  /* additive: the reference constructor with this implementation's options */
  micropp(const micropp_params_t &params, const micropp_options_t &options);
  ~micropp();
  micropp(const micropp &) = delete;
  micropp &operator=(const micropp &) = delete;

  /* The most important functions */
  void set_strain(const int gp_id, const double *strain);
  void get_stress(const int gp_id, double *stress) const;
  void get_ctan(const int gp_id, double *ctan) const;
  void homogenize();
  void homogenize_linear();

  /* Extras */
  int is_non_linear(const int gp_id) const;
  int get_non_linear_gps(void) const;
  int get_cost(int gp_id) const;
  bool has_converged(int gp_id) const;
  bool has_subiterated(int gp_id) const;
  void output(int gp_id, const char *filename);
  void output2(const int gp_id, const int elem_global, const int time_step);
  void update_vars();
  void write_restart(const int restart_id) const;
  void read_restart(const int restart_id) const;
  void print_info() const;

  /* Extensions (no reference counterpart): diagnostics of the GP state.
   * which = 0 u_n, 1 u_k, 2 vars_n, 3 vars_k; false if the array is absent. */
  bool get_state(int gp_id, int which, double *out) const;
  void get_ctan_lin(double *ctan) const;
  void get_volume_fractions(double *vf) const;

  /* Implementation state (GPU buffers, per-GP host mirrors). */
  struct impl;

 private:
  impl *p_;
};

extern template class micropp<3>;
