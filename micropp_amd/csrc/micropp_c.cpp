// micropp_c.cpp -- the C interface (include/micropp_c.h), replacing the
// reference src/micropp_c.cpp:29-142 entry for entry, and material_set /
// material_print (src/material.c:26-45).
#include "micropp_c.h"

#include <cstdio>
#include <cstring>
#include <exception>

#include "micropp.hpp"

extern "C" {

void material_set(struct material_base *self, const int _type, const double _E, const double _nu, const double _Ka,
                  const double _Sy, const double _Xt) {
  self->E = _E;
  self->nu = _nu;
  self->Ka = _Ka;
  self->Sy = _Sy;
  self->Xt = _Xt;
  self->type = _type;
  self->k = _E / (3. * (1. - 2. * _nu));
  self->mu = _E / (2. * (1. + _nu));
  self->lambda = _nu * _E / ((1. + _nu) * (1. - 2. * _nu));
}

void material_print(const struct material_base *self) {
  printf("E = %e\nnu = %e\nKa = %e\nSy = %e\nXt = %e\nk = %e\nmu = %e\nlambda = %e\ntype = %1d\n", self->E, self->nu,
         self->Ka, self->Sy, self->Xt, self->k, self->mu, self->lambda, self->type);
}

// micropp3_new (src/micropp_c.cpp:33-56): subiterations on, A0 and the log
// off, as in the reference.  A failure leaves self->ptr == NULL and prints the
// reason (the reference has no error channel here).
void micropp3_new(struct micropp3 *self, int ngp, const int size[3], const int type, const double *geo_params,
                  const struct material_base *materials, const int *coupling, const int nsubiterations,
                  const int mpi_rank) {
  micropp_params_t params;
  params.ngp = ngp;
  memcpy(params.size, size, 3 * sizeof(int));
  params.type = type;
  memcpy(params.geo_params, geo_params, 4 * sizeof(double));
  for (int i = 0; i < MAX_MATERIALS; ++i) params.materials[i] = materials[i];
  int *cpl = nullptr;
  if (coupling) {
    cpl = new int[ngp];
    memcpy(cpl, coupling, ngp * sizeof(int));
  }
  params.coupling = cpl;
  params.subiterations = true;
  params.nsubiterations = nsubiterations;
  params.mpi_rank = mpi_rank;
  params.use_A0 = false;
  params.its_with_A0 = 1;
  params.write_log = false;
  self->ptr = nullptr;
  try {
    self->ptr = new micropp<3>(params);
  } catch (const std::exception &e) {
    fprintf(stderr, "micropp3_new: %s\n", e.what());
  }
  delete[] cpl;
}

#define M3(self) (static_cast<micropp<3> *>((self)->ptr))

void micropp3_free(struct micropp3 *self) {
  delete M3(self);
  self->ptr = nullptr;
}
void micropp3_set_strain(struct micropp3 *self, const int gp_id, const double *strain) { M3(self)->set_strain(gp_id, strain); }
void micropp3_get_stress(const struct micropp3 *self, const int gp_id, double *stress) { M3(self)->get_stress(gp_id, stress); }
void micropp3_get_ctan(const struct micropp3 *self, const int gp_id, double *ctan) { M3(self)->get_ctan(gp_id, ctan); }
void micropp3_homogenize(struct micropp3 *self) { M3(self)->homogenize(); }
void micropp3_homogenize_linear(struct micropp3 *self) { M3(self)->homogenize_linear(); }
void micropp3_update_vars(struct micropp3 *self) { M3(self)->update_vars(); }
bool micropp3_is_non_linear(const struct micropp3 *self, const int gp_id) { return M3(self)->is_non_linear(gp_id); }
int micropp3_get_cost(const struct micropp3 *self, int gp_id) { return M3(self)->get_cost(gp_id); }
bool micropp3_has_converged(const struct micropp3 *self, int gp_id) { return M3(self)->has_converged(gp_id); }
bool micropp3_has_subiterated(const struct micropp3 *self, int gp_id) { return M3(self)->has_subiterated(gp_id); }
void micropp3_output(struct micropp3 *self, const int gp_id, const char *filename) { M3(self)->output(gp_id, filename); }
void micropp3_output2(struct micropp3 *self, const int gp_id, const int elem_global, const int time_step) {
  M3(self)->output2(gp_id, elem_global, time_step);
}
void micropp3_print_info(struct micropp3 *self) { M3(self)->print_info(); }
int micropp3_get_non_linear_gps(const struct micropp3 *self) { return M3(self)->get_non_linear_gps(); }
void micropp3_write_restart(const struct micropp3 *self, const int restart_id) { M3(self)->write_restart(restart_id); }
void micropp3_read_restart(const struct micropp3 *self, const int restart_id) { M3(self)->read_restart(restart_id); }

void micropp3_params_default(struct micropp3_params *p) {
  const micropp_params_t d{};
  memset(p, 0, sizeof(*p));
  p->ngp = d.ngp;
  p->type = d.type;
  for (int i = 0; i < 4; ++i) p->geo_params[i] = d.geo_params[i];
  p->coupling = nullptr;
  p->subiterations = d.subiterations;
  p->nsubiterations = d.nsubiterations;
  p->mpi_rank = d.mpi_rank;
  p->nr_max_its = d.nr_max_its;
  p->nr_max_tol = d.nr_max_tol;
  p->nr_rel_tol = d.nr_rel_tol;
  p->cg_max_its = d.cg_max_its;
  p->cg_abs_tol = d.cg_abs_tol;
  p->cg_rel_tol = d.cg_rel_tol;
  p->calc_ctan_lin = d.calc_ctan_lin;
  p->use_A0 = d.use_A0;
  p->its_with_A0 = d.its_with_A0;
  p->lin_stress = d.lin_stress;
  p->write_log = d.write_log;
  p->mixed_precision = micropp_options_t{}.mixed_precision;
}

void micropp3_new_ext(struct micropp3 *self, const struct micropp3_params *p) {
  micropp_params_t params;
  params.ngp = p->ngp;
  memcpy(params.size, p->size, 3 * sizeof(int));
  params.type = p->type;
  memcpy(params.geo_params, p->geo_params, 4 * sizeof(double));
  for (int i = 0; i < MAX_MATERIALS; ++i) params.materials[i] = p->materials[i];
  params.coupling = const_cast<int *>(p->coupling);
  params.subiterations = p->subiterations != 0;
  params.nsubiterations = p->nsubiterations;
  params.mpi_rank = p->mpi_rank;
  params.nr_max_its = p->nr_max_its;
  params.nr_max_tol = p->nr_max_tol;
  params.nr_rel_tol = p->nr_rel_tol;
  params.cg_max_its = p->cg_max_its;
  params.cg_abs_tol = p->cg_abs_tol;
  params.cg_rel_tol = p->cg_rel_tol;
  params.calc_ctan_lin = p->calc_ctan_lin != 0;
  params.use_A0 = p->use_A0 != 0;
  params.its_with_A0 = p->its_with_A0;
  params.lin_stress = p->lin_stress != 0;
  params.write_log = p->write_log != 0;
  micropp_options_t options;
  options.mixed_precision = p->mixed_precision != 0;
  self->ptr = nullptr;
  try {
    self->ptr = new micropp<3>(params, options);
  } catch (const std::exception &e) {
    fprintf(stderr, "micropp3_new_ext: %s\n", e.what());
  }
}

int micropp3_get_state(const struct micropp3 *self, const int gp_id, const int which, double *out) {
  return M3(self)->get_state(gp_id, which, out) ? 1 : 0;
}
void micropp3_get_ctan_lin(const struct micropp3 *self, double *ctan) { M3(self)->get_ctan_lin(ctan); }
void micropp3_get_volume_fractions(const struct micropp3 *self, double *vf) { M3(self)->get_volume_fractions(vf); }

}  // extern "C"
