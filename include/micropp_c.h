/*
 * micropp_c.h -- C interface of micropp_amd (the FFI boundary the macro-scale
 * code binds: Alya through Fortran bind(C), C drivers directly).
 *
 * Every entry point keeps the name, signature and meaning of the reference
 * include/micropp_c.h:33-70 (implementation src/micropp_c.cpp:29-142); each
 * line below cites the reference declaration it replaces.  The handle owns a
 * micropp<3> whose solves run on the GPU; no device pointer crosses this
 * interface and all arrays are caller-owned host memory.
 */
#ifndef MICROPP_AMD_MICROPP_C_H
#define MICROPP_AMD_MICROPP_C_H

#include "material_base.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference include/micropp_c.h:33-35 */
struct micropp3 {
  void *ptr;
};

/* reference include/micropp_c.h:37-39; as there, subiterations are switched
 * on, use_A0 off, and coupling must hold ngp entries. */
void micropp3_new(struct micropp3 *, int ngp, const int size[3], const int micro_type, const double *micro_params,
                  const struct material_base *materials, const int *coupling, const int nsubiterations,
                  const int mpi_rank);

void micropp3_free(struct micropp3 *in);                                                    /* :41 */
void micropp3_set_strain(struct micropp3 *self, const int gp_id, const double *strain);    /* :43 */
void micropp3_get_stress(const struct micropp3 *self, const int gp_id, double *stress);    /* :45 */
void micropp3_get_ctan(const struct micropp3 *self, const int gp_id, double *ctan);        /* :47 */
void micropp3_homogenize(struct micropp3 *self);                                          /* :49 */
void micropp3_homogenize_linear(struct micropp3 *self);                                   /* :51 */
void micropp3_update_vars(struct micropp3 *self);                                         /* :53 */
bool micropp3_is_non_linear(const struct micropp3 *self, const int gp_id);                /* :55 */
int micropp3_get_cost(const struct micropp3 *self, int gp_id);                            /* :57 */
bool micropp3_has_converged(const struct micropp3 *self, int gp_id);                      /* :59 */
bool micropp3_has_subiterated(const struct micropp3 *self, int gp_id);                    /* :61 */
void micropp3_output(struct micropp3 *self, const int gp_id, const char *filename);       /* :63 */
void micropp3_print_info(struct micropp3 *self);                                          /* :65 */
int micropp3_get_non_linear_gps(const struct micropp3 *self);                             /* :67 */
void micropp3_write_restart(const struct micropp3 *self, const int restart_id);          /* :69 */
void micropp3_read_restart(const struct micropp3 *self, const int restart_id);           /* :70 */
/* defined in reference src/micropp_c.cpp:113-116 and src/micropp.f95:148-156
 * but missing from its header */
void micropp3_output2(struct micropp3 *self, const int gp_id, const int elem_global, const int time_step);

/* ---- extensions (no reference counterpart) ----
 * The whole micropp_params_t (reference include/types.hpp:43-83) as a plain C
 * struct, so C, Fortran and ctypes callers can set what micropp3_new fixes
 * (lin_stress, tolerances, A0, log).  Field meaning and defaults are those of
 * micropp_params_t; booleans are ints. */
struct micropp3_params {
  int ngp;
  int size[3];
  int type;
  double geo_params[4];
  struct material_base materials[3];
  const int *coupling; /* NULL: every GP FE_ONE_WAY */
  int subiterations;
  int nsubiterations;
  int mpi_rank;
  int nr_max_its;
  double nr_max_tol;
  double nr_rel_tol;
  int cg_max_its;
  double cg_abs_tol;
  double cg_rel_tol;
  int calc_ctan_lin;
  int use_A0;
  int its_with_A0;
  int lin_stress;
  int write_log;
  int mixed_precision; /* micropp_options_t::mixed_precision */
};
void micropp3_params_default(struct micropp3_params *p);
/* self->ptr is NULL on failure (message on stderr) */
void micropp3_new_ext(struct micropp3 *self, const struct micropp3_params *p);
/* Diagnostics: which = 0 u_n, 1 u_k (3 nx ny nz doubles), 2 vars_n, 3 vars_k
 * (56 nelem doubles); returns 0 when that array does not exist for the GP. */
int micropp3_get_state(const struct micropp3 *self, const int gp_id, const int which, double *out);
/* ctan of the linear FE models (the constructor's 6 solves), 36 doubles */
void micropp3_get_ctan_lin(const struct micropp3 *self, double *ctan);
/* volume fractions {matrix, inclusion} */
void micropp3_get_volume_fractions(const struct micropp3 *self, double *vf);

#ifdef __cplusplus
}
#endif
#endif
