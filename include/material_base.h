/*
 * material_base.h -- material description shared by the C, C++ and Fortran
 * interfaces of micropp_amd.
 *
 * Drop-in for the reference include/material_base.h:27-53: same enum values,
 * same constants, same struct field order (the Fortran bind(C) mirror in
 * src/material.f95:30-35 depends on it) and the same material_set signature.
 */
#ifndef MICROPP_AMD_MATERIAL_BASE_H
#define MICROPP_AMD_MATERIAL_BASE_H

/* material kinds (reference include/material_base.h:27) */
enum { MATERIAL_ELASTIC = 0, MATERIAL_PLASTIC, MATERIAL_DAMAGE };

/* forward-difference step of the material tangent (reference :29) */
#define D_EPS_CTAN 1.0e-8
/* truncated sqrt(2/3) of the J2 law (reference :30) */
#define SQRT_2DIV3 0.816496581

#ifdef __cplusplus
#include <cstdio>
extern "C" {
#else
#include <stdbool.h>
#include <stdio.h>
#endif

/* reference include/material_base.h:40-45 */
struct material_base {
  double E, nu, Ka, Sy;
  double k, mu, lambda;
  double Xt;
  int type;
};

/* Fills a material and its derived constants k, mu, lambda
 * (replaces reference src/material.c:26-38). */
void material_set(struct material_base *self, const int _type, const double E, const double nu, const double Ka,
                  const double Sy, const double Xt);

/* Prints the material (replaces reference src/material.c:40-45). */
void material_print(const struct material_base *self);

#ifdef __cplusplus
}
#endif
#endif
