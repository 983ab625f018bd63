/* c_driver.c -- a plain C caller of the micropp3_* interface
 * (include/micropp_c.h), the way a macro code written in C binds it: one
 * instance with two Gauss points (FE_FULL and FE_ONE_WAY) of a 6x6x6-node
 * sphere RVE with a J2-plastic matrix, three load steps.  For every step and
 * GP it prints the flags, the stress and the ctan as text. */
#include <stdio.h>

#include "micropp_c.h"

enum { NGP = 2, STEPS = 3 };

int main(void) {
  const int size[3] = {6, 6, 6};
  const double geo[4] = {0.25, 0.0, 0.0, 0.0};
  const int coupling[NGP] = {2 /* FE_FULL */, 1 /* FE_ONE_WAY */};
  struct material_base mats[3];
  material_set(&mats[0], MATERIAL_PLASTIC, 1.0e6, 0.3, 1.0e5, 1.0e3, 0.0);
  material_set(&mats[1], MATERIAL_ELASTIC, 1.0e7, 0.25, 0.0, 0.0, 0.0);
  material_set(&mats[2], MATERIAL_ELASTIC, 1.0e7, 0.25, 0.0, 0.0, 0.0);
  struct micropp3 m;
  micropp3_new(&m, NGP, size, 1 /* MIC_SPHERE */, geo, mats, coupling, 4, 0);
  if (!m.ptr) return 1;
  for (int t = 0; t < STEPS; ++t) {
    for (int g = 0; g < NGP; ++g) {
      double eps[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      eps[0] = 1.0e-3 * (t + 1) * (g + 1);
      eps[3] = -4.0e-4 * (t + 1);
      micropp3_set_strain(&m, g, eps);
    }
    micropp3_homogenize(&m);
    for (int g = 0; g < NGP; ++g) {
      double sig[6], ctan[36];
      micropp3_get_stress(&m, g, sig);
      micropp3_get_ctan(&m, g, ctan);
      printf("flags %d %d %d %d %d %d\n", t, g, (int)micropp3_is_non_linear(&m, g), micropp3_get_cost(&m, g),
             (int)micropp3_has_converged(&m, g), (int)micropp3_has_subiterated(&m, g));
      printf("sig");
      for (int i = 0; i < 6; ++i) printf(" %.17e", sig[i]);
      printf("\nctan");
      for (int i = 0; i < 36; ++i) printf(" %.17e", ctan[i]);
      printf("\n");
    }
    micropp3_update_vars(&m);
  }
  micropp3_free(&m);
  return 0;
}
