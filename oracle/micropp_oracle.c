/*
 * micropp_oracle.c -- CPU restatement of Micropp's RVE homogenization hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline).  The product
 * library never links or calls this file; its kernels are in
 * micropp_amd/csrc.  Build: oracle/Makefile (gcc -O2 -ffp-contract=off
 * -fopenmp).  Contraction must stay off: the reference is built without FMA
 * and the restatement reproduces it bit for bit.
 *
 * Each function names the reference function it follows (paths relative to
 * the Micropp source tree).  Loop orders, constants and the order of every
 * floating-point expression follow the reference; the code itself is written
 * from scratch for this repository.
 *
 * Pins (checks independent of this file): tests/test_oracle.py compares it
 *   - bit for bit with the compiled reference (oracle/_ref) for cols, phase
 *     maps, bmat, vals, rhs, CG iterates, sigma and ctan;
 *   - with the golden tables in the reference's own tests (benchmark-elastic,
 *     -plastic, -damage; test_ell_1; test_get_elem_nodes);
 *   - with closed forms (homogeneous patch test, symmetry, Hill bounds).
 * The random-sphere micro type (MIC3D_RANDOM_SPHERES) has no reference
 * counterpart; it is pinned by an independent centroid computation in the
 * tests.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Compile-time constants of the reference (include/params.hpp:26-41,
 * include/material_base.h:29-30, include/ell.hpp:33-35). */
#define ORC_NVAR 7
#define ORC_D_EPS_CTAN_AVE 1.0e-8
#define ORC_GAUSS 0.577350269189626
#define ORC_D_EPS_CTAN 1.0e-8
#define ORC_SQRT_2DIV3 0.816496581

enum { ORC_FE_LINEAR = 0, ORC_FE_ONE_WAY = 1, ORC_FE_FULL = 2, ORC_MIX_RULE_CHAMIS = 3 };
enum {
  ORC_MIC_HOMOGENEOUS = 0,
  ORC_MIC_SPHERE,
  ORC_MIC_LAYER_Y,
  ORC_MIC_CILI_FIB_X,
  ORC_MIC_CILI_FIB_Z,
  ORC_MIC_CILI_FIB_XZ,
  ORC_MIC_QUAD_FIB_XYZ,
  ORC_MIC_QUAD_FIB_XZ,
  ORC_MIC_QUAD_FIB_XZ_BROKEN_X,
  ORC_MIC3D_SPHERES,
  ORC_MIC3D_8,
  ORC_MIC3D_FIBS_20_ORDER,
  ORC_MIC3D_FIBS_20_DISORDER,
  ORC_MIC3D_RANDOM_SPHERES /* extension: seeded sphere list, no reference counterpart */
};

void orc_params_default(orc_params *p) {
  memset(p, 0, sizeof(*p));
  p->ngp = 1;
  p->geo[0] = p->geo[1] = p->geo[2] = p->geo[3] = 0.1;
  p->coupling = NULL;
  p->subiterations = 0;
  p->nsubiterations = 10;
  p->nr_max_its = 4;
  p->nr_max_tol = 1.0e-10;
  p->nr_rel_tol = 1.0e-3;
  p->cg_max_its = 1000;
  p->cg_abs_tol = 1.0e-50;
  p->cg_rel_tol = 1.0e-5;
  p->calc_ctan_lin = 1;
  p->use_A0 = 0;
  p->its_with_A0 = 1;
  p->lin_stress = 1;
}

/* ===================================================================== */
/* Constitutive models                                                   */
/* ===================================================================== */

/* material_set (src/material.c:26-38) */
void orc_material_set(orc_material *m, int type, double E, double nu, double Ka, double Sy, double Xt) {
  m->E = E;
  m->nu = nu;
  m->Ka = Ka;
  m->Sy = Sy;
  m->Xt = Xt;
  m->type = type;
  m->k = E / (3. * (1. - 2. * nu));
  m->mu = E / (2. * (1. + nu));
  m->lambda = nu * E / ((1. + nu) * (1. - 2. * nu));
}

/* Linear isotropic stress, shared by material_elastic::get_stress
 * (src/material.cpp:76-82) and the linear trial of damage_law (:242-245). */
static void lin_stress(const orc_material *m, const double *e, double *s) {
  const double tr = e[0] + e[1] + e[2];
  const double two_mu = 2 * m->mu;
  for (int i = 0; i < 3; ++i) s[i] = m->lambda * tr + two_mu * e[i];
  for (int i = 3; i < 6; ++i) s[i] = m->mu * e[i];
}

/* get_dev_tensor (src/material.cpp:65-69) */
static void deviator(const double *t, double *d) {
  const double third_tr = (1 / 3.0) * (t[0] + t[1] + t[2]);
  for (int i = 0; i < 6; ++i) d[i] = t[i];
  for (int i = 0; i < 3; ++i) d[i] -= third_tr;
}

/* material_plastic::plastic_law (src/material.cpp:110-148): J2 radial return
 * trial state with linear isotropic hardening. */
static int j2_law(const orc_material *m, const double *eps, const double *ep_old, const double *alpha_old_p,
                  double *dl, double *nrm, double *s_tr) {
  static const double zero6[6] = {0, 0, 0, 0, 0, 0};
  const double alpha_old = alpha_old_p ? *alpha_old_p : 0;
  const double *ep = ep_old ? ep_old : zero6;
  double ed[6], epd[6];
  deviator(ep, epd);
  deviator(eps, ed);
  const double two_mu = 2 * m->mu;
  for (int i = 0; i < 3; ++i) s_tr[i] = two_mu * (ed[i] - epd[i]);
  for (int i = 3; i < 6; ++i) s_tr[i] = m->mu * (ed[i] - epd[i]);
  double ss = 0.0;
  for (int i = 0; i < 6; ++i) ss += s_tr[i] * s_tr[i];
  const double s_norm = sqrt(ss);
  const double f = s_norm - ORC_SQRT_2DIV3 * (m->Sy + m->Ka * alpha_old);
  if (f > 0) {
    for (int i = 0; i < 6; ++i) nrm[i] = s_tr[i] / s_norm;
    *dl = f / (2. * m->mu * (1. + m->Ka / (3. * m->mu)));
    return 1;
  }
  for (int i = 0; i < 6; ++i) nrm[i] = 0.0;
  *dl = 0;
  return 0;
}

/* material_damage::hardening_law (src/material.cpp:206-225); the hard-coded
 * Ey, H0, H1 are part of the reference law. */
static double damage_hardening(const orc_material *m, double r) {
  const double Ey = 10.0e4;
  const double Ey_inf = 10. * Ey;
  const double H0 = 10.0, H1 = 5.0;
  const double sqE = sqrt(m->E);
  const double r0 = Ey / sqE;
  const double q0 = r0;
  const double q1 = Ey_inf / sqE;
  const double r1 = r0 + (q1 - q0) / H0;
  if (r < r0) return 0.0;
  if (r < r1) return q0 + H0 * (r - r0);
  return q1 + H1 * (r - r1);
}

/* material_damage::damage_law (src/material.cpp:227-269) */
static int damage_law(const orc_material *m, const double *eps, double r_old, double D_old, double *r_out,
                      double *D_out, double *stress) {
  double local[6] = {0, 0, 0, 0, 0, 0};
  double *s = stress ? stress : local;
  lin_stress(m, eps, s);
  double prod = 0.0;
  for (int i = 0; i < 6; ++i) prod += s[i] * eps[i];
  const double r = (prod >= 0) ? sqrt(prod) : 0;
  const double thr = m->Xt / sqrt(m->E);
  const double r_adm = (r_old < thr) ? thr : r_old;
  if (r <= r_adm) {
    *r_out = r_adm;
    *D_out = D_old;
    return 0;
  }
  const double q = damage_hardening(m, r);
  *r_out = r;
  *D_out = 1. - q / r;
  return 1;
}

/* material_*::get_stress (src/material.cpp:76-82, 150-163, 271-287) */
void orc_mat_stress(const orc_material *m, const double eps[6], const double *hist, double sig[6]) {
  if (m->type == 0) {
    lin_stress(m, eps, sig);
  } else if (m->type == 1) {
    double dl, nrm[6], s_tr[6];
    j2_law(m, eps, hist ? hist : NULL, hist ? hist + 6 : NULL, &dl, nrm, s_tr);
    const double tr = eps[0] + eps[1] + eps[2];
    for (int i = 0; i < 6; ++i) sig[i] = s_tr[i];
    for (int i = 0; i < 3; ++i) sig[i] += m->k * tr;
    const double two_mu_dl = 2 * m->mu * dl;
    for (int i = 0; i < 6; ++i) sig[i] -= two_mu_dl * nrm[i];
  } else {
    const double r_old = hist ? hist[0] : 0.0;
    const double D_old = hist ? hist[1] : 0.0;
    double D, r;
    damage_law(m, eps, r_old, D_old, &r, &D, sig);
    for (int i = 0; i < 6; ++i) sig[i] *= (1 - D);
  }
}

/* material_elastic::get_ctan (src/material.cpp:84-95) and the forward
 * perturbation tangent material_t::apply_perturbation (:48-63). */
void orc_mat_ctan(const orc_material *m, const double eps[6], const double *hist, double ctan[36]) {
  if (m->type == 0) {
    for (int i = 0; i < 36; ++i) ctan[i] = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) ctan[i * 6 + j] += m->lambda;
    for (int i = 0; i < 3; ++i) ctan[i * 6 + i] += 2 * m->mu;
    for (int i = 3; i < 6; ++i) ctan[i * 6 + i] = m->mu;
    return;
  }
  double s0[6];
  orc_mat_stress(m, eps, hist, s0);
  for (int c = 0; c < 6; ++c) {
    double e1[6], s1[6];
    for (int i = 0; i < 6; ++i) e1[i] = eps[i];
    e1[c] += ORC_D_EPS_CTAN;
    orc_mat_stress(m, e1, hist, s1);
    for (int r = 0; r < 6; ++r) ctan[r * 6 + c] = (s1[r] - s0[r]) / ORC_D_EPS_CTAN;
  }
}

/* material_*::evolute (src/material.cpp:97-100, 170-186, 289-302).  As in
 * the reference, the plastic law writes nothing when there is no old state. */
int orc_mat_evolute(const orc_material *m, const double eps[6], const double *hist_old, double *hist_new) {
  if (m->type == 0) return 0;
  if (m->type == 1) {
    double dl, nrm[6], s_tr[6];
    const int nl = j2_law(m, eps, hist_old ? hist_old : NULL, hist_old ? hist_old + 6 : NULL, &dl, nrm, s_tr);
    if (hist_old && hist_new)
      for (int i = 0; i < 6; ++i) hist_new[i] = hist_old[i] + dl * nrm[i];
    if (hist_old && hist_new) hist_new[6] = hist_old[6] + ORC_SQRT_2DIV3 * dl + 0;
    return nl;
  }
  const double r_old = hist_old ? hist_old[0] : 0;
  const double D_old = hist_old ? hist_old[1] : 0;
  return damage_law(m, eps, r_old, D_old, &hist_new[0], &hist_new[1], NULL);
}

/* ===================================================================== */
/* Geometry                                                              */
/* ===================================================================== */

/* Natural coordinates of the 8 hex nodes; the Gauss points use the same
 * sign pattern scaled by CONSTXG (include/micropp.hpp:86-89,
 * src/common.cpp:30-41). */
static const int hex_sign[8][3] = {{-1, -1, -1}, {+1, -1, -1}, {+1, +1, -1}, {-1, +1, -1},
                                   {-1, -1, +1}, {+1, -1, +1}, {+1, +1, +1}, {-1, +1, +1}};

/* calc_bmat (src/micro3D.cpp:80-120): trilinear shape-function derivatives at
 * Gauss point gp of a dx*dy*dz brick, Voigt rows [xx,yy,zz,xy,xz,yz]. */
static void bmat_gp(int gp, double dx, double dy, double dz, double b[6][24]) {
  const double xi[3] = {hex_sign[gp][0] * ORC_GAUSS, hex_sign[gp][1] * ORC_GAUSS, hex_sign[gp][2] * ORC_GAUSS};
  for (int a = 0; a < 8; ++a) {
    /* (1 + s_a * x) for each natural direction; for s_a = -1 this is 1 - x */
    double f[3];
    for (int d = 0; d < 3; ++d) f[d] = (hex_sign[a][d] > 0) ? (1 + xi[d]) : (1 - xi[d]);
    double g[3];
    g[0] = hex_sign[a][0] * f[1] * f[2] / 8. * 2. / dx;
    g[1] = hex_sign[a][1] * f[0] * f[2] / 8. * 2. / dy;
    g[2] = hex_sign[a][2] * f[0] * f[1] / 8. * 2. / dz;
    double *c0 = &b[0][a * 3];
    for (int v = 0; v < 6; ++v) b[v][a * 3 + 0] = b[v][a * 3 + 1] = b[v][a * 3 + 2] = 0;
    (void)c0;
    b[0][a * 3 + 0] = g[0];
    b[1][a * 3 + 1] = g[1];
    b[2][a * 3 + 2] = g[2];
    b[3][a * 3 + 0] = g[1];
    b[3][a * 3 + 1] = g[0];
    b[4][a * 3 + 0] = g[2];
    b[4][a * 3 + 2] = g[0];
    b[5][a * 3 + 1] = g[2];
    b[5][a * 3 + 2] = g[1];
  }
}

void orc_bmat(int nx, int ny, int nz, double bmat[8 * 6 * 24]) {
  const double dx = 1.0 / (nx - 1), dy = 1.0 / (ny - 1), dz = 1.0 / (nz - 1);
  for (int gp = 0; gp < 8; ++gp) bmat_gp(gp, dx, dy, dz, (double(*)[24])(bmat + gp * 144));
}

/* norm<double,3> and dot_prod<double,3> (include/util.hpp:78-113) */
static double norm3(const double *v) {
  double s = 0;
  for (int i = 0; i < 3; ++i) s += v[i] * v[i];
  return sqrt(s);
}
static double dot3(const double *a, const double *b) {
  double s = 0;
  for (int i = 0; i < 3; ++i) s += a[i] * b[i];
  return s;
}
/* point_inside_sphere (include/util.hpp:115-124): strict < */
static int in_sphere(const double *c, double r, const double *p) {
  const double v[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
  return norm3(v) < r;
}
/* point_inside_cilinder_inf (include/util.hpp:126-143): distance to the axis
 * through sqrt(1 - cos^2); NaN (point on the axis point) tests false. */
static int in_cylinder(const double *dir, const double *c, double r, const double *p) {
  const double v[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
  const double dv = dot3(dir, v);
  const double nd = norm3(dir);
  const double nv = norm3(v);
  const double cs = dv / (nd * nv);
  const double sn = sqrt(1 - cs * cs);
  const double dist = nv * sn;
  return dist <= r;
}

/* Sphere tables of MIC3D_SPHERES and the fibre tables of MIC3D_FIBS_20_*
 * (src/micropp.cpp:427-457, :505-541): fixed data of the reference
 * micro-structures, reproduced as data. */
static const double sph40_c[40][3] = {
    {.8663, .0689, .1568}, {.2305, .4008, .2093}, {.1987, .8423, .2126}, {.6465, .7095, .4446},
    {.6151, .9673, .2257}, {.1311, .9739, .4129}, {.8433, .0738, .2233}, {.5124, .2111, .0369},
    {.4094, .7030, .6241}, {.1215, .8289, .3812}, {.1125, .9266, .6872}, {.7422, .4030, .6000},
    {.0791, .3819, .9297}, {.1201, .7712, .0069}, {.9339, .2177, .1976}, {.3880, .1331, .9032},
    {.7319, .7146, .8150}, {.8796, .2858, .6701}, {.1427, .9309, .9830}, {.1388, .3126, .8054},
    {.0729, .4754, .4950}, {.5038, .7755, .7333}, {.0242, .2185, .9884}, {.5371, .5372, .8269},
    {.4564, .0039, .1715}, {.7410, .0799, .8775}, {.2627, .9047, .5681}, {.7894, .1908, .2993},
    {.1373, .5370, .9916}, {.0207, .8020, .9283}, {.6540, .2359, .0286}, {.7801, .3568, .5086},
    {.7322, .5706, .1114}, {.5479, .1493, .1267}, {.6722, .1530, .1003}, {.7659, .3426, .9181},
    {.4582, .4636, .6310}, {.1811, .9665, .4713}, {.0834, .6066, .6936}, {.4865, .7584, .3635}};
static const double sph40_rf[40] = {.5741, .1735, .5065, .8565, .9735, .7087, .9585, .2843, .4029, .2574,
                                    .1575, .3316, .1874, .6300, .7049, .7258, .8002, .8176, .9970, .4736,
                                    .6351, .3140, .3440, .2444, .3894, .5234, .1658, .4020, .6813, .2060,
                                    .1427, .9332, .5406, .9765, .0956, .6432, .9998, .4166, .6907, .0404};
static const double fib20_dyz[20][2] = {{.5741, .8515}, {.1735, .1103}, {.5065, .4600}, {.8565, .9045},
                                        {.9735, .6313}, {.7087, .1547}, {.9585, .0220}, {.2843, .4062},
                                        {.4029, .8095}, {.2574, .4742}, {.1575, .0768}, {.3316, .0320},
                                        {.1874, .6364}, {.6300, .2688}, {.7049, .5137}, {.7258, .2799},
                                        {.8002, .4794}, {.8176, .4142}, {.9970, .2189}, {.4736, .6202}};

/* splitmix64 stream (Steele, Lea & Flood 2014) for the seeded sphere list of
 * MIC3D_RANDOM_SPHERES: geo = {seed, count, r_min, r_max}; per sphere the
 * draws are cx, cy, cz ~ U[0,1), r = r_min + (r_max - r_min) * U[0,1). */
static uint64_t sm64_next(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static double sm64_unit(uint64_t *s) { return (double)(sm64_next(s) >> 11) * (1.0 / 9007199254740992.0); }

int orc_random_spheres(const double geo[4], double *out, int max_spheres) {
  uint64_t s = (uint64_t)geo[0];
  int n = (int)geo[1];
  if (n < 0) n = 0;
  if (n > max_spheres) n = max_spheres;
  for (int i = 0; i < n; ++i) {
    out[i * 4 + 0] = sm64_unit(&s);
    out[i * 4 + 1] = sm64_unit(&s);
    out[i * 4 + 2] = sm64_unit(&s);
    out[i * 4 + 3] = geo[2] + (geo[3] - geo[2]) * sm64_unit(&s);
  }
  return n;
}

/* micropp<3>::get_elem_type (src/micropp.cpp:338-544): phase of an element
 * from its centroid.  lx = ly = lz = 1. */
static int phase_of(int type, const double *geo, const double *spheres, int nsph, const double *c) {
  const double L = 1.0;
  const double mid[3] = {L / 2, L / 2, L / 2};
  switch (type) {
    case ORC_MIC_HOMOGENEOUS:
      return 0;
    case ORC_MIC_SPHERE:
      return in_sphere(mid, geo[0], c);
    case ORC_MIC_LAYER_Y:
      return c[1] < geo[0];
    case ORC_MIC_CILI_FIB_X: {
      const double d[3] = {1, 0, 0};
      return in_cylinder(d, mid, geo[0], c);
    }
    case ORC_MIC_CILI_FIB_Z: {
      const double d[3] = {0, 0, 1};
      return in_cylinder(d, mid, geo[0], c);
    }
    case ORC_MIC_CILI_FIB_XZ: {
      const double c1[3] = {L / 2., L * .75, L / 2.};
      const double c2[3] = {L / 2., L * .25, L / 2.};
      const double dx_[3] = {1, 0, 0}, dz_[3] = {0, 0, 1};
      return (in_cylinder(dz_, c1, geo[0], c) || in_cylinder(dx_, c2, geo[0], c)) ? 1 : 0;
    }
    case ORC_MIC_QUAD_FIB_XYZ: {
      const double w = geo[0];
      const int ax = fabs(c[0] - mid[0]) < w, ay = fabs(c[1] - mid[1]) < w, az = fabs(c[2] - mid[2]) < w;
      return ((ax && ay) || (ax && az) || (ay && az)) ? 1 : 0;
    }
    case ORC_MIC_QUAD_FIB_XZ: {
      const double w = geo[0];
      const int ax = fabs(c[0] - mid[0]) < w, ay = fabs(c[1] - mid[1]) < w, az = fabs(c[2] - mid[2]) < w;
      return ((ax && ay) || (ay && az)) ? 1 : 0;
    }
    case ORC_MIC_QUAD_FIB_XZ_BROKEN_X: {
      const double w = geo[0];
      const int ax = fabs(c[0] - mid[0]) < w, ay = fabs(c[1] - mid[1]) < w, az = fabs(c[2] - mid[2]) < w;
      if (ax && ay) return 1;
      return (ay && az && (c[0] < L * .8 || c[0] > L * .9)) ? 1 : 0;
    }
    case ORC_MIC3D_SPHERES:
      for (int i = 0; i < 40; ++i)
        if (in_sphere(sph40_c[i], 0.1 * sph40_rf[i], c)) return 1;
      return 0;
    case ORC_MIC3D_8: {
      const double rc = 0.1, w_flat = 0.02, w_cyl = 0.01;
      const double c1[3] = {L * .25, L * .75, 0.0}, c2[3] = {L * .75, L * .75, 0.0};
      const double c3[3] = {0.0, L * .25, L * .25}, c4[3] = {0.0, L * .25, L * .75};
      const double dx_[3] = {1, 0, 0}, dz_[3] = {0, 0, 1};
      if (in_cylinder(dz_, c1, rc, c) || in_cylinder(dz_, c2, rc, c) || in_cylinder(dx_, c3, rc, c) ||
          in_cylinder(dx_, c4, rc, c))
        return 1;
      if (in_cylinder(dz_, c1, rc + w_cyl, c) || in_cylinder(dz_, c2, rc + w_cyl, c) ||
          in_cylinder(dx_, c3, rc + w_cyl, c) || in_cylinder(dx_, c4, rc + w_cyl, c) || fabs(c[1] - L / 2) < w_flat)
        return 2;
      return 0;
    }
    case ORC_MIC3D_FIBS_20_ORDER: {
      const double d[3] = {1, 0, 0};
      const double hz = 1.0 / (5 + 1), hy = 1.0 / (4 + 1);
      for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 4; ++j) {
          const double cc[3] = {0.0, (j + 1) * hy, (i + 1) * hz};
          if (in_cylinder(d, cc, 0.05, c)) return 1;
        }
      return 0;
    }
    case ORC_MIC3D_FIBS_20_DISORDER:
      for (int i = 0; i < 20; ++i) {
        const double d[3] = {1, fib20_dyz[i][0], fib20_dyz[i][1]};
        if (in_cylinder(d, sph40_c[i], 0.05, c)) return 1;
      }
      return 0;
    case ORC_MIC3D_RANDOM_SPHERES:
      for (int i = 0; i < nsph; ++i)
        if (in_sphere(&spheres[i * 4], spheres[i * 4 + 3], c)) return 1;
      return 0;
    default:
      return -1;
  }
}

int orc_elem_types(int micro_type, const double geo[4], int nx, int ny, int nz, int *et) {
  const int nex = nx - 1, ney = ny - 1, nez = nz - 1;
  const double dx = 1.0 / nex, dy = 1.0 / ney, dz = 1.0 / nez;
  double sph[64 * 4];
  int nsph = 0;
  if (micro_type == ORC_MIC3D_RANDOM_SPHERES) nsph = orc_random_spheres(geo, sph, 64);
  int bad = 0;
  for (int ez = 0; ez < nez; ++ez)
    for (int ey = 0; ey < ney; ++ey)
      for (int ex = 0; ex < nex; ++ex) {
        const double c[3] = {ex * dx + dx / 2., ey * dy + dy / 2., ez * dz + dz / 2.};
        const int t = phase_of(micro_type, geo, sph, nsph, c);
        et[(ez * ney + ey) * nex + ex] = t;
        if (t < 0) bad = 1;
      }
  return bad ? -1 : 0;
}

/* ===================================================================== */
/* ELL matrix                                                            */
/* ===================================================================== */

/* ell_init, dim 3 (src/ell-common.cpp:98-131): row = node*3 + fi, slot =
 * nb*3 + fj, nb = (dz+1)*9 + (dy+1)*3 + (dx+1); a neighbour outside the grid
 * points at column fj (node 0). */
void orc_ell_cols3d(int nx, int ny, int nz, int *cols) {
  const int nxny = nx * ny;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int node = k * nxny + j * nx + i;
        int nbr[27];
        for (int dz = -1; dz <= 1; ++dz)
          for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
              const int in = (i + dx >= 0 && i + dx < nx) && (j + dy >= 0 && j + dy < ny) && (k + dz >= 0 && k + dz < nz);
              nbr[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)] = in ? node + dz * nxny + dy * nx + dx : 0;
            }
        for (int fi = 0; fi < 3; ++fi) {
          int *row = cols + (size_t)(node * 3 + fi) * 81;
          for (int n = 0; n < 27; ++n)
            for (int fj = 0; fj < 3; ++fj) row[n * 3 + fj] = nbr[n] * 3 + fj;
        }
      }
}

/* ell_init, dim 2, nfield 1 (src/ell-common.cpp:71-95) */
void orc_ell_cols2d(int nx, int ny, int *cols) {
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const int node = j * nx + i;
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int in = (i + dx >= 0 && i + dx < nx) && (j + dy >= 0 && j + dy < ny);
          cols[node * 9 + (dy + 1) * 3 + (dx + 1)] = in ? node + dy * nx + dx : 0;
        }
    }
}

/* Local element node positions (src/common.cpp:30-41) and the neighbour
 * slot of local node b seen from local node a (the cols_row tables of
 * ell_add_2D / ell_add_3D, src/ell-common.cpp:147,175-178). */
static const int hex_pos[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
static int nbr_slot3(int a, int b) {
  return (hex_pos[b][2] - hex_pos[a][2] + 1) * 9 + (hex_pos[b][1] - hex_pos[a][1] + 1) * 3 +
         (hex_pos[b][0] - hex_pos[a][0] + 1);
}
static int nbr_slot2(int a, int b) {
  return (hex_pos[b][1] - hex_pos[a][1] + 1) * 3 + (hex_pos[b][0] - hex_pos[a][0] + 1);
}

/* ell_add_2D for every element, ex outer (test/test_ell_1.cpp:76-78,
 * src/ell-common.cpp:141-164), nfield 1. */
void orc_ell_add2d_all(int nx, int ny, const double Ae[16], double *vals) {
  for (int ex = 0; ex < nx - 1; ++ex)
    for (int ey = 0; ey < ny - 1; ++ey) {
      const int n0 = ey * nx + ex;
      const int nodes[4] = {n0, n0 + 1, n0 + nx + 1, n0 + nx};
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) vals[nodes[a] * 9 + nbr_slot2(a, b)] += Ae[a * 4 + b];
    }
}

/* ell_set_bc_2D (src/ell-common.cpp:202-236), nfield 1 */
void orc_ell_bc2d(int nx, int ny, double *vals) {
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      if (!(i == 0 || j == 0 || i == nx - 1 || j == ny - 1)) continue;
      double *row = vals + (j * nx + i) * 9;
      for (int s = 0; s < 9; ++s) row[s] = 0;
      row[4] = 1;
    }
}

/* ell_mvp (src/ell.cpp:35-44): one sequential sum per row */
void orc_ell_mvp(int nrow, int nnz, const int *cols, const double *vals, const double *x, double *y) {
  for (int i = 0; i < nrow; ++i) {
    const int *c = cols + (size_t)i * nnz;
    const double *v = vals + (size_t)i * nnz;
    double acc = 0;
    for (int s = 0; s < nnz; ++s) acc += v[s] * x[c[s]];
    y[i] = acc;
  }
}

/* get_dot (src/ell.cpp:52-56): sequential left-to-right */
static double seq_dot(const double *a, const double *b, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* ell_solve_cgpd (src/ell.cpp:66-122): Jacobi-preconditioned CG from x = 0,
 * stopping test at the top of the loop on ||D^-1 r||; *err = r.z. */
int orc_ell_cg(int nn, int nfield, int nnz, int shift, const int *cols, const double *vals, const double *b, double *x,
               double min_err, double rel_err, int max_its, double *err) {
  const int n = nn * nfield;
  double *kinv = malloc(sizeof(double) * n), *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
  double *p = malloc(sizeof(double) * n), *Ap = malloc(sizeof(double) * n);
  for (int i = 0; i < nn; ++i)
    for (int d = 0; d < nfield; ++d)
      kinv[i * nfield + d] = 1 / vals[(size_t)i * nfield * nnz + shift * nfield + d * nnz + d];
  for (int i = 0; i < n; ++i) x[i] = 0.0;
  orc_ell_mvp(n, nnz, cols, vals, x, r);
  for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
  for (int i = 0; i < n; ++i) z[i] = kinv[i] * r[i];
  for (int i = 0; i < n; ++i) p[i] = z[i];
  double rz = seq_dot(r, z, n);
  const double pn0 = sqrt(seq_dot(z, z, n));
  double pn = pn0;
  int its = 0;
  for (; its < max_its; ++its) {
    if (pn < min_err || pn < pn0 * rel_err) break;
    orc_ell_mvp(n, nnz, cols, vals, p, Ap);
    const double alpha = rz / seq_dot(p, Ap, n);
    for (int i = 0; i < n; ++i) x[i] += alpha * p[i];
    for (int i = 0; i < n; ++i) r[i] -= alpha * Ap[i];
    for (int i = 0; i < n; ++i) z[i] = kinv[i] * r[i];
    pn = sqrt(seq_dot(z, z, n));
    const double rz_new = seq_dot(r, z, n);
    const double beta = rz_new / rz;
    for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    rz = rz_new;
  }
  *err = rz;
  free(kinv);
  free(r);
  free(z);
  free(p);
  free(Ap);
  return its;
}

/* ===================================================================== */
/* One RVE                                                               */
/* ===================================================================== */

struct orc_rve {
  int nx, ny, nz, nn, nndim, nex, ney, nez, nelem;
  double dx, dy, dz, wg, vol_tot, evol;
  double bmat[8][6][24];
  orc_material mat[3];
  int *et;
};

/* make_material (src/material.cpp:26-46) rebuilds the derived constants from
 * E and nu with the constructor expressions (include/material.hpp:69-77). */
static orc_material make_material(const orc_material *in) {
  orc_material m;
  orc_material_set(&m, in->type, in->E, in->nu, in->Ka, in->Sy, in->Xt);
  return m;
}

orc_rve *orc_rve_new(int nx, int ny, int nz, int micro_type, const double geo[4], const orc_material mat[3]) {
  orc_rve *r = calloc(1, sizeof(orc_rve));
  r->nx = nx;
  r->ny = ny;
  r->nz = nz;
  r->nn = nx * ny * nz;
  r->nndim = r->nn * 3;
  r->nex = nx - 1;
  r->ney = ny - 1;
  r->nez = nz - 1;
  r->nelem = r->nex * r->ney * r->nez;
  /* micropp ctor (src/micropp.cpp:29-71): lx = ly = lz = 1 */
  r->dx = 1.0 / r->nex;
  r->dy = 1.0 / r->ney;
  r->dz = 1.0 / r->nez;
  r->wg = r->dx * r->dy * r->dz / 8;
  r->vol_tot = 1.0 * 1.0 * 1.0;
  r->evol = r->dx * r->dy * r->dz;
  for (int gp = 0; gp < 8; ++gp) bmat_gp(gp, r->dx, r->dy, r->dz, r->bmat[gp]);
  for (int i = 0; i < 3; ++i) r->mat[i] = make_material(&mat[i]);
  r->et = malloc(sizeof(int) * r->nelem);
  if (orc_elem_types(micro_type, geo, nx, ny, nz, r->et) != 0) {
    free(r->et);
    free(r);
    return NULL;
  }
  return r;
}

void orc_rve_free(orc_rve *r) {
  if (!r) return;
  free(r->et);
  free(r);
}

int orc_rve_elem_type(const orc_rve *r, int e) { return r->et[e]; }

/* get_elem_nodes (src/common.cpp:28-41) */
static void elem_nodes(const orc_rve *r, int ex, int ey, int ez, int n[8]) {
  const int nxny = r->nx * r->ny;
  const int n0 = ez * nxny + ey * r->nx + ex;
  for (int a = 0; a < 8; ++a) n[a] = n0 + hex_pos[a][2] * nxny + hex_pos[a][1] * r->nx + hex_pos[a][0];
}

/* get_elem_displ + get_strain (src/common.cpp:43-72) */
static void gp_strain(const orc_rve *r, const double *u, int gp, int ex, int ey, int ez, double eps[6]) {
  int n[8];
  elem_nodes(r, ex, ey, ez, n);
  double ue[24];
  for (int a = 0; a < 8; ++a)
    for (int d = 0; d < 3; ++d) ue[a * 3 + d] = u[n[a] * 3 + d];
  for (int v = 0; v < 6; ++v) {
    double s = 0;
    for (int i = 0; i < 24; ++i) s += r->bmat[gp][v][i] * ue[i];
    eps[v] = s;
  }
}

void orc_elem_nodes(int nx, int ny, int ex, int ey, int ez, int n[8]) {
  const int nxny = nx * ny;
  const int n0 = ez * nxny + ey * nx + ex;
  for (int a = 0; a < 8; ++a) n[a] = n0 + hex_pos[a][2] * nxny + hex_pos[a][1] * nx + hex_pos[a][0];
}

static int elem_id(const orc_rve *r, int ex, int ey, int ez) { return (ez * r->ney + ey) * r->nex + ex; }

/* micropp<3>::get_stress (src/micropp.cpp:599-607) */
static void gp_stress(const orc_rve *r, const double *vars, int e, int gp, const double eps[6], double sig[6]) {
  const double *h = vars ? vars + ((size_t)e * 8 + gp) * ORC_NVAR : NULL;
  orc_mat_stress(&r->mat[r->et[e]], eps, h, sig);
}

/* set_displ_bc (src/micro3D.cpp:26-78): affine u = eps_t x on the six faces,
 * in the order z=0, z=lz, y=0, y=ly, x=0, x=lx (later faces overwrite the
 * shared edges and so fix the coordinates those nodes use). */
void orc_set_displ_bc(const orc_rve *r, const double eps[6], double *u) {
  const double et[3][3] = {{eps[0], 0.5 * eps[3], 0.5 * eps[4]},
                           {0.5 * eps[3], eps[1], 0.5 * eps[5]},
                           {0.5 * eps[4], 0.5 * eps[5], eps[2]}};
  const int nx = r->nx, ny = r->ny, nz = r->nz;
  const double L = 1.0;
  for (int face = 0; face < 6; ++face) {
    const int axis = 2 - face / 2; /* faces z, y, x */
    const int hi = face & 1;
    const int na = (axis == 0) ? ny : nx, nb = (axis == 2) ? ny : nz;
    for (int a = 0; a < na; ++a)
      for (int b = 0; b < nb; ++b) {
        int i, j, k;
        double x[3];
        if (axis == 2) {
          i = a;
          j = b;
          k = hi ? nz - 1 : 0;
          x[0] = i * r->dx;
          x[1] = j * r->dy;
          x[2] = hi ? L : 0;
        } else if (axis == 1) {
          i = a;
          k = b;
          j = hi ? ny - 1 : 0;
          x[0] = i * r->dx;
          x[1] = hi ? L : 0;
          x[2] = k * r->dz;
        } else {
          j = a;
          k = b;
          i = hi ? nx - 1 : 0;
          x[0] = hi ? L : 0;
          x[1] = j * r->dy;
          x[2] = k * r->dz;
        }
        double *un = u + ((size_t)(k * ny + j) * nx + i) * 3;
        for (int d = 0; d < 3; ++d) {
          double s = 0.0;
          for (int c = 0; c < 3; ++c) s += et[d][c] * x[c];
          un[d] = s;
        }
      }
  }
}

static int on_boundary(const orc_rve *r, int i, int j, int k) {
  return i == 0 || j == 0 || k == 0 || i == r->nx - 1 || j == r->ny - 1 || k == r->nz - 1;
}

/* get_elem_rhs (src/assembly.cpp:123-138) */
static void elem_rhs(const orc_rve *r, const double *u, const double *vars, int ex, int ey, int ez, double be[24]) {
  const int e = elem_id(r, ex, ey, ez);
  for (int i = 0; i < 24; ++i) be[i] = 0;
  for (int gp = 0; gp < 8; ++gp) {
    double eps[6], sig[6];
    gp_strain(r, u, gp, ex, ey, ez, eps);
    gp_stress(r, vars, e, gp, eps, sig);
    for (int i = 0; i < 24; ++i)
      for (int v = 0; v < 6; ++v) be[i] += r->bmat[gp][v][i] * sig[v] * r->wg;
  }
}

/* assembly_rhs (src/assembly.cpp:27-103): element loop ez, ey, ex; Dirichlet
 * DOFs zeroed; negated; returns ||b||_2. */
double orc_assembly_rhs(const orc_rve *r, const double *u, const double *vars, double *b) {
  for (int i = 0; i < r->nndim; ++i) b[i] = 0.;
  for (int ez = 0; ez < r->nez; ++ez)
    for (int ey = 0; ey < r->ney; ++ey)
      for (int ex = 0; ex < r->nex; ++ex) {
        double be[24];
        int n[8];
        elem_nodes(r, ex, ey, ez, n);
        elem_rhs(r, u, vars, ex, ey, ez, be);
        for (int a = 0; a < 8; ++a)
          for (int d = 0; d < 3; ++d) b[n[a] * 3 + d] += be[a * 3 + d];
      }
  for (int k = 0; k < r->nz; ++k)
    for (int j = 0; j < r->ny; ++j)
      for (int i = 0; i < r->nx; ++i)
        if (on_boundary(r, i, j, k))
          for (int d = 0; d < 3; ++d) b[((k * r->ny + j) * r->nx + i) * 3 + d] = 0.;
  for (int i = 0; i < r->nndim; ++i) b[i] = -b[i];
  double s = 0.0;
  for (int i = 0; i < r->nndim; ++i) s += b[i] * b[i];
  return sqrt(s);
}

/* get_elem_mat (src/assembly.cpp:140-178): Ke = sum_gp B^T C B wg, with the
 * reference's grouping (C B wg first, then B^T times it). */
static void elem_mat(const orc_rve *r, const double *u, const double *vars, int ex, int ey, int ez, double *Ke) {
  const int e = elem_id(r, ex, ey, ez);
  const orc_material *m = &r->mat[r->et[e]];
  for (int i = 0; i < 576; ++i) Ke[i] = 0.0;
  for (int gp = 0; gp < 8; ++gp) {
    double eps[6], C[36], cb[6][24];
    gp_strain(r, u, gp, ex, ey, ez, eps);
    orc_mat_ctan(m, eps, vars ? vars + ((size_t)e * 8 + gp) * ORC_NVAR : NULL, C);
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 24; ++j) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += C[i * 6 + k] * r->bmat[gp][k][j];
        cb[i][j] = s * r->wg;
      }
    for (int v = 0; v < 6; ++v)
      for (int i = 0; i < 24; ++i) {
        const double bvi = r->bmat[gp][v][i];
        for (int j = 0; j < 24; ++j) Ke[i * 24 + j] += bvi * cb[v][j];
      }
  }
}

/* assembly_mat (src/assembly.cpp:105-121) + ell_add_3D
 * (src/ell-common.cpp:166-198) + ell_set_bc_3D (:238-297): element loop ex,
 * ey, ez (ez fastest), so each entry sums its elements in that order. */
void orc_assembly_mat(const orc_rve *r, const double *u, const double *vars, double *vals) {
  const size_t nv = (size_t)r->nndim * 81;
  for (size_t i = 0; i < nv; ++i) vals[i] = 0;
  double Ke[576];
  for (int ex = 0; ex < r->nex; ++ex)
    for (int ey = 0; ey < r->ney; ++ey)
      for (int ez = 0; ez < r->nez; ++ez) {
        int n[8];
        elem_nodes(r, ex, ey, ez, n);
        elem_mat(r, u, vars, ex, ey, ez, Ke);
        for (int a = 0; a < 8; ++a)
          for (int b = 0; b < 8; ++b) {
            const int sl = nbr_slot3(a, b);
            for (int fi = 0; fi < 3; ++fi)
              for (int fj = 0; fj < 3; ++fj)
                vals[((size_t)n[a] * 3 + fi) * 81 + sl * 3 + fj] += Ke[(a * 3 + fi) * 24 + b * 3 + fj];
          }
      }
  for (int k = 0; k < r->nz; ++k)
    for (int j = 0; j < r->ny; ++j)
      for (int i = 0; i < r->nx; ++i)
        if (on_boundary(r, i, j, k))
          for (int d = 0; d < 3; ++d) {
            double *row = vals + ((size_t)((k * r->ny + j) * r->nx + i) * 3 + d) * 81;
            for (int s = 0; s < 81; ++s) row[s] = 0;
            row[13 * 3 + d] = 1;
          }
}

/* calc_ave_stress (src/average.cpp:57-82): per-element partial, then sum in
 * element order ez, ey, ex. */
void orc_ave_stress(const orc_rve *r, const double *u, const double *vars, double sig[6]) {
  for (int v = 0; v < 6; ++v) sig[v] = 0;
  for (int ez = 0; ez < r->nez; ++ez)
    for (int ey = 0; ey < r->ney; ++ey)
      for (int ex = 0; ex < r->nex; ++ex) {
        const int e = elem_id(r, ex, ey, ez);
        double part[6] = {0, 0, 0, 0, 0, 0};
        for (int gp = 0; gp < 8; ++gp) {
          double eps[6], s[6];
          gp_strain(r, u, gp, ex, ey, ez, eps);
          gp_stress(r, vars, e, gp, eps, s);
          for (int v = 0; v < 6; ++v) part[v] += s[v] * r->wg;
        }
        for (int v = 0; v < 6; ++v) sig[v] += part[v];
      }
  for (int v = 0; v < 6; ++v) sig[v] /= r->vol_tot;
}

/* calc_vars_new (src/update.cpp:32-56) */
int orc_vars_new(const orc_rve *r, const double *u, const double *vars_old, double *vars_new) {
  int nl = 0;
  for (int ez = 0; ez < r->nez; ++ez)
    for (int ey = 0; ey < r->ney; ++ey)
      for (int ex = 0; ex < r->nex; ++ex) {
        const int e = elem_id(r, ex, ey, ez);
        const orc_material *m = &r->mat[r->et[e]];
        for (int gp = 0; gp < 8; ++gp) {
          const size_t off = ((size_t)e * 8 + gp) * ORC_NVAR;
          double eps[6];
          gp_strain(r, u, gp, ex, ey, ez, eps);
          nl |= orc_mat_evolute(m, eps, vars_old ? vars_old + off : NULL, vars_new + off);
        }
      }
  return nl;
}

/* calc_fields (src/average.cpp:84-112): per element, the Gauss-point strain
 * and stress summed with weight wg, times ivol = 1 / (wg * 8) (src/micropp.cpp:59). */
void orc_calc_fields(const orc_rve *r, const double *u, const double *vars, double *elem_strain, double *elem_stress) {
  const double ivol = 1.0 / (r->wg * 8);
  for (int ez = 0; ez < r->nez; ++ez)
    for (int ey = 0; ey < r->ney; ++ey)
      for (int ex = 0; ex < r->nex; ++ex) {
        double ea[6] = {0, 0, 0, 0, 0, 0}, sa[6] = {0, 0, 0, 0, 0, 0};
        const int e = elem_id(r, ex, ey, ez);
        for (int gp = 0; gp < 8; ++gp) {
          double eps[6], sig[6];
          gp_strain(r, u, gp, ex, ey, ez, eps);
          gp_stress(r, vars, e, gp, eps, sig);
          for (int v = 0; v < 6; ++v) {
            ea[v] += eps[v] * r->wg;
            sa[v] += sig[v] * r->wg;
          }
        }
        for (int v = 0; v < 6; ++v) {
          elem_strain[e * 6 + v] = ea[v] * ivol;
          elem_stress[e * 6 + v] = sa[v] * ivol;
        }
      }
}

/* One cell array of write_vtu: the mean over the 8 Gauss points of history
 * slot v (or, kind 1, of |eps_p|^2 plus its square root, as the reference's
 * "plasticity" array computes it, src/output.cpp:163-176). */
static double vtu_cell_mean(const double *vars, int e, int kind, int v) {
  double s = 0.0;
  if (kind == 1) {
    if (vars)
      for (int gp = 0; gp < 8; ++gp)
        for (int c = 0; c < 6; ++c) {
          const double x = vars[((size_t)e * 8 + gp) * ORC_NVAR + c];
          s += x * x;
        }
    s += sqrt(s);
  } else if (vars) {
    for (int gp = 0; gp < 8; ++gp) s += vars[((size_t)e * 8 + gp) * ORC_NVAR + v];
  }
  return s / 8;
}

/* write_vtu (src/output.cpp:67-214) after calc_fields: the ASCII
 * UnstructuredGrid file <filename>.vtu with positions, hexahedra (VTK 12),
 * displacements, element strain and stress, phase ids and the history
 * summaries.  Reals are printed as "%e" (the reference's std::scientific at
 * the default precision).  Returns 0, or -1 if the file cannot be opened. */
int orc_write_vtu(const orc_rve *r, const double *u, const double *vars, const char *filename) {
  char path[4096];
  snprintf(path, sizeof(path), "%s.vtu", filename);
  FILE *f = fopen(path, "w");
  if (!f) return -1;
  const int ne = r->nelem;
  double *es = malloc(sizeof(double) * 6 * (size_t)ne), *ss = malloc(sizeof(double) * 6 * (size_t)ne);
  orc_calc_fields(r, u, vars, es, ss);
  fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" "
             "byte_order=\"LittleEndian\">\n<UnstructuredGrid>\n");
  fprintf(f, "<Piece NumberOfPoints=\"%d\" NumberOfCells=\"%d\">\n<Points>\n", r->nn, ne);
  fprintf(f, "<DataArray type=\"Float64\" Name=\"Position\" NumberOfComponents=\"3\" format=\"ascii\">\n");
  for (int k = 0; k < r->nz; ++k)
    for (int j = 0; j < r->ny; ++j)
      for (int i = 0; i < r->nx; ++i) fprintf(f, "%e %e %e \n", i * r->dx, j * r->dy, k * r->dz);
  fprintf(f, "</DataArray>\n</Points>\n<Cells>\n\n");
  fprintf(f, "<DataArray type=\"Int32\" Name=\"connectivity\" NumberOfComponents=\"1\" format=\"ascii\">\n");
  for (int ez = 0; ez < r->nez; ++ez)
    for (int ey = 0; ey < r->ney; ++ey)
      for (int ex = 0; ex < r->nex; ++ex) {
        int n[8];
        orc_elem_nodes(r->nx, r->ny, ex, ey, ez, n);
        for (int a = 0; a < 8; ++a) fprintf(f, "%d ", n[a]);
        fprintf(f, "\n");
      }
  fprintf(f, "</DataArray>\n");
  fprintf(f, "<DataArray type=\"Int32\" Name=\"offsets\" NumberOfComponents=\"1\" format=\"ascii\">\n");
  for (int e = 0; e < ne; ++e) fprintf(f, "%d ", 8 * (e + 1));
  fprintf(f, "\n</DataArray>\n");
  fprintf(f, "<DataArray type=\"UInt8\"  Name=\"types\" NumberOfComponents=\"1\" format=\"ascii\">\n");
  for (int e = 0; e < ne; ++e) fprintf(f, "12 ");
  fprintf(f, "\n</DataArray>\n</Cells>\n");
  fprintf(f, "<PointData Vectors=\"displ\" >\n");
  fprintf(f, "<DataArray type=\"Float64\" Name=\"displ\" NumberOfComponents=\"3\" format=\"ascii\" >\n");
  for (int n = 0; n < r->nn; ++n) fprintf(f, "%e %e %e \n", u[n * 3], u[n * 3 + 1], u[n * 3 + 2]);
  fprintf(f, "</DataArray>\n</PointData>\n<CellData>\n");
  const char *names[2] = {"strain", "stress"};
  const double *arr[2] = {es, ss};
  for (int q = 0; q < 2; ++q) {
    fprintf(f, "<DataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"6\" format=\"ascii\">\n", names[q]);
    for (int e = 0; e < ne; ++e) {
      for (int v = 0; v < 6; ++v) fprintf(f, "%e ", arr[q][e * 6 + v]);
      fprintf(f, "\n");
    }
    fprintf(f, "</DataArray>\n");
  }
  fprintf(f, "<DataArray type=\"Int32\" Name=\"elem_type\" NumberOfComponents=\"1\" format=\"ascii\">\n");
  for (int e = 0; e < ne; ++e) fprintf(f, "%d ", r->et[e]);
  fprintf(f, "\n</DataArray>\n");
  const char *hn[4] = {"plasticity", "damage_e", "damage_D", "hardening"};
  const int kind[4] = {1, 0, 0, 0}, slot[4] = {0, 0, 1, 6};
  for (int q = 0; q < 4; ++q) {
    fprintf(f, "<DataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"1\" format=\"ascii\">\n", hn[q]);
    for (int e = 0; e < ne; ++e) fprintf(f, "%e ", vtu_cell_mean(vars, e, kind[q], slot[q]));
    fprintf(f, "\n</DataArray>\n");
  }
  fprintf(f, "</CellData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n");
  free(es);
  free(ss);
  return fclose(f) == 0 ? 0 : -1;
}

/* newton_raphson (src/solve.cpp:28-82).  The convergence test sits at the top
 * of the loop, so leaving at its == nr_max_its reports "not converged" even
 * when the last residual passes. */
static void newton_ws(const orc_rve *r, const double eps[6], const double *vars_old, double *u, int nr_max_its,
                      double nr_max_tol, double nr_rel_tol, int cg_max_its, double cg_abs_tol, double cg_rel_tol,
                      const double *A0_vals, int its_with_A0, int out[3], const int *cols, double *vals, double *b,
                      double *du) {
  int its = 0, solver_its = 0, conv = 0;
  orc_set_displ_bc(r, eps, u);
  double nrm = orc_assembly_rhs(r, u, vars_old, b);
  const double nrm0 = nrm;
  while (its < nr_max_its) {
    if (nrm < nr_max_tol || nrm < nrm0 * nr_rel_tol) {
      conv = 1;
      break;
    }
    const double *A;
    if (!A0_vals || its > its_with_A0 - 1) {
      orc_assembly_mat(r, u, vars_old, vals);
      A = vals;
    } else {
      A = A0_vals;
    }
    double err;
    solver_its += orc_ell_cg(r->nn, 3, 81, 13, cols, A, b, du, cg_abs_tol, cg_rel_tol, cg_max_its, &err);
    for (int i = 0; i < r->nndim; ++i) u[i] += du[i];
    nrm = orc_assembly_rhs(r, u, vars_old, b);
    its++;
  }
  out[0] = its;
  out[1] = solver_its;
  out[2] = conv;
}

void orc_newton(const orc_rve *r, const double eps[6], const double *vars_old, double *u, int nr_max_its,
                double nr_max_tol, double nr_rel_tol, int cg_max_its, double cg_abs_tol, double cg_rel_tol,
                const double *A0_vals, int its_with_A0, int out[3]) {
  int *cols = malloc(sizeof(int) * (size_t)r->nndim * 81);
  double *vals = malloc(sizeof(double) * (size_t)r->nndim * 81);
  double *b = calloc(r->nndim, sizeof(double)), *du = calloc(r->nndim, sizeof(double));
  orc_ell_cols3d(r->nx, r->ny, r->nz, cols);
  newton_ws(r, eps, vars_old, u, nr_max_its, nr_max_tol, nr_rel_tol, cg_max_its, cg_abs_tol, cg_rel_tol, A0_vals,
            its_with_A0, out, cols, vals, b, du);
  free(cols);
  free(vals);
  free(b);
  free(du);
}

/* ===================================================================== */
/* micropp<3> instance                                                   */
/* ===================================================================== */

typedef struct {
  double strain_old[6], strain[6], stress[6], ctan[36];
  int allocated;
  double *vars_n, *vars_k, *u_n, *u_k;
  long cost;
  int converged, subiterated, coupling;
} orc_gp;

struct orc_micro {
  orc_params p;
  orc_rve *r;
  orc_gp *gp;
  int nvars;
  double ctan_lin[36];
  double Vm, Vf;
  double *A0;
};

/* invert_3x3 (include/util.hpp:145-163) */
static double inv3(const double m[3][3], double o[3][3]) {
  const double det = m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
                     m[0][1] * (m[1][0] * m[2][2] - m[2][0] * m[1][2]) +
                     m[0][2] * (m[1][0] * m[2][1] - m[2][0] * m[1][1]);
  o[0][0] = +(m[1][1] * m[2][2] - m[2][1] * m[1][2]) / det;
  o[0][1] = -(m[1][0] * m[2][2] - m[2][0] * m[1][2]) / det;
  o[0][2] = +(m[1][0] * m[2][1] - m[2][0] * m[1][1]) / det;
  o[1][0] = -(m[0][1] * m[2][2] - m[2][1] * m[0][2]) / det;
  o[1][1] = +(m[0][0] * m[2][2] - m[2][0] * m[0][2]) / det;
  o[1][2] = -(m[0][0] * m[2][1] - m[2][0] * m[0][1]) / det;
  o[2][0] = +(m[0][1] * m[1][2] - m[1][1] * m[0][2]) / det;
  o[2][1] = -(m[0][0] * m[1][2] - m[1][0] * m[0][2]) / det;
  o[2][2] = +(m[0][0] * m[1][1] - m[1][0] * m[0][1]) / det;
  return det;
}

/* calc_ctan_lin_mix_rule_Chamis (src/micropp.cpp:286-331) */
static void chamis(const orc_micro *m, double c[36]) {
  const orc_material *mm = &m->r->mat[0], *ff = &m->r->mat[1];
  const double Em = mm->E, nu_m = mm->nu, Gm = Em / (2 * (1 + nu_m));
  const double Ef = ff->E, nu_f = ff->nu, Gf = Ef / (2 * (1 + nu_f));
  const double Vf = m->Vf, Vm = m->Vm;
  const double E11 = Vf * Ef + Vm * Em;
  const double E22 = Em / (1 - sqrt(Vf) * (1 - Em / Ef));
  const double nu12 = Vf * nu_f + Vm * nu_m;
  const double G12 = Gm / (1 - sqrt(Vf) * (1 - Gm / Gf));
  const double nu23 = nu12;
  const double S[3][3] = {{1 / E11, -nu12 / E11, -nu12 / E11},
                          {-nu12 / E11, 1 / E22, -nu23 / E22},
                          {-nu12 / E11, -nu23 / E22, 1 / E22}};
  double Si[3][3];
  inv3(S, Si);
  for (int i = 0; i < 36; ++i) c[i] = 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) c[i * 6 + j] = Si[i][j];
  for (int i = 3; i < 6; ++i) c[i * 6 + i] = G12;
}

typedef struct {
  int *cols;
  double *vals, *b, *du, *u;
} orc_ws;

static void ws_init(orc_ws *w, const orc_rve *r) {
  w->cols = malloc(sizeof(int) * (size_t)r->nndim * 81);
  w->vals = malloc(sizeof(double) * (size_t)r->nndim * 81);
  w->b = calloc(r->nndim, sizeof(double));
  w->du = calloc(r->nndim, sizeof(double));
  w->u = calloc(r->nndim, sizeof(double));
  orc_ell_cols3d(r->nx, r->ny, r->nz, w->cols);
}
static void ws_free(orc_ws *w) {
  free(w->cols);
  free(w->vals);
  free(w->b);
  free(w->du);
  free(w->u);
}

static void micro_newton(const orc_micro *m, orc_ws *w, const double eps[6], const double *vars_old, int out[3]) {
  const orc_params *p = &m->p;
  newton_ws(m->r, eps, vars_old, w->u, p->nr_max_its, p->nr_max_tol, p->nr_rel_tol, p->cg_max_its, p->cg_abs_tol,
            p->cg_rel_tol, p->use_A0 ? m->A0 : NULL, p->its_with_A0, out, w->cols, w->vals, w->b, w->du);
}

/* calc_ctan_lin_fe_models (src/micropp.cpp:255-284) */
static void ctan_lin_fe(orc_micro *m) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = 0; i < 6; ++i) {
    orc_ws w;
    ws_init(&w, m->r);
    double eps[6] = {0, 0, 0, 0, 0, 0}, sig[6];
    eps[i] += ORC_D_EPS_CTAN_AVE;
    int out[3];
    micro_newton(m, &w, eps, NULL, out);
    orc_ave_stress(m->r, w.u, NULL, sig);
    for (int v = 0; v < 6; ++v) m->ctan_lin[v * 6 + i] = sig[v] / ORC_D_EPS_CTAN_AVE;
    ws_free(&w);
  }
}

orc_micro *orc_micro_new(const orc_params *p) {
  orc_micro *m = calloc(1, sizeof(orc_micro));
  m->p = *p;
  m->r = orc_rve_new(p->size[0], p->size[1], p->size[2], p->type, p->geo, p->mat);
  if (!m->r) {
    free(m);
    return NULL;
  }
  m->p.coupling = NULL;
  m->nvars = m->r->nelem * 8 * ORC_NVAR;
  m->gp = calloc(p->ngp, sizeof(orc_gp));
  int n_fe = 0;
  for (int g = 0; g < p->ngp; ++g) {
    orc_gp *q = &m->gp[g];
    q->coupling = p->coupling ? p->coupling[g] : ORC_FE_ONE_WAY;
    q->converged = 1;
    if (q->coupling == ORC_FE_LINEAR || q->coupling == ORC_FE_ONE_WAY || q->coupling == ORC_FE_FULL) n_fe++;
    if (!p->coupling || q->coupling == ORC_FE_ONE_WAY || q->coupling == ORC_FE_FULL) {
      q->u_n = calloc(m->r->nndim, sizeof(double));
      q->u_k = calloc(m->r->nndim, sizeof(double));
    }
  }
  /* calc_volume_fractions (src/micropp.cpp:609-627) */
  m->Vm = 0.0;
  m->Vf = 0.0;
  for (int e = 0; e < m->r->nelem; ++e) {
    if (m->r->et[e] == 0)
      m->Vm += m->r->evol;
    else if (m->r->et[e] >= 1)
      m->Vf += m->r->evol;
  }
  m->Vm /= m->r->vol_tot;
  m->Vf /= m->r->vol_tot;
  if (p->use_A0) {
    double *u0 = calloc(m->r->nndim, sizeof(double));
    m->A0 = malloc(sizeof(double) * (size_t)m->r->nndim * 81);
    orc_assembly_mat(m->r, u0, NULL, m->A0);
    free(u0);
  }
  for (int i = 0; i < 36; ++i) m->ctan_lin[i] = 0.0;
  if (p->calc_ctan_lin && n_fe > 0) ctan_lin_fe(m);
  for (int g = 0; g < p->ngp; ++g) {
    orc_gp *q = &m->gp[g];
    if (q->coupling == ORC_MIX_RULE_CHAMIS)
      chamis(m, q->ctan);
    else
      memcpy(q->ctan, m->ctan_lin, sizeof(q->ctan));
  }
  return m;
}

void orc_micro_free(orc_micro *m) {
  if (!m) return;
  for (int g = 0; g < m->p.ngp; ++g) {
    free(m->gp[g].u_n);
    free(m->gp[g].u_k);
    free(m->gp[g].vars_n);
    free(m->gp[g].vars_k);
  }
  free(m->gp);
  free(m->A0);
  orc_rve_free(m->r);
  free(m);
}

void orc_set_strain(orc_micro *m, int gp, const double eps[6]) { memcpy(m->gp[gp].strain, eps, 6 * sizeof(double)); }
void orc_get_stress(const orc_micro *m, int gp, double sig[6]) { memcpy(sig, m->gp[gp].stress, 6 * sizeof(double)); }
void orc_get_ctan(const orc_micro *m, int gp, double c[36]) { memcpy(c, m->gp[gp].ctan, 36 * sizeof(double)); }

/* homogenize_linear(gp) (src/homogenize.cpp:101-109) */
static void lin_homog(orc_gp *q) {
  for (int i = 0; i < 6; ++i) q->stress[i] = 0.0;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) q->stress[i] += q->ctan[i * 6 + j] * q->strain[j];
}

/* homogenize_fe_one_way / homogenize_fe_full (src/homogenize.cpp:111-282) */
static void fe_homog(orc_micro *m, orc_gp *q, int full) {
  const orc_rve *r = m->r;
  orc_ws w;
  ws_init(&w, r);
  double *aux = calloc(m->nvars, sizeof(double));
  double *vnew = q->allocated ? q->vars_k : aux;
  q->cost = 0;
  q->subiterated = 0;
  memcpy(w.u, q->u_n, sizeof(double) * r->nndim);
  int nt[3];
  micro_newton(m, &w, q->strain, q->vars_n, nt);
  memcpy(q->u_k, w.u, sizeof(double) * r->nndim);
  q->cost += nt[1];
  q->converged = nt[2];
  if (!q->converged && m->p.subiterations) {
    double es[6], de[6];
    memcpy(w.u, q->u_n, sizeof(double) * r->nndim);
    memcpy(es, q->strain_old, sizeof(es));
    q->subiterated = 1;
    for (int i = 0; i < 6; ++i) de[i] = (q->strain[i] - q->strain_old[i]) / m->p.nsubiterations;
    for (int s = 0; s < m->p.nsubiterations; ++s) {
      for (int j = 0; j < 6; ++j) es[j] += de[j];
      micro_newton(m, &w, es, q->vars_n, nt);
      q->cost += nt[1];
    }
    q->converged = nt[2];
    memcpy(q->u_k, w.u, sizeof(double) * r->nndim);
  }
  if (m->p.lin_stress)
    lin_homog(q);
  else
    orc_ave_stress(r, q->u_k, q->vars_n, q->stress);
  const int nl = orc_vars_new(r, q->u_k, q->vars_n, vnew);
  if (nl && !q->allocated) {
    q->vars_n = calloc(m->nvars, sizeof(double));
    q->vars_k = calloc(m->nvars, sizeof(double));
    q->allocated = 1;
    memcpy(q->vars_k, vnew, sizeof(double) * m->nvars);
  }
  if (full && q->allocated) {
    double s0[6], s1[6], e1[6];
    memcpy(w.u, q->u_k, sizeof(double) * r->nndim);
    memcpy(s0, q->stress, sizeof(s0));
    for (int i = 0; i < 6; ++i) {
      memcpy(e1, q->strain, sizeof(e1));
      e1[i] += ORC_D_EPS_CTAN_AVE;
      int nt2[3];
      micro_newton(m, &w, e1, q->vars_n, nt2);
      q->cost += nt[1]; /* the reference adds the stale count of the sigma solve */
      orc_ave_stress(r, w.u, q->vars_n, s1);
      for (int v = 0; v < 6; ++v) q->ctan[v * 6 + i] = (s1[v] - s0[v]) / ORC_D_EPS_CTAN_AVE;
    }
  }
  ws_free(&w);
  free(aux);
}

/* homogenize (src/homogenize.cpp:68-99): OpenMP over GPs */
void orc_homogenize(orc_micro *m) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int g = 0; g < m->p.ngp; ++g) {
    orc_gp *q = &m->gp[g];
    if (q->coupling == ORC_FE_LINEAR || q->coupling == ORC_MIX_RULE_CHAMIS)
      lin_homog(q);
    else if (q->coupling == ORC_FE_ONE_WAY)
      fe_homog(m, q, 0);
    else if (q->coupling == ORC_FE_FULL)
      fe_homog(m, q, 1);
  }
}

void orc_homogenize_linear(orc_micro *m) {
  for (int g = 0; g < m->p.ngp; ++g) lin_homog(&m->gp[g]);
}

/* gp_t::update_vars (include/gp.hpp:95-105) */
void orc_update_vars(orc_micro *m) {
  for (int g = 0; g < m->p.ngp; ++g) {
    orc_gp *q = &m->gp[g];
    double *t = q->vars_n;
    q->vars_n = q->vars_k;
    q->vars_k = t;
    t = q->u_n;
    q->u_n = q->u_k;
    q->u_k = t;
    memcpy(q->strain_old, q->strain, sizeof(q->strain));
  }
}

int orc_is_non_linear(const orc_micro *m, int gp) { return m->gp[gp].allocated; }
int orc_get_non_linear_gps(const orc_micro *m) {
  int c = 0;
  for (int g = 0; g < m->p.ngp; ++g) c += m->gp[g].allocated ? 1 : 0;
  return c;
}
long orc_get_cost(const orc_micro *m, int gp) { return m->gp[gp].cost; }
int orc_has_converged(const orc_micro *m, int gp) { return m->gp[gp].converged; }
int orc_has_subiterated(const orc_micro *m, int gp) { return m->gp[gp].subiterated; }
void orc_get_ctan_lin(const orc_micro *m, double c[36]) { memcpy(c, m->ctan_lin, sizeof(m->ctan_lin)); }
void orc_get_volume_fractions(const orc_micro *m, double vf[2]) {
  vf[0] = m->Vm;
  vf[1] = m->Vf;
}

int orc_get_state(const orc_micro *m, int gp, int which, double *out) {
  const orc_gp *q = &m->gp[gp];
  const double *src = which == 0 ? q->u_n : which == 1 ? q->u_k : which == 2 ? q->vars_n : q->vars_k;
  const int n = which < 2 ? m->r->nndim : m->nvars;
  if (!src) return 0;
  memcpy(out, src, sizeof(double) * n);
  return 1;
}

int orc_nthreads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
