// physics.hpp -- device-side constitutive laws, shape functions and
// micro-structure phase tests of the RVE hot path (gfx950).
//
// Every function is evaluated with FP contraction off and IEEE division and
// square root, in the reference's expression order, so that assembled
// matrices, residuals, stresses and phase maps are bit-identical to the
// reference CPU code.  Reference citations are relative to the Micropp tree.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "material_base.h"

namespace mpp {

constexpr int NVAR = 7;                     // include/params.hpp:27
constexpr double GAUSS = 0.577350269189626; // include/params.hpp:34 (truncated 1/sqrt(3))

// Node sign pattern of the hex (src/common.cpp:30-41); the Gauss points use
// the same pattern scaled by GAUSS (include/micropp.hpp:86-89).
__device__ __constant__ static const signed char kHexSign[8][3] = {
    {-1, -1, -1}, {+1, -1, -1}, {+1, +1, -1}, {-1, +1, -1}, {-1, -1, +1}, {+1, -1, +1}, {+1, +1, +1}, {-1, +1, +1}};
// Local node positions inside an element, 0/1 per axis.
__device__ __constant__ static const signed char kHexPos[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                                                  {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

// Local index of the node at element position (px, py, pz) in {0,1}^3.
__device__ __forceinline__ int local_node(int px, int py, int pz) {
  // inverse of kHexPos: z layer * 4 + counter-clockwise index in the layer
  const int in_layer = (py == 0) ? px : (px == 0 ? 3 : 2);
  return pz * 4 + in_layer;
}

// ---------------------------------------------------------------------------
// Constitutive laws (src/material.cpp).  material_base layout is shared with
// the host API; k, mu, lambda are rebuilt from E, nu on the host exactly as
// the reference's make_material does.
// ---------------------------------------------------------------------------
__device__ inline void lin_stress(const material_base &m, const double *e, double *s) {
#pragma clang fp contract(off)
  const double tr = e[0] + e[1] + e[2];
  const double two_mu = 2 * m.mu;
  for (int i = 0; i < 3; ++i) s[i] = m.lambda * tr + two_mu * e[i];
  for (int i = 3; i < 6; ++i) s[i] = m.mu * e[i];
}

// get_dev_tensor (src/material.cpp:65-69)
__device__ inline void deviator(const double *t, double *d) {
#pragma clang fp contract(off)
  const double third_tr = (1 / 3.0) * (t[0] + t[1] + t[2]);
  for (int i = 0; i < 6; ++i) d[i] = t[i];
  for (int i = 0; i < 3; ++i) d[i] -= third_tr;
}

// J2 trial state and return increment, plastic_law (src/material.cpp:110-148)
__device__ inline int j2_law(const material_base &m, const double *eps, const double *hist, double *dl, double *nrm,
                             double *s_tr) {
#pragma clang fp contract(off)
  const double alpha_old = hist ? hist[6] : 0;
  double ed[6], epd[6];
  if (hist) {
    deviator(hist, epd);
  } else {
    for (int i = 0; i < 6; ++i) epd[i] = 0.0;
  }
  deviator(eps, ed);
  const double two_mu = 2 * m.mu;
  for (int i = 0; i < 3; ++i) s_tr[i] = two_mu * (ed[i] - epd[i]);
  for (int i = 3; i < 6; ++i) s_tr[i] = m.mu * (ed[i] - epd[i]);
  double ss = 0.0;
  for (int i = 0; i < 6; ++i) ss += s_tr[i] * s_tr[i];
  const double s_norm = sqrt(ss);
  const double f = s_norm - SQRT_2DIV3 * (m.Sy + m.Ka * alpha_old);
  if (f > 0) {
    for (int i = 0; i < 6; ++i) nrm[i] = s_tr[i] / s_norm;
    *dl = f / (2. * m.mu * (1. + m.Ka / (3. * m.mu)));
    return 1;
  }
  for (int i = 0; i < 6; ++i) nrm[i] = 0.0;
  *dl = 0;
  return 0;
}

// hardening_law (src/material.cpp:206-225), constants of the reference law
__device__ inline double damage_hardening(const material_base &m, double r) {
#pragma clang fp contract(off)
  const double Ey = 10.0e4;
  const double Ey_inf = 10. * Ey;
  const double H0 = 10.0, H1 = 5.0;
  const double sqE = sqrt(m.E);
  const double r0 = Ey / sqE;
  const double q0 = r0;
  const double q1 = Ey_inf / sqE;
  const double r1 = r0 + (q1 - q0) / H0;
  if (r < r0) return 0.0;
  if (r < r1) return q0 + H0 * (r - r0);
  return q1 + H1 * (r - r1);
}

// damage_law (src/material.cpp:227-269); s receives the linear stress
__device__ inline int damage_law(const material_base &m, const double *eps, double r_old, double D_old, double *r_out,
                                 double *D_out, double *s) {
#pragma clang fp contract(off)
  lin_stress(m, eps, s);
  double prod = 0.0;
  for (int i = 0; i < 6; ++i) prod += s[i] * eps[i];
  const double r = (prod >= 0) ? sqrt(prod) : 0;
  const double thr = m.Xt / sqrt(m.E);
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

// material_*::get_stress (src/material.cpp:76-82, 150-163, 271-287)
__device__ inline void mat_stress(const material_base &m, const double *eps, const double *hist, double *sig) {
#pragma clang fp contract(off)
  if (m.type == MATERIAL_ELASTIC) {
    lin_stress(m, eps, sig);
  } else if (m.type == MATERIAL_PLASTIC) {
    double dl, nrm[6], s_tr[6];
    j2_law(m, eps, hist, &dl, nrm, s_tr);
    const double tr = eps[0] + eps[1] + eps[2];
    for (int i = 0; i < 6; ++i) sig[i] = s_tr[i];
    for (int i = 0; i < 3; ++i) sig[i] += m.k * tr;
    const double two_mu_dl = 2 * m.mu * dl;
    for (int i = 0; i < 6; ++i) sig[i] -= two_mu_dl * nrm[i];
  } else {
    const double r_old = hist ? hist[0] : 0.0;
    const double D_old = hist ? hist[1] : 0.0;
    double D, r;
    damage_law(m, eps, r_old, D_old, &r, &D, sig);
    for (int i = 0; i < 6; ++i) sig[i] *= (1 - D);
  }
}

// material_elastic::get_ctan (src/material.cpp:84-95) and the forward
// perturbation tangent apply_perturbation (:48-63) for the nonlinear laws.
__device__ inline void mat_ctan(const material_base &m, const double *eps, const double *hist, double *C) {
#pragma clang fp contract(off)
  if (m.type == MATERIAL_ELASTIC) {
    for (int i = 0; i < 36; ++i) C[i] = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) C[i * 6 + j] += m.lambda;
    for (int i = 0; i < 3; ++i) C[i * 6 + i] += 2 * m.mu;
    for (int i = 3; i < 6; ++i) C[i * 6 + i] = m.mu;
    return;
  }
  double s0[6];
  mat_stress(m, eps, hist, s0);
  for (int c = 0; c < 6; ++c) {
    double e1[6], s1[6];
    for (int i = 0; i < 6; ++i) e1[i] = eps[i];
    e1[c] += D_EPS_CTAN;
    mat_stress(m, e1, hist, s1);
    for (int r = 0; r < 6; ++r) C[r * 6 + c] = (s1[r] - s0[r]) / D_EPS_CTAN;
  }
}

// material_*::evolute (src/material.cpp:97-100, 170-186, 289-302).  As in the
// reference, the J2 law leaves hist_new untouched when there is no old state.
__device__ inline int mat_evolute(const material_base &m, const double *eps, const double *hist_old,
                                  double *hist_new) {
#pragma clang fp contract(off)
  if (m.type == MATERIAL_ELASTIC) return 0;
  if (m.type == MATERIAL_PLASTIC) {
    double dl, nrm[6], s_tr[6];
    const int nl = j2_law(m, eps, hist_old, &dl, nrm, s_tr);
    if (hist_old && hist_new) {
      for (int i = 0; i < 6; ++i) hist_new[i] = hist_old[i] + dl * nrm[i];
      hist_new[6] = hist_old[6] + SQRT_2DIV3 * dl + 0;
    }
    return nl;
  }
  const double r_old = hist_old ? hist_old[0] : 0;
  const double D_old = hist_old ? hist_old[1] : 0;
  double s[6], r_new, D_new;
  const int nl = damage_law(m, eps, r_old, D_old, &r_new, &D_new, s);
  if (hist_new) {
    hist_new[0] = r_new;
    hist_new[1] = D_new;
  }
  return nl;
}

// ---------------------------------------------------------------------------
// Shape functions: calc_bmat (src/micro3D.cpp:80-120)
// ---------------------------------------------------------------------------
// B of Gauss point gp of a dx*dy*dz brick, rows [xx,yy,zz,xy,xz,yz]
// (engineering shear), columns node*3 + d.
__device__ inline void bmat_gp(int gp, double dx, double dy, double dz, double *b /* [6][24] */) {
#pragma clang fp contract(off)
  const double xi[3] = {kHexSign[gp][0] * GAUSS, kHexSign[gp][1] * GAUSS, kHexSign[gp][2] * GAUSS};
  for (int i = 0; i < 144; ++i) b[i] = 0;
  for (int a = 0; a < 8; ++a) {
    double f[3];
    for (int d = 0; d < 3; ++d) f[d] = (kHexSign[a][d] > 0) ? (1 + xi[d]) : (1 - xi[d]);
    const double gx = kHexSign[a][0] * f[1] * f[2] / 8. * 2. / dx;
    const double gy = kHexSign[a][1] * f[0] * f[2] / 8. * 2. / dy;
    const double gz = kHexSign[a][2] * f[0] * f[1] / 8. * 2. / dz;
    b[0 * 24 + a * 3 + 0] = gx;
    b[1 * 24 + a * 3 + 1] = gy;
    b[2 * 24 + a * 3 + 2] = gz;
    b[3 * 24 + a * 3 + 0] = gy;
    b[3 * 24 + a * 3 + 1] = gx;
    b[4 * 24 + a * 3 + 0] = gz;
    b[4 * 24 + a * 3 + 2] = gx;
    b[5 * 24 + a * 3 + 1] = gz;
    b[5 * 24 + a * 3 + 2] = gy;
  }
}

// ---------------------------------------------------------------------------
// Micro-structure phase tests (src/micropp.cpp:338-544, include/util.hpp)
// ---------------------------------------------------------------------------
__device__ inline double norm3(const double *v) {
#pragma clang fp contract(off)
  double s = 0;
  for (int i = 0; i < 3; ++i) s += v[i] * v[i];
  return sqrt(s);
}
__device__ inline double dot3(const double *a, const double *b) {
#pragma clang fp contract(off)
  double s = 0;
  for (int i = 0; i < 3; ++i) s += a[i] * b[i];
  return s;
}
// point_inside_sphere (include/util.hpp:115-124): strict <
__device__ inline bool in_sphere(const double *c, double r, const double *p) {
#pragma clang fp contract(off)
  const double v[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
  return norm3(v) < r;
}
// point_inside_cilinder_inf (include/util.hpp:126-143): sqrt(1 - cos^2)
// form; a centroid on the axis point yields NaN, which tests false.
__device__ inline bool in_cylinder(const double *dir, const double *c, double r, const double *p) {
#pragma clang fp contract(off)
  const double v[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
  const double dv = dot3(dir, v);
  const double nd = norm3(dir);
  const double nv = norm3(v);
  const double cs = dv / (nd * nv);
  const double sn = sqrt(1 - cs * cs);
  return nv * sn <= r;
}

// Tables of the reference micro-structures MIC3D_SPHERES
// (src/micropp.cpp:427-457) and MIC3D_FIBS_20_DISORDER (:505-541).
__device__ __constant__ static const double kSph40C[40][3] = {
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
__device__ __constant__ static const double kSph40R[40] = {
    .5741, .1735, .5065, .8565, .9735, .7087, .9585, .2843, .4029, .2574, .1575, .3316, .1874, .6300,
    .7049, .7258, .8002, .8176, .9970, .4736, .6351, .3140, .3440, .2444, .3894, .5234, .1658, .4020,
    .6813, .2060, .1427, .9332, .5406, .9765, .0956, .6432, .9998, .4166, .6907, .0404};
__device__ __constant__ static const double kFib20Dyz[20][2] = {
    {.5741, .8515}, {.1735, .1103}, {.5065, .4600}, {.8565, .9045}, {.9735, .6313}, {.7087, .1547}, {.9585, .0220},
    {.2843, .4062}, {.4029, .8095}, {.2574, .4742}, {.1575, .0768}, {.3316, .0320}, {.1874, .6364}, {.6300, .2688},
    {.7049, .5137}, {.7258, .2799}, {.8002, .4794}, {.8176, .4142}, {.9970, .2189}, {.4736, .6202}};

enum {
  kMicHomogeneous = 0,
  kMicSphere,
  kMicLayerY,
  kMicCiliFibX,
  kMicCiliFibZ,
  kMicCiliFibXZ,
  kMicQuadFibXYZ,
  kMicQuadFibXZ,
  kMicQuadFibXZBrokenX,
  kMic3dSpheres,
  kMic3d8,
  kMic3dFibs20Order,
  kMic3dFibs20Disorder,
  kMic3dRandomSpheres,
  kMicCount
};

// Phase of an element from its centroid c (lx = ly = lz = 1).
__device__ inline int phase_of(int type, const double *geo, const double *spheres, int nsph, const double *c) {
#pragma clang fp contract(off)
  const double L = 1.0;
  const double mid[3] = {L / 2, L / 2, L / 2};
  switch (type) {
    case kMicHomogeneous:
      return 0;
    case kMicSphere:
      return in_sphere(mid, geo[0], c);
    case kMicLayerY:
      return c[1] < geo[0];
    case kMicCiliFibX: {
      const double d[3] = {1, 0, 0};
      return in_cylinder(d, mid, geo[0], c);
    }
    case kMicCiliFibZ: {
      const double d[3] = {0, 0, 1};
      return in_cylinder(d, mid, geo[0], c);
    }
    case kMicCiliFibXZ: {
      const double c1[3] = {L / 2., L * .75, L / 2.};
      const double c2[3] = {L / 2., L * .25, L / 2.};
      const double dxv[3] = {1, 0, 0}, dzv[3] = {0, 0, 1};
      return (in_cylinder(dzv, c1, geo[0], c) || in_cylinder(dxv, c2, geo[0], c)) ? 1 : 0;
    }
    case kMicQuadFibXYZ:
    case kMicQuadFibXZ:
    case kMicQuadFibXZBrokenX: {
      const double w = geo[0];
      const bool ax = fabs(c[0] - mid[0]) < w, ay = fabs(c[1] - mid[1]) < w, az = fabs(c[2] - mid[2]) < w;
      if (type == kMicQuadFibXYZ) return ((ax && ay) || (ax && az) || (ay && az)) ? 1 : 0;
      if (type == kMicQuadFibXZ) return ((ax && ay) || (ay && az)) ? 1 : 0;
      if (ax && ay) return 1;
      return (ay && az && (c[0] < L * .8 || c[0] > L * .9)) ? 1 : 0;
    }
    case kMic3dSpheres:
      for (int i = 0; i < 40; ++i)
        if (in_sphere(kSph40C[i], 0.1 * kSph40R[i], c)) return 1;
      return 0;
    case kMic3d8: {
      const double rc = 0.1, w_flat = 0.02, w_cyl = 0.01;
      const double c1[3] = {L * .25, L * .75, 0.0}, c2[3] = {L * .75, L * .75, 0.0};
      const double c3[3] = {0.0, L * .25, L * .25}, c4[3] = {0.0, L * .25, L * .75};
      const double dxv[3] = {1, 0, 0}, dzv[3] = {0, 0, 1};
      if (in_cylinder(dzv, c1, rc, c) || in_cylinder(dzv, c2, rc, c) || in_cylinder(dxv, c3, rc, c) ||
          in_cylinder(dxv, c4, rc, c))
        return 1;
      if (in_cylinder(dzv, c1, rc + w_cyl, c) || in_cylinder(dzv, c2, rc + w_cyl, c) ||
          in_cylinder(dxv, c3, rc + w_cyl, c) || in_cylinder(dxv, c4, rc + w_cyl, c) || fabs(c[1] - L / 2) < w_flat)
        return 2;
      return 0;
    }
    case kMic3dFibs20Order: {
      const double d[3] = {1, 0, 0};
      const double hz = 1.0 / (5 + 1), hy = 1.0 / (4 + 1);
      for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 4; ++j) {
          const double cc[3] = {0.0, (j + 1) * hy, (i + 1) * hz};
          if (in_cylinder(d, cc, 0.05, c)) return 1;
        }
      return 0;
    }
    case kMic3dFibs20Disorder:
      for (int i = 0; i < 20; ++i) {
        const double d[3] = {1, kFib20Dyz[i][0], kFib20Dyz[i][1]};
        if (in_cylinder(d, kSph40C[i], 0.05, c)) return 1;
      }
      return 0;
    case kMic3dRandomSpheres:
      for (int i = 0; i < nsph; ++i)
        if (in_sphere(&spheres[i * 4], spheres[i * 4 + 3], c)) return 1;
      return 0;
    default:
      return -1;
  }
}

// splitmix64 (Steele, Lea & Flood 2014): the seeded stream of the
// MIC3D_RANDOM_SPHERES micro-structure.
__device__ inline uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__device__ inline double splitmix_unit(uint64_t *s) {
  return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace mpp
