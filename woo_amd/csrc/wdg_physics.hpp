// Contact physics parameters (Cp2 functors) as device functions.
// They depend only on radii, materials and masses, so the contact kernel
// recomputes them per step instead of storing a CPhys per contact.
#pragma once
#include "wdg_internal.hpp"
#include "wdg_math.hpp"

namespace wdg {

struct PhysVals {
  double kn, kt, tanPhi;          // FrictPhys (kn,kt of Hertz are per step, from uN)
  double K, R, kt0, gamma, alpha; // HertzPhys
  double alpha_sqrtMK;
};

// L6Geom.lens and contA fixed at contact creation (L6Geom.cpp:61-67), iniLensTouch=true
__device__ __forceinline__ void lens_area(double r1, double r2, double& l0, double& l1, double& A) {
  l0 = fabs(r1);
  l1 = fabs(r2);
  if (l0 < 0) l0 = l1;
  if (l1 < 0) l1 = l0;
  double rr = (r1 > 0 && r2 > 0) ? fmin(r1, r2) : (r1 > 0 ? r1 : r2);
  A = M_PI * (rr * rr);
}

// the radii the geometry functor passes to handleSpheresLikeContact: both radii for two
// spheres (Sphere.cpp:117), (-r, r) for a wall (Wall.cpp:82), (max(halfThick, r), r) for a
// facet (Facet.cpp:389)
__device__ __forceinline__ void contact_radii(uint32_t info, uint32_t pA, uint32_t pB, const double* rad,
                                              const double* fgeo, double& r1, double& r2) {
  const uint32_t t = info & CI_TYPE_MASK;
  r2 = rad[pB];
  if (t == CT_SS) r1 = rad[pA];
  else if (t == CT_FS) r1 = fmax(fgeo[(int64_t)FGEO * pA + 12], r2);
  else r1 = -r2;
}

// Cp2_FrictMat_FrictPhys::updateFrictPhys (FrictMat.cpp:20-44)
__device__ __forceinline__ void phys_frict(const wdg_material& m1, const wdg_material& m2, double l0, double l1,
                                           double A, PhysVals& ph) {
  ph.kn = 1 / (1 / (m1.young * A / l0) + 1 / (m2.young * A / l1));
  ph.kt = .5 * (m1.ktDivKn + m2.ktDivKn) * ph.kn;
  ph.tanPhi = fmin(m1.tanPhi, m2.tanPhi);
}

// Cp2_HertzMat_HertzPhys::go (Hertz.cpp:34-77); viscAlpha is the en-only factor computed on the host
__device__ __forceinline__ void phys_hertz(const wdg_material& m1, const wdg_material& m2, double r1, double r2,
                                           double massA, double massB, double poisson, double en, double viscAlpha,
                                           PhysVals& ph) {
  const double E1 = m1.young, E2 = m2.young;
  const double nu1 = poisson, nu2 = poisson;
  ph.K = (4 / 3.) * 1. / ((1 - nu1 * nu1) / E1 + (1 - nu2 * nu2) / E2);
  ph.R = 1. / (1. / r1 + 1. / r2);
  double kn0 = ph.K * sqrt(ph.R);
  ph.gamma = fmin(m1.surfEnergy, m2.surfEnergy);
  ph.alpha = .5 * (m1.alpha + m2.alpha);
  double G1 = E1 / (2 * (1 + nu1)), G2 = E2 / (2 * (1 + nu2));
  double G = .5 * (G1 + G2);
  double nu = .5 * (nu1 + nu2);
  ph.kt0 = 2 * sqrt(4 * ph.R) * G / (2 - nu);
  ph.tanPhi = fmin(m1.tanPhi, m2.tanPhi);
  if (en > 0 && en < 1.) {
    double mbar = (massA <= 0 && massB > 0) ? massB : ((massA > 0 && massB <= 0) ? massA : (massA * massB) / (massA + massB));
    ph.alpha_sqrtMK = fmax(0., viscAlpha * sqrt(mbar * kn0));
  } else {
    ph.alpha_sqrtMK = 0.0;
  }
}

}  // namespace wdg
