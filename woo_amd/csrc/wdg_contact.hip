// Fused contact kernel: geometry (Cg2) + physics (Cp2) + constitutive law (Law2)
// + rotation of the local force to global coordinates, one thread per contact
// slot. Replaces the per-contact body of ContactLoop::run (ContactLoop.cpp:134-231).
//
// Per real contact the persistent state is 13 doubles (L6Geom.trsf columns 0
// and 1, contact point, local force, uN) in SoA arrays aligned with the slot
// list, so every access is coalesced; column 2 of trsf is always col0 x col1
// and the Cp2 parameters are functions of radii, materials and masses.
#include "wdg_internal.hpp"
#include "wdg_math.hpp"
#include "wdg_physics.hpp"

namespace wdg {

struct ContactArgs {
  int64_t nc;
  const uint32_t *A, *B, *info;
  const uint64_t* dist;
  uint32_t* real;
  double *N0, *N1, *N2, *Y0, *Y1, *Y2, *P0, *P1, *P2, *F0, *F1, *F2, *uN;
  double *G0, *G1, *G2;  // global force on pA
  // particles
  const double *px, *py, *pz, *vx, *vy, *vz, *wx, *wy, *wz, *rad, *mass;
  const uint32_t* flags;
  const int32_t* mat;
  const wdg_material* mats;
  // parameters
  double L[3];
  double absDf;
  double dt;
  int renorm;  // 1 if trsf is renormalised this step
  double poisson, en, viscAlpha;
  int noSlip, noBreak, noFrict, noAttraction;
  uint32_t* dflags;
};

template <bool PERIODIC, int LAW>
__global__ __launch_bounds__(256) void k_contact(ContactArgs a) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.nc) return;
  const uint32_t info = a.info[c];
  const bool real = a.real[c] != 0u;
  if (!real && !(info & CI_NEAR)) return;
  const uint32_t iA = a.A[c], iB = a.B[c];
  const bool swap = (info & CI_SWAP) != 0u;
  const uint32_t pA = swap ? iB : iA, pB = swap ? iA : iB;
  const int type = (int)(info & CI_TYPE_MASK);

  d3 shift2 = mk3(0., 0., 0.);
  if (PERIODIC) {
    uint64_t cd = a.dist[c];
    shift2 = mk3(a.L[0] * unpack_cd(cd, 0), a.L[1] * unpack_cd(cd, 1), a.L[2] * unpack_cd(cd, 2));
  }
  const d3 posA = mk3(a.px[pA], a.py[pA], a.pz[pA]);
  const d3 posB = mk3(a.px[pB], a.py[pB], a.pz[pB]);

  d3 normal, contPt, pos1, pos2;
  double uN, r1, r2;
  if (type == CT_SS) {
    // Cg2_Sphere_Sphere_L6Geom::go (Sphere.cpp:92-117)
    r1 = a.rad[pA];
    r2 = a.rad[pB];
    pos1 = posA;
    pos2 = add3(posB, shift2);
    d3 relPos = sub3(pos2, posA);
    double unDistSq = sqn3(relPos) - p2(a.absDf * (r1 + r2));
    if (unDistSq > 0 && !real) return;
    double dist = sqrt(sqn3(relPos));
    uN = dist - (r1 + r2);
    normal = div3(relPos, dist);
    contPt = add3(posA, mul3(normal, r1 + 0.5 * uN));
  } else {
    // Cg2_Wall_Sphere_L6Geom::go (Wall.cpp:50-84), pA is the wall
    const uint32_t fl = a.flags[pA];
    const int ax = flag_axis(fl), sense = flag_sense(fl);
    const double radius = a.rad[pB];
    pos1 = posA;
    pos2 = add3(posB, shift2);
    double dist = comp3(pos2, ax) - comp3(posA, ax);
    if (!real && fabs(dist) > radius) return;
    contPt = pos2;
    setcomp3(contPt, ax, comp3(posA, ax));
    double nv;
    if (sense == 0) nv = !real ? (dist > 0 ? 1. : -1.) : comp3(mk3(a.N0[c], a.N1[c], a.N2[c]), ax);
    else nv = (sense == 1 ? 1. : -1.);
    normal = mk3(0., 0., 0.);
    setcomp3(normal, ax, nv);
    uN = nv * dist - radius;
    r1 = -radius;
    r2 = radius;
  }
  const d3 vel1 = mk3(a.vx[pA], a.vy[pA], a.vz[pA]), vel2 = mk3(a.vx[pB], a.vy[pB], a.vz[pB]);
  const d3 av1 = mk3(a.wx[pA], a.wy[pA], a.wz[pA]), av2 = mk3(a.wx[pB], a.wy[pB], a.wz[pB]);

  // Cg2_Any_Any_L6Geom__Base::handleSpheresLikeContact (L6Geom.cpp:49-216)
  d3 T0, T1, T2, gvel;
  if (!real) {
    // setInitialLocalCoords (L6Geom.cpp:25-36)
    d3 locX = normal;
    d3 e = fabs(locX.y) < fabs(locX.z) ? mk3(0., 1., 0.) : mk3(0., 0., 1.);
    d3 locY = cross3(locX, e);
    locY = sub3(locY, mul3(locX, dot3(locY, locX)));
    locY = normalized3(locY);
    T0 = locX;
    T1 = locY;
    T2 = cross3(locX, locY);
    gvel = mk3(0., 0., 0.);
  } else {
    const double dt = a.dt;
    const d3 prevNormal = mk3(a.N0[c], a.N1[c], a.N2[c]);
    const d3 prevY = mk3(a.Y0[c], a.Y1[c], a.Y2[c]);
    const d3 prevContPt = mk3(a.P0[c], a.P1[c], a.P2[c]);
    const d3& currNormal = normal;
    d3 normRotVec = cross3(prevNormal, currNormal);
    d3 midNormal = mul3(add3(prevNormal, currNormal), .5);
    if (sqrt(sqn3(midNormal)) == 0) midNormal = currNormal;
    else midNormal = normalized3(midNormal);
    double twist = dot3(midNormal, add3(av1, av2));
    d3 normTwistVec = mul3(mul3(mul3(midNormal, dt), .5), twist);
    d3 rt = add3(normRotVec, normTwistVec);
    d3 M0 = midNormal;
    d3 M1 = sub3(prevY, div3(cross3(prevY, rt), 2.));
    d3 M2 = cross3(M0, M1);
    T0 = currNormal;
    T1 = sub3(prevY, cross3(M1, rt));
    T2 = cross3(T0, T1);
    if (a.renorm) {
      T0 = normalized3(T0);
      T1 = sub3(T1, mul3(T0, dot3(T1, T0)));
      T1 = normalized3(T1);
      T2 = cross3(T0, T1);
    }
    d3 midContPt = mul3(add3(prevContPt, contPt), .5);
    d3 midPos1 = sub3(pos1, mul3(vel1, dt / 2.));
    d3 midPos2 = sub3(pos2, mul3(vel2, dt / 2.));
    d3 c1x = sub3(midContPt, midPos1);
    d3 c2x = sub3(midContPt, midPos2);
    d3 relVel = sub3(add3(vel2, cross3(av2, c2x)), add3(vel1, cross3(av1, c1x)));
    gvel = mk3(dot3(M0, relVel), dot3(M1, relVel), dot3(M2, relVel));
  }

  // Cp2 (recomputed; constant over the life of the contact)
  double l0, l1, A;
  lens_area(r1, r2, l0, l1, A);
  const wdg_material mA = a.mats[a.mat[pA]], mB = a.mats[a.mat[pB]];
  double Fn, Ft0 = real ? a.F1[c] : 0., Ft1 = real ? a.F2[c] : 0.;
  if (LAW == WDG_LAW_IDEALELPL) {
    // Law2_L6Geom_FrictPhys_IdealElPl::go (IdealElPl.cpp:15-125)
    PhysVals ph;
    phys_frict(mA, mB, l0, l1, A, ph);
    if (uN > 0 && !a.noBreak) {
      if (real) a.real[c] = 0u;
      return;
    }
    Fn = ph.kn * uN;
    if (a.noFrict || ph.tanPhi == 0.) {
      Ft0 = 0.;
      Ft1 = 0.;
    } else {
      double dk = a.dt * ph.kt;
      Ft0 = Ft0 + dk * gvel.y;
      Ft1 = Ft1 + dk * gvel.z;
      double maxFt = fabs(Fn) * ph.tanPhi;
      if (Ft0 * Ft0 + Ft1 * Ft1 > maxFt * maxFt && !a.noSlip) {
        double FtNorm = sqrt(Ft0 * Ft0 + Ft1 * Ft1);
        double ratio = maxFt / FtNorm;
        if (FtNorm == 0.) ratio = 0.;
        Ft0 = Ft0 * ratio;
        Ft1 = Ft1 * ratio;
      }
      if (isnan(Fn) || isnan(Ft0) || isnan(Ft1)) atomicOr(&a.dflags[1], ERR_NAN_FORCE);
    }
  } else {
    // Law2_L6Geom_HertzPhys_DMT::go (Hertz.cpp:85-210), DMT branch
    PhysVals ph;
    phys_hertz(mA, mB, l0, l1, a.mass[pA], a.mass[pB], a.poisson, a.en, a.viscAlpha, ph);
    const double velN = gvel.x;
    double kn0 = ph.K * sqrt(ph.R);
    if (uN > 0) {
      if (real) a.real[c] = 0u;
      return;
    }
    double s = sqrt(-uN);
    double Fne = -kn0 * ((s * s) * s);
    double Fna = 4 * M_PI * ph.R * ph.gamma;
    double eta = (ph.alpha_sqrtMK > 0 ? ph.alpha_sqrtMK * sqrt(sqrt(fmax(-uN, 0.))) : 0.);
    double Fnv = eta * velN;
    if (ph.alpha == 0. && a.noAttraction && Fne + Fnv > 0) Fnv = -Fne;
    Fn = Fne + Fna + Fnv;
    double kt = ph.kt0 * sqrt(uN < 0 ? -uN : 0);
    double dk = a.dt * kt;
    Ft0 = Ft0 + dk * gvel.y;
    Ft1 = Ft1 + dk * gvel.z;
    double maxFt = fabs(fmin(0., Fn) * ph.tanPhi);
    if (Ft0 * Ft0 + Ft1 * Ft1 > p2(maxFt)) {
      double FtNorm = sqrt(Ft0 * Ft0 + Ft1 * Ft1);
      double ratio = maxFt / FtNorm;
      Ft0 = Ft0 * ratio;
      Ft1 = Ft1 * ratio;
    } else {
      Ft0 = Ft0 + eta * gvel.y;
      Ft1 = Ft1 + eta * gvel.z;
    }
  }
  if (!real) a.real[c] = 1u;
  a.N0[c] = T0.x; a.N1[c] = T0.y; a.N2[c] = T0.z;
  a.Y0[c] = T1.x; a.Y1[c] = T1.y; a.Y2[c] = T1.z;
  a.P0[c] = contPt.x; a.P1[c] = contPt.y; a.P2[c] = contPt.z;
  a.F0[c] = Fn; a.F1[c] = Ft0; a.F2[c] = Ft1;
  a.uN[c] = uN;
  // Contact::getForceTorqueBranch: F = node.ori * force, node.ori = Quaternion(trsf)
  q4 q = quat_from_cols(T0, T1, T2);
  d3 G = quat_rot(q, mk3(Fn, Ft0, Ft1));
  a.G0[c] = G.x; a.G1[c] = G.y; a.G2[c] = G.z;
}

void contact_loop(Ctx& c, const wdg_contact_params* p) {
  Contacts& C = c.con();
  if (c.mats.empty() || !c.dMats.p) throw Error(WDG_E_STATE, "no materials set (wdg_set_materials)");
  if (p->lawKind == WDG_LAW_HERTZDMT && p->physKind != WDG_PHYS_HERTZ)
    throw Error(WDG_E_ARG, "Law2_L6Geom_HertzPhys_DMT needs Cp2_HertzMat_HertzPhys");
  if (p->lawKind == WDG_LAW_IDEALELPL && p->physKind != WDG_PHYS_FRICT)
    throw Error(WDG_E_ARG, "Law2_L6Geom_FrictPhys_IdealElPl needs Cp2_FrictMat_FrictPhys");
  if (p->lawKind == WDG_LAW_HERTZDMT)
    for (const auto& m : c.mats)
      if (!(m.alpha == 0. || m.surfEnergy == 0.))
        throw Error(WDG_E_ARG, "Schwarz adhesion (HertzMat.alpha and surfEnergy both non-zero) is not supported");
  if (c.forcesPending) flush_forces(c);
  if (C.n == 0 || c.n == 0) {
    c.forcesPending = false;
    return;
  }
  ContactArgs a{};
  a.nc = C.n;
  a.A = C.A.p; a.B = C.B.p; a.info = C.info.p; a.dist = C.dist.p; a.real = C.real.p;
  a.N0 = C.hN[0].p; a.N1 = C.hN[1].p; a.N2 = C.hN[2].p;
  a.Y0 = C.hY[0].p; a.Y1 = C.hY[1].p; a.Y2 = C.hY[2].p;
  a.P0 = C.hP[0].p; a.P1 = C.hP[1].p; a.P2 = C.hP[2].p;
  a.F0 = C.hF[0].p; a.F1 = C.hF[1].p; a.F2 = C.hF[2].p; a.uN = C.huN.p;
  a.G0 = C.F[0].p; a.G1 = C.F[1].p; a.G2 = C.F[2].p;
  a.px = c.pos[0].p; a.py = c.pos[1].p; a.pz = c.pos[2].p;
  a.vx = c.vel[0].p; a.vy = c.vel[1].p; a.vz = c.vel[2].p;
  a.wx = c.angVel[0].p; a.wy = c.angVel[1].p; a.wz = c.angVel[2].p;
  a.rad = c.radius.p; a.mass = c.mass.p; a.flags = c.flags.p; a.mat = c.mat.p; a.mats = c.dMats.p;
  for (int k = 0; k < 3; k++) a.L[k] = c.field.cell[k];
  a.absDf = std::fabs(c.field.distFactor);
  a.dt = p->dt;
  a.renorm = (p->trsfRenorm > 0 && (p->step % p->trsfRenorm) == 0) ? 1 : 0;
  a.poisson = p->poisson;
  a.en = p->en;
  a.viscAlpha = (p->en > 0 && p->en < 1.)
                    ? -std::sqrt(5) * std::log(p->en) / (std::sqrt(std::log(p->en) * std::log(p->en) + M_PI * M_PI))
                    : 0.;
  c.lastPhysKind = p->physKind;
  c.lastPoisson = p->poisson;
  c.lastEn = p->en;
  c.lastViscAlpha = a.viscAlpha;
  a.noSlip = p->noSlip; a.noBreak = p->noBreak; a.noFrict = p->noFrict; a.noAttraction = p->noAttraction;
  a.dflags = c.dflags.p;
  const bool per = c.field.periodic != 0;
  dim3 grid(nblocks(C.n, 256)), block(256);
  if (p->lawKind == WDG_LAW_HERTZDMT) {
    if (per) hipLaunchKernelGGL((k_contact<true, WDG_LAW_HERTZDMT>), grid, block, 0, c.stream, a);
    else hipLaunchKernelGGL((k_contact<false, WDG_LAW_HERTZDMT>), grid, block, 0, c.stream, a);
  } else {
    if (per) hipLaunchKernelGGL((k_contact<true, WDG_LAW_IDEALELPL>), grid, block, 0, c.stream, a);
    else hipLaunchKernelGGL((k_contact<false, WDG_LAW_IDEALELPL>), grid, block, 0, c.stream, a);
  }
  WDG_LAUNCH_CHECK();
  c.forcesPending = p->applyForces != 0;
}

}  // namespace wdg
