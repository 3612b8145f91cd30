// Per-particle kernels: deterministic force/torque gather over the contact
// slots (no float atomics), fused into the leapfrog integrator
// (Leapfrog.cpp:104-297, spherical integrator, gradV zero), ForceResetter
// (Leapfrog.cpp:15-26), DynDt stiffness (DynDt.cpp:16-71) and observables
// (Funcs.cpp:84-114).
//
// Each particle walks its adjacency list, which holds its slots in increasing
// order of the other particle's user id: the order of Particle::contacts (a
// std::map keyed by the other id), so sums are reproducible bit for bit.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "wdg_internal.hpp"
#include "wdg_math.hpp"
#include "wdg_physics.hpp"

namespace wdg {

struct GatherArgs {
  int64_t n, nW;
  const uint32_t *adj, *adjOff;
  const uint32_t* sidx;  // state index of each real slot (null: found from the real bits)
  const uint32_t* eidx;  // per adjacency entry its slot when real, else NO_STATE (null: not kept)
  const uint32_t* flags;
  const uint32_t* gstop;  // graph replay: a set stop word skips the kernel
  RealView rv;
  const double* gp;
  const double* gtq;  // global phys torque per real contact (rolling resistance), else null
  double L[3];
};

// F32: the FP32 contact form's record (global force, branch vectors from pA and from pB's
// image as floats); the sums stay FP64
template <bool PERIODIC, bool F32>
__device__ __forceinline__ void gather_one(const GatherArgs& g, uint32_t e, d3 pos, d3& F, d3& T) {
  {
    const uint32_t slot = e & ~ADJ_PA;
    if (!rv_real(g.rv, slot)) return;
    if (F32) {
      const float4* r =
          reinterpret_cast<const float4*>(reinterpret_cast<const float*>(g.gp) + (int64_t)GP32 * rv_index(g.rv, slot));
      const float4 r0 = r[0], r1 = r[1], r2 = r[2];
      const bool isPA = (e & ADJ_PA) != 0u;
      const d3 Fc = mul3(mk3(r0.x, r0.y, r0.z), isPA ? 1. : -1.);
      const d3 xc = isPA ? mk3(r0.w, r1.x, r1.y) : mk3(r1.z, r1.w, r2.x);
      F = add3(F, Fc);
      T = add3(T, cross3(xc, Fc));
      return;
    }
    const double2* r = reinterpret_cast<const double2*>(g.gp + (int64_t)GP * rv_index(g.rv, slot));
    const double2 r0 = r[0], r1 = r[1], r2 = r[2];
    const bool isPA = (e & ADJ_PA) != 0u;
    const double sign = isPA ? 1. : -1.;
    const d3 Fc = mul3(mk3(r0.x, r0.y, r1.x), sign);
    d3 shift = mk3(0., 0., 0.);
    if (PERIODIC && !isPA) {
      const uint64_t cd = (uint64_t)__double_as_longlong(r[3].x);
      shift = mk3(g.L[0] * unpack_cd(cd, 0), g.L[1] * unpack_cd(cd, 1), g.L[2] * unpack_cd(cd, 2));
    }
    // Contact::getForceTorqueBranch: xc = contact point - (node pos + periodic shift of pB)
    const d3 xc = sub3(mk3(r1.y, r2.x, r2.y), add3(pos, shift));
    F = add3(F, Fc);
    if (g.gtq) {
      // applyForceUninodal: xc x F + T, T the phys torque (rolling resistance) times the sign
      const double* t = g.gtq + (int64_t)3 * rv_index(g.rv, slot);
      T = add3(T, add3(cross3(xc, Fc), mul3(mk3(t[0], t[1], t[2]), sign)));
    } else {
      T = add3(T, cross3(xc, Fc));
    }
  }
}

// one contact's force and torque on the particle at pos: the record at state index q, the
// particle being the contact's pA or pB
template <bool PERIODIC, bool F32>
__device__ __forceinline__ void gather_record(const GatherArgs& g, uint32_t q, bool isPA, d3 pos, d3& F, d3& T) {
  if (F32) {
    const float4* r = reinterpret_cast<const float4*>(reinterpret_cast<const float*>(g.gp) + (int64_t)GP32 * q);
    const float4 r0 = r[0], r1 = r[1], r2 = r[2];
    const d3 Fc = mul3(mk3(r0.x, r0.y, r0.z), isPA ? 1. : -1.);
    const d3 xc = isPA ? mk3(r0.w, r1.x, r1.y) : mk3(r1.z, r1.w, r2.x);
    F = add3(F, Fc);
    T = add3(T, cross3(xc, Fc));
  } else {
    const double2* r = reinterpret_cast<const double2*>(g.gp + (int64_t)GP * q);
    const double2 r0 = r[0], r1 = r[1], r2 = r[2];
    const d3 Fc = mul3(mk3(r0.x, r0.y, r1.x), isPA ? 1. : -1.);
    d3 shift = mk3(0., 0., 0.);
    if (PERIODIC && !isPA) {
      const uint64_t cd = (uint64_t)__double_as_longlong(r[3].x);
      shift = mk3(g.L[0] * unpack_cd(cd, 0), g.L[1] * unpack_cd(cd, 1), g.L[2] * unpack_cd(cd, 2));
    }
    // Contact::getForceTorqueBranch: xc = contact point - (node pos + periodic shift of pB)
    const d3 xc = sub3(mk3(r1.y, r2.x, r2.y), add3(pos, shift));
    F = add3(F, Fc);
    if (g.gtq) {
      // applyForceUninodal: xc x F + T, T the phys torque (rolling resistance) times the sign
      const double* t = g.gtq + (int64_t)3 * q;
      T = add3(T, add3(cross3(xc, Fc), mul3(mk3(t[0], t[1], t[2]), isPA ? 1. : -1.)));
    } else {
      T = add3(T, cross3(xc, Fc));
    }
  }
}

// With the per-entry real slots the list is one contiguous array: EIDX_BATCH of its entries
// are loaded together, then the state indices of the real ones (usually none or one), then
// their records, summed in list order (the order of the entry-by-entry walk, so the sums are
// bit-identical to it).
constexpr int EIDX_BATCH = 4;
template <bool PERIODIC, bool F32>
__device__ __forceinline__ void gather_all(const GatherArgs& g, int64_t i, d3 pos, d3& F, d3& T) {
  const uint32_t k0 = g.adjOff[i], k1 = g.adjOff[i + 1];
  if (g.eidx && g.sidx) {
    for (uint32_t k = k0; k < k1; k += EIDX_BATCH) {
      uint32_t v[EIDX_BATCH], q[EIDX_BATCH];
#pragma unroll
      for (int j = 0; j < EIDX_BATCH; j++) v[j] = k + j < k1 ? g.eidx[k + j] : NO_STATE;
#pragma unroll
      for (int j = 0; j < EIDX_BATCH; j++) q[j] = v[j] != NO_STATE ? g.sidx[v[j] & ~EIDX_PA] : 0u;
#pragma unroll
      for (int j = 0; j < EIDX_BATCH; j++)
        if (v[j] != NO_STATE) gather_record<PERIODIC, F32>(g, q[j], (v[j] & EIDX_PA) != 0u, pos, F, T);
    }
    return;
  }
  // without them (a real set not made by the contact loop): entry by entry through the real bits
  for (uint32_t k = k0; k < k1; k++) gather_one<PERIODIC, F32>(g, g.adj[k], pos, F, T);
}

// Contact forces on the boundary bodies other than facets (storage [0, nW), listed in
// gbody), which may touch up to every sphere (a floor under a spread pile): each body's list is
// cut into `parts` equal parts (set at each refresh from the largest such list), one workgroup
// each, whose threads sum every 256th entry of their part in order, then a fixed-order tree;
// k_gather_bnd_sum adds the parts in order to the value in the force arrays (which always hold
// the current force: the reset value, or the gathered one). The spheres' gathers (k_gather,
// k_leapfrog) skip these bodies. Deterministic; the summation order differs from the
// sequential one, by rounding.
constexpr int BND_PARTS_MAX = 64;
template <bool PERIODIC, bool F32>
__global__ __launch_bounds__(256) void k_gather_bnd(GatherArgs g, const uint32_t* __restrict__ gbody, int parts,
                                                     double* part, const double* __restrict__ px,
                                                     const double* __restrict__ py, const double* __restrict__ pz) {
  if (g.gstop && *g.gstop) return;
  const int64_t b = blockIdx.x / parts;
  const int p = (int)(blockIdx.x % parts);
  const int64_t i = gbody[b];
  const uint32_t k0 = g.adjOff[i], k1 = g.adjOff[i + 1];
  const uint32_t chunk = (k1 - k0 + parts - 1) / parts;
  const uint32_t a0 = k0 + (uint32_t)p * chunk, a1 = a0 + chunk < k1 ? a0 + chunk : k1;
  const d3 pos = mk3(px[i], py[i], pz[i]);
  d3 F = mk3(0., 0., 0.), T = mk3(0., 0., 0.);
  for (uint32_t k = a0 + threadIdx.x; k < a1; k += blockDim.x) gather_one<PERIODIC, F32>(g, g.adj[k], pos, F, T);
  double v[6] = {F.x, F.y, F.z, T.x, T.y, T.z};
  block_sum_store<6, 256>(v, part + ((int64_t)b * parts + p) * 6);
}
__global__ void k_gather_bnd_sum(int64_t nb, const uint32_t* __restrict__ gbody, int parts,
                                 const uint32_t* __restrict__ adjOff, const uint32_t* gstop,
                                 const double* __restrict__ part, double* fx, double* fy, double* fz, double* tx,
                                 double* ty, double* tz) {
  if (gstop && *gstop) return;
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int64_t i = gbody[b];
  if (adjOff[i] == adjOff[i + 1]) return;
  double v[6] = {0., 0., 0., 0., 0., 0.};
  for (int p = 0; p < parts; p++)
    for (int q = 0; q < 6; q++) v[q] += part[((int64_t)b * parts + p) * 6 + q];
  fx[i] = fx[i] + v[0]; fy[i] = fy[i] + v[1]; fz[i] = fz[i] + v[2];
  tx[i] = tx[i] + v[3]; ty[i] = ty[i] + v[4]; tz[i] = tz[i] + v[5];
}

template <bool PERIODIC, bool F32>
__global__ __launch_bounds__(256) void k_gather(GatherArgs g, double* fx, double* fy, double* fz, double* tx,
                                                 double* ty, double* tz, const double* __restrict__ px,
                                                 const double* __restrict__ py, const double* __restrict__ pz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  if (g.adjOff[i] == g.adjOff[i + 1]) return;
  // boundary bodies: k_gather_bnd; facets take no force from the contact loop (multinodal)
  if (i < g.nW) return;
  d3 F = mk3(fx[i], fy[i], fz[i]), T = mk3(tx[i], ty[i], tz[i]);
  gather_all<PERIODIC, F32>(g, i, mk3(px[i], py[i], pz[i]), F, T);
  fx[i] = F.x; fy[i] = F.y; fz[i] = F.z;
  tx[i] = T.x; ty[i] = T.y; tz[i] = T.z;
}

struct LeapArgs {
  int64_t n;
  double *px, *py, *pz, *vx, *vy, *vz, *wx, *wy, *wz, *qw, *qx, *qy, *qz, *fx, *fy, *fz, *tx, *ty, *tz;
  const double *mass, *ix, *iy, *iz;
  const double *lx, *ly, *lz;
  const uint32_t* flags;
  double dt, damping;
  int reset;
  double g[3];
  int hasGravity;
  int gather;
  double maxD2;
  const double* vdS;  // per-scene Verlet distances (else null: maxD2)
  int checkDirty;
  int clean;  // force/torque of integrated particles hold the reset value
  uint32_t* dirty;      // set when a particle left its Verlet skin
  uint32_t* dirtyNext;  // cleared for the next launch
  const double* dtScene;  // batched scenes: dt of each particle's scene (else null, dt above)
  const uint32_t* scene;
  const uint32_t* gstop;  // graph replay: a set stop word skips the kernel
  double* enPart;         // energy tracking: kinetic, grav, nonviscDamp partials per workgroup
  double* enP;            // energy tracking per scene: the three terms of each particle, else null
};

// std::max's comparison, so that a NaN propagates as in the reference's isinf(maxCoeff)
__device__ __forceinline__ double max_cmp(double x, double y) { return x < y ? y : x; }

// ENERGY: Leapfrog::doDampingDissipation, doGravityWork and doKineticEnergy (Leapfrog.cpp:50-98,
// called at Leapfrog.cpp:156-159 and 218) summed over the workgroup into a.enPart[3 * blockIdx.x ..]
template <bool PERIODIC, bool F32, bool ENERGY>
__global__ __launch_bounds__(256) void k_leapfrog(LeapArgs a, GatherArgs g) {
  if (a.gstop && *a.gstop) return;
  const int64_t i = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  bool dirty = false;
  double en[3] = {0., 0., 0.};
  if (i < a.n) {
    const uint32_t fl = a.flags[i];
    const bool integrated = (fl & F_INTEGRATED) != 0u;
    const bool grav = a.hasGravity && !(fl & F_GRAVITY_SKIP);
    // a boundary body's contact forces were summed into its arrays by k_gather_bnd
    const bool pre = a.gather && i < g.nW;
    d3 pos = mk3(a.px[i], a.py[i], a.pz[i]);
    d3 F, T;
    if (a.clean && integrated && !pre) {
      // the value the last reset stored (ForceResetter / Leapfrog.reset)
      const double m = a.mass[i];
      F = grav ? mk3(m * a.g[0], m * a.g[1], m * a.g[2]) : mk3(0., 0., 0.);
      T = mk3(0., 0., 0.);
    } else {
      F = mk3(a.fx[i], a.fy[i], a.fz[i]);
      T = mk3(a.tx[i], a.ty[i], a.tz[i]);
    }
    if (a.gather && !pre) gather_all<PERIODIC, F32>(g, i, pos, F, T);
    if (!integrated) {
      if (a.gather && !pre) {
        a.fx[i] = F.x; a.fy[i] = F.y; a.fz[i] = F.z;
        a.tx[i] = T.x; a.ty[i] = T.y; a.tz[i] = T.z;
      }
    } else {
      const double dt = a.dtScene ? a.dtScene[a.scene[i]] : a.dt, damping = a.damping;
      const bool damp = (damping != 0. && !(fl & F_DAMPING_SKIP));
      const uint32_t bl = fl & F_DOF_ALL;
      const double m = a.mass[i];
      d3 vel = mk3(a.vx[i], a.vy[i], a.vz[i]);
      d3 av = mk3(a.wx[i], a.wy[i], a.wz[i]);
      // on-step velocities v(t) = v(t-dt/2) + a dt/2 for the kinetic energy (zero when all
      // DOFs are blocked)
      d3 cv = mk3(0., 0., 0.), cw = mk3(0., 0., 0.);
      if (ENERGY) {
        if (damp && !(fl & F_ENERGY_SKIP)) {
          const double fv = (fabs(vel.x) * fabs(F.x) + fabs(vel.y) * fabs(F.y)) + fabs(vel.z) * fabs(F.z);
          const double tw = (fabs(av.x) * fabs(T.x) + fabs(av.y) * fabs(T.y)) + fabs(av.z) * fabs(T.z);
          en[2] = fv * damping * dt + tw * damping * dt;
        }
        if (grav) {
          if (bl == 0) {
            en[1] = -((a.g[0] * vel.x + a.g[1] * vel.y) + a.g[2] * vel.z) * m * dt;
          } else {
            const double v[3] = {vel.x, vel.y, vel.z};
            double gr = 0.;
            for (int k = 0; k < 3; k++)
              if (!(bl & (1u << k))) gr -= a.g[k] * v[k] * m * dt;
            en[1] = gr;
          }
        }
      }
      if (bl != F_DOF_ALL) {
        double f[3] = {F.x, F.y, F.z}, t[3] = {T.x, T.y, T.z};
        double v[3] = {vel.x, vel.y, vel.z}, w[3] = {av.x, av.y, av.z};
        double I[3] = {a.ix[i], a.iy[i], a.iz[i]};
        double la[3], aa[3];
        for (int k = 0; k < 3; k++) {
          if (bl == 0) la[k] = f[k] / m;
          else la[k] = (bl & (1u << k)) ? 0. : 0. + f[k] / m;
        }
        if (damp)
          for (int k = 0; k < 3; k++) la[k] *= 1 - damping * sgn(f[k] * (v[k] + 0.5 * dt * la[k]));
        for (int k = 0; k < 3; k++) v[k] += dt * la[k];
        if (!(I[0] == 0 && I[1] == 0 && I[2] == 0)) {
          for (int k = 0; k < 3; k++) {
            if (bl == 0) aa[k] = t[k] / I[k];
            else aa[k] = (bl & (8u << k)) ? 0. : 0. + t[k] / I[k];
          }
          if (damp)
            for (int k = 0; k < 3; k++) aa[k] *= 1 - damping * sgn(t[k] * (w[k] + 0.5 * dt * aa[k]));
          for (int k = 0; k < 3; k++) w[k] += dt * aa[k];
        } else {
          aa[0] = aa[1] = aa[2] = 0.;
        }
        if (ENERGY) {
          cv = mk3(vel.x + .5 * dt * la[0], vel.y + .5 * dt * la[1], vel.z + .5 * dt * la[2]);
          cw = mk3(av.x + .5 * dt * aa[0], av.y + .5 * dt * aa[1], av.z + .5 * dt * aa[2]);
        }
        vel = mk3(v[0], v[1], v[2]);
        av = mk3(w[0], w[1], w[2]);
      }
      if (ENERGY && !(fl & F_ENERGY_SKIP)) {
        const double Ix = a.ix[i], Iy = a.iy[i], Iz = a.iz[i];
        const double Et = .5 * m * ((cv.x * cv.x + cv.y * cv.y) + cv.z * cv.z);
        double Er = 0.5 * ((cw.x * (Ix * cw.x) + cw.y * (Iy * cw.y)) + cw.z * (Iz * cw.z));
        if (isnan(Er) || isinf(max_cmp(max_cmp(Ix, Iy), Iz))) Er = 0;
        en[0] = Et + Er;
      }
      pos = mk3(pos.x + vel.x * dt, pos.y + vel.y * dt, pos.z + vel.z * dt);
      a.px[i] = pos.x; a.py[i] = pos.y; a.pz[i] = pos.z;
      a.vx[i] = vel.x; a.vy[i] = vel.y; a.vz[i] = vel.z;
      a.wx[i] = av.x; a.wy[i] = av.y; a.wz[i] = av.z;
      // Leapfrog.cpp:240-244; when the arrays already hold the reset value nothing is stored
      if (a.reset) {
        if (!a.clean || pre) {
          F = grav ? mk3(m * a.g[0], m * a.g[1], m * a.g[2]) : mk3(0., 0., 0.);
          a.fx[i] = F.x; a.fy[i] = F.y; a.fz[i] = F.z;
          a.tx[i] = 0.; a.ty[i] = 0.; a.tz[i] = 0.;
        }
      } else {
        a.fx[i] = F.x; a.fy[i] = F.y; a.fz[i] = F.z;
        a.tx[i] = T.x; a.ty[i] = T.y; a.tz[i] = T.z;
      }
      if (a.checkDirty) {
        d3 d = sub3(mk3(a.lx[i], a.ly[i], a.lz[i]), pos);
        double d2 = 0;
        d2 = fmax(d2, sqn3(d));
        double m2 = a.maxD2;
        if (a.vdS) {
          const double v = a.vdS[a.scene[i]];
          m2 = v * v;
        }
        dirty = d2 + 0. > m2;
      }
      // leapfrogSphericalRotate (Leapfrog.cpp:280-297)
      q4 ori = q4{a.qw[i], a.qx[i], a.qy[i], a.qz[i]};
      if (!(av.x == 0 && av.y == 0 && av.z == 0)) {
        double angle = sqrt(sqn3(av));
        d3 axis = div3(av, angle);
        double ha = 0.5 * (angle * dt);
        double sh, ch;
        sincos(ha, &sh, &ch);
        q4 q = q4{ch, sh * axis.x, sh * axis.y, sh * axis.z};
        ori = quat_mul(q, ori);
      }
      double qn = sqrt(ori.w * ori.w + ori.x * ori.x + ori.y * ori.y + ori.z * ori.z);
      a.qw[i] = ori.w / qn; a.qx[i] = ori.x / qn; a.qy[i] = ori.y / qn; a.qz[i] = ori.z / qn;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *a.dirtyNext = 0u;
  if (__ballot(dirty) != 0ull && (threadIdx.x & 63) == 0) atomicOr(a.dirty, 1u);
  if (ENERGY && a.enP) {
    if (i < a.n) {
      double* e = a.enP + 3 * i;
      e[0] = en[0]; e[1] = en[1]; e[2] = en[2];
    }
  } else if (ENERGY) {
    block_sum_store<3, 256>(en, a.enPart + (int64_t)3 * blockIdx.x);
  }
}

__global__ void k_force_reset(int64_t n, const uint32_t* __restrict__ flags, const double* __restrict__ mass,
                              double g0, double g1, double g2, int hasGravity, double* fx, double* fy, double* fz,
                              double* tx, double* ty, double* tz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t fl = flags[i];
  if (!(fl & F_INTEGRATED)) return;
  const double m = mass[i];
  const bool grav = hasGravity && !(fl & F_GRAVITY_SKIP);
  fx[i] = grav ? m * g0 : 0.;
  fy[i] = grav ? m * g1 : 0.;
  fz[i] = grav ? m * g2 : 0.;
  tx[i] = 0.; ty[i] = 0.; tz[i] = 0.;
}

// IntraForce with In2_Facet (Facet.cpp:301-328, bary=true): one workgroup per facet spreads
// each real contact's force over the facet's 3 nodes, weighted by the barycentric coordinates
// (CompUtils::triangleBarycentrics, CompUtils.cpp:94-112: the 2x2 system of x - A = u (B - A)
// + v (C - A) dotted with B - A and C - A, solved by the adjugate over the determinant) of the
// contact point projected onto the facet plane (normal of Facet::getNormal, Facet.cpp:74-77).
// Contact::getForceTorqueBranch (Contact.cpp:24-34) per node: F = the force on the facet,
// xc = contact point - node position (minus the periodic shift when the facet is pB); node i
// gets w_i F and xc x (w_i F) (the laws here have no contact torque). The workgroup takes its
// facet's entries FN_BS at a time, each thread one entry, and 18 threads then add one
// component each over them in list order: the sums run in the order of Particle::contacts
// (increasing other id), as in the sequential loop, so they are bit-identical to it. A facet
// without real contacts gets zeros (the reset value of its massless nodes).
constexpr int FN_BS = 256;
template <bool PERIODIC, bool F32>
__global__ __launch_bounds__(FN_BS) void k_facet_nodes(GatherArgs g, const double* __restrict__ fgeo,
                                                        const double* __restrict__ px, const double* __restrict__ py,
                                                        const double* __restrict__ pz, double* __restrict__ out) {
  __shared__ double sh[18][FN_BS];
  if (g.gstop && *g.gstop) return;
  const int64_t i = blockIdx.x;
  if (flag_shape(g.flags[i]) != WDG_FACET) return;
  const int t = threadIdx.x;
  const uint32_t k0 = g.adjOff ? g.adjOff[i] : 0u, k1 = g.adjOff ? g.adjOff[i + 1] : 0u;
  const double* fg = fgeo + (int64_t)FGEO * i;
  const d3 v[3] = {mk3(fg[0], fg[1], fg[2]), mk3(fg[3], fg[4], fg[5]), mk3(fg[6], fg[7], fg[8])};
  const d3 vB = sub3(v[1], v[0]), vC = sub3(v[2], v[0]);
  const d3 normal = normalized3(cross3(vB, vC));
  const double m00 = dot3(vB, vB), m01 = dot3(vC, vB), m10 = dot3(vB, vC), m11 = dot3(vC, vC);
  const double invdet = 1. / (m00 * m11 - m10 * m01);
  const double a00 = m11 * invdet, a01 = -m01 * invdet, a10 = -m10 * invdet, a11 = m00 * invdet;
  double acc = 0.;
  for (uint32_t base = k0; base < k1; base += FN_BS) {
    double cv[18];
#pragma unroll
    for (int q = 0; q < 18; q++) cv[q] = 0.;
    const uint32_t k = base + (uint32_t)t;
    bool real = false, isPA = true;
    uint32_t q = 0;
    if (k < k1) {
      if (g.eidx && g.sidx) {
        const uint32_t e = g.eidx[k];
        if (e != NO_STATE) {
          real = true;
          isPA = (e & EIDX_PA) != 0u;
          q = g.sidx[e & ~EIDX_PA];
        }
      } else {
        const uint32_t e = g.adj[k], slot = e & ~ADJ_PA;
        if (rv_real(g.rv, slot)) {
          real = true;
          isPA = (e & ADJ_PA) != 0u;
          q = rv_index(g.rv, slot);
        }
      }
    }
    if (real) {
      d3 Fc, cp, shift = mk3(0., 0., 0.), Tc = mk3(0., 0., 0.);
      bool hasT = false;
      if (F32) {
        // the FP32 record: force on pA, and the branches from pA and from pB's image
        const float4* r = reinterpret_cast<const float4*>(reinterpret_cast<const float*>(g.gp) + (int64_t)GP32 * q);
        const float4 r0 = r[0], r1 = r[1], r2 = r[2];
        Fc = mul3(mk3(r0.x, r0.y, r0.z), isPA ? 1. : -1.);
        const d3 br = isPA ? mk3(r0.w, r1.x, r1.y) : mk3(r1.z, r1.w, r2.x);
        cp = add3(mk3(px[i], py[i], pz[i]), br);
      } else {
        const double2* r = reinterpret_cast<const double2*>(g.gp + (int64_t)GP * q);
        const double2 r0 = r[0], r1 = r[1], r2 = r[2];
        Fc = mul3(mk3(r0.x, r0.y, r1.x), isPA ? 1. : -1.);
        cp = mk3(r1.y, r2.x, r2.y);
        if (PERIODIC && !isPA) {
          const uint64_t cd = (uint64_t)__double_as_longlong(r[3].x);
          shift = mk3(g.L[0] * unpack_cd(cd, 0), g.L[1] * unpack_cd(cd, 1), g.L[2] * unpack_cd(cd, 2));
        }
        if (g.gtq) {
          // the phys torque (rolling resistance): node i gets w_i T besides xc x (w_i F)
          const double* tq = g.gtq + (int64_t)3 * q;
          hasT = !(tq[0] == 0. && tq[1] == 0. && tq[2] == 0.);
          Tc = mul3(mk3(tq[0], tq[1], tq[2]), isPA ? 1. : -1.);
        }
      }
      const d3 p = sub3(cp, mul3(normal, dot3(sub3(cp, v[0]), normal)));
      const d3 vX = sub3(p, v[0]);
      const double r0 = dot3(vX, vB), r1 = dot3(vX, vC);
      const double u = a00 * r0 + a01 * r1, w2 = a10 * r0 + a11 * r1;
      const double w[3] = {1. - u - w2, u, w2};
#pragma unroll
      for (int n = 0; n < 3; n++) {
        const d3 xc = sub3(cp, add3(v[n], shift));
        const d3 F = mul3(Fc, w[n]);
        const d3 T = hasT ? add3(cross3(xc, F), mul3(Tc, w[n])) : cross3(xc, F);
        cv[3 * n] = F.x; cv[3 * n + 1] = F.y; cv[3 * n + 2] = F.z;
        cv[9 + 3 * n] = T.x; cv[9 + 3 * n + 1] = T.y; cv[9 + 3 * n + 2] = T.z;
      }
    }
#pragma unroll
    for (int c = 0; c < 18; c++) sh[c][t] = cv[c];
    __syncthreads();
    if (t < 18) {
      const uint32_t m = k1 - base < (uint32_t)FN_BS ? k1 - base : (uint32_t)FN_BS;
      for (uint32_t j = 0; j < m; j++) acc += sh[t][j];
    }
    __syncthreads();
  }
  if (t < 18) out[18 * i + t] = acc;
}

static GatherArgs gather_args(Ctx& c) {
  Contacts& C = c.con();
  GatherArgs g{};
  g.n = c.n;
  g.adj = C.adj.p; g.adjOff = C.adjOff.p; g.rv = c.rview(); g.gp = c.reals().gp.p; g.flags = c.flags.p;
  if (c.rollActive) g.gtq = c.reals().gtq.p;
  g.sidx = c.reals().sidxValid ? c.reals().sidx.p : nullptr;
  g.eidx = c.eidxValid ? C.eidx.p : nullptr;
  g.nW = c.nW;
  if (c.graphCapture) g.gstop = c.gstate.p;
  for (int k = 0; k < 3; k++) g.L[k] = c.field.cell[k];
  return g;
}

static void gather_boundary(Ctx& c, const GatherArgs& g) {
  const int64_t nb = c.nGBody;
  if (nb == 0) return;
  const int parts = std::min(std::max(c.bndParts, 1), BND_PARTS_MAX);
  auto* kern = c.precision == 32 ? k_gather_bnd<false, true>
                                 : (c.field.periodic ? k_gather_bnd<true, false> : k_gather_bnd<false, false>);
  hipLaunchKernelGGL(kern, dim3((unsigned)(nb * parts)), dim3(256), 0, c.stream, g, (const uint32_t*)c.gbody.p, parts,
                     c.bndPart.p, (const double*)c.pos[0].p, (const double*)c.pos[1].p, (const double*)c.pos[2].p);
  WDG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_gather_bnd_sum, dim3(nblocks(nb, 64)), dim3(64), 0, c.stream, nb, (const uint32_t*)c.gbody.p,
                     parts, g.adjOff, g.gstop, (const double*)c.bndPart.p, c.force[0].p, c.force[1].p, c.force[2].p,
                     c.torque[0].p, c.torque[1].p, c.torque[2].p);
  WDG_LAUNCH_CHECK();
}

void intra_force(Ctx& c) {
  if (c.nFacet == 0) return;
  GatherArgs g = gather_args(c);
  if (c.con().n == 0) g.adjOff = nullptr;  // no slot list yet: every facet gets zeros
  auto* kern = c.precision == 32 ? k_facet_nodes<false, true>
                                 : (c.field.periodic ? k_facet_nodes<true, false> : k_facet_nodes<false, false>);
  hipLaunchKernelGGL(kern, dim3((unsigned)c.nW), dim3(FN_BS), 0, c.stream, g, (const double*)c.fgeo.p,
                     (const double*)c.pos[0].p, (const double*)c.pos[1].p, (const double*)c.pos[2].p, c.fnode.p);
  WDG_LAUNCH_CHECK();
}

void flush_forces(Ctx& c) {
  if (!c.forcesPending) return;
  c.forcesPending = false;
  c.forceClean = false;
  if (c.con().n == 0 || c.n == 0) return;
  GatherArgs g = gather_args(c);
  gather_boundary(c, g);
  dim3 grid(nblocks(c.n, 256)), block(256);
  auto* kern = c.precision == 32 ? k_gather<false, true> : (c.field.periodic ? k_gather<true, false> : k_gather<false, false>);
  hipLaunchKernelGGL(kern, grid, block, 0, c.stream, g, c.force[0].p, c.force[1].p, c.force[2].p, c.torque[0].p,
                     c.torque[1].p, c.torque[2].p, c.pos[0].p, c.pos[1].p, c.pos[2].p);
  WDG_LAUNCH_CHECK();
}

void leapfrog(Ctx& c, const wdg_leapfrog_params* p) {
  if (c.n == 0) return;
  LeapArgs a{};
  a.n = c.n;
  a.px = c.pos[0].p; a.py = c.pos[1].p; a.pz = c.pos[2].p;
  a.vx = c.vel[0].p; a.vy = c.vel[1].p; a.vz = c.vel[2].p;
  a.wx = c.angVel[0].p; a.wy = c.angVel[1].p; a.wz = c.angVel[2].p;
  a.qw = c.ori[0].p; a.qx = c.ori[1].p; a.qy = c.ori[2].p; a.qz = c.ori[3].p;
  a.fx = c.force[0].p; a.fy = c.force[1].p; a.fz = c.force[2].p;
  a.tx = c.torque[0].p; a.ty = c.torque[1].p; a.tz = c.torque[2].p;
  a.mass = c.mass.p; a.ix = c.inertia[0].p; a.iy = c.inertia[1].p; a.iz = c.inertia[2].p;
  a.lx = c.lastPos[0].p; a.ly = c.lastPos[1].p; a.lz = c.lastPos[2].p;
  a.flags = c.flags.p;
  a.dt = p->dt; a.damping = p->damping; a.reset = p->reset;
  for (int k = 0; k < 3; k++) a.g[k] = c.field.gravity[k];
  a.hasGravity = !(c.field.gravity[0] == 0 && c.field.gravity[1] == 0 && c.field.gravity[2] == 0);
  a.gather = (c.forcesPending && c.con().n > 0) ? 1 : 0;
  a.checkDirty = c.boundsValid && c.verletDist > 0 ? 1 : 0;
  a.maxD2 = c.verletDist * c.verletDist;
  a.vdS = c.vdSceneActive ? c.vdScene.p : nullptr;
  a.clean = c.forceClean ? 1 : 0;
  a.dtScene = c.nScenes > 0 ? c.dtScene.p : nullptr;
  a.scene = c.scene.p;
  if (c.graphCapture) a.gstop = c.gstate.p;
  // the flag word of this launch was cleared by the previous launch (or at creation)
  const int s = c.lfParity;
  a.dirty = c.dflags.p + DIRTY_PP + s;
  a.dirtyNext = c.dflags.p + DIRTY_PP + (1 - s);
  GatherArgs g = gather_args(c);
  if (a.gather) gather_boundary(c, g);
  dim3 grid(nblocks(c.n, 256)), block(256);
  // Leapfrog.cpp:132: a periodic scene is not tracked at step 0 (gradV is zero here)
  const bool track = c.trackEnergy && (!c.field.periodic || c.stepNow > 0);
  if (track && energy_per_scene(c)) {
    energy_scenes_prepare(c, c.con().n);
    a.enP = c.pe.p;
    auto* kern = c.precision == 32 ? k_leapfrog<false, true, true>
                                   : (c.field.periodic ? k_leapfrog<true, false, true> : k_leapfrog<false, false, true>);
    hipLaunchKernelGGL(kern, grid, block, 0, c.stream, a, g);
    WDG_LAUNCH_CHECK();
    energy_scenes_particles(c);
  } else if (track) {
    c.enPart.ensure((size_t)3 * grid.x);
    a.enPart = c.enPart.p;
    auto* kern = c.precision == 32 ? k_leapfrog<false, true, true>
                                   : (c.field.periodic ? k_leapfrog<true, false, true> : k_leapfrog<false, false, true>);
    hipLaunchKernelGGL(kern, grid, block, 0, c.stream, a, g);
    WDG_LAUNCH_CHECK();
    // kinetic is resettable: this step's sum replaces it (the same as the reset at the step's
    // start and one add); grav and nonviscDamp accumulate
    static const int tgt[3] = {WDG_EN_KINETIC, WDG_EN_GRAV, WDG_EN_NONVISCDAMP};
    energy_add(c, grid.x, 3, tgt, 1);
  } else {
    auto* kern = c.precision == 32 ? k_leapfrog<false, true, false>
                                   : (c.field.periodic ? k_leapfrog<true, false, false> : k_leapfrog<false, false, false>);
    hipLaunchKernelGGL(kern, grid, block, 0, c.stream, a, g);
    WDG_LAUNCH_CHECK();
  }
  c.lfParity = 1 - s;
  c.dirtyIdx = DIRTY_PP + s;
  c.forcesPending = false;
  // after a resetting step the integrated particles hold the reset value
  c.forceClean = p->reset != 0;
}

void force_reset(Ctx& c) {
  if (c.n == 0) return;
  if (c.forcesPending) flush_forces(c);
  hipLaunchKernelGGL(k_force_reset, dim3(nblocks(c.n, 256)), dim3(256), 0, c.stream, c.n, c.flags.p, c.mass.p,
                     c.field.gravity[0], c.field.gravity[1], c.field.gravity[2],
                     (int)!(c.field.gravity[0] == 0 && c.field.gravity[1] == 0 && c.field.gravity[2] == 0),
                     c.force[0].p, c.force[1].p, c.force[2].p, c.torque[0].p, c.torque[1].p, c.torque[2].p);
  WDG_LAUNCH_CHECK();
  c.forceClean = true;
}

// ------------------------------------------------------------------ DynDt
struct DynArgs {
  int64_t n;
  const uint32_t *A, *B, *info, *adj, *adjOff;
  RealView rv;
  const double *N0, *N1, *N2, *Y0, *Y1, *Y2, *uN;
  const double *rad, *mass, *ix, *iy, *iz;
  const uint32_t* flags;
  const int32_t* mat;
  const wdg_material* mats;
  const uint8_t* hE;  // each real contact's material epoch (null: one table)
  int nMats;
  const double* fgeo;
  int hertz;
  double poisson, en, viscAlpha;
  unsigned long long* out;
  const uint32_t* scene;  // batched scenes: the minimum per scene, out[scene] (else one, out[0])
};

template <int PREC>
__device__ __forceinline__ void dyn_one(const DynArgs& a, uint32_t e, d3& kt, d3& kr) {
  const uint32_t s = e & ~ADJ_PA;
  if (!rv_real(a.rv, s)) return;
  const uint32_t r = rv_index(a.rv, s);
  const uint32_t inf = a.info[s];
  const bool swap = (inf & CI_SWAP) != 0u;
  const uint32_t pA = swap ? a.B[s] : a.A[s], pB = swap ? a.A[s] : a.B[s];
  const int ix = (e & ADJ_PA) ? 0 : 1;
  double r1, r2;
  contact_radii(inf, pA, pB, a.rad, a.fgeo, r1, r2);
  double l0, l1, A;
  lens_area(r1, r2, l0, l1, A);
  const uint32_t eb = a.hE ? (uint32_t)a.hE[r] * (uint32_t)a.nMats : 0u;
  const wdg_material mA = a.mats[eb + a.mat[pA]], mB = a.mats[eb + a.mat[pB]];
  PhysVals ph;
  double kn, ktt;
  if (a.hertz) {
    phys_hertz(mA, mB, l0, l1, a.mass[pA], a.mass[pB], a.poisson, a.en, a.viscAlpha, ph);
    double kn0 = ph.K * sqrt(ph.R);
    double uN = ldr<PREC>(a.uN, r);
    kn = (3 / 2.) * kn0 * sqrt(-uN);
    ktt = ph.kt0 * sqrt(uN < 0 ? -uN : 0);
  } else {
    phys_frict(mA, mB, l0, l1, A, ph);
    kn = ph.kn;
    ktt = ph.kt;
  }
  d3 T0 = mk3(ldr<PREC>(a.N0, r), ldr<PREC>(a.N1, r), ldr<PREC>(a.N2, r)),
     T1 = mk3(ldr<PREC>(a.Y0, r), ldr<PREC>(a.Y1, r), ldr<PREC>(a.Y2, r));
  q4 q = quat_from_cols(T0, T1, cross3(T0, T1));
  d3 n = quat_rot(q, mk3(1., 0., 0.));
  d3 n2 = mk3(n.x * n.x, n.y * n.y, n.z * n.z);
  kt = mk3(kt.x + (n2.x * (kn - ktt) + ktt), kt.y + (n2.y * (kn - ktt) + ktt), kt.z + (n2.z * (kn - ktt) + ktt));
  double l = ix == 0 ? l0 : l1;
  double lk = (l * l) * ktt;
  kr = mk3(kr.x + lk * (n2.y + n2.z), kr.y + lk * (n2.z + n2.x), kr.z + lk * (n2.x + n2.y));
}

template <int PREC>
__global__ __launch_bounds__(256) void k_dyndt(DynArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double ret = INFINITY;
  if (i < a.n) {
    const uint32_t fl = a.flags[i];
    if ((fl & F_INTEGRATED) && (fl & F_DOF_ALL) != F_DOF_ALL) {
      d3 kt = mk3(0., 0., 0.), kr = mk3(0., 0., 0.);
      for (uint32_t k = a.adjOff[i]; k < a.adjOff[i + 1]; k++) dyn_one<PREC>(a, a.adj[k], kt, kr);
      const double m = a.mass[i];
      const double I[3] = {a.ix[i], a.iy[i], a.iz[i]};
      const double ktv[3] = {kt.x, kt.y, kt.z}, krv[3] = {kr.x, kr.y, kr.z};
      for (int k = 0; k < 3; k++)
        if (ktv[k] != 0 && m > 0. && !(fl & (1u << k))) ret = fmin(ret, m / fabs(ktv[k]));
      for (int k = 0; k < 3; k++)
        if (krv[k] != 0 && I[k] > 0. && !(fl & (8u << k))) ret = fmin(ret, I[k] / fabs(krv[k]));
      ret = 2 * ret;
    }
  }
  // non-negative doubles order like their bit patterns
  unsigned long long b = (unsigned long long)__double_as_longlong(ret);
  uint32_t sc = 0u;
  if (a.scene) {
    // scenes lie in runs of storage order: a wavefront of one scene reduces first, a mixed one
    // (or the tail past n) takes one atomic per particle
    sc = i < a.n ? a.scene[i] : 0xFFFFFFFFu;
    const uint32_t s0 = __shfl(sc, 0);
    if (!__all(sc == s0)) {
      if (i < a.n && ret != INFINITY) atomicMin(&a.out[sc], b);
      return;
    }
    if (s0 == 0xFFFFFFFFu) return;
  }
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(b, off);
    b = o < b ? o : b;
  }
  if ((threadIdx.x & 63) == 0) atomicMin(&a.out[sc], b);
}

static DynArgs dyn_args(Ctx& c) {
  Contacts& C = c.con();
  DynArgs a{};
  a.n = c.n;
  Reals& R = c.reals();
  a.A = C.A.p; a.B = C.B.p; a.info = C.info.p; a.rv = c.rview(); a.adj = C.adj.p; a.adjOff = C.adjOff.p;
  a.N0 = R.hN[0].p; a.N1 = R.hN[1].p; a.N2 = R.hN[2].p;
  a.Y0 = R.hY[0].p; a.Y1 = R.hY[1].p; a.Y2 = R.hY[2].p; a.uN = R.huN.p;
  a.rad = c.radius.p; a.mass = c.mass.p; a.ix = c.inertia[0].p; a.iy = c.inertia[1].p; a.iz = c.inertia[2].p;
  a.flags = c.flags.p; a.mat = c.mat.p; a.mats = c.dMats.p; a.fgeo = c.fgeo.p;
  a.nMats = (int)c.mats.size();
  if (c.matEpochs > 1) a.hE = R.hE.p;
  a.hertz = c.lastPhysKind == WDG_PHYS_HERTZ;
  a.poisson = c.lastPoisson; a.en = c.lastEn; a.viscAlpha = c.lastViscAlpha;
  return a;
}

double crit_dt(Ctx& c) {
  if (c.nScenes > 0) throw Error(WDG_E_STATE, "batched scenes: DynDt's timestep is per scene (wdg_crit_dt_scenes)");
  if (c.n == 0) return INFINITY;
  if (c.con().n == 0) return INFINITY;
  DynArgs a = dyn_args(c);
  unsigned long long* out = reinterpret_cast<unsigned long long*>(c.red.p);
  unsigned long long init = (unsigned long long)0x7FF0000000000000ull;  // +inf
  WDG_HIP(hipMemcpyAsync(out, &init, sizeof(init), hipMemcpyHostToDevice, c.stream));
  a.out = out;
  hipLaunchKernelGGL(c.precision == 32 ? k_dyndt<32> : k_dyndt<64>, dim3(nblocks(c.n, 256)), dim3(256), 0, c.stream,
                     a);
  WDG_LAUNCH_CHECK();
  unsigned long long hb;
  WDG_HIP(hipMemcpyAsync(&hb, out, sizeof(hb), hipMemcpyDeviceToHost, c.stream));
  WDG_HIP(hipStreamSynchronize(c.stream));
  double r;
  std::memcpy(&r, &hb, 8);
  return std::sqrt(r);
}

// DynDt::critDt_stiffness of each scene of a batched context: the same per-particle values,
// their minimum taken per scene (a minimum does not depend on the order)
void crit_dt_scenes(Ctx& c, std::vector<double>& out) {
  if (c.nScenes <= 0) throw Error(WDG_E_STATE, "not a batched context (wdg_set_scenes)");
  out.assign((size_t)c.nScenes, INFINITY);
  if (c.n == 0 || c.con().n == 0) return;
  DynArgs a = dyn_args(c);
  DBuf<unsigned long long> d;
  d.ensure((size_t)c.nScenes);
  std::vector<unsigned long long> h((size_t)c.nScenes, 0x7FF0000000000000ull);  // +inf
  WDG_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * 8, hipMemcpyHostToDevice, c.stream));
  a.out = d.p;
  a.scene = c.scene.p;
  hipLaunchKernelGGL(c.precision == 32 ? k_dyndt<32> : k_dyndt<64>, dim3(nblocks(c.n, 256)), dim3(256), 0, c.stream,
                     a);
  WDG_LAUNCH_CHECK();
  WDG_HIP(hipMemcpyAsync(h.data(), d.p, h.size() * 8, hipMemcpyDeviceToHost, c.stream));
  WDG_HIP(hipStreamSynchronize(c.stream));
  d.release();
  for (size_t s = 0; s < h.size(); s++) {
    double r;
    std::memcpy(&r, &h[s], 8);
    out[s] = std::sqrt(r);
  }
}

// ------------------------------------------------------------------ observables
__global__ void k_count_contacts(int64_t n, RealView rv, const uint32_t* __restrict__ adj,
                                 const uint32_t* __restrict__ adjOff, uint32_t* cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t k = 0;
  for (uint32_t j = adjOff[i]; j < adjOff[i + 1]; j++) k += rv_real(rv, adj[j] & ~ADJ_PA) ? 1u : 0u;
  cnt[i] = k;
}

// block partial sums in fixed order (deterministic): [C2, N0, N1, N, Vs]
__global__ __launch_bounds__(256) void k_obs_partial(int64_t n, const uint32_t* __restrict__ cnt,
                                                     const uint32_t* __restrict__ flags, const double* __restrict__ rad,
                                                     const double* __restrict__ px, const double* __restrict__ py,
                                                     const double* __restrict__ pz, double lo0, double lo1, double lo2,
                                                     double hi0, double hi1, double hi2, double* part) {
  __shared__ double s[5][256];
  const int tid = threadIdx.x;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + tid;
  double v[5] = {0, 0, 0, 0, 0};
  if (i < n) {
    uint32_t k = cnt[i];
    v[0] = k;
    v[1] = k == 0;
    v[2] = k == 1;
    v[3] = 1;
    if (flag_shape(flags[i]) == WDG_SPHERE) {
      double x = px[i], y = py[i], z = pz[i];
      bool in = x >= lo0 && x <= hi0 && y >= lo1 && y <= hi1 && z >= lo2 && z <= hi2;
      double r = rad[i];
      if (in) v[4] = (4 / 3.) * M_PI * (r * r * r);
    }
  }
  for (int q = 0; q < 5; q++) s[q][tid] = v[q];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st)
      for (int q = 0; q < 5; q++) s[q][tid] += s[q][tid + st];
    __syncthreads();
  }
  if (tid == 0)
    for (int q = 0; q < 5; q++) part[(int64_t)q * gridDim.x + blockIdx.x] = s[q][0];
}

void observables(Ctx& c, const double lo[3], const double hi[3], double out[3]) {
  if (c.forcesPending) flush_forces(c);
  Contacts& C = c.con();
  c.scount.ensure(c.n + 1);
  if (C.n > 0 && c.n > 0) {
    hipLaunchKernelGGL(k_count_contacts, dim3(nblocks(c.n, 256)), dim3(256), 0, c.stream, c.n, c.rview(), C.adj.p,
                       C.adjOff.p, c.scount.p);
    WDG_LAUNCH_CHECK();
  } else {
    WDG_HIP(hipMemsetAsync(c.scount.p, 0, (c.n + 1) * sizeof(uint32_t), c.stream));
  }
  unsigned nb = nblocks(c.n, 256);
  std::vector<double> part((size_t)5 * nb);
  DBuf<double> dpart;
  dpart.ensure(part.size());
  hipLaunchKernelGGL(k_obs_partial, dim3(nb), dim3(256), 0, c.stream, c.n, c.scount.p, c.flags.p, c.radius.p,
                     c.pos[0].p, c.pos[1].p, c.pos[2].p, lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], dpart.p);
  WDG_LAUNCH_CHECK();
  WDG_HIP(hipMemcpyAsync(part.data(), dpart.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  WDG_HIP(hipStreamSynchronize(c.stream));
  dpart.release();
  double S[5] = {0, 0, 0, 0, 0};
  for (int q = 0; q < 5; q++)
    for (unsigned b = 0; b < nb; b++) S[q] += part[(size_t)q * nb + b];
  const double C2 = S[0], N0 = S[1], N1 = S[2], N = S[3], Vs = S[4];
  out[0] = C2 / N;
  out[1] = (C2 - N1) / (N - N0 - N1);
  double V = ((hi[0] - lo[0]) * (hi[1] - lo[1])) * (hi[2] - lo[2]);
  out[2] = Vs / V;
}

// ------------------------------------------------------------------ stress, kinetic energy
// DemFuncs::stressStiffness, stress part (Funcs.cpp:183-229): per real contact
// d0 = |posA - shift - posB|, n = ori*e_x, F = ori*force (the gather record's global force),
// fN = F.n, fT = F - n fN, sigma_ij += d0 (fN n_i n_j + (fT_i n_j + fT_j n_i)/2).
// Block partial sums in slot order, added on the host in block order.
struct StressArgs {
  int64_t nc;
  const uint32_t *A, *B, *info;
  const uint64_t* dist;
  RealView rv;
  const double *N0, *N1, *N2, *Y0, *Y1, *Y2, *gp;
  const double *px, *py, *pz;
  double L[3];
  int periodic;
  double* part;
};
template <int PREC>
__global__ __launch_bounds__(256) void k_stress(StressArgs a) {
  __shared__ double sh[9][256];
  const int tid = threadIdx.x;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + tid;
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (c < a.nc && rv_real(a.rv, (uint32_t)c)) {
    const uint32_t r = rv_index(a.rv, (uint32_t)c);
    const uint32_t inf = a.info[c];
    const bool swap = (inf & CI_SWAP) != 0u;
    const uint32_t pA = swap ? a.B[c] : a.A[c], pB = swap ? a.A[c] : a.B[c];
    d3 posA = mk3(a.px[pA], a.py[pA], a.pz[pA]);
    const d3 posB = mk3(a.px[pB], a.py[pB], a.pz[pB]);
    if (a.periodic) {
      const uint64_t cd = a.dist[c];
      posA = sub3(posA, mk3(a.L[0] * unpack_cd(cd, 0), a.L[1] * unpack_cd(cd, 1), a.L[2] * unpack_cd(cd, 2)));
    }
    // a multinodal pA (facet) is represented by the contact point (skipMultinodal=false)
    if ((inf & CI_TYPE_MASK) == CT_FS) {
      if (PREC == 32) {
        const float* g = reinterpret_cast<const float*>(a.gp) + (int64_t)GP32 * r;
        posA = add3(posA, mk3(g[3], g[4], g[5]));
      } else {
        const double* g = a.gp + (int64_t)GP * r;
        posA = mk3(g[3], g[4], g[5]);
      }
    }
    const double d0 = sqrt(sqn3(sub3(posA, posB)));
    const d3 T0 = mk3(ldr<PREC>(a.N0, r), ldr<PREC>(a.N1, r), ldr<PREC>(a.N2, r)),
             T1 = mk3(ldr<PREC>(a.Y0, r), ldr<PREC>(a.Y1, r), ldr<PREC>(a.Y2, r));
    const q4 q = quat_from_cols(T0, T1, cross3(T0, T1));
    const d3 n = quat_rot(q, mk3(1., 0., 0.));
    d3 F;
    if (PREC == 32) {
      const float* g = reinterpret_cast<const float*>(a.gp) + (int64_t)GP32 * r;
      F = mk3(g[0], g[1], g[2]);
    } else {
      const double* g = a.gp + (int64_t)GP * r;
      F = mk3(g[0], g[1], g[2]);
    }
    const double fN = dot3(F, n);
    const d3 fT = sub3(F, mul3(n, fN));
    const double nn[3] = {n.x, n.y, n.z}, tt[3] = {fT.x, fT.y, fT.z};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) v[3 * i + j] = d0 * (fN * nn[i] * nn[j] + .5 * (tt[i] * nn[j] + tt[j] * nn[i]));
  }
  for (int q = 0; q < 9; q++) sh[q][tid] = v[q];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st)
      for (int q = 0; q < 9; q++) sh[q][tid] += sh[q][tid + st];
    __syncthreads();
  }
  if (tid == 0)
    for (int q = 0; q < 9; q++) a.part[(int64_t)blockIdx.x * 9 + q] = sh[q][0];
}

void stress(Ctx& c, double volume, double out[9]) {
  for (int k = 0; k < 9; k++) out[k] = 0;
  if (volume <= 0) {
    if (!c.field.periodic) throw Error(WDG_E_ARG, "a positive volume must be given for aperiodic scenes");
    volume = c.field.cell[0] * c.field.cell[1] * c.field.cell[2];
  }
  Contacts& C = c.con();
  if (C.n == 0) return;
  Reals& R = c.reals();
  const unsigned nb = nblocks(C.n, 256);
  DBuf<double> part;
  part.ensure((size_t)nb * 9);
  StressArgs a{};
  a.nc = C.n; a.A = C.A.p; a.B = C.B.p; a.info = C.info.p; a.dist = C.dist.p; a.rv = c.rview();
  a.N0 = R.hN[0].p; a.N1 = R.hN[1].p; a.N2 = R.hN[2].p; a.Y0 = R.hY[0].p; a.Y1 = R.hY[1].p; a.Y2 = R.hY[2].p;
  a.gp = R.gp.p;
  a.px = c.pos[0].p; a.py = c.pos[1].p; a.pz = c.pos[2].p;
  for (int k = 0; k < 3; k++) a.L[k] = c.field.cell[k];
  a.periodic = c.field.periodic;
  a.part = part.p;
  hipLaunchKernelGGL(c.precision == 32 ? k_stress<32> : k_stress<64>, dim3(nb), dim3(256), 0, c.stream, a);
  WDG_LAUNCH_CHECK();
  std::vector<double> h((size_t)nb * 9);
  WDG_HIP(hipMemcpyAsync(h.data(), part.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  WDG_HIP(hipStreamSynchronize(c.stream));
  part.release();
  for (unsigned b = 0; b < nb; b++)
    for (int k = 0; k < 9; k++) out[k] += h[(size_t)b * 9 + k];
  for (int k = 0; k < 9; k++) out[k] /= volume;
}

// sum over integrated nodes of m v.v/2 + w.(I w)/2
__global__ __launch_bounds__(256) void k_ek(int64_t n, const uint32_t* __restrict__ flags, const double* __restrict__ m,
                                          const double* __restrict__ vx, const double* __restrict__ vy,
                                          const double* __restrict__ vz, const double* __restrict__ wx,
                                          const double* __restrict__ wy, const double* __restrict__ wz,
                                          const double* __restrict__ ix, const double* __restrict__ iy,
                                          const double* __restrict__ iz, double* part) {
  __shared__ double sh[256];
  const int tid = threadIdx.x;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + tid;
  double e = 0;
  if (i < n && (flags[i] & F_INTEGRATED)) {
    const d3 v = mk3(vx[i], vy[i], vz[i]), w = mk3(wx[i], wy[i], wz[i]);
    const double et = .5 * m[i] * sqn3(v);
    const double er = 0.5 * dot3(w, mk3(ix[i] * w.x, iy[i] * w.y, iz[i] * w.z));
    e = et + er;
  }
  sh[tid] = e;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) sh[tid] += sh[tid + st];
    __syncthreads();
  }
  if (tid == 0) part[blockIdx.x] = sh[0];
}

double kinetic_energy(Ctx& c) {
  if (c.n == 0) return 0;
  const unsigned nb = nblocks(c.n, 256);
  DBuf<double> part;
  part.ensure(nb);
  hipLaunchKernelGGL(k_ek, dim3(nb), dim3(256), 0, c.stream, c.n, c.flags.p, c.mass.p, c.vel[0].p, c.vel[1].p,
                     c.vel[2].p, c.angVel[0].p, c.angVel[1].p, c.angVel[2].p, c.inertia[0].p, c.inertia[1].p,
                     c.inertia[2].p, part.p);
  WDG_LAUNCH_CHECK();
  std::vector<double> h(nb);
  WDG_HIP(hipMemcpyAsync(h.data(), part.p, nb * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  WDG_HIP(hipStreamSynchronize(c.stream));
  part.release();
  double s = 0;
  for (double x : h) s += x;
  return s;
}

// ------------------------------------------------------------------ kinetic energy, no allocation
// For callers that cannot allocate device memory (code under HIP graph capture, serving loops):
// k_ek's per-workgroup partials go to a file-scope device array instead of a fresh buffer, and
// one thread adds them in block order, left to right, as kinetic_energy does on the host, so
// both give the same bits. Every GPU holds its own copy of the array until the process ends;
// its address on each device is looked up once, with that device current, at context creation.
// All calls on one GPU share the array: calls that overlap on one GPU race. The partials are
// written in full before they are read, so nothing needs clearing.
constexpr int64_t EK_SCRATCH_MAX = int64_t(1) << 18;  // workgroup partials: up to 2^26 particles
constexpr int SCRATCH_DEVICES = 64;
static __device__ __align__(256) double scratch[EK_SCRATCH_MAX];
static double* g_scratch[SCRATCH_DEVICES];
static std::mutex g_scratch_mu;

void scratch_setup(int device) {
  if (device < 0 || device >= SCRATCH_DEVICES) throw Error(WDG_E_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return;
  void* p = nullptr;
  WDG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch[device] = static_cast<double*>(p);
}

__global__ void k_sum_in_order(const double* __restrict__ part, int64_t nb, double* out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double s = 0;
  for (int64_t b = 0; b < nb; b++) s += part[b];
  *out = s;
}

void kinetic_energy_noalloc(Ctx& c, double* dOut) {
  const int64_t nb = nblocks(c.n, 256);
  if (nb > EK_SCRATCH_MAX)
    throw Error(WDG_E_ARG, "wdg_kinetic_energy_noalloc: " + std::to_string(c.n) + " particles exceed its maximum of " +
                               std::to_string(EK_SCRATCH_MAX * 256));
  double* part = (c.device >= 0 && c.device < SCRATCH_DEVICES) ? g_scratch[c.device] : nullptr;
  if (!part) throw Error(WDG_E_STATE, "the scratch array of this device was not set up");
  if (nb > 0) {
    hipLaunchKernelGGL(k_ek, dim3((unsigned)nb), dim3(256), 0, c.stream, c.n, c.flags.p, c.mass.p, c.vel[0].p,
                       c.vel[1].p, c.vel[2].p, c.angVel[0].p, c.angVel[1].p, c.angVel[2].p, c.inertia[0].p,
                       c.inertia[1].p, c.inertia[2].p, part);
    WDG_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_sum_in_order, dim3(1), dim3(64), 0, c.stream, part, nb, dOut);
  WDG_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ in-memory snapshots
// Scene::saveTmp / loadTmp (src/core/Scene.cpp:291-294) for the kinematic state: positions,
// velocities, angular velocities, orientations, forces and torques, kept on the device by user
// id (so a snapshot survives the re-sorts of the broadphase). Loading one restarts the contact
// history, as loading a saved scene without its contacts would.
constexpr int SNAP_FIELDS = 19;
struct SnapArgs {
  double* f[SNAP_FIELDS];
  double* snap;
  const uint32_t* uid;
  int64_t n;
};
static SnapArgs snap_args(Ctx& c, int slot) {
  SnapArgs a{};
  int q = 0;
  for (int k = 0; k < 3; k++) a.f[q++] = c.pos[k].p;
  for (int k = 0; k < 3; k++) a.f[q++] = c.vel[k].p;
  for (int k = 0; k < 3; k++) a.f[q++] = c.angVel[k].p;
  for (int k = 0; k < 4; k++) a.f[q++] = c.ori[k].p;
  for (int k = 0; k < 3; k++) a.f[q++] = c.force[k].p;
  for (int k = 0; k < 3; k++) a.f[q++] = c.torque[k].p;
  a.snap = c.snap[slot].p;
  a.uid = c.uid.p;
  a.n = c.n;
  return a;
}
__global__ void k_snap_save(SnapArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int64_t u = a.uid[i];
#pragma unroll
  for (int q = 0; q < SNAP_FIELDS; q++) a.snap[q * a.n + u] = a.f[q][i];
}
__global__ void k_snap_load(SnapArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int64_t u = a.uid[i];
#pragma unroll
  for (int q = 0; q < SNAP_FIELDS; q++) a.f[q][i] = a.snap[q * a.n + u];
}
// largest squared distance between the particles [i0, n) and their snapshot positions
__global__ __launch_bounds__(256) void k_snap_dist(SnapArgs a, int64_t i0, unsigned long long* out) {
  __shared__ double wm[4];
  double m = 0.;
  for (int64_t i = i0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t u = a.uid[i];
    const double dx = a.f[0][i] - a.snap[u], dy = a.f[1][i] - a.snap[a.n + u], dz = a.f[2][i] - a.snap[2 * a.n + u];
    m = fmax(m, (dx * dx + dy * dy) + dz * dz);
  }
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0)
    atomicMax(out, (unsigned long long)__double_as_longlong(fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3]))));
}

static void snap_check(Ctx& c, int slot, bool mustExist) {
  if (slot < 0 || slot >= Ctx::NSNAP) throw Error(WDG_E_ARG, "snapshot slot must be 0.." + std::to_string(Ctx::NSNAP - 1));
  if (mustExist && (c.snapN[slot] != c.n || c.snapUpload[slot] != c.uploads))
    throw Error(WDG_E_STATE, "no snapshot of the loaded particles in slot " + std::to_string(slot));
}

void snapshot_save(Ctx& c, int slot) {
  snap_check(c, slot, false);
  if (c.forcesPending) flush_forces(c);
  c.snap[slot].ensure((size_t)std::max<int64_t>(c.n, 1) * SNAP_FIELDS);
  if (c.n > 0) {
    hipLaunchKernelGGL(k_snap_save, dim3(nblocks(c.n, 256)), dim3(256), 0, c.stream, snap_args(c, slot));
    WDG_LAUNCH_CHECK();
  }
  c.snapN[slot] = c.n;
  c.snapUpload[slot] = c.uploads;
  snapshot_contacts_save(c, slot);
}

void snapshot_load(Ctx& c, int slot) {
  snap_check(c, slot, true);
  if (c.forcesPending) flush_forces(c);
  if (c.n > 0) {
    hipLaunchKernelGGL(k_snap_load, dim3(nblocks(c.n, 256)), dim3(256), 0, c.stream, snap_args(c, slot));
    WDG_LAUNCH_CHECK();
  }
  drop_contacts(c);
  c.forceClean = false;
  c.dirtyKnown = false;
  // without saved contacts the next collider call resolves its Verlet distance again, as after
  // an upload; with them the slot list is rebuilt now at the snapshot's Verlet distance
  c.verletDist = NAN;
  snapshot_contacts_load(c, slot);
}

double snapshot_distance(Ctx& c, int slot) {
  snap_check(c, slot, true);
  const int64_t m = c.n - c.nW;
  if (m <= 0) return 0.;
  unsigned long long* d = reinterpret_cast<unsigned long long*>(c.red.p);
  WDG_HIP(hipMemsetAsync(d, 0, 8, c.stream));
  hipLaunchKernelGGL(k_snap_dist, dim3((unsigned)std::min<int64_t>(nblocks(m, 256), 256)), dim3(256), 0, c.stream,
                     snap_args(c, slot), c.nW, d);
  WDG_LAUNCH_CHECK();
  unsigned long long hb = 0;
  WDG_HIP(hipMemcpyAsync(&hb, d, 8, hipMemcpyDeviceToHost, c.stream));
  WDG_HIP(hipStreamSynchronize(c.stream));
  double d2;
  memcpy(&d2, &hb, 8);
  return sqrt(d2);
}

int64_t count_real(Ctx& c) {
  Contacts& C = c.con();
  if (C.n == 0) return 0;
  std::vector<uint32_t> g((size_t)ngroups(C.n));
  WDG_HIP(hipMemcpyAsync(g.data(), c.reals().gcnt.p, g.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
  WDG_HIP(hipStreamSynchronize(c.stream));
  int64_t s = 0;
  for (uint32_t x : g) s += x;
  return s;
}

// ------------------------------------------------------------------ energy tracking (EnergyTracker)
// one workgroup: thread t sums the partials of workgroups t, t+256, ... in order, then the
// threads' sums are added in a fixed tree; run to run the result is the same
// (bit q of setMask: the sum replaces energy[target q] instead of being added to it)
__global__ __launch_bounds__(256) void k_energy_add(const double* __restrict__ part, int64_t nBlocks, int nTerms,
                                                    int4 target, int setMask, double* energy, const uint32_t* gstop) {
  // a stopped graph replay: the step did not happen
  if (gstop && *gstop) return;
  __shared__ double sh[4][256];
  const int t = threadIdx.x;
  double acc[4] = {0., 0., 0., 0.};
  // all terms of a partial at once (their loads are independent); per term the order is
  // workgroups t, t+256, ...
#pragma unroll 4
  for (int64_t b = t; b < nBlocks; b += 256) {
    const double* p = part + b * nTerms;
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (q < nTerms) acc[q] += p[q];
  }
#pragma unroll
  for (int q = 0; q < 4; q++) sh[q][t] = acc[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
#pragma unroll
      for (int q = 0; q < 4; q++) sh[q][t] += sh[q][t + w];
    __syncthreads();
  }
  if (t < nTerms) {
    const int e = t == 0 ? target.x : t == 1 ? target.y : t == 2 ? target.z : target.w;
    energy[e] = ((setMask >> t) & 1) ? sh[t][0] : energy[e] + sh[t][0];
  }
}

// Scene.cpp:416 (EnergyTracker::resetResettables): kinetic and elast start every step at zero
__global__ void k_energy_reset(double* energy, int clearAll) {
  const int t = threadIdx.x;
  if (t < WDG_EN_N && (clearAll || t == WDG_EN_KINETIC || t == WDG_EN_ELAST)) energy[t] = 0.;
}

// the same for every scene's terms (energy[WDG_EN_N * s + term])
__global__ void k_energy_reset_scenes(double* energy, int64_t ns, int clearAll) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns * WDG_EN_N) return;
  const int t = (int)(i % WDG_EN_N);
  if (clearAll || t == WDG_EN_KINETIC || t == WDG_EN_ELAST) energy[i] = 0.;
}

static void energy_scenes_reset(Ctx& c, int clearAll) {
  const int64_t m = (int64_t)c.nScenes * WDG_EN_N;
  hipLaunchKernelGGL(k_energy_reset_scenes, dim3(nblocks(m, 256)), dim3(256), 0, c.stream, c.energyS.p,
                     (int64_t)c.nScenes, clearAll);
  WDG_LAUNCH_CHECK();
}

void energy_enable(Ctx& c, bool on) {
  if (on && !c.energy.p) {
    c.energy.ensure(WDG_EN_N);
    energy_clear(c);
  }
  c.trackEnergy = on;
}

void energy_clear(Ctx& c) {
  if (!c.energy.p) return;
  hipLaunchKernelGGL(k_energy_reset, dim3(1), dim3(64), 0, c.stream, c.energy.p, 1);
  WDG_LAUNCH_CHECK();
  if (c.nScenes > 1) {
    c.energyS.ensure((size_t)c.nScenes * WDG_EN_N);
    energy_scenes_reset(c, 1);
    c.energySScenes = c.nScenes;
  }
}

// one term to zero, skipped by a stopped graph replay (a step of a contact loop without slots:
// its elast is the empty sum)
__global__ void k_energy_zero(double* energy, int term, const uint32_t* gstop) {
  if (gstop && *gstop) return;
  if (threadIdx.x == 0) energy[term] = 0.;
}

// the term of every scene
__global__ void k_energy_zero_scenes(double* energy, int64_t ns, int term, const uint32_t* gstop) {
  if (gstop && *gstop) return;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < ns) energy[WDG_EN_N * s + term] = 0.;
}

void energy_zero(Ctx& c, int term) {
  if (energy_per_scene(c)) {
    energy_scenes_prepare(c, c.con().n);
    hipLaunchKernelGGL(k_energy_zero_scenes, dim3(nblocks(c.nScenes, 256)), dim3(256), 0, c.stream, c.energyS.p,
                       (int64_t)c.nScenes, term, c.graphCapture ? (const uint32_t*)c.gstate.p : nullptr);
    WDG_LAUNCH_CHECK();
    return;
  }
  hipLaunchKernelGGL(k_energy_zero, dim3(1), dim3(64), 0, c.stream, c.energy.p, term,
                     c.graphCapture ? (const uint32_t*)c.gstate.p : nullptr);
  WDG_LAUNCH_CHECK();
}

void energy_step_begin(Ctx& c) {
  if (energy_per_scene(c)) {
    energy_scenes_prepare(c, c.con().n);
    energy_scenes_reset(c, 0);
    return;
  }
  hipLaunchKernelGGL(k_energy_reset, dim3(1), dim3(64), 0, c.stream, c.energy.p, 0);
  WDG_LAUNCH_CHECK();
}

void energy_add(Ctx& c, int64_t nBlocks, int nTerms, const int* target, int setMask) {
  if (nBlocks <= 0) return;
  const int4 t = make_int4(target[0], nTerms > 1 ? target[1] : 0, nTerms > 2 ? target[2] : 0, nTerms > 3 ? target[3] : 0);
  hipLaunchKernelGGL(k_energy_add, dim3(1), dim3(256), 0, c.stream, c.enPart.p, nBlocks, nTerms, t, setMask,
                     c.energy.p, c.graphCapture ? (const uint32_t*)c.gstate.p : nullptr);
  WDG_LAUNCH_CHECK();
}

void energy_read(Ctx& c, double out[7]) {
  if (c.nScenes > 1 && c.energyS.p) {
    // a batched context: the sum over its scenes, in scene order
    std::vector<double> s;
    energy_scenes_read(c, s);
    for (int q = 0; q < WDG_EN_N; q++) out[q] = 0.;
    for (int64_t k = 0; k < c.nScenes; k++)
      for (int q = 0; q < WDG_EN_N; q++) out[q] += s[(size_t)k * WDG_EN_N + q];
    return;
  }
  if (!c.energy.p) {
    for (int q = 0; q < WDG_EN_N; q++) out[q] = 0.;
    return;
  }
  WDG_HIP(hipMemcpyAsync(out, c.energy.p, WDG_EN_N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  WDG_HIP(hipStreamSynchronize(c.stream));
}

// ------------------------------------------------------------------ energy of batched scenes
// EnergyTracker per scene (Scene.energy of each scene of the batch). The leapfrog writes its
// three terms per particle (pe), the contact loop its four per contact (enC); a contact's terms
// are counted at its pB, a sphere (boundary bodies are always pA), by walking the sphere's
// adjacency (pc); then the workgroups of a scene sum equal parts of its particles' terms in
// storage order with a fixed tree, and the parts are added in order. Deterministic; the order of the additions differs from that of the scene
// stepped alone, so the sums agree to rounding.

constexpr int ENS_PARTS_MAX = 64;

__global__ void k_scene_keys(int64_t n, const uint32_t* __restrict__ scene, uint32_t* key, uint32_t* idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    key[i] = scene[i];
    idx[i] = (uint32_t)i;
  }
}

// off[s] = the first position of the sorted keys at or above s (off[ns] = n)
__global__ void k_scene_off(int64_t n, int64_t ns, const uint32_t* __restrict__ key, uint32_t* off) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s > ns) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)key[mid] < s) lo = mid + 1;
    else hi = mid;
  }
  off[s] = (uint32_t)lo;
}

void energy_scenes_prepare(Ctx& c, int64_t nc) {
  const int64_t n = c.n, ns = c.nScenes;
  c.energyS.ensure((size_t)ns * WDG_EN_N);
  if (c.energySScenes != ns) {
    // another set of scenes (wdg_set_scenes): their terms start at zero
    if (c.graphCapture) throw Error(WDG_E_STATE, "the scenes' energy terms are not set up while capturing");
    energy_scenes_reset(c, 1);
    c.energySScenes = ns;
  }
  c.pe.ensure((size_t)3 * std::max<int64_t>(n, 1));
  c.pc.ensure((size_t)4 * std::max<int64_t>(n, 1));
  c.enC.ensure((size_t)4 * std::max<int64_t>(ngroups(nc) * GROUP, 1));
  c.sceneOff.ensure((size_t)ns + 1);
  c.enSPart.ensure((size_t)4 * ns * ENS_PARTS_MAX);
  if (c.sceneListGen == c.gen) return;
  if (c.graphCapture) throw Error(WDG_E_STATE, "the scenes' particle lists are stale while capturing");
  c.sceneKey.ensure((size_t)std::max<int64_t>(n, 1));
  c.sceneKeyTmp.ensure((size_t)std::max<int64_t>(n, 1));
  c.sceneList.ensure((size_t)std::max<int64_t>(n, 1));
  c.sceneListTmp.ensure((size_t)std::max<int64_t>(n, 1));
  if (n > 0) {
    hipLaunchKernelGGL(k_scene_keys, dim3(nblocks(n, 256)), dim3(256), 0, c.stream, n, (const uint32_t*)c.scene.p,
                       c.sceneKey.p, c.sceneList.p);
    WDG_LAUNCH_CHECK();
  }
  int bits = 1;
  while (bits < 32 && ((int64_t)1 << bits) < ns) bits++;
  // stable: within a scene the particles stay in storage order
  radix_sort_pairs<uint32_t>(c, c.sceneKey.p, c.sceneList.p, c.sceneKeyTmp.p, c.sceneListTmp.p, n, bits);
  hipLaunchKernelGGL(k_scene_off, dim3(nblocks(ns + 1, 256)), dim3(256), 0, c.stream, n, ns,
                     (const uint32_t*)c.sceneKey.p, c.sceneOff.p);
  WDG_LAUNCH_CHECK();
  c.sceneListGen = c.gen;
}

// pc[4 i + q] = the terms of the real contacts whose pB is particle i, in adjacency order (zero
// for the boundary bodies, storage [0, nW): they are never pB, and their long lists are not walked)
__global__ __launch_bounds__(256) void k_energy_at_pb(int64_t n, int64_t nW, const uint32_t* __restrict__ adj,
                                                      const uint32_t* __restrict__ adjOff, RealView rv,
                                                      const double* __restrict__ enC, double* pc,
                                                      const uint32_t* gstop) {
  if (gstop && *gstop) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double e[4] = {0., 0., 0., 0.};
  const uint32_t k1 = i < nW ? adjOff[i] : adjOff[i + 1];
  for (uint32_t k = adjOff[i]; k < k1; k++) {
    const uint32_t v = adj[k];
    if (v & ADJ_PA) continue;
    const uint32_t slot = v & ~ADJ_PA;
    if (!rv_real(rv, slot)) continue;
    const double* t = enC + 4 * (int64_t)rv_index(rv, slot);
#pragma unroll
    for (int q = 0; q < 4; q++) e[q] += t[q];
  }
  double* o = pc + 4 * i;
#pragma unroll
  for (int q = 0; q < 4; q++) o[q] = e[q];
}

// `parts` workgroups per scene, each over an equal contiguous part of the scene's list: nTerms
// terms of stride `stride` per particle, thread t summing the part's particles t, t+256, ... in
// list order, then a fixed tree, into sp[4 (s parts + part) + q]
__global__ __launch_bounds__(256) void k_energy_scenes(const double* __restrict__ part, int stride, int nTerms,
                                                       int parts, const uint32_t* __restrict__ list,
                                                       const uint32_t* __restrict__ off, double* sp,
                                                       const uint32_t* gstop) {
  if (gstop && *gstop) return;
  __shared__ double sh[4][256];
  const int64_t s = blockIdx.x / parts;
  const int pi = (int)(blockIdx.x % parts);
  const int t = threadIdx.x;
  const uint32_t o0 = off[s], o1 = off[s + 1];
  const uint32_t chunk = (o1 - o0 + parts - 1) / parts;
  const uint32_t a0 = o0 + (uint32_t)pi * chunk < o1 ? o0 + (uint32_t)pi * chunk : o1;
  const uint32_t a1 = a0 + chunk < o1 ? a0 + chunk : o1;
  double acc[4] = {0., 0., 0., 0.};
  for (uint32_t k = a0 + t; k < a1; k += 256) {
    const double* p = part + (int64_t)stride * list[k];
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (q < nTerms) acc[q] += p[q];
  }
#pragma unroll
  for (int q = 0; q < 4; q++) sh[q][t] = acc[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
#pragma unroll
      for (int q = 0; q < 4; q++) sh[q][t] += sh[q][t + w];
    __syncthreads();
  }
  if (t < nTerms) sp[4 * blockIdx.x + t] = sh[t][0];
}

// per scene and term: the parts' sums in order; bit q of setMask: the sum replaces
// energy[WDG_EN_N s + target q] instead of being added to it
__global__ void k_energy_scenes_sum(const double* __restrict__ sp, int64_t ns, int parts, int nTerms, int4 target,
                                    int setMask, double* energy, const uint32_t* gstop) {
  if (gstop && *gstop) return;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  for (int q = 0; q < nTerms; q++) {
    double v = 0.;
    for (int p = 0; p < parts; p++) v += sp[4 * (s * parts + p) + q];
    const int e = q == 0 ? target.x : q == 1 ? target.y : q == 2 ? target.z : target.w;
    double* d = energy + (int64_t)WDG_EN_N * s + e;
    *d = ((setMask >> q) & 1) ? v : *d + v;
  }
}

static void energy_scenes_sum(Ctx& c, const double* part, int stride, int nTerms, int4 tgt) {
  const uint32_t* gstop = c.graphCapture ? (const uint32_t*)c.gstate.p : nullptr;
  const int64_t ns = c.nScenes;
  // about 2048 workgroups over the scenes, so that a batch of few scenes still fills the chip
  const int parts = (int)std::min<int64_t>(ENS_PARTS_MAX, std::max<int64_t>(1, (2048 + ns - 1) / ns));
  hipLaunchKernelGGL(k_energy_scenes, dim3((unsigned)(ns * parts)), dim3(256), 0, c.stream, part, stride, nTerms,
                     parts, (const uint32_t*)c.sceneList.p, (const uint32_t*)c.sceneOff.p, c.enSPart.p, gstop);
  WDG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_energy_scenes_sum, dim3(nblocks(ns, 256)), dim3(256), 0, c.stream, (const double*)c.enSPart.p,
                     ns, parts, nTerms, tgt, 1, c.energyS.p, gstop);
  WDG_LAUNCH_CHECK();
}

void energy_scenes_contacts(Ctx& c) {
  Contacts& C = c.con();
  const uint32_t* gstop = c.graphCapture ? (const uint32_t*)c.gstate.p : nullptr;
  hipLaunchKernelGGL(k_energy_at_pb, dim3(nblocks(c.n, 256)), dim3(256), 0, c.stream, c.n, c.nW, (const uint32_t*)C.adj.p,
                     (const uint32_t*)C.adjOff.p, c.rview(), (const double*)c.enC.p, c.pc.p, gstop);
  WDG_LAUNCH_CHECK();
  // EnergyTracker: elast is resettable (this step's sum replaces it), the dissipations accumulate
  energy_scenes_sum(c, c.pc.p, 4, 4, make_int4(WDG_EN_ELAST, WDG_EN_PLAST, WDG_EN_VISCN, WDG_EN_VISCT));
}

void energy_scenes_particles(Ctx& c) {
  // kinetic is resettable; grav and nonviscDamp accumulate
  energy_scenes_sum(c, c.pe.p, 3, 3, make_int4(WDG_EN_KINETIC, WDG_EN_GRAV, WDG_EN_NONVISCDAMP, 0));
}

void energy_scenes_read(Ctx& c, std::vector<double>& out) {
  if (c.nScenes <= 1) throw Error(WDG_E_STATE, "not a batched context of several scenes (wdg_set_scenes)");
  out.assign((size_t)c.nScenes * WDG_EN_N, 0.);
  // none tracked yet for this set of scenes: zero
  if (!c.energyS.p || c.energySScenes != c.nScenes) return;
  WDG_HIP(hipMemcpyAsync(out.data(), c.energyS.p, out.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  WDG_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace wdg
