// Contact loop: geometry (Cg2) + physics (Cp2) + constitutive law (Law2) +
// rotation of the local force to global coordinates. Replaces the per-contact
// body of ContactLoop::run (ContactLoop.cpp:134-231).
//
// Two passes over the slot list of the last refresh:
//  1. k_cfilter, one workgroup per group of slots, one thread per slot: the
//     geometric test and the law's break test decide which slots are real after
//     this step (a contact that fails the geometry test while not real is not
//     created; a real one whose law breaks it is removed). One ballot per
//     wavefront writes the real bits of 64 slots; the workgroup turns the
//     wavefront counts into tile prefixes and the list of real slots of the group.
//  2. k_cforce, one thread per real contact (dense wavefronts): full
//     L6Geom update, Cp2 and Law2, with the previous state read at the slot's
//     old index and the new state written at its new index. Both follow slot
//     order, so state reads and writes stream through HBM.
// The state of a real contact is L6Geom.trsf columns 0 and 1 (column 2 is
// always col0 x col1), local force, uN, and the 64-byte gather record (global
// force on pA, contact point, cellDist) that the force gather reads.
#include "wdg_internal.hpp"
#include "wdg_math.hpp"
#include "wdg_math32.hpp"
#include "wdg_physics.hpp"

namespace wdg {

struct ContactArgs {
  int64_t nc;
  const uint32_t *A, *B, *info;
  const uint64_t* dist;
  RealView ov, nv;  // real contacts before / after this step
  const double *oN0, *oN1, *oN2, *oY0, *oY1, *oY2, *oF1, *oF2, *ogp;
  double *N0, *N1, *N2, *Y0, *Y1, *Y2, *F0, *F1, *F2, *uN, *gp;
  unsigned long long* nbits;
  uint32_t *ntpre, *ngcnt;
  uint32_t* act;
  uint32_t* sidx;  // state index of each real slot of the new set
  const uint32_t* osidx;  // the same for the old set, when its filter made it (else null)
  uint4* desc;           // per new state index: info, pA, pB, old state index (or NO_STATE)
  uint64_t* descDist;    // per new state index: packed cellDist (periodic cells)
  // particles
  const double *px, *py, *pz, *vx, *vy, *vz, *wx, *wy, *wz, *rad, *mass;
  const uint32_t* flags;
  const int32_t* mat;
  const wdg_material* mats;
  int nMats;
  // material epochs (more than one table kept): each real contact's epoch in the old and new
  // state, the current epoch, and whether every contact takes the current one (updatePhys)
  const uint8_t* oE;
  uint8_t* nE;
  uint32_t epoch;
  int updAll;
  const double* fgeo;  // facet geometry per boundary body
  // batched scenes: dt of each particle's scene (else null, dt above)
  const double* dtScene;
  const uint32_t* scene;
  // graph replay: a set stop word skips the kernel; the step's renorm flag from the device
  const uint32_t* gstop;
  const uint32_t* renormPtr;
  int cfBlocks;  // workgroups per group of slots in the force kernel
  // parameters
  double L[3];
  double absDf;
  double dt;
  int renorm;  // 1 if trsf is renormalised this step
  double poisson, en, viscAlpha;
  int noSlip, noBreak, noFrict, noAttraction;
  // IdealElPl rolling and twisting resistance (the LAW_IEP_ROLL instantiation): the parameters
  // and the global phys torque of each real contact (3 doubles per state index)
  double relTwistStiff, rollTanPhi;
  double* gtq;
  uint32_t* dflags;
  double* enPart;  // energy tracking: 4 partial sums per force workgroup
  double* enC;     // energy tracking per scene: the 4 terms of each contact (state index), else null
  uint32_t* eidx;          // per adjacency entry: its slot when real, else NO_STATE (null: not kept)
  const uint32_t* entPos;  // per slot: positions of its two entries
};

template <bool PERIODIC>
__device__ __forceinline__ d3 slot_shift(const ContactArgs& a, int64_t c, uint64_t& cd) {
  cd = 0;
  if (!PERIODIC) return mk3(0., 0., 0.);
  cd = a.dist[c];
  return mk3(a.L[0] * unpack_cd(cd, 0), a.L[1] * unpack_cd(cd, 1), a.L[2] * unpack_cd(cd, 2));
}

// CompUtils::closestSegmentPt (CompUtils.cpp:59-65)
template <class V, class T>
__device__ __forceinline__ V closest_segment_pt(V P, V A, V B) {
  const V BA = sub3(B, A);
  if (sqn3(BA) == T(0)) return A;
  const T u = (dot3(P, BA) - dot3(A, BA)) / sqn3(BA);
  return add3(A, mul3(BA, fmin(T(1), fmax(T(0), u))));
}

// Closest point of a facet to the sphere centre sC (Cg2_Facet_Sphere_L6Geom::go,
// Facet.cpp:336-360, with Facet::getNormal 74-77 and getOuterVectors 113-118): the facet
// normal, the signed plane distance, and the closest point by Voronoi region of the
// triangle v0 v1 v2. Returns false for the impossible region (outside all three edges).
template <class V, class T>
__device__ __forceinline__ bool facet_closest(V v0, V v1, V v2, V sC, V& fNormal, T& planeDist, V& contPt,
                                              int* region = nullptr) {
  const V nn = cross3(sub3(v1, v0), sub3(v2, v0));
  fNormal = normalized3(nn);
  planeDist = dot3(sub3(sC, v0), fNormal);
  const V fC = sub3(sC, mul3(fNormal, planeDist));
  const T l0 = dot3(cross3(sub3(v1, v0), nn), sub3(fC, v0));
  const T l1 = dot3(cross3(sub3(v2, v1), nn), sub3(fC, v1));
  const T l2 = dot3(cross3(sub3(v0, v2), nn), sub3(fC, v2));
  const int w = (l0 > 0 ? 1 : 0) + (l1 > 0 ? 2 : 0) + (l2 > 0 ? 4 : 0);
  switch (w) {
    case 0: contPt = fC; break;
    case 1: contPt = closest_segment_pt<V, T>(fC, v0, v1); break;
    case 2: contPt = closest_segment_pt<V, T>(fC, v1, v2); break;
    case 4: contPt = closest_segment_pt<V, T>(fC, v2, v0); break;
    case 3: contPt = v1; break;
    case 5: contPt = v0; break;
    case 6: contPt = v2; break;
    default: contPt = fC; return false;
  }
  if (region) *region = w;
  return true;
}

// Facet::adjustBoundaryContactGeom (Facet.cpp:220-237) and its _impl (239-265): an edge or
// vertex contact (region w of facet_closest) of a facet with Facet.n21lim (per edge, the dot
// product of its normal with the edge neighbour's; NaN: none) has its normal held within the
// neighbour's, and its contact point moved to the sphere centre's projection along it; a vertex
// interpolates the outer vector and the limit between its two edges (edgeNums, Facet.cpp:372)
template <class V, class T>
__device__ __forceinline__ void facet_adjust(const V v[3], int w, V fN, V sC, const T lim[3], V& normal, V& contPt) {
  int e0, e1;
  switch (w) {
    case 1: e0 = 0; e1 = -1; break;
    case 2: e0 = 1; e1 = -1; break;
    case 3: e0 = 0; e1 = 1; break;
    case 4: e0 = 2; e1 = -1; break;
    case 5: e0 = 2; e1 = 0; break;
    case 6: e0 = 1; e1 = 2; break;
    default: return;
  }
  if (isnan(lim[e0])) return;
  // Facet::getOuterVectors (Facet.cpp:113-118): edge i is (v_i, v_i+1), its outer vector
  // (v_i+1 - v_i) x nn
  const V nn = cross3(sub3(v[1], v[0]), sub3(v[2], v[0]));
  const V out0 = normalized3(cross3(sub3(v[(e0 + 1) % 3], v[e0]), nn));
  V uo = out0, edgePt = v[e0];
  T nnLim = lim[e0];
  if (e1 >= 0) {
    if (isnan(lim[e1])) return;
    const V out1 = normalized3(cross3(sub3(v[(e1 + 1) % 3], v[e1]), nn));
    const T p0 = dot3(out0, normal), pMin = dot3(out0, out1), pMax = T(1);
    const T t01 = (p0 - pMin) / (pMax - pMin);
    uo = add3(mul3(out0, t01), mul3(out1, T(1) - t01));
    nnLim = t01 * lim[e0] + (T(1) - t01) * lim[e1];
    edgePt = v[e1];
  }
  const T nnv = dot3(normal, fN);
  if ((nnv >= T(0) && nnLim > T(0) && nnv >= nnLim) || (nnv < T(0) && nnLim < T(0) && nnv <= nnLim)) return;
  const T noLim = (fabs(nnLim) >= T(1) ? T(1) : sqrt(T(1) - nnLim * nnLim));
  T nno0, nno1;
  if ((nnLim >= T(0) && nnv <= -noLim) || (nnLim <= T(0) && nnv >= noLim)) {
    nno0 = -nnLim;
    nno1 = T(0);
  } else {
    nno0 = nnLim;
    nno1 = noLim;
  }
  normal = add3(mul3(fN, nno0), mul3(uo, nno1));
  contPt = sub3(sC, mul3(normal, dot3(sub3(sC, edgePt), normal)));
}

// Law2 break test: IdealElPl.cpp:26-33 (unless noBreak), Hertz.cpp:105-108
template <int LAW>
__device__ __forceinline__ bool law_keeps(double uN, int noBreak) {
  if (LAW == WDG_LAW_IDEALELPL) return !(uN > 0 && !noBreak);
  return !(uN > 0);
}

// What the force pass needs of a kept slot, found by the filter: its type and role bits, the
// two particles in role order (pA, pB), the state index of its previous step (NO_STATE when
// it was not real) and, in periodic cells, its packed cellDist. The filter writes it at the
// slot's new state index, so the force pass starts from one coalesced record.
struct SlotD {
  uint32_t info, pA, pB, ko;
  uint64_t cd;
};

// Group kernel tail, one workgroup of GROUP_BS threads per group of GROUP slots, called
// once per pass of GROUP_BS slots: real bits of the pass, tile prefixes, and (when act
// is given) the slot of every state index, the state index of every real slot and the
// force pass's record. running = real slots of the group so far.
__device__ __forceinline__ void group_pass(int64_t c, bool keep, int64_t base, uint32_t& running,
                                           unsigned long long* bits, uint32_t* tpre, uint32_t* act,
                                           uint32_t* sidx, const SlotD& d = SlotD{}, uint4* desc = nullptr,
                                           uint64_t* descDist = nullptr) {
  __shared__ uint32_t wc[GROUP_BS / 64];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const unsigned long long b = __ballot(keep);
  if (lane == 0) {
    bits[c >> 6] = b;
    wc[wv] = (uint32_t)__popcll(b);
  }
  __syncthreads();
  uint32_t wpre = 0, tot = 0;
  for (int w = 0; w < GROUP_BS / 64; w++) {
    const uint32_t x = wc[w];
    wpre += w < wv ? x : 0u;
    tot += x;
  }
  if (lane == 0 && (wv & 3) == 0) tpre[c >> 8] = running + wpre;
  if (keep && act) {
    const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    const uint32_t st = state_index((uint32_t)(base / GROUP), running + wpre + (uint32_t)__popcll(b & lt));
    act[st] = (uint32_t)c;
    sidx[c] = st;
    desc[st] = make_uint4(d.info, d.pA, d.pB, d.ko);
    if (descDist) descDist[st] = d.cd;
  }
  running += tot;
  __syncthreads();
}

// The per-entry real slots of the filter's slots, written after its last barrier (the loads
// of the entry positions delay no other wavefront): the entries of a slot that became real or
// stopped being real, and after a reset (EIDX_FULL) those of every real slot; the entries of
// the other slots hold their value already. Only changes are written: scattered stores are
// costly, and a contact's slot, unlike its state index, stays the same from step to step.
// wk: bit p = slot of pass p kept, bit 8+p = it was real (one register across the passes).
template <int NP>
__device__ __forceinline__ void eidx_tail(const ContactArgs& a, int64_t base, uint32_t wk) {
  const bool full = a.dflags[EIDX_FULL] != 0u;
#pragma unroll
  for (int pass = 0; pass < NP; pass++) {
    const int64_t c = base + pass * GROUP_BS + threadIdx.x;
    const bool keep = (wk >> pass) & 1u, was = (wk >> (8 + pass)) & 1u;
    if (keep ? (full || !was) : was) {
      const uint2 ep = reinterpret_cast<const uint2*>(a.entPos)[c];
      a.eidx[ep.x] = keep ? ((uint32_t)c | EIDX_PA) : NO_STATE;
      a.eidx[ep.y] = keep ? (uint32_t)c : NO_STATE;
    }
  }
}

// state index of slot c in the old real set (a real slot)
__device__ __forceinline__ uint32_t old_state(const ContactArgs& a, int64_t c) {
  return a.osidx ? a.osidx[c] : rv_index(a.ov, (uint32_t)c);
}

template <bool PERIODIC, int LAW>
__device__ __forceinline__ bool filter_slot(const ContactArgs& a, int64_t c, SlotD& d) {
  bool keep = false;
  d.ko = NO_STATE;
  if (c < a.nc) {
    const uint32_t info = a.info[c];
    const bool real = rv_real(a.ov, (uint32_t)c);
    if (real || (info & CI_NEAR)) {
      const bool swap = (info & CI_SWAP) != 0u;
      const uint32_t pA = swap ? a.B[c] : a.A[c], pB = swap ? a.A[c] : a.B[c];
      uint64_t cd;
      const d3 shift2 = slot_shift<PERIODIC>(a, c, cd);
      d.info = info; d.pA = pA; d.pB = pB; d.cd = cd;
      d.ko = real ? old_state(a, c) : NO_STATE;

      const d3 posA = mk3(a.px[pA], a.py[pA], a.pz[pA]);
      const d3 pos2 = add3(mk3(a.px[pB], a.py[pB], a.pz[pB]), shift2);
      bool pass;
      double uN;
      if ((info & CI_TYPE_MASK) == CT_SS) {
        // Cg2_Sphere_Sphere_L6Geom::go (Sphere.cpp:92-105)
        const double r1 = a.rad[pA], r2 = a.rad[pB];
        const d3 relPos = sub3(pos2, posA);
        const double unDistSq = sqn3(relPos) - p2(a.absDf * (r1 + r2));
        pass = !(unDistSq > 0 && !real);
        uN = sqrt(sqn3(relPos)) - (r1 + r2);
      } else if ((info & CI_TYPE_MASK) == CT_FS) {
        // Cg2_Facet_Sphere_L6Geom::go (Facet.cpp:333-363, 373), pA is the facet
        const double* g = a.fgeo + (int64_t)FGEO * pA;
        const double radius = a.rad[pB], ht = g[12];
        d3 fNormal, contPt;
        double planeDist;
        const bool ok = facet_closest<d3, double>(mk3(g[0], g[1], g[2]), mk3(g[3], g[4], g[5]), mk3(g[6], g[7], g[8]),
                                                  pos2, fNormal, planeDist, contPt);
        const bool near = !(fabs(planeDist) > (radius + ht) && !real);
        if (near && !ok) atomicOr(&a.dflags[1], ERR_FACET);
        const d3 nv = sub3(pos2, contPt);
        pass = near && ok && !(sqn3(nv) > p2(radius + ht) && !real);
        uN = (sqrt(sqn3(nv)) - radius) - ht;
      } else {
        // Cg2_Wall_Sphere_L6Geom::go (Wall.cpp:50-71)
        const uint32_t fl = a.flags[pA];
        const int ax = flag_axis(fl), sense = flag_sense(fl);
        const double radius = a.rad[pB];
        const double dist = comp3(pos2, ax) - comp3(posA, ax);
        pass = !(!real && fabs(dist) > radius);
        double nvv;
        if (sense == 0) {
          if (!real) nvv = dist > 0 ? 1. : -1.;
          else {
            const uint32_t ko = d.ko;
            nvv = ax == 0 ? a.oN0[ko] : (ax == 1 ? a.oN1[ko] : a.oN2[ko]);
          }
        } else {
          nvv = (sense == 1 ? 1. : -1.);
        }
        uN = nvv * dist - radius;
      }
      keep = pass && law_keeps<LAW>(uN, a.noBreak);
    }
  }
  return keep;
}

template <bool PERIODIC, int LAW>
__global__ __launch_bounds__(GROUP_BS) void k_cfilter(ContactArgs a) {
  if (a.gstop && *a.gstop) return;
  const uint32_t g = xcd_block(blockIdx.x, gridDim.x);
  const int64_t base = (int64_t)g * GROUP;
  // all slots of the thread first (their loads overlap), then the bookkeeping passes
  constexpr int NP = GROUP / GROUP_BS;
  bool keep[NP];
  SlotD d[NP];
#pragma unroll
  for (int pass = 0; pass < NP; pass++)
    keep[pass] = filter_slot<PERIODIC, LAW>(a, base + pass * GROUP_BS + threadIdx.x, d[pass]);
  uint32_t wk = 0;
#pragma unroll
  for (int pass = 0; pass < NP; pass++)
    wk |= (keep[pass] ? 1u : 0u) << pass | (d[pass].ko != NO_STATE ? 1u : 0u) << (8 + pass);
  uint32_t running = 0;
#pragma unroll
  for (int pass = 0; pass < NP; pass++)
    group_pass(base + pass * GROUP_BS + threadIdx.x, keep[pass], base, running, a.nbits, a.ntpre, a.act, a.sidx,
               d[pass], a.desc, PERIODIC ? a.descDist : nullptr);
  if (threadIdx.x == 0) a.ngcnt[g] = running;
  if (a.eidx) eidx_tail<NP>(a, base, wk);
}

// a.cfBlocks workgroups of CF_BS per group of slots
constexpr int CF_BS = 256;

// FACETS: the scene has facets (a separate instantiation keeps the facet geometry's
// registers out of the other scenes' kernel)
// ENERGY: the law's energy terms (elast, plast, viscN, viscT) of the workgroup's contacts are
// summed into a.enPart[4 * blockIdx.x ..]; every thread reaches the sum (the only early exit,
// a stopped graph replay, is taken by the whole workgroup)
// materials staged in LDS by the force kernels (a larger table is read from HBM beyond this)
constexpr int MAT_LDS = 32;
__device__ __forceinline__ void stage_materials(const ContactArgs& a, wdg_material* smat) {
  const int nt = a.nMats * (int)(a.epoch + 1);
  const int nm = nt < MAT_LDS ? nt : MAT_LDS;
  for (int t = threadIdx.x; t < nm; t += blockDim.x) smat[t] = a.mats[t];
  __syncthreads();
}
__device__ __forceinline__ wdg_material material(const ContactArgs& a, const wdg_material* smat, uint32_t m) {
  return m < (uint32_t)MAT_LDS ? smat[m] : a.mats[m];
}

// Law2_L6Geom_FrictPhys_IdealElPl with rolling and twisting resistance (an internal law code)
constexpr int LAW_IEP_ROLL = 2;
// IdealElPl.cpp:93-123: the phys torque is zeroed at every step (IdealElPl.cpp:44), so it is this
// step's dt kr angVel, kr = kn charLen, capped at |Fn| rollTanPhi charLen (twist scaled by
// relTwistStiff, capped alone); the operations in the reference's order
__device__ __forceinline__ d3 rolling_torque(const ContactArgs& a, double kn, double Fn, double charLen, double dt,
                                             d3 lav) {
  const double kr = kn * charLen;
  double t0 = 0., t1 = 0., t2 = 0.;
  if (a.relTwistStiff > 0) {
    t0 = t0 + dt * a.relTwistStiff * kr * lav.x;
    const double maxTt = fabs(Fn * a.rollTanPhi * charLen);
    if (fabs(t0) > maxTt) t0 = copysign(maxTt, t0);
  }
  const double dtkr = dt * kr;
  t1 = t1 + dtkr * lav.y;
  t2 = t2 + dtkr * lav.z;
  const double maxTr = fmax(0., fabs(Fn) * a.rollTanPhi * charLen);
  if (t1 * t1 + t2 * t2 > maxTr * maxTr) {
    const double TrNorm = sqrt(t1 * t1 + t2 * t2);
    double ratio = maxTr / TrNorm;
    if (TrNorm == 0.) ratio = 0.;
    t1 = t1 * ratio;
    t2 = t2 * ratio;
  }
  return mk3(t0, t1, t2);
}

// a law's energy term q of the contact at state index k (each term is added at most once per
// contact, so the sums are those of adding it to en alone); per scene it is kept per contact too
__device__ __forceinline__ void en_add(const ContactArgs& a, double* en, int64_t k, int q, double x) {
  en[q] += x;
  if (a.enC) a.enC[4 * k + q] = x;
}
__device__ __forceinline__ void en_clear(const ContactArgs& a, int64_t k) {
  double2* e2 = reinterpret_cast<double2*>(a.enC + 4 * k);
  e2[0] = make_double2(0., 0.);
  e2[1] = make_double2(0., 0.);
}

template <bool PERIODIC, int LAW, bool FACETS, bool ENERGY>
__global__ __launch_bounds__(CF_BS, FACETS ? 3 : 4) void k_cforce(ContactArgs a) {
  __shared__ wdg_material smat[MAT_LDS];
  if (a.gstop && *a.gstop) return;
  stage_materials(a, smat);
  const int renorm = a.renormPtr ? (int)*a.renormPtr : a.renorm;
  const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t nblk = (uint32_t)a.cfBlocks;
  const int64_t g = lb / nblk;
  const uint32_t cnt = a.ngcnt[g];
  double en[4] = {0., 0., 0., 0.};
  // the filter has written the entries of every real slot
  if (a.eidx && blockIdx.x == 0 && threadIdx.x == 0) a.dflags[EIDX_FULL] = 0u;
  for (uint32_t loc = (lb % nblk) * CF_BS + threadIdx.x; loc < cnt; loc += CF_BS * nblk) {
    const int64_t k = state_index((uint32_t)g, loc);
    if (ENERGY && a.enC) en_clear(a, k);
    // the filter's record of the slot: type bits, particles in role order, old state index
    const uint4 dk = a.desc[k];
    const uint32_t info = dk.x, pA = dk.y, pB = dk.z;
    const bool real = dk.w != NO_STATE;
    const uint32_t ko = real ? dk.w : 0u;
    const int type = (int)(info & CI_TYPE_MASK);
    // every load of the contact at once: both particles and the previous state (read at
    // index 0 for a new contact and not used)
    const d3 posA = mk3(a.px[pA], a.py[pA], a.pz[pA]);
    const d3 posB = mk3(a.px[pB], a.py[pB], a.pz[pB]);
    const double radA = a.rad[pA], radB = a.rad[pB];
    const d3 velA = mk3(a.vx[pA], a.vy[pA], a.vz[pA]), avA = mk3(a.wx[pA], a.wy[pA], a.wz[pA]);
    const d3 vel2 = mk3(a.vx[pB], a.vy[pB], a.vz[pB]);
    const d3 av2 = mk3(a.wx[pB], a.wy[pB], a.wz[pB]);
    const double massA = a.mass[pA], massB = a.mass[pB];
    // material rows of the contact's epoch (the one it was created in, or the current one):
    // folded into the indices here, so that no extra register lives through the geometry
    uint32_t matA = (uint32_t)a.mat[pA], matB = (uint32_t)a.mat[pB];
    if (a.nE) {
      const uint32_t ep = (real && !a.updAll) ? (uint32_t)a.oE[ko] : a.epoch;
      a.nE[k] = (uint8_t)ep;
      matA += ep * (uint32_t)a.nMats;
      matB += ep * (uint32_t)a.nMats;
    }
    const d3 prevNormalL = mk3(a.oN0[ko], a.oN1[ko], a.oN2[ko]);
    const d3 prevYL = mk3(a.oY0[ko], a.oY1[ko], a.oY2[ko]);
    const double oFt0 = a.oF1[ko], oFt1 = a.oF2[ko];
    const double* og = a.ogp + (int64_t)GP * ko;
    const double2 ogc = reinterpret_cast<const double2*>(og)[2];
    const d3 prevContPtL = mk3(og[3], ogc.x, ogc.y);
    uint64_t cd = 0;
    if (PERIODIC) cd = a.descDist[k];
    const double dt = a.dtScene ? a.dtScene[a.scene[pA]] : a.dt;
    d3 shift2 = mk3(0., 0., 0.);
    if (PERIODIC) shift2 = mk3(a.L[0] * unpack_cd(cd, 0), a.L[1] * unpack_cd(cd, 1), a.L[2] * unpack_cd(cd, 2));
    const d3 prevNormal = real ? prevNormalL : mk3(0., 0., 0.);

    d3 normal, contPt, pos1, pos2;
    double uN, r1, r2;
    d3 vel1 = velA, av1 = avA;
    if (type == CT_SS) {
      // Cg2_Sphere_Sphere_L6Geom::go (Sphere.cpp:92-117)
      r1 = radA;
      r2 = radB;
      pos1 = posA;
      pos2 = add3(posB, shift2);
      const d3 relPos = sub3(pos2, posA);
      const double dist = sqrt(sqn3(relPos));
      uN = dist - (r1 + r2);
      normal = div3(relPos, dist);
      contPt = add3(posA, mul3(normal, r1 + 0.5 * uN));
    } else if (FACETS && type == CT_FS) {
      // Cg2_Facet_Sphere_L6Geom::go (Facet.cpp:333-390), pA is the facet; n21lim is NaN
      // (adjustBoundaryContactGeom returns at once); static nodes: the velocity at the contact
      // point is fakeVel (zero when its x is NaN), the angular velocity zero
      const double* g = a.fgeo + (int64_t)FGEO * pA;
      const double radius = radB, ht = g[12];
      pos2 = add3(posB, shift2);
      d3 fNormal;
      double planeDist;
      const d3 fv[3] = {mk3(g[0], g[1], g[2]), mk3(g[3], g[4], g[5]), mk3(g[6], g[7], g[8])};
      int w = 0;
      facet_closest<d3, double>(fv[0], fv[1], fv[2], pos2, fNormal, planeDist, contPt, &w);
      const d3 nv = sub3(pos2, contPt);
      const double dist = sqrt(sqn3(nv));
      if (dist != 0) normal = div3(nv, dist);
      else if (real) {
        // the previous normal as node.ori * UnitX
        const d3 pY = prevYL;
        normal = quat_rot(quat_from_cols(prevNormal, pY, cross3(prevNormal, pY)), mk3(1., 0., 0.));
      } else normal = mk3(1., 0., 0.);
      if (w > 0) {
        const double lim[3] = {g[13], g[14], g[15]};
        facet_adjust<d3, double>(fv, w, fNormal, pos2, lim, normal, contPt);
      }
      if (ht > 0) contPt = add3(contPt, mul3(normal, ht));
      uN = (dist - radius) - ht;
      pos1 = contPt;
      r1 = fmax(ht, radius);
      r2 = radius;
      vel1 = isnan(g[9]) ? mk3(0., 0., 0.) : mk3(g[9], g[10], g[11]);
      av1 = mk3(0., 0., 0.);
    } else {
      // Cg2_Wall_Sphere_L6Geom::go (Wall.cpp:50-84), pA is the wall
      const uint32_t fl = a.flags[pA];
      const int ax = flag_axis(fl), sense = flag_sense(fl);
      const double radius = radB;
      pos1 = posA;
      pos2 = add3(posB, shift2);
      const double dist = comp3(pos2, ax) - comp3(posA, ax);
      contPt = pos2;
      setcomp3(contPt, ax, comp3(posA, ax));
      double nvv;
      if (sense == 0) nvv = !real ? (dist > 0 ? 1. : -1.) : comp3(prevNormal, ax);
      else nvv = (sense == 1 ? 1. : -1.);
      normal = mk3(0., 0., 0.);
      setcomp3(normal, ax, nvv);
      uN = nvv * dist - radius;
      r1 = -radius;
      r2 = radius;
    }

    // Cg2_Any_Any_L6Geom__Base::handleSpheresLikeContact (L6Geom.cpp:49-216)
    d3 T0, T1, T2, gvel;
    // L6Geom.angVel (L6Geom.cpp:211), read by the rolling resistance only; zero when fresh
    d3 lav = mk3(0., 0., 0.);
    if (!real) {
      // setInitialLocalCoords (L6Geom.cpp:25-36)
      const d3 locX = normal;
      const d3 e = fabs(locX.y) < fabs(locX.z) ? mk3(0., 1., 0.) : mk3(0., 0., 1.);
      d3 locY = cross3(locX, e);
      locY = sub3(locY, mul3(locX, dot3(locY, locX)));
      locY = normalized3(locY);
      T0 = locX;
      T1 = locY;
      T2 = cross3(locX, locY);
      gvel = mk3(0., 0., 0.);
    } else {
      const d3 prevY = prevYL;
      const d3 prevContPt = prevContPtL;
      const d3& currNormal = normal;
      const d3 normRotVec = cross3(prevNormal, currNormal);
      d3 midNormal = mul3(add3(prevNormal, currNormal), .5);
      if (sqrt(sqn3(midNormal)) == 0) midNormal = currNormal;
      else midNormal = normalized3(midNormal);
      const double twist = dot3(midNormal, add3(av1, av2));
      const d3 normTwistVec = mul3(mul3(mul3(midNormal, dt), .5), twist);
      const d3 rt = add3(normRotVec, normTwistVec);
      const d3 M0 = midNormal;
      const d3 M1 = sub3(prevY, div3(cross3(prevY, rt), 2.));
      const d3 M2 = cross3(M0, M1);
      T0 = currNormal;
      T1 = sub3(prevY, cross3(M1, rt));
      T2 = cross3(T0, T1);
      if (renorm) {
        T0 = normalized3(T0);
        T1 = sub3(T1, mul3(T0, dot3(T1, T0)));
        T1 = normalized3(T1);
        T2 = cross3(T0, T1);
      }
      const d3 midContPt = mul3(add3(prevContPt, contPt), .5);
      const d3 midPos1 = sub3(pos1, mul3(vel1, dt / 2.));
      const d3 midPos2 = sub3(pos2, mul3(vel2, dt / 2.));
      const d3 c1x = sub3(midContPt, midPos1);
      const d3 c2x = sub3(midContPt, midPos2);
      const d3 relVel = sub3(add3(vel2, cross3(av2, c2x)), add3(vel1, cross3(av1, c1x)));
      gvel = mk3(dot3(M0, relVel), dot3(M1, relVel), dot3(M2, relVel));
      if (LAW == LAW_IEP_ROLL) {
        const d3 dw = sub3(av2, av1);
        lav = mk3(dot3(M0, dw), dot3(M1, dw), dot3(M2, dw));
      }
    }

    // Cp2 (recomputed; constant over the life of the contact)
    double l0, l1, A;
    lens_area(r1, r2, l0, l1, A);
    const wdg_material mA = material(a, smat, matA), mB = material(a, smat, matB);
    double Fn, Ft0 = real ? oFt0 : 0., Ft1 = real ? oFt1 : 0.;
    d3 tq = mk3(0., 0., 0.);  // CPhys.torque (local), rolling resistance only
    if (LAW != WDG_LAW_HERTZDMT) {
      // Law2_L6Geom_FrictPhys_IdealElPl::go (IdealElPl.cpp:15-125); breaking decided by k_cfilter
      PhysVals ph;
      phys_frict(mA, mB, l0, l1, A, ph);
      Fn = ph.kn * uN;
      if (a.noFrict || ph.tanPhi == 0.) {
        Ft0 = 0.;
        Ft1 = 0.;
      } else {
        const double dk = dt * ph.kt;
        Ft0 = Ft0 + dk * gvel.y;
        Ft1 = Ft1 + dk * gvel.z;
        const double maxFt = fabs(Fn) * ph.tanPhi;
        if (Ft0 * Ft0 + Ft1 * Ft1 > maxFt * maxFt && !a.noSlip) {
          const double FtNorm = sqrt(Ft0 * Ft0 + Ft1 * Ft1);
          double ratio = maxFt / FtNorm;
          if (FtNorm == 0.) ratio = 0.;
          // IdealElPl.cpp:62-73: slip dissipation, while Ft is the trial value
          if (ENERGY) en_add(a, en, k, 1, (ph.kt != 0. ? maxFt * (FtNorm - maxFt) / ph.kt : 0.));
          Ft0 = Ft0 * ratio;
          Ft1 = Ft1 * ratio;
        }
        if (isnan(Fn) || isnan(Ft0) || isnan(Ft1)) atomicOr(&a.dflags[1], ERR_NAN_FORCE);
      }
      if (ENERGY) {
        // IdealElPl.cpp:85-92: elastic potential, 0 when it is NaN
        double el = 0.5 * (p2(Fn) / ph.kn + (ph.kt != 0. ? (Ft0 * Ft0 + Ft1 * Ft1) / ph.kt : 0.));
        if (isnan(el)) el = 0.;
        en_add(a, en, k, 0, el);
      }
      if (LAW == LAW_IEP_ROLL) tq = rolling_torque(a, ph.kn, Fn, l0 + l1, dt, lav);
    } else {
      // Law2_L6Geom_HertzPhys_DMT::go (Hertz.cpp:85-210), DMT branch; breaking decided by k_cfilter
      PhysVals ph;
      phys_hertz(mA, mB, l0, l1, massA, massB, a.poisson, a.en, a.viscAlpha, ph);
      const double velN = gvel.x;
      const double kn0 = ph.K * sqrt(ph.R);
      const double s = sqrt(-uN);
      const double Fne = -kn0 * ((s * s) * s);
      const double Fna = 4 * M_PI * ph.R * ph.gamma;
      const double eta = (ph.alpha_sqrtMK > 0 ? ph.alpha_sqrtMK * sqrt(sqrt(fmax(-uN, 0.))) : 0.);
      double Fnv = eta * velN;
      if (ph.alpha == 0. && a.noAttraction && Fne + Fnv > 0) Fnv = -Fne;
      Fn = Fne + Fna + Fnv;
      if (ENERGY) en_add(a, en, k, 2, Fnv * velN * dt);  // Hertz.cpp:181
      const double kt = ph.kt0 * sqrt(uN < 0 ? -uN : 0);
      const double dk = dt * kt;
      Ft0 = Ft0 + dk * gvel.y;
      Ft1 = Ft1 + dk * gvel.z;
      const double maxFt = fabs(fmin(0., Fn) * ph.tanPhi);
      if (Ft0 * Ft0 + Ft1 * Ft1 > p2(maxFt)) {
        const double FtNorm = sqrt(Ft0 * Ft0 + Ft1 * Ft1);
        const double ratio = maxFt / FtNorm;
        if (ENERGY) en_add(a, en, k, 1, (.5 * (FtNorm - maxFt) + maxFt) * (FtNorm - maxFt) / kt);  // Hertz.cpp:193
        Ft0 = Ft0 * ratio;
        Ft1 = Ft1 * ratio;
      } else {
        Ft0 = Ft0 + eta * gvel.y;
        Ft1 = Ft1 + eta * gvel.z;
        if (ENERGY) en_add(a, en, k, 3, eta * (gvel.y * gvel.y + gvel.z * gvel.z) * dt);  // Hertz.cpp:199
      }
      if (ENERGY && uN < 0) {
        // Hertz.cpp:203-206: kn0 (2/5) (-uN)^(5/2) with pown's square-and-multiply
        const double q = s * s;
        en_add(a, en, k, 0, kn0 * (2 / 5.) * ((q * q) * s) + 0.5 * (Ft0 * Ft0 + Ft1 * Ft1) / kt);
      }
    }
    a.N0[k] = T0.x; a.N1[k] = T0.y; a.N2[k] = T0.z;
    a.Y0[k] = T1.x; a.Y1[k] = T1.y; a.Y2[k] = T1.z;
    a.F0[k] = Fn; a.F1[k] = Ft0; a.F2[k] = Ft1;
    a.uN[k] = uN;
    // Contact::getForceTorqueBranch: F = node.ori * force, node.ori = Quaternion(trsf)
    const q4 q = quat_from_cols(T0, T1, T2);
    const d3 G = quat_rot(q, mk3(Fn, Ft0, Ft1));
    double2* g2 = reinterpret_cast<double2*>(a.gp + (int64_t)GP * k);
    g2[0] = make_double2(G.x, G.y);
    g2[1] = make_double2(G.z, contPt.x);
    g2[2] = make_double2(contPt.y, contPt.z);
    g2[3] = make_double2(__longlong_as_double((long long)cd), 0.);
    if (LAW == LAW_IEP_ROLL) {
      // Contact::getForceTorqueBranch: T = node.ori * torque, zero when the torque is zero
      const bool tz = tq.x == 0. && tq.y == 0. && tq.z == 0.;
      const d3 TG = tz ? mk3(0., 0., 0.) : quat_rot(q, tq);
      double* gt = a.gtq + (int64_t)3 * k;
      gt[0] = TG.x; gt[1] = TG.y; gt[2] = TG.z;
    }
  }
  if (ENERGY && !a.enC) block_sum_store<4, CF_BS>(en, a.enPart + (int64_t)4 * blockIdx.x);
}

// ================================================================== FP32 contact form
// Contact geometry, physics and law in float; positions, the position/velocity update and
// the per-particle force sums stay FP64. The geometry starts from the FP64 relative position
// of pB's image (no cancellation in float); the contact point is kept relative to pA, and
// the gather record carries both branch vectors (from pA, from pB's image), so the gather
// reads no positions. The contact point's update uses its previous value relative to pA:
// midContPt - midPos1 = (prev + current)/2 when pA moved by vel1*dt since the last step.
// Float state per real contact: trsf columns 0-1 (6), local force (3), uN (1) in the
// state arrays (their first halves), and a 12-float record: global force on pA (3), branch
// from pA (3), branch from pB (3), unused (3).
__device__ __forceinline__ f3 tof3(d3 v) { return mkf3((float)v.x, (float)v.y, (float)v.z); }

template <bool PERIODIC, int LAW>
__device__ __forceinline__ bool filter_slot32(const ContactArgs& a, int64_t c, SlotD& d) {
  bool keep = false;
  d.ko = NO_STATE;
  if (c < a.nc) {
    const uint32_t info = a.info[c];
    const bool real = rv_real(a.ov, (uint32_t)c);
    if (real || (info & CI_NEAR)) {
      const bool swap = (info & CI_SWAP) != 0u;
      const uint32_t pA = swap ? a.B[c] : a.A[c], pB = swap ? a.A[c] : a.B[c];
      uint64_t cd;
      const d3 shift2 = slot_shift<PERIODIC>(a, c, cd);
      d.info = info; d.pA = pA; d.pB = pB; d.cd = cd;
      d.ko = real ? old_state(a, c) : NO_STATE;

      const f3 r = tof3(sub3(add3(mk3(a.px[pB], a.py[pB], a.pz[pB]), shift2), mk3(a.px[pA], a.py[pA], a.pz[pA])));
      bool pass;
      float uN;
      if ((info & CI_TYPE_MASK) == CT_SS) {
        const float r1 = (float)a.rad[pA], r2 = (float)a.rad[pB];
        const float reach = (float)a.absDf * (r1 + r2);
        pass = !(sqn3(r) - reach * reach > 0.f && !real);
        uN = sqrtf(sqn3(r)) - (r1 + r2);
      } else if ((info & CI_TYPE_MASK) == CT_FS) {
        // facet vertices relative to vertex 0 (the facet's node)
        const double* g = a.fgeo + (int64_t)FGEO * pA;
        const float radius = (float)a.rad[pB], ht = (float)g[12];
        const f3 e1 = tof3(mk3(g[3] - g[0], g[4] - g[1], g[5] - g[2]));
        const f3 e2 = tof3(mk3(g[6] - g[0], g[7] - g[1], g[8] - g[2]));
        f3 fNormal, contPt;
        float planeDist;
        const bool ok = facet_closest<f3, float>(mkf3(0.f, 0.f, 0.f), e1, e2, r, fNormal, planeDist, contPt);
        const bool near = !(fabsf(planeDist) > (radius + ht) && !real);
        if (near && !ok) atomicOr(&a.dflags[1], ERR_FACET);
        const f3 nv = sub3(r, contPt);
        pass = near && ok && !(sqn3(nv) > (radius + ht) * (radius + ht) && !real);
        uN = (sqrtf(sqn3(nv)) - radius) - ht;
      } else {
        const uint32_t fl = a.flags[pA];
        const int ax = flag_axis(fl), sense = flag_sense(fl);
        const float radius = (float)a.rad[pB];
        const float dist = comp3(r, ax);
        pass = !(!real && fabsf(dist) > radius);
        float nvv;
        if (sense == 0) {
          if (!real) nvv = dist > 0.f ? 1.f : -1.f;
          else {
            const uint32_t ko = d.ko;
            const float* on = reinterpret_cast<const float*>(ax == 0 ? a.oN0 : (ax == 1 ? a.oN1 : a.oN2));
            nvv = on[ko];
          }
        } else {
          nvv = sense == 1 ? 1.f : -1.f;
        }
        uN = nvv * dist - radius;
      }
      keep = pass && law_keeps<LAW>((double)uN, a.noBreak);
    }
  }
  return keep;
}

template <bool PERIODIC, int LAW>
__global__ __launch_bounds__(GROUP_BS) void k_cfilter32(ContactArgs a) {
  if (a.gstop && *a.gstop) return;
  const uint32_t g = xcd_block(blockIdx.x, gridDim.x);
  const int64_t base = (int64_t)g * GROUP;
  constexpr int NP = GROUP / GROUP_BS;
  bool keep[NP];
  SlotD d[NP];
#pragma unroll
  for (int pass = 0; pass < NP; pass++)
    keep[pass] = filter_slot32<PERIODIC, LAW>(a, base + pass * GROUP_BS + threadIdx.x, d[pass]);
  uint32_t wk = 0;
#pragma unroll
  for (int pass = 0; pass < NP; pass++)
    wk |= (keep[pass] ? 1u : 0u) << pass | (d[pass].ko != NO_STATE ? 1u : 0u) << (8 + pass);
  uint32_t running = 0;
#pragma unroll
  for (int pass = 0; pass < NP; pass++)
    group_pass(base + pass * GROUP_BS + threadIdx.x, keep[pass], base, running, a.nbits, a.ntpre, a.act, a.sidx,
               d[pass], a.desc, PERIODIC ? a.descDist : nullptr);
  if (threadIdx.x == 0) a.ngcnt[g] = running;
  if (a.eidx) eidx_tail<NP>(a, base, wk);
}

template <bool PERIODIC, int LAW, bool FACETS, bool ENERGY>
__global__ __launch_bounds__(CF_BS, 4) void k_cforce32(ContactArgs a) {
  const float *oN0 = reinterpret_cast<const float*>(a.oN0), *oN1 = reinterpret_cast<const float*>(a.oN1),
              *oN2 = reinterpret_cast<const float*>(a.oN2), *oY0 = reinterpret_cast<const float*>(a.oY0),
              *oY1 = reinterpret_cast<const float*>(a.oY1), *oY2 = reinterpret_cast<const float*>(a.oY2),
              *oF1 = reinterpret_cast<const float*>(a.oF1), *oF2 = reinterpret_cast<const float*>(a.oF2),
              *ogp = reinterpret_cast<const float*>(a.ogp);
  float *N0 = reinterpret_cast<float*>(a.N0), *N1 = reinterpret_cast<float*>(a.N1),
        *N2 = reinterpret_cast<float*>(a.N2), *Y0 = reinterpret_cast<float*>(a.Y0), *Y1 = reinterpret_cast<float*>(a.Y1),
        *Y2 = reinterpret_cast<float*>(a.Y2), *F0 = reinterpret_cast<float*>(a.F0), *F1 = reinterpret_cast<float*>(a.F1),
        *F2 = reinterpret_cast<float*>(a.F2), *UN = reinterpret_cast<float*>(a.uN), *gp = reinterpret_cast<float*>(a.gp);
  if (a.gstop && *a.gstop) return;
  const int renorm = a.renormPtr ? (int)*a.renormPtr : a.renorm;
  const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t nblk = (uint32_t)a.cfBlocks;
  const int64_t g = lb / nblk;
  const uint32_t cnt = a.ngcnt[g];
  // energy terms from the float law's values, evaluated and summed in FP64
  double en[4] = {0., 0., 0., 0.};
  if (a.eidx && blockIdx.x == 0 && threadIdx.x == 0) a.dflags[EIDX_FULL] = 0u;
  for (uint32_t loc = (lb % nblk) * CF_BS + threadIdx.x; loc < cnt; loc += CF_BS * nblk) {
    const int64_t k = state_index((uint32_t)g, loc);
    if (ENERGY && a.enC) en_clear(a, k);
    // the filter's record of the slot: type bits, particles in role order, old state index
    const uint4 dk = a.desc[k];
    const uint32_t info = dk.x, pA = dk.y, pB = dk.z;
    const bool real = dk.w != NO_STATE;
    const uint32_t ko = real ? dk.w : 0u;
    // material rows of the contact's epoch (folded into the particles' material indices)
    uint32_t eb = 0u;
    if (a.nE) {
      const uint32_t ep = (real && !a.updAll) ? (uint32_t)a.oE[ko] : a.epoch;
      a.nE[k] = (uint8_t)ep;
      eb = ep * (uint32_t)a.nMats;
    }
    const float dt = (float)(a.dtScene ? a.dtScene[a.scene[pA]] : a.dt);
    uint64_t cd = 0;
    d3 shift2 = mk3(0., 0., 0.);
    if (PERIODIC) {
      cd = a.descDist[k];
      shift2 = mk3(a.L[0] * unpack_cd(cd, 0), a.L[1] * unpack_cd(cd, 1), a.L[2] * unpack_cd(cd, 2));
    }
    const f3 r = tof3(sub3(add3(mk3(a.px[pB], a.py[pB], a.pz[pB]), shift2), mk3(a.px[pA], a.py[pA], a.pz[pA])));
    f3 prevNormal = mkf3(0.f, 0.f, 0.f);
    if (real) prevNormal = mkf3(oN0[ko], oN1[ko], oN2[ko]);
    f3 normal, cA;
    float uN, r1, r2;
    f3 vel1 = tof3(mk3(a.vx[pA], a.vy[pA], a.vz[pA])), av1 = tof3(mk3(a.wx[pA], a.wy[pA], a.wz[pA]));
    f3 vmove1 = vel1;  // the velocity pA's node moved with over the last step
    if ((info & CI_TYPE_MASK) == CT_SS) {
      r1 = (float)a.rad[pA];
      r2 = (float)a.rad[pB];
      const float dist = sqrtf(sqn3(r));
      uN = dist - (r1 + r2);
      normal = div3(r, dist);
      cA = mul3(normal, r1 + 0.5f * uN);
    } else if (FACETS && (info & CI_TYPE_MASK) == CT_FS) {
      const double* g = a.fgeo + (int64_t)FGEO * pA;
      const float radius = (float)a.rad[pB], ht = (float)g[12];
      const f3 e1 = tof3(mk3(g[3] - g[0], g[4] - g[1], g[5] - g[2]));
      const f3 e2 = tof3(mk3(g[6] - g[0], g[7] - g[1], g[8] - g[2]));
      f3 fNormal;
      float planeDist;
      int w = 0;
      facet_closest<f3, float>(mkf3(0.f, 0.f, 0.f), e1, e2, r, fNormal, planeDist, cA, &w);
      const f3 nv = sub3(r, cA);
      const float dist = sqrtf(sqn3(nv));
      if (dist != 0.f) normal = div3(nv, dist);
      else if (real) {
        const f3 pY = mkf3(oY0[ko], oY1[ko], oY2[ko]);
        normal = quat_rot(quat_from_cols(prevNormal, pY, cross3(prevNormal, pY)), mkf3(1.f, 0.f, 0.f));
      } else normal = mkf3(1.f, 0.f, 0.f);
      if (w > 0) {
        const f3 fv[3] = {mkf3(0.f, 0.f, 0.f), e1, e2};
        const float lim[3] = {(float)g[13], (float)g[14], (float)g[15]};
        facet_adjust<f3, float>(fv, w, fNormal, r, lim, normal, cA);
      }
      if (ht > 0.f) cA = add3(cA, mul3(normal, ht));
      uN = (dist - radius) - ht;
      r1 = fmaxf(ht, radius);
      r2 = radius;
      vel1 = isnan(g[9]) ? mkf3(0.f, 0.f, 0.f) : tof3(mk3(g[9], g[10], g[11]));
      av1 = mkf3(0.f, 0.f, 0.f);
      vmove1 = mkf3(0.f, 0.f, 0.f);
    } else {
      const uint32_t fl = a.flags[pA];
      const int ax = flag_axis(fl), sense = flag_sense(fl);
      const float radius = (float)a.rad[pB];
      const float dist = comp3(r, ax);
      // the sphere centre projected on the wall plane, relative to the wall node
      cA = r;
      setcomp3(cA, ax, 0.f);
      float nvv;
      if (sense == 0) nvv = !real ? (dist > 0.f ? 1.f : -1.f) : comp3(prevNormal, ax);
      else nvv = sense == 1 ? 1.f : -1.f;
      normal = mkf3(0.f, 0.f, 0.f);
      setcomp3(normal, ax, nvv);
      uN = nvv * dist - radius;
      r1 = -radius;
      r2 = radius;
    }
    const f3 cB = sub3(cA, r);
    const f3 vel2 = tof3(mk3(a.vx[pB], a.vy[pB], a.vz[pB]));
    const f3 av2 = tof3(mk3(a.wx[pB], a.wy[pB], a.wz[pB]));
    f3 T0, T1, T2, gvel;
    if (!real) {
      const f3 locX = normal;
      const f3 e = fabsf(locX.y) < fabsf(locX.z) ? mkf3(0.f, 1.f, 0.f) : mkf3(0.f, 0.f, 1.f);
      f3 locY = cross3(locX, e);
      locY = sub3(locY, mul3(locX, dot3(locY, locX)));
      locY = normalized3(locY);
      T0 = locX;
      T1 = locY;
      T2 = cross3(locX, locY);
      gvel = mkf3(0.f, 0.f, 0.f);
    } else {
      const f3 prevY = mkf3(oY0[ko], oY1[ko], oY2[ko]);
      const float* og = ogp + (int64_t)GP32 * ko;
      const f3 prevCA = mkf3(og[3], og[4], og[5]);
      const f3 normRotVec = cross3(prevNormal, normal);
      f3 midNormal = mul3(add3(prevNormal, normal), .5f);
      if (sqrtf(sqn3(midNormal)) == 0.f) midNormal = normal;
      else midNormal = normalized3(midNormal);
      const float twist = dot3(midNormal, add3(av1, av2));
      const f3 normTwistVec = mul3(mul3(mul3(midNormal, dt), .5f), twist);
      const f3 rt = add3(normRotVec, normTwistVec);
      const f3 M0 = midNormal;
      const f3 M1 = sub3(prevY, div3(cross3(prevY, rt), 2.f));
      const f3 M2 = cross3(M0, M1);
      T0 = normal;
      T1 = sub3(prevY, cross3(M1, rt));
      T2 = cross3(T0, T1);
      if (renorm) {
        T0 = normalized3(T0);
        T1 = sub3(T1, mul3(T0, dot3(T1, T0)));
        T1 = normalized3(T1);
        T2 = cross3(T0, T1);
      }
      // midContPt - midPos1 and midContPt - midPos2 relative to pA's node, which moved by
      // vmove1*dt since the last step (a facet's does not). pos1 is pA's node (sphere, wall),
      // or the contact point for a facet
      const f3 mid = mul3(add3(prevCA, cA), .5f);
      const f3 c1x = FACETS && (info & CI_TYPE_MASK) == CT_FS ? add3(sub3(mid, cA), mul3(vel1, .5f * dt)) : mid;
      const f3 c2x = add3(sub3(mid, r), mul3(sub3(vel2, vmove1), .5f * dt));
      const f3 relVel = sub3(add3(vel2, cross3(av2, c2x)), add3(vel1, cross3(av1, c1x)));
      gvel = mkf3(dot3(M0, relVel), dot3(M1, relVel), dot3(M2, relVel));
    }
    // Cp2 parameters in FP64 (functions of radii, materials and masses), the law in float
    double l0, l1, A;
    lens_area((double)r1, (double)r2, l0, l1, A);
    const wdg_material mA = a.mats[eb + a.mat[pA]], mB = a.mats[eb + a.mat[pB]];
    float Fn, Ft0 = real ? oF1[ko] : 0.f, Ft1 = real ? oF2[ko] : 0.f;
    if (LAW == WDG_LAW_IDEALELPL) {
      PhysVals ph;
      phys_frict(mA, mB, l0, l1, A, ph);
      const float kn = (float)ph.kn, kt = (float)ph.kt, tanPhi = (float)ph.tanPhi;
      Fn = kn * uN;
      if (a.noFrict || tanPhi == 0.f) {
        Ft0 = 0.f;
        Ft1 = 0.f;
      } else {
        const float dk = dt * kt;
        Ft0 = Ft0 + dk * gvel.y;
        Ft1 = Ft1 + dk * gvel.z;
        const float maxFt = fabsf(Fn) * tanPhi;
        if (Ft0 * Ft0 + Ft1 * Ft1 > maxFt * maxFt && !a.noSlip) {
          const float FtNorm = sqrtf(Ft0 * Ft0 + Ft1 * Ft1);
          float ratio = maxFt / FtNorm;
          if (FtNorm == 0.f) ratio = 0.f;
          if (ENERGY) en_add(a, en, k, 1, (kt != 0.f ? (double)maxFt * ((double)FtNorm - maxFt) / kt : 0.));
          Ft0 = Ft0 * ratio;
          Ft1 = Ft1 * ratio;
        }
        if (isnan(Fn) || isnan(Ft0) || isnan(Ft1)) atomicOr(&a.dflags[1], ERR_NAN_FORCE);
      }
      if (ENERGY) {
        const double f0 = Fn, f1 = Ft0, f2 = Ft1;
        double el = 0.5 * (f0 * f0 / kn + (kt != 0.f ? (f1 * f1 + f2 * f2) / kt : 0.));
        if (isnan(el)) el = 0.;
        en_add(a, en, k, 0, el);
      }
    } else {
      PhysVals ph;
      phys_hertz(mA, mB, l0, l1, a.mass[pA], a.mass[pB], a.poisson, a.en, a.viscAlpha, ph);
      const float velN = gvel.x;
      const float kn0 = (float)(ph.K * sqrt(ph.R));
      const float s = sqrtf(-uN);
      const float Fne = -kn0 * ((s * s) * s);
      const float Fna = (float)(4 * M_PI * ph.R * ph.gamma);
      const float aMK = (float)ph.alpha_sqrtMK;
      const float eta = (aMK > 0.f ? aMK * sqrtf(sqrtf(fmaxf(-uN, 0.f))) : 0.f);
      float Fnv = eta * velN;
      if (ph.alpha == 0. && a.noAttraction && Fne + Fnv > 0.f) Fnv = -Fne;
      Fn = Fne + Fna + Fnv;
      if (ENERGY) en_add(a, en, k, 2, (double)Fnv * velN * dt);
      const float kt = (float)ph.kt0 * sqrtf(uN < 0.f ? -uN : 0.f);
      const float dk = dt * kt;
      Ft0 = Ft0 + dk * gvel.y;
      Ft1 = Ft1 + dk * gvel.z;
      const float maxFt = fabsf(fminf(0.f, Fn) * (float)ph.tanPhi);
      if (Ft0 * Ft0 + Ft1 * Ft1 > maxFt * maxFt) {
        const float FtNorm = sqrtf(Ft0 * Ft0 + Ft1 * Ft1);
        const float ratio = maxFt / FtNorm;
        if (ENERGY) {
          const double d = (double)FtNorm - maxFt;
          en_add(a, en, k, 1, (.5 * d + maxFt) * d / kt);
        }
        Ft0 = Ft0 * ratio;
        Ft1 = Ft1 * ratio;
      } else {
        Ft0 = Ft0 + eta * gvel.y;
        Ft1 = Ft1 + eta * gvel.z;
        if (ENERGY) en_add(a, en, k, 3, (double)eta * ((double)gvel.y * gvel.y + (double)gvel.z * gvel.z) * dt);
      }
      if (ENERGY && uN < 0.f) {
        const double sd = sqrt(-(double)uN), q = sd * sd, f1 = Ft0, f2 = Ft1;
        en_add(a, en, k, 0, ph.K * sqrt(ph.R) * (2 / 5.) * ((q * q) * sd) + 0.5 * (f1 * f1 + f2 * f2) / kt);
      }
    }
    N0[k] = T0.x; N1[k] = T0.y; N2[k] = T0.z;
    Y0[k] = T1.x; Y1[k] = T1.y; Y2[k] = T1.z;
    F0[k] = Fn; F1[k] = Ft0; F2[k] = Ft1;
    UN[k] = uN;
    const f3 G = quat_rot(quat_from_cols(T0, T1, T2), mkf3(Fn, Ft0, Ft1));
    float4* g4 = reinterpret_cast<float4*>(gp + (int64_t)GP32 * k);
    g4[0] = make_float4(G.x, G.y, G.z, cA.x);
    g4[1] = make_float4(cA.y, cA.z, cB.x, cB.y);
    g4[2] = make_float4(cB.z, 0.f, 0.f, 0.f);
  }
  if (ENERGY && !a.enC) block_sum_store<4, CF_BS>(en, a.enPart + (int64_t)4 * blockIdx.x);
}

static ContactArgs contact_args(Ctx& c, int from, int to) {
  Contacts& C = c.con();
  Reals& O = c.R[from];
  Reals& N = c.R[to];
  ContactArgs a{};
  a.nc = C.n;
  a.A = C.A.p; a.B = C.B.p; a.info = C.info.p; a.dist = C.dist.p;
  a.ov = c.rview(from);
  a.nv = c.rview(to);
  a.oN0 = O.hN[0].p; a.oN1 = O.hN[1].p; a.oN2 = O.hN[2].p;
  a.oY0 = O.hY[0].p; a.oY1 = O.hY[1].p; a.oY2 = O.hY[2].p;
  a.oF1 = O.hF[1].p; a.oF2 = O.hF[2].p; a.ogp = O.gp.p;
  a.N0 = N.hN[0].p; a.N1 = N.hN[1].p; a.N2 = N.hN[2].p;
  a.Y0 = N.hY[0].p; a.Y1 = N.hY[1].p; a.Y2 = N.hY[2].p;
  a.F0 = N.hF[0].p; a.F1 = N.hF[1].p; a.F2 = N.hF[2].p; a.uN = N.huN.p; a.gp = N.gp.p;
  a.nbits = N.bits.p;
  a.ntpre = N.tpre.p;
  a.ngcnt = N.gcnt.p;
  a.act = c.act.p;
  a.sidx = N.sidx.p;
  a.osidx = O.sidxValid ? O.sidx.p : nullptr;
  a.desc = c.desc.p;
  a.descDist = c.descDist.p;
  a.px = c.pos[0].p; a.py = c.pos[1].p; a.pz = c.pos[2].p;
  a.vx = c.vel[0].p; a.vy = c.vel[1].p; a.vz = c.vel[2].p;
  a.wx = c.angVel[0].p; a.wy = c.angVel[1].p; a.wz = c.angVel[2].p;
  a.rad = c.radius.p; a.mass = c.mass.p; a.flags = c.flags.p; a.mat = c.mat.p; a.mats = c.dMats.p;
  a.nMats = (int)c.mats.size();
  a.epoch = (uint32_t)(c.matEpochs - 1);
  a.updAll = c.updatePhys != 0;
  if (c.matEpochs > 1) {
    a.oE = O.hE.p;
    a.nE = N.hE.p;
  }
  a.fgeo = c.fgeo.p;
  a.dtScene = c.nScenes > 0 ? c.dtScene.p : nullptr;
  a.scene = c.scene.p;
  a.cfBlocks = c.tune.cfBlocks;
  if (c.graphCapture) {
    a.gstop = c.gstate.p;
    a.renormPtr = c.gstate.p + 2;
  }
  for (int k = 0; k < 3; k++) a.L[k] = c.field.cell[k];
  a.absDf = std::fabs(c.field.distFactor);
  a.dflags = c.dflags.p;
  return a;
}

__global__ void k_fill_u32(int64_t n, uint32_t* p, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------------------------ ContactLoop.stress
// ContactLoop.cpp:216-229: per real contact F branch^T, F = node.ori * force (the gather
// record's global force), branch = Contact::dPos = pB + periodic shift - pA, at the contact
// loop's positions. One workgroup per group of slots (its real contacts are state_index(g, 0 ..
// gcnt[g]-1)), a fixed-order workgroup sum per group; k_cl_stress_sum adds the groups in order.
template <bool PERIODIC>
__global__ __launch_bounds__(256) void k_cl_stress(ContactArgs a, int64_t ng, double* part) {
  if (a.gstop && *a.gstop) return;
  const int64_t g = blockIdx.x;
  const uint32_t cnt = a.ngcnt[g];
  double v[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
  for (uint32_t loc = threadIdx.x; loc < cnt; loc += blockDim.x) {
    const int64_t k = state_index((uint32_t)g, loc);
    const uint4 dk = a.desc[k];
    const uint32_t pA = dk.y, pB = dk.z;
    d3 b = sub3(mk3(a.px[pB], a.py[pB], a.pz[pB]), mk3(a.px[pA], a.py[pA], a.pz[pA]));
    if (PERIODIC) {
      const uint64_t cd = a.descDist[k];
      b = add3(b, mk3(a.L[0] * unpack_cd(cd, 0), a.L[1] * unpack_cd(cd, 1), a.L[2] * unpack_cd(cd, 2)));
    }
    const double2* r = reinterpret_cast<const double2*>(a.gp + (int64_t)GP * k);
    const double2 r0 = r[0], r1 = r[1];
    const double F[3] = {r0.x, r0.y, r1.x}, B[3] = {b.x, b.y, b.z};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) v[3 * i + j] += F[i] * B[j];
  }
  block_sum_store<9, 256>(v, part + 9 * g);
}
__global__ __launch_bounds__(256) void k_cl_stress_sum(int64_t ng, const double* __restrict__ part, double vol,
                                                       const uint32_t* gstop, double* out) {
  if (gstop && *gstop) return;
  double v[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
  for (int64_t g = threadIdx.x; g < ng; g += blockDim.x)
#pragma unroll
    for (int q = 0; q < 9; q++) v[q] += part[9 * g + q];
  block_sum_store<9, 256>(v, out);
  __syncthreads();
  if (threadIdx.x < 9) out[threadIdx.x] = out[threadIdx.x] / vol;
}

// ------------------------------------------------------------------ material epochs
// which epochs the real contacts refer to (256 bits); a workgroup per group of slots, whose
// real contacts hold the state indices state_index(g, 0 .. gcnt[g]-1)
__global__ void k_epoch_used(int64_t ng, const uint32_t* __restrict__ gcnt, const uint8_t* __restrict__ hE,
                             uint32_t* used) {
  __shared__ uint32_t sm[8];
  if (threadIdx.x < 8) sm[threadIdx.x] = 0u;
  __syncthreads();
  for (int64_t g = blockIdx.x; g < ng; g += gridDim.x)
    for (uint32_t loc = threadIdx.x; loc < gcnt[g]; loc += blockDim.x) {
      const uint32_t e = hE[state_index((uint32_t)g, loc)];
      atomicOr(&sm[e >> 5], 1u << (e & 31u));
    }
  __syncthreads();
  if (threadIdx.x < 8 && sm[threadIdx.x]) atomicOr(&used[threadIdx.x], sm[threadIdx.x]);
}
struct EpochMap {
  uint8_t m[256];
};
// the real contacts' epochs renumbered
__global__ void k_epoch_map(int64_t ng, const uint32_t* __restrict__ gcnt, uint8_t* hE, EpochMap map) {
  for (int64_t g = blockIdx.x; g < ng; g += gridDim.x)
    for (uint32_t loc = threadIdx.x; loc < gcnt[g]; loc += blockDim.x) {
      const uint32_t k = state_index((uint32_t)g, loc);
      hE[k] = map.m[hE[k]];
    }
}

// ContactLoop.updatePhys (ContactLoop.hpp:117, ContactLoop.cpp:180): a contact keeps the phys
// its creation computed unless the loop updates it. Phys here are recomputed from the materials
// at every step, so 'never' keeps, for each real contact, the material table it was created
// with: a changed table while contacts exist becomes a new epoch after the tables (of the same
// length) the real contacts still refer to, and the contact loop reads each contact's epoch.
static bool same_material(const wdg_material& a, const wdg_material& b) {
  return a.kind == b.kind && a.density == b.density && a.young == b.young && a.ktDivKn == b.ktDivKn &&
         a.tanPhi == b.tanPhi && a.surfEnergy == b.surfEnergy && a.alpha == b.alpha;
}

void set_materials(Ctx& c, const wdg_material* mats, int n) {
  const int nOld = (int)c.mats.size();
  bool same = n == nOld && n > 0;
  for (int m = 0; same && m < n; m++) same = same_material(c.mats[m], mats[m]);
  if (same) return;
  const int64_t nReal = (c.n > 0 && c.con().n > 0) ? count_real(c) : 0;
  std::vector<wdg_material> tab;
  int epochs = 1;
  if (nReal > 0 && nOld > 0 && n >= nOld) {
    Reals& R = c.reals();
    const int64_t ng = ngroups(c.con().n);
    const unsigned nb = (unsigned)std::min<int64_t>(ng, 1024);
    // the state arrays span every group's range (state_index)
    if (c.matEpochs == 1) WDG_HIP(hipMemsetAsync(R.hE.p, 0, (size_t)ng * GROUP, c.stream));
    // the epochs still referred to, renumbered in order; the new table is the last one
    DBuf<uint32_t> used;
    used.ensure(8);
    WDG_HIP(hipMemsetAsync(used.p, 0, 8 * sizeof(uint32_t), c.stream));
    hipLaunchKernelGGL(k_epoch_used, dim3(nb), dim3(256), 0, c.stream, ng, (const uint32_t*)R.gcnt.p,
                       (const uint8_t*)R.hE.p, used.p);
    WDG_LAUNCH_CHECK();
    uint32_t hu[8];
    WDG_HIP(hipMemcpyAsync(hu, used.p, sizeof(hu), hipMemcpyDeviceToHost, c.stream));
    WDG_HIP(hipStreamSynchronize(c.stream));
    used.release();
    EpochMap map{};
    int kept = 0;
    for (int e = 0; e < c.matEpochs; e++)
      if (hu[e >> 5] >> (e & 31) & 1u) {
        if (kept == 255) throw Error(WDG_E_CAPACITY, "material tables: more than 255 epochs referred to by contacts");
        map.m[e] = (uint8_t)kept++;
        // an older table extended by the materials added since (no contact of it refers to them)
        for (int m = 0; m < n; m++) tab.push_back(m < nOld ? c.epTab[(size_t)e * nOld + m] : mats[m]);
      }
    hipLaunchKernelGGL(k_epoch_map, dim3(nb), dim3(256), 0, c.stream, ng, (const uint32_t*)R.gcnt.p, R.hE.p, map);
    WDG_LAUNCH_CHECK();
    epochs = kept + 1;
  }
  tab.insert(tab.end(), mats, mats + n);
  // kernels still queued on the engine's stream may read the table
  WDG_HIP(hipStreamSynchronize(c.stream));
  c.mats.assign(mats, mats + n);
  c.epTab = tab;
  const wdg_material* before = c.dMats.p;
  c.dMats.ensure(std::max<size_t>(tab.size(), 1));
  if (!tab.empty()) WDG_HIP(hipMemcpy(c.dMats.p, tab.data(), tab.size() * sizeof(wdg_material), hipMemcpyHostToDevice));
  // a captured step graph holds the table's address and the epoch arguments
  if (c.dMats.p != before || epochs != c.matEpochs) c.gen++;
  c.matEpochs = epochs;
  c.matGen++;
}

void eidx_clear(Ctx& c) {
  Contacts& C = c.con();
  if (C.n == 0) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(nblocks(2 * C.n, 256)), dim3(256), 0, c.stream, 2 * C.n, C.eidx.p, NO_STATE);
  WDG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c.stream, (int64_t)1, c.dflags.p + EIDX_FULL, 1u);
  WDG_LAUNCH_CHECK();
}

template <bool E>
static void launch_cforce(Ctx& c, const ContactArgs& a, dim3 grid, bool f32, bool hertz, bool roll, bool per,
                          bool fac) {
  const dim3 block(CF_BS);
  if (f32) {
    if (hertz) {
      if (per) hipLaunchKernelGGL((k_cforce32<true, WDG_LAW_HERTZDMT, false, E>), grid, block, 0, c.stream, a);
      else if (fac) hipLaunchKernelGGL((k_cforce32<false, WDG_LAW_HERTZDMT, true, E>), grid, block, 0, c.stream, a);
      else hipLaunchKernelGGL((k_cforce32<false, WDG_LAW_HERTZDMT, false, E>), grid, block, 0, c.stream, a);
    } else {
      if (per) hipLaunchKernelGGL((k_cforce32<true, WDG_LAW_IDEALELPL, false, E>), grid, block, 0, c.stream, a);
      else if (fac) hipLaunchKernelGGL((k_cforce32<false, WDG_LAW_IDEALELPL, true, E>), grid, block, 0, c.stream, a);
      else hipLaunchKernelGGL((k_cforce32<false, WDG_LAW_IDEALELPL, false, E>), grid, block, 0, c.stream, a);
    }
  } else {
    if (hertz) {
      if (per) hipLaunchKernelGGL((k_cforce<true, WDG_LAW_HERTZDMT, false, E>), grid, block, 0, c.stream, a);
      else if (fac) hipLaunchKernelGGL((k_cforce<false, WDG_LAW_HERTZDMT, true, E>), grid, block, 0, c.stream, a);
      else hipLaunchKernelGGL((k_cforce<false, WDG_LAW_HERTZDMT, false, E>), grid, block, 0, c.stream, a);
    } else if (roll) {
      if (per) hipLaunchKernelGGL((k_cforce<true, LAW_IEP_ROLL, false, E>), grid, block, 0, c.stream, a);
      else if (fac) hipLaunchKernelGGL((k_cforce<false, LAW_IEP_ROLL, true, E>), grid, block, 0, c.stream, a);
      else hipLaunchKernelGGL((k_cforce<false, LAW_IEP_ROLL, false, E>), grid, block, 0, c.stream, a);
    } else {
      if (per) hipLaunchKernelGGL((k_cforce<true, WDG_LAW_IDEALELPL, false, E>), grid, block, 0, c.stream, a);
      else if (fac) hipLaunchKernelGGL((k_cforce<false, WDG_LAW_IDEALELPL, true, E>), grid, block, 0, c.stream, a);
      else hipLaunchKernelGGL((k_cforce<false, WDG_LAW_IDEALELPL, false, E>), grid, block, 0, c.stream, a);
    }
  }
  WDG_LAUNCH_CHECK();
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
  const bool roll = rolling(*p);
  if (roll && c.precision == 32)
    throw Error(WDG_E_ARG, "rolling resistance (IdealElPl.relRollStiff, rollTanPhi) needs the FP64 contact form");
  if (c.forcesPending) flush_forces(c);
  const double viscAlpha = (p->en > 0 && p->en < 1.)
                               ? -std::sqrt(5) * std::log(p->en) /
                                     (std::sqrt(std::log(p->en) * std::log(p->en) + M_PI * M_PI))
                               : 0.;
  c.stepNow = p->step;
  c.lastPhysKind = p->physKind;
  c.lastPoisson = p->poisson;
  c.lastEn = p->en;
  c.lastViscAlpha = viscAlpha;
  if (C.n == 0 || c.n == 0) {
    // no contact: this step's elast is zero (EnergyTracker resets it every step)
    if (c.trackEnergy) energy_zero(c, WDG_EN_ELAST);
    c.forcesPending = false;
    c.rollActive = false;
    c.clStressOn = false;  // no contact: the stress is zero
    return;
  }
  const int from = c.rcur, to = 1 - c.rcur;
  c.R[to].ensure(C.n);
  if (roll) c.R[to].ensureRoll(C.n);
  const int64_t ng = ngroups(C.n);
  c.act.ensure(ng * GROUP);
  c.desc.ensure(ng * GROUP);
  if (c.field.periodic) c.descDist.ensure(ng * GROUP);
  ContactArgs a = contact_args(c, from, to);
  // per-entry real slots for the gather (slot indices must stay below EIDX_PA); the filter
  // writes the entries of slots whose real state changes (all real ones after a reset), so
  // the others must already hold their value
  if (ng * GROUP < (int64_t)EIDX_PA) {
    a.eidx = C.eidx.p;
    a.entPos = C.entPos.p;
    if (!c.eidxValid && !c.eidxClean) eidx_clear(c);
  }
  a.dt = p->dt;
  a.renorm = (p->trsfRenorm > 0 && (p->step % p->trsfRenorm) == 0) ? 1 : 0;
  a.poisson = p->poisson;
  a.en = p->en;
  a.viscAlpha = viscAlpha;
  a.noSlip = p->noSlip; a.noBreak = p->noBreak; a.noFrict = p->noFrict; a.noAttraction = p->noAttraction;
  if (roll) {
    a.relTwistStiff = p->relTwistStiff;
    a.rollTanPhi = p->rollTanPhi;
    a.gtq = c.R[to].gtq.p;
  }
  const bool per = c.field.periodic != 0;
  const bool fac = c.nFacet > 0;
  const bool hertz = p->lawKind == WDG_LAW_HERTZDMT;
  const dim3 gfil((unsigned)ng), bfil(GROUP_BS);
  if (c.precision == 32) {
    if (hertz) {
      if (per) hipLaunchKernelGGL((k_cfilter32<true, WDG_LAW_HERTZDMT>), gfil, bfil, 0, c.stream, a);
      else hipLaunchKernelGGL((k_cfilter32<false, WDG_LAW_HERTZDMT>), gfil, bfil, 0, c.stream, a);
    } else {
      if (per) hipLaunchKernelGGL((k_cfilter32<true, WDG_LAW_IDEALELPL>), gfil, bfil, 0, c.stream, a);
      else hipLaunchKernelGGL((k_cfilter32<false, WDG_LAW_IDEALELPL>), gfil, bfil, 0, c.stream, a);
    }
  } else {
    if (hertz) {
      if (per) hipLaunchKernelGGL((k_cfilter<true, WDG_LAW_HERTZDMT>), gfil, bfil, 0, c.stream, a);
      else hipLaunchKernelGGL((k_cfilter<false, WDG_LAW_HERTZDMT>), gfil, bfil, 0, c.stream, a);
    } else {
      if (per) hipLaunchKernelGGL((k_cfilter<true, WDG_LAW_IDEALELPL>), gfil, bfil, 0, c.stream, a);
      else hipLaunchKernelGGL((k_cfilter<false, WDG_LAW_IDEALELPL>), gfil, bfil, 0, c.stream, a);
    }
  }
  WDG_LAUNCH_CHECK();
  const dim3 gforce((unsigned)(ng * a.cfBlocks));
  if (energy_per_scene(c)) {
    // per contact, summed per scene once the new real set is current (below)
    energy_scenes_prepare(c, C.n);
    a.enC = c.enC.p;
    launch_cforce<true>(c, a, gforce, c.precision == 32, hertz, roll, per, fac);
  } else if (c.trackEnergy) {
    c.enPart.ensure((size_t)4 * gforce.x);
    a.enPart = c.enPart.p;
    launch_cforce<true>(c, a, gforce, c.precision == 32, hertz, roll, per, fac);
    // EnergyTracker: elast is resettable (this step's sum replaces it), the dissipations accumulate
    static const int tgt[4] = {WDG_EN_ELAST, WDG_EN_PLAST, WDG_EN_VISCN, WDG_EN_VISCT};
    energy_add(c, gforce.x, 4, tgt, 1);
  } else {
    launch_cforce<false>(c, a, gforce, c.precision == 32, hertz, roll, per, fac);
  }
  // ContactLoop.evalStress (ContactLoop.cpp:123): the stress of this loop's real contacts, in
  // periodic cells (gradV is zero, so there is no gradV work to add)
  c.clStressOn = c.evalStress && c.field.periodic;
  if (c.clStressOn) {
    if (c.nFacet > 0) throw Error(WDG_E_ARG, "ContactLoop.evalStress with facets (multi-nodal particles)");
    c.clPart.ensure((size_t)9 * ng);
    c.clStress.ensure(9);
    hipLaunchKernelGGL(k_cl_stress<true>, dim3((unsigned)ng), dim3(256), 0, c.stream, a, ng, c.clPart.p);
    WDG_LAUNCH_CHECK();
    const double vol = c.field.cell[0] * c.field.cell[1] * c.field.cell[2];
    hipLaunchKernelGGL(k_cl_stress_sum, dim3(1), dim3(256), 0, c.stream, ng, (const double*)c.clPart.p, vol,
                       c.graphCapture ? (const uint32_t*)c.gstate.p : nullptr, c.clStress.p);
    WDG_LAUNCH_CHECK();
  }
  c.R[to].sidxValid = true;
  c.eidxValid = a.eidx != nullptr;
  c.eidxClean = false;
  c.rcur = to;
  c.rollActive = roll;
  if (energy_per_scene(c)) energy_scenes_contacts(c);
  c.forcesPending = p->applyForces != 0;
  // ContactLoop::run's epilogue (ContactLoop.cpp:265-266): 'once' updates one step only
  if (c.updatePhys == 2 && !c.graphCapture) c.updatePhys = 0;
}

// ------------------------------------------------------------------ real set from per-slot flags
__global__ __launch_bounds__(GROUP_BS) void k_flag_bits(int64_t nc, const uint32_t* __restrict__ flag,
                                                        unsigned long long* bits, uint32_t* tpre, uint32_t* gcnt) {
  const int64_t base = (int64_t)blockIdx.x * GROUP;
  uint32_t running = 0;
  for (int pass = 0; pass < GROUP / GROUP_BS; pass++) {
    const int64_t c = base + pass * GROUP_BS + threadIdx.x;
    group_pass(c, c < nc && flag[c] != 0u, base, running, bits, tpre, nullptr, nullptr);
  }
  if (threadIdx.x == 0) gcnt[blockIdx.x] = running;
}

void reals_from_flags(Ctx& c, const uint32_t* flag, int64_t nslots, int dst) {
  Reals& N = c.R[dst];
  N.ensure(nslots);
  N.sidxValid = false;
  c.eidxValid = false;
  c.eidxClean = false;
  const int64_t ng = ngroups(nslots);
  if (ng > 0) {
    hipLaunchKernelGGL(k_flag_bits, dim3((unsigned)ng), dim3(GROUP_BS), 0, c.stream, nslots, flag, N.bits.p,
                       N.tpre.p, N.gcnt.p);
    WDG_LAUNCH_CHECK();
  }
}

}  // namespace wdg
