// CPU FP64 oracle of Woo's DEM step — TEST INFRASTRUCTURE ONLY.
//
// This file restates the reference algorithm (woodem/woo, see oracle.h for the
// full list of restated files) in plain C++17 so that the HIP engine can be
// checked against it and so that a CPU baseline can be timed beside it. It is
// never linked into the product library; the product must fail loudly when
// its HIP extension is missing instead of falling back to this code.
//
// Arithmetic conventions (shared with the engine it checks, by specification,
// not by shared code):
//  * every 3-sum is evaluated left to right: (a0+a1)+a2;
//  * cross(a,b) = (a1*b2-a2*b1, a2*b0-a0*b2, a0*b1-a1*b0);
//  * normalize divides each component by sqrt(squaredNorm) when that is >0;
//  * built with -ffp-contract=off: no fused multiply-add anywhere.
// The reference evaluates the same expressions through Eigen 3.3; whether
// Eigen's vectorised reductions use the same association is not pinned (the
// reference cannot be built here), so bit-level agreement with the real Woo
// binary is "parity unpinned"; the physics is pinned by closed-form tests.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;
const double Inf = std::numeric_limits<double>::infinity();
const double NaN = std::numeric_limits<double>::quiet_NaN();

// ---------------------------------------------------------------- vector math
struct V3 {
  double v[3];
  double& operator[](int i) { return v[i]; }
  double operator[](int i) const { return v[i]; }
};
inline V3 mk(double a, double b, double c) { return V3{{a, b, c}}; }
inline V3 zero3() { return mk(0., 0., 0.); }
inline V3 add(const V3& a, const V3& b) { return mk(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
inline V3 sub(const V3& a, const V3& b) { return mk(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline V3 mul(const V3& a, double s) { return mk(a[0] * s, a[1] * s, a[2] * s); }
inline V3 smul(double s, const V3& a) { return mk(s * a[0], s * a[1], s * a[2]); }
inline V3 divs(const V3& a, double s) { return mk(a[0] / s, a[1] / s, a[2] / s); }
inline double dot(const V3& a, const V3& b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
inline double sqn(const V3& a) { return dot(a, a); }
inline double nrm(const V3& a) { return std::sqrt(sqn(a)); }
inline V3 cross(const V3& a, const V3& b) {
  return mk(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
inline void normalize(V3& a) {
  double z = sqn(a);
  if (z > 0.) {
    double s = std::sqrt(z);
    a[0] = a[0] / s; a[1] = a[1] / s; a[2] = a[2] / s;
  }
}
inline V3 ld3(const double* p) { return mk(p[0], p[1], p[2]); }
inline void st3(double* p, const V3& a) { p[0] = a[0]; p[1] = a[1]; p[2] = a[2]; }

// quaternion stored (w,x,y,z)
struct Q4 { double w, x, y, z; };
// Quaternion from a rotation matrix given by its columns (Shoemake 1987,
// "Quaternion Calculus and Fast Animation": trace branch, else largest
// diagonal element). m(r,c) = col[c][r].
inline Q4 quatFromCols(const V3& c0, const V3& c1, const V3& c2) {
  const V3* col[3] = {&c0, &c1, &c2};
  auto m = [&](int r, int c) { return (*col[c])[r]; };
  Q4 q;
  double t = (m(0, 0) + m(1, 1)) + m(2, 2);
  if (t > 0.) {
    t = std::sqrt(t + 1.0);
    q.w = 0.5 * t;
    t = 0.5 / t;
    q.x = (m(2, 1) - m(1, 2)) * t;
    q.y = (m(0, 2) - m(2, 0)) * t;
    q.z = (m(1, 0) - m(0, 1)) * t;
  } else {
    int i = 0;
    if (m(1, 1) > m(0, 0)) i = 1;
    if (m(2, 2) > m(i, i)) i = 2;
    int j = (i + 1) % 3, k = (j + 1) % 3;
    t = std::sqrt(((m(i, i) - m(j, j)) - m(k, k)) + 1.0);
    double v[3];
    v[i] = 0.5 * t;
    t = 0.5 / t;
    q.w = (m(k, j) - m(j, k)) * t;
    v[j] = (m(j, i) + m(i, j)) * t;
    v[k] = (m(k, i) + m(i, k)) * t;
    q.x = v[0]; q.y = v[1]; q.z = v[2];
  }
  return q;
}
// q*v = v + 2w(q×v) + q×(2 q×v)
inline V3 quatRot(const Q4& q, const V3& v) {
  V3 qv = mk(q.x, q.y, q.z);
  V3 uv = cross(qv, v);
  uv = add(uv, uv);
  V3 r = add(v, smul(q.w, uv));
  return add(r, cross(qv, uv));
}
inline Q4 quatMul(const Q4& a, const Q4& b) {
  Q4 r;
  r.w = ((a.w * b.w - a.x * b.x) - a.y * b.y) - a.z * b.z;
  r.x = ((a.w * b.x + a.x * b.w) + a.y * b.z) - a.z * b.y;
  r.y = ((a.w * b.y + a.y * b.w) + a.z * b.x) - a.x * b.z;
  r.z = ((a.w * b.z + a.z * b.w) + a.x * b.y) - a.y * b.x;
  return r;
}
inline double signum(double f) { return f < 0 ? -1. : (f > 0 ? 1. : 0.); }
inline double pow2(double x) { return x * x; }
// Math.hpp:166-175 pown, square-and-multiply
inline double pown(double x, unsigned n) {
  unsigned d = n >> 1, r = n & 1;
  double x2d = d == 0 ? 1. : pown(x * x, d);
  double xr = r == 0 ? 1. : x;
  return x2d * xr;
}

// ------------------------------------------------------------------ entities
struct Material { int kind; double density, young, ktDivKn, tanPhi, surfEnergy, alpha; };

enum : uint32_t {
  DOF_X = 1, DOF_Y = 2, DOF_Z = 4, DOF_RX = 8, DOF_RY = 16, DOF_RZ = 32, DOF_ALL = 63,
  ENERGY_SKIP = 256, GRAVITY_SKIP = 512, DAMPING_SKIP = 2048
};

struct Particle {
  int shape = ORC_SPHERE;
  double radius = NaN;
  int axis = 0, sense = 0;
  int mat = 0;
  uint32_t mask = 0;
  // the single node (spheres and walls are uninodal)
  double pos[3] = {0, 0, 0}, vel[3] = {0, 0, 0}, angVel[3] = {0, 0, 0}, ori[4] = {1, 0, 0, 0};
  double force[3] = {0, 0, 0}, torque[3] = {0, 0, 0};
  double mass = 0, inertia[3] = {0, 0, 0};
  uint32_t flags = 0;
  bool integrated = false;  // member of DemField.nodes
  // Facet: static vertices (node 0 is pos), Facet.fakeVel, Facet.halfThick
  double fv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double fakeVel[3] = {0, 0, 0};
  double halfThick = 0;
  // Facet.n21lim: per edge, the facet normal's dot product with the edge-neighbour's (NaN: none)
  double n21lim[3] = {NaN, NaN, NaN};
  // force and torque on the facet's 3 nodes from the last IntraForce (In2_Facet)
  double nodeF[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, nodeT[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  // Aabb
  bool hasBound = false;
  double bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
  double lastPos[3] = {0, 0, 0};
  double maxD2 = NaN;
  // Particle::contacts: other id -> index into the contact vector
  std::map<int, int> contacts;
  // energy terms of this node in the last leapfrog: kinetic, grav, nonviscDamp
  double en[3] = {0, 0, 0};
};

struct Contact {
  int pA = -1, pB = -1;
  int cellDist[3] = {0, 0, 0};
  int64_t stepCreated = -1;
  double minDist00Sq = -1;
  bool geom = false, phys = false;
  // L6Geom
  V3 trsf[3];            // columns
  V3 nodePos;            // contact point
  Q4 nodeOri;
  double uN = NaN;
  double lens[2] = {0, 0};
  double contA = NaN;
  V3 vel = {{0, 0, 0}}, angVel = {{0, 0, 0}};
  // FrictPhys / HertzPhys
  V3 force = {{0, 0, 0}}, torque = {{0, 0, 0}};
  double kn = NaN, kt = NaN, tanPhi = NaN;
  double kt0 = 0, alpha_sqrtMK = 0, R = 0, K = 0, gamma = 0, alpha = 0;
  // energy terms of this contact in the last contact loop: elast, plast, viscN, viscT
  double en[4] = {0, 0, 0, 0};
  bool isReal() const { return geom && phys; }
};

inline uint64_t pairKey(int a, int b) {
  int lo = std::min(a, b), hi = std::max(a, b);
  return (uint64_t(uint32_t(lo)) << 32) | uint32_t(hi);
}

struct Bound {
  double coord;
  int id;
  bool isMin, hasBB;
  bool operator<(const Bound& b) const {
    if (id == b.id && coord == b.coord) return isMin;
    return coord < b.coord;
  }
  bool operator>(const Bound& b) const {
    if (id == b.id && coord == b.coord) return !isMin;
    return coord > b.coord;
  }
};

}  // namespace

struct orc_scene {
  orc_params prm;
  std::vector<Material> mats;
  std::vector<Particle> par;
  std::vector<Contact> con;  // ContactContainer::linView
  std::vector<std::pair<int, int>> pending;  // requestRemoval'd (pA,pB)
  int64_t step = 0;
  double time = 0;
  double nextDt = NaN;
  // collider state
  double verletDist = NaN;
  std::vector<Bound> BB[3];
  bool boundsInit = false;
  int nRefresh = 0;
  // DynDt PeriodicEngine bookkeeping
  int64_t dynStepLast = -1, dynStepPrev = -1;
  std::vector<std::mutex> locks;
  // Scene.energy, ORC_E_* order
  double energy[ORC_E_N] = {0, 0, 0, 0, 0, 0, 0};
  double clStress[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // ContactLoop.stress of the last contact loop

  // ---------------------------------------------------------- contacts
  int find(int a, int b) const {
    if (a < 0 || b < 0 || a >= (int)par.size()) return -1;
    auto it = par[a].contacts.find(b);
    return it == par[a].contacts.end() ? -1 : it->second;
  }
  // ContactContainer::addMaybe_fast
  void addMaybe(const Contact& c) {
    if (par[c.pA].contacts.count(c.pB)) return;
    con.push_back(c);
    int ix = (int)con.size() - 1;
    par[c.pA].contacts[c.pB] = ix;
    par[c.pB].contacts[c.pA] = ix;
  }
  // ContactContainer::removeMaybe_fast + linView_remove
  void removeAt(int ix) {
    Contact& c = con[ix];
    par[c.pA].contacts.erase(c.pB);
    par[c.pB].contacts.erase(c.pA);
    int last = (int)con.size() - 1;
    if (ix < last) {
      con[ix] = con[last];
      par[con[ix].pA].contacts[con[ix].pB] = ix;
      par[con[ix].pB].contacts[con[ix].pA] = ix;
    }
    con.pop_back();
  }
  void removePair(int a, int b) {
    int ix = find(a, b);
    if (ix >= 0) removeAt(ix);
  }

  // ---------------------------------------------------------- helpers
  V3 shiftPos(const int cd[3]) const {
    if (!prm.periodic) return zero3();
    return mk(prm.cell[0] * cd[0], prm.cell[1] * cd[1], prm.cell[2] * cd[2]);
  }
  bool isFresh(const Contact& c) const { return c.stepCreated == step; }

  // Collision.cpp:21-43 mayCollide (uninodal, no clumps)
  bool mayCollide(int a, int b) const {
    if (a == b) return false;
    const Particle &A = par[a], &B = par[b];
    if (!(A.mask & B.mask)) return false;
    if ((A.mask & B.mask & prm.loneMask) != 0) return false;
    return true;
  }
  // canonical roles: functor order for mixed types (Wall or Facet before Sphere), min id first otherwise
  void canonicalRoles(int a, int b, int& pA, int& pB) const {
    if (par[a].shape != par[b].shape) {
      if (par[a].shape != ORC_SPHERE) { pA = a; pB = b; }
      else { pA = b; pB = a; }
    } else {
      pA = std::min(a, b); pB = std::max(a, b);
    }
  }
  // nearest periodic image of B relative to A (from positions at the refresh)
  void nearestImage(int pA, int pB, int cd[3]) const {
    for (int k = 0; k < 3; k++) {
      if (!prm.periodic) { cd[k] = 0; continue; }
      double d = (par[pB].lastPos[k] - par[pA].lastPos[k]) / prm.cell[k];
      cd[k] = -(int)std::floor(d + 0.5);
    }
  }
  // AABB overlap for roles (pA,pB) with B shifted by cd; InsertionSortCollider.cpp:67-73
  bool aabbOverlapShift(int a, int b, const int cd[3]) const {
    const Particle &A = par[a], &B = par[b];
    for (int k = 0; k < 3; k++) {
      double s = prm.periodic ? prm.cell[k] * cd[k] : 0.;
      if (!(A.bmin[k] <= B.bmax[k] + s && A.bmax[k] >= B.bmin[k] + s)) return false;
    }
    return true;
  }
  // number of overlapping periodic images along one axis among nearest and its two neighbours
  int imagesOverlapping(int a, int b, const int cd[3]) const {
    int prod = 1;
    for (int k = 0; k < 3; k++) {
      int cnt = 0;
      for (int dd = -1; dd <= 1; dd++) {
        double s = prm.cell[k] * (cd[k] + dd);
        if (par[a].bmin[k] <= par[b].bmax[k] + s && par[a].bmax[k] >= par[b].bmin[k] + s) cnt++;
      }
      prod *= cnt;
    }
    return prod;
  }
  bool contactAabbOverlap(const Contact& c) const {
    if (!par[c.pA].hasBound || !par[c.pB].hasBound) return false;
    return aabbOverlapShift(c.pA, c.pB, c.cellDist);
  }

  // ---------------------------------------------------------- collider
  // Collision.cpp:164-184
  void setVerletDist() {
    if (!(verletDist != verletDist) && verletDist >= 0) return;
    double vd = prm.verletDist;
    if (vd >= 0) { verletDist = vd; return; }
    double minR = Inf;
    for (const auto& p : par)
      if (p.shape == ORC_SPHERE) minR = std::min(p.radius, minR);
    verletDist = std::isinf(minR) ? 0. : std::fabs(vd) * minR;
  }
  // Collision.cpp:80-123 (spheres and walls: maxRot<0, uninodal)
  bool aabbIsDirty(const Particle& p) const {
    if (!p.hasBound) return true;
    double mx = std::max(std::max(p.bmin[0], p.bmin[1]), p.bmin[2]);
    double Mx = std::max(std::max(p.bmax[0], p.bmax[1]), p.bmax[2]);
    if (std::isnan(mx) || std::isnan(Mx)) return true;
    double d2 = 0;
    d2 = std::max(d2, sqn(sub(ld3(p.lastPos), ld3(p.pos))));
    return d2 + 0. > p.maxD2;
  }
  // Bo1_Sphere_Aabb (Sphere.cpp:52-82), Bo1_Wall_Aabb (Wall.cpp:25-38), Bo1_Facet_Aabb (Facet.cpp:275-298,
  // aperiodic), updateAabb (Collision.cpp:125-162)
  void updateAabb(Particle& p) {
    if (p.shape == ORC_SPHERE) {
      double hs = (prm.distFactor > 0 ? prm.distFactor : 1.) * p.radius;
      for (int k = 0; k < 3; k++) { p.bmin[k] = p.pos[k] - hs; p.bmax[k] = p.pos[k] + hs; }
    } else if (p.shape == ORC_FACET) {
      for (int k = 0; k < 3; k++) {
        p.bmin[k] = p.fv[k] - p.halfThick;
        p.bmax[k] = p.fv[k] + p.halfThick;
      }
      for (int i = 1; i < 3; i++)
        for (int k = 0; k < 3; k++) {
          p.bmin[k] = std::min(p.bmin[k], p.fv[3 * i + k] - p.halfThick);
          p.bmax[k] = std::max(p.bmax[k], p.fv[3 * i + k] + p.halfThick);
        }
    } else {
      for (int k = 0; k < 3; k++) { p.bmin[k] = -Inf; p.bmax[k] = Inf; }
      p.bmin[p.axis] = p.bmax[p.axis] = p.pos[p.axis];
    }
    p.hasBound = true;
    for (int k = 0; k < 3; k++) p.lastPos[k] = p.pos[k];
    p.maxD2 = pow2(verletDist);
    if (verletDist > 0) {
      for (int k = 0; k < 3; k++) { p.bmax[k] = p.bmax[k] + verletDist; p.bmin[k] = p.bmin[k] - verletDist; }
    }
  }
  Contact newContact(int a, int b, const int cd[3]) const {
    Contact c;
    c.pA = a; c.pB = b;
    for (int k = 0; k < 3; k++) c.cellDist[k] = cd[k];
    c.stepCreated = step;
    return c;
  }
  // InsertionSortCollider::handleBoundInversion (aperiodic), with "later" lists
  void handleInversion(int id1, int id2, bool separating, std::vector<Contact>& mk_, std::vector<std::pair<int, int>>& rm) {
    const int zeroCd[3] = {0, 0, 0};
    bool overlap = separating ? false : aabbOverlapShift(id1, id2, zeroCd);
    int ix = find(id1, id2);
    bool hasCon = ix >= 0;
    if (!overlap && !hasCon) return;
    if (overlap && hasCon) return;
    if (overlap) {
      if (!mayCollide(id1, id2)) return;
      int a, b;
      canonicalRoles(id1, id2, a, b);
      mk_.push_back(newContact(a, b, zeroCd));
      return;
    }
    if (!con[ix].isReal()) rm.push_back({con[ix].pA, con[ix].pB});
  }
  // InsertionSortCollider::insertionSort_part
  void insertionSortPart(std::vector<Bound>& v, long iBegin, long iEnd, long iStart,
                         std::vector<Contact>& mk_, std::vector<std::pair<int, int>>& rm) {
    const bool earlyStop = (iBegin != iStart);
    for (long i = std::max(iStart, iBegin + 1); i < iEnd; i++) {
      if (!(v[i - 1] > v[i])) {
        if (earlyStop) return;
        continue;
      }
      const Bound vi = v[i];
      long j = i - 1;
      while (j >= iBegin && v[j] > vi) {
        v[j + 1] = v[j];
        if (vi.isMin != v[j].isMin && vi.hasBB && v[j].hasBB && vi.id != v[j].id)
          handleInversion(vi.id, v[j].id, !vi.isMin, mk_, rm);
        j--;
      }
      v[j + 1] = vi;
    }
  }
  // InsertionSortCollider::insertionSort, chunked OpenMP version
  void insertionSort(std::vector<Bound>& v, std::vector<std::vector<Contact>>& mkT,
                     std::vector<std::vector<std::pair<int, int>>>& rmT) {
    long size = (long)v.size();
    if (size == 0) return;
    int TH = 1;
#ifdef _OPENMP
    TH = omp_get_max_threads();
#endif
    long chunks = TH;
    if (prm.deterministic || TH == 1) chunks = 1;
    long minPerCore = 2 * 1000;
    if (chunks > 1 && size / chunks < minPerCore) chunks = TH * (long)std::ceil(std::max(1L, size / minPerCore) * 1. / TH);
    long chunkSize = chunks ? size / chunks : size;
    if (chunks <= 1 || chunkSize < 100) { insertionSortPart(v, 0, size, 0, mkT[0], rmT[0]); return; }
    std::vector<long> s0(chunks + 1), s1(chunks);
    for (long i = 0; i < chunks; i++) { s0[i] = i * chunkSize; s1[i] = i * chunkSize + chunkSize / 2; }
    s0[chunks] = size;
    bool ok = false;
    for (long pass = 0; !ok; pass++) {
      if (pass == 100L * TH) throw std::runtime_error("InsertionSortCollider: sort not done after too many partial sort passes");
      bool even = (pass % 2) == 0;
      const std::vector<long>& s = even ? s0 : s1;
      long nch = even ? chunks : chunks - 1;
#pragma omp parallel for schedule(static)
      for (long ch = 0; ch < nch; ch++) {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        long start = pass == 0 ? s[ch] : (even ? s1[ch] : s0[ch + 1]);
        insertionSortPart(v, s[ch], s[ch + 1], start, mkT[t], rmT[t]);
      }
      ok = true;
      for (long ch = (even ? 1 : 0); ch < chunks; ch++) {
        if (v[s[ch] - 1] > v[s[ch]]) { ok = false; break; }
      }
    }
  }
  // InsertionSortCollider::shouldBeRemoved + ContactContainer::removePending
  void removePending() {
    for (const auto& pr : pending) {
      int ix = find(pr.first, pr.second);
      if (ix < 0) continue;
      if (con[ix].isReal()) continue;  // became real again in the meantime: not possible without a collider run
      if (!contactAabbOverlap(con[ix])) removeAt(ix);
    }
    pending.clear();
  }
  // enumerate enlarged-AABB-overlapping pairs (canonical roles, with cellDist), cell grid for spheres
  void overlapPairs(std::vector<Contact>& out) const {
    int n = (int)par.size();
    std::vector<int> spheres, walls;
    for (int i = 0; i < n; i++) (par[i].shape == ORC_SPHERE ? spheres : walls).push_back(i);
    // walls and facets against every sphere
    for (int w : walls)
      for (int s : spheres) {
        if (!mayCollide(w, s)) continue;
        int a, b, cd[3];
        canonicalRoles(w, s, a, b);
        nearestImage(a, b, cd);
        if (aabbOverlapShift(a, b, cd)) out.push_back(newContact(a, b, cd));
      }
    if (spheres.empty()) return;
    // grid with cell >= largest AABB extent
    double ext = 0, lo[3] = {Inf, Inf, Inf}, hi[3] = {-Inf, -Inf, -Inf};
    for (int s : spheres)
      for (int k = 0; k < 3; k++) {
        ext = std::max(ext, par[s].bmax[k] - par[s].bmin[k]);
        lo[k] = std::min(lo[k], par[s].lastPos[k]);
        hi[k] = std::max(hi[k], par[s].lastPos[k]);
      }
    int nc[3];
    double h[3], org[3];
    bool brute = spheres.size() < 64;
    for (int k = 0; k < 3; k++) {
      if (prm.periodic) {
        nc[k] = (int)std::floor(prm.cell[k] / ext);
        if (nc[k] < 3) brute = true;
        nc[k] = std::max(nc[k], 1);
        h[k] = prm.cell[k] / nc[k];
        org[k] = 0;
      } else {
        double span = hi[k] - lo[k];
        nc[k] = std::max(1, std::min(1 << 10, (int)std::floor(span / ext) + 1));
        h[k] = std::max(ext, span / nc[k] * (1 + 1e-12));
        nc[k] = (int)std::floor(span / h[k]) + 1;
        org[k] = lo[k];
      }
    }
    auto test = [&](int i, int j) {
      if (!mayCollide(i, j)) return;
      int a, b, cd[3];
      canonicalRoles(i, j, a, b);
      nearestImage(a, b, cd);
      if (!aabbOverlapShift(a, b, cd)) return;
      if (prm.periodic && imagesOverlapping(a, b, cd) > 1)
        throw std::runtime_error("InsertionSortCollider: particles ##" + std::to_string(a) + "+" + std::to_string(b) +
                                 " overlap through two periodic images (particle larger than half of the cell?)");
      out.push_back(newContact(a, b, cd));
    };
    if (brute || (double)nc[0] * nc[1] * nc[2] > 64.0 * spheres.size() + 1e6) {
      for (size_t x = 0; x < spheres.size(); x++)
        for (size_t y = x + 1; y < spheres.size(); y++) test(spheres[x], spheres[y]);
      return;
    }
    auto cellOf = [&](int s, int c[3]) {
      for (int k = 0; k < 3; k++) {
        double x = par[s].lastPos[k] - org[k];
        if (prm.periodic) x = x - std::floor(x / prm.cell[k]) * prm.cell[k];
        int ci = (int)std::floor(x / h[k]);
        c[k] = std::min(std::max(ci, 0), nc[k] - 1);
      }
    };
    long ncell = (long)nc[0] * nc[1] * nc[2];
    std::vector<std::vector<int>> cells(ncell);
    for (int s : spheres) {
      int c[3];
      cellOf(s, c);
      cells[((long)c[2] * nc[1] + c[1]) * nc[0] + c[0]].push_back(s);
    }
    for (int s : spheres) {
      int c[3];
      cellOf(s, c);
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++) {
            int q[3] = {c[0] + dx, c[1] + dy, c[2] + dz};
            bool skip = false;
            for (int k = 0; k < 3; k++) {
              if (q[k] < 0 || q[k] >= nc[k]) {
                if (!prm.periodic) { skip = true; break; }
                q[k] = (q[k] + nc[k]) % nc[k];
              }
            }
            if (skip) continue;
            for (int o : cells[((long)q[2] * nc[1] + q[1]) * nc[0] + q[0]])
              if (o > s) test(s, o);
          }
    }
  }
  // full rebuild at refresh: potential set := overlap set ∪ real contacts
  void fullRebuild() {
    std::vector<Contact> fresh;
    overlapPairs(fresh);
    std::unordered_map<uint64_t, int> want;
    want.reserve(fresh.size() * 2 + 16);
    for (size_t i = 0; i < fresh.size(); i++) want[pairKey(fresh[i].pA, fresh[i].pB)] = (int)i;
    // remove non-real contacts that lost overlap
    for (int ix = (int)con.size() - 1; ix >= 0; ix--) {
      if (ix >= (int)con.size()) continue;
      const Contact& c = con[ix];
      if (c.isReal()) continue;
      if (!want.count(pairKey(c.pA, c.pB))) removeAt(ix);
    }
    for (const Contact& c : fresh) addMaybe(c);
  }
  // InsertionSortCollider::run
  void colliderRun() {
    setVerletDist();
    bool fullRun = !boundsInit || BB[0].size() != 2 * par.size();
    bool recompute = verletDist == 0;
    if (!recompute)
      for (const auto& p : par)
        if (aabbIsDirty(p)) { recompute = true; break; }
    if (recompute) {
#pragma omp parallel for schedule(static)
      for (long i = 0; i < (long)par.size(); i++) updateAabb(par[i]);
      fullRun = true;
    }
    if (!fullRun) { removePending(); return; }
    nRefresh++;
    if (prm.periodic || prm.colliderMode == 1) {
      removePending();
      fullRebuild();
      boundsInit = true;
      for (int ax = 0; ax < 3; ax++) BB[ax].assign(2 * par.size(), Bound{0, 0, true, true});
      return;
    }
    long nPar = (long)par.size();
    bool doInitSort = false;
    if ((long)BB[0].size() != 2 * nPar) {
      for (int ax = 0; ax < 3; ax++) {
        BB[ax].clear();
        for (long id = 0; id < nPar; id++) {
          BB[ax].push_back(Bound{0, (int)id, true, true});
          BB[ax].push_back(Bound{0, (int)id, false, true});
        }
      }
      doInitSort = true;
    }
    for (int ax = 0; ax < 3; ax++) {
      for (auto& b : BB[ax]) {
        const Particle& p = par[b.id];
        b.hasBB = p.hasBound;
        b.coord = p.hasBound ? (b.isMin ? p.bmin[ax] : p.bmax[ax]) : p.pos[ax];
      }
    }
    removePending();
    int TH = 1;
#ifdef _OPENMP
    TH = omp_get_max_threads();
#endif
    std::vector<std::vector<Contact>> mkT(TH);
    std::vector<std::vector<std::pair<int, int>>> rmT(TH);
    if (!doInitSort) {
      for (int ax = 0; ax < 3; ax++) insertionSort(BB[ax], mkT, rmT);
    } else {
      for (int ax = 0; ax < 3; ax++) std::sort(BB[ax].begin(), BB[ax].end());
      const std::vector<Bound>& V = BB[0];
      for (long i = 0; i < 2 * nPar; i++) {
        if (!(V[i].isMin && V[i].hasBB)) continue;
        int iid = V[i].id;
        for (long j = i + 1; j < 2 * nPar && V[j].id != iid; j++) {
          if (!V[j].isMin) continue;
          handleInversion(iid, V[j].id, false, mkT[0], rmT[0]);
        }
      }
    }
    // makeRemoveContactLater_process
    for (auto& rm : rmT)
      for (auto& pr : rm) {
        int ix = find(pr.first, pr.second);
        if (ix >= 0) removeAt(ix);
      }
    for (auto& mk_ : mkT)
      for (auto& c : mk_) addMaybe(c);
    boundsInit = true;
  }

  // ---------------------------------------------------------- geometry
  // L6Geom::setInitialLocalCoords (L6Geom.cpp:25-36)
  static void initialLocalCoords(Contact& c, const V3& locX) {
    V3 ax = std::fabs(locX[1]) < std::fabs(locX[2]) ? mk(0, 1, 0) : mk(0, 0, 1);
    V3 locY = cross(locX, ax);
    locY = sub(locY, mul(locX, dot(locY, locX)));
    normalize(locY);
    V3 locZ = cross(locX, locY);
    c.trsf[0] = locX; c.trsf[1] = locY; c.trsf[2] = locZ;
  }
  // Cg2_Any_Any_L6Geom__Base::handleSpheresLikeContact (L6Geom.cpp:49-216), approxMask=0, noRatch=false
  void handleSpheresLike(Contact& c, const V3& pos1, const V3& vel1, const V3& angVel1, const V3& pos2,
                         const V3& vel2, const V3& angVel2, const V3& normal, const V3& contPt, double uN,
                         double r1, double r2) const {
    if (!c.geom) {
      c.geom = true;
      initialLocalCoords(c, normal);
      c.uN = uN;
      c.lens[0] = std::fabs(r1); c.lens[1] = std::fabs(r2);
      if (c.lens[0] < 0) c.lens[0] = c.lens[1];
      if (c.lens[1] < 0) c.lens[1] = c.lens[0];
      c.contA = M_PI * pow2((r1 > 0 && r2 > 0) ? std::min(r1, r2) : (r1 > 0 ? r1 : r2));
      c.nodePos = contPt;
      c.nodeOri = quatFromCols(c.trsf[0], c.trsf[1], c.trsf[2]);
      c.vel = zero3(); c.angVel = zero3();
      return;
    }
    const V3& currNormal = normal;
    const V3 prevNormal = c.trsf[0];
    const V3 prevContPt = c.nodePos;
    const double dt = prm.dt;
    // shiftVel2 is zero: gradV is zero in this engine
    V3 normRotVec = cross(prevNormal, currNormal);
    V3 midNormal = smul(.5, add(prevNormal, currNormal));
    if (nrm(midNormal) == 0) midNormal = currNormal;
    else normalize(midNormal);
    double twist = dot(midNormal, add(angVel1, angVel2));
    V3 normTwistVec = mul(mul(mul(midNormal, dt), .5), twist);
    const V3 prevY = c.trsf[1];
    V3 rt = add(normRotVec, normTwistVec);
    V3 midT[3], curT[3];
    midT[0] = midNormal;
    midT[1] = sub(prevY, divs(cross(prevY, rt), 2.));
    midT[2] = cross(midT[0], midT[1]);
    curT[0] = currNormal;
    curT[1] = sub(prevY, cross(midT[1], rt));
    curT[2] = cross(curT[0], curT[1]);
    if (prm.trsfRenorm > 0 && (step % prm.trsfRenorm) == 0) {
      normalize(curT[0]);
      curT[1] = sub(curT[1], mul(curT[0], dot(curT[1], curT[0])));
      normalize(curT[1]);
      curT[2] = cross(curT[0], curT[1]);
    }
    V3 midContPt = smul(.5, add(prevContPt, contPt));
    V3 midPos1 = sub(pos1, smul(dt / 2., vel1));
    V3 midPos2 = sub(pos2, smul(dt / 2., vel2));
    V3 c1x = sub(midContPt, midPos1);
    V3 c2x = sub(midContPt, midPos2);
    V3 relVel = sub(add(vel2, cross(angVel2, c2x)), add(vel1, cross(angVel1, c1x)));
    c.trsf[0] = curT[0]; c.trsf[1] = curT[1]; c.trsf[2] = curT[2];
    V3 dav = sub(angVel2, angVel1);
    c.vel = mk(dot(midT[0], relVel), dot(midT[1], relVel), dot(midT[2], relVel));
    c.angVel = mk(dot(midT[0], dav), dot(midT[1], dav), dot(midT[2], dav));
    c.uN = uN;
    c.nodePos = contPt;
    c.nodeOri = quatFromCols(c.trsf[0], c.trsf[1], c.trsf[2]);
  }
  // Cg2_Sphere_Sphere_L6Geom::go (Sphere.cpp:92-117)
  bool geomSphereSphere(Contact& c, const V3& shift2) const {
    const Particle &s1 = par[c.pA], &s2 = par[c.pB];
    double r1 = s1.radius, r2 = s2.radius;
    V3 p1 = ld3(s1.pos);
    V3 p2 = add(ld3(s2.pos), shift2);
    V3 relPos = sub(p2, p1);
    double unDistSq = sqn(relPos) - pow2(std::fabs(prm.distFactor) * (r1 + r2));
    if (unDistSq > 0 && !c.isReal()) return false;
    double dist = nrm(relPos);
    double uN = dist - (r1 + r2);
    V3 normal = divs(relPos, dist);
    V3 contPt = add(p1, smul(r1 + 0.5 * uN, normal));
    handleSpheresLike(c, p1, ld3(s1.vel), ld3(s1.angVel), p2, ld3(s2.vel), ld3(s2.angVel), normal, contPt, uN, r1, r2);
    return true;
  }
  // Cg2_Wall_Sphere_L6Geom::go (Wall.cpp:50-84)
  bool geomWallSphere(Contact& c, const V3& shift2) const {
    const Particle &w = par[c.pA], &s = par[c.pB];
    double radius = s.radius;
    int ax = w.axis, sense = w.sense;
    V3 wallPos = ld3(w.pos);
    V3 spherePos = add(ld3(s.pos), shift2);
    double dist = spherePos[ax] - wallPos[ax];
    if (!c.isReal() && std::fabs(dist) > radius) return false;
    V3 contPt = spherePos;
    contPt[ax] = wallPos[ax];
    V3 normal = zero3();
    if (sense == 0) {
      if (!c.geom) normal[ax] = dist > 0 ? 1. : -1.;
      else normal[ax] = c.trsf[0][ax];
    } else normal[ax] = (sense == 1 ? 1. : -1.);
    double uN = normal[ax] * dist - radius;
    handleSpheresLike(c, wallPos, ld3(w.vel), ld3(w.angVel), spherePos, ld3(s.vel), ld3(s.angVel), normal, contPt, uN,
                      -radius, radius);
    return true;
  }
  // CompUtils::closestSegmentPt (CompUtils.cpp:59-65)
  static V3 closestSegmentPt(const V3& P, const V3& A, const V3& B) {
    V3 BA = sub(B, A);
    if (sqn(BA) == 0.) return A;
    double u = (dot(P, BA) - dot(A, BA)) / sqn(BA);
    return add(A, smul(std::min(1., std::max(0., u)), BA));
  }
  // Cg2_Facet_Sphere_L6Geom::go (Facet.cpp:333-390) for a facet with static vertices:
  // Facet::getNormal (74-77), getOuterVectors (113-118), interpolatePtLinAngVel (161-177) with
  // zero node velocities (linear velocity = fakeVel, or zero when fakeVel.x is NaN; angular
  // velocity zero). Facet.n21lim is NaN (no neighbour angles), so adjustBoundaryContactGeom
  // (220-265) returns at once and is not restated.
  // Facet::adjustBoundaryContactGeom (Facet.cpp:220-237): an edge contact uses the edge's
  // n21lim and outer vector; a vertex contact interpolates both between its two edges
  static void adjustBoundaryContactGeom(const Particle& f, const V3 v[3], const int e0e1[2], const V3& fNormal,
                                        const V3 outVec[3], const V3& otherCenter, V3& normal, V3& contPt) {
    const int e0 = e0e1[0], e1 = e0e1[1];
    if (std::isnan(f.n21lim[e0])) return;
    V3 out0 = outVec[e0];
    normalize(out0);
    if (e1 < 0) {
      adjustBoundaryContactGeom_impl(fNormal, v[e0], out0, otherCenter, f.n21lim[e0], normal, contPt);
      return;
    }
    if (std::isnan(f.n21lim[e1])) return;
    V3 out1 = outVec[e1];
    normalize(out1);
    const double p0 = dot(out0, normal), pMin = dot(out0, out1), pMax = 1.;
    const double t01 = (p0 - pMin) / (pMax - pMin);
    const V3 uo = add(mul(out0, t01), mul(out1, 1 - t01));
    adjustBoundaryContactGeom_impl(fNormal, v[e1], uo, otherCenter, t01 * f.n21lim[e0] + (1 - t01) * f.n21lim[e1],
                                   normal, contPt);
  }
  // Facet::adjustBoundaryContactGeom_impl (Facet.cpp:239-265)
  static void adjustBoundaryContactGeom_impl(const V3& fN, const V3& edgePt, const V3& unitOutVec,
                                             const V3& otherCenter, double nnLim, V3& normal, V3& contPt) {
    if (std::isnan(nnLim)) return;
    const double nn = dot(normal, fN);
    if ((nn >= 0 && nnLim > 0 && nn >= nnLim) || (nn < 0 && nnLim < 0 && nn <= nnLim)) return;
    const double noLim = (std::fabs(nnLim) >= 1. ? 1. : std::sqrt(1. - nnLim * nnLim));
    double nno0, nno1;
    if ((nnLim >= 0 && nn <= -noLim) || (nnLim <= 0 && nn >= noLim)) {
      nno0 = -nnLim;
      nno1 = 0;
    } else {
      nno0 = nnLim;
      nno1 = noLim;
    }
    normal = add(mul(fN, nno0), mul(unitOutVec, nno1));
    contPt = sub(otherCenter, mul(normal, dot(sub(otherCenter, edgePt), normal)));
  }
  bool geomFacetSphere(Contact& c, const V3& shift2) const {
    const Particle &f = par[c.pA], &s = par[c.pB];
    const V3 v[3] = {ld3(f.fv), ld3(f.fv + 3), ld3(f.fv + 6)};
    const V3 sC = add(ld3(s.pos), shift2);
    const V3 nn = cross(sub(v[1], v[0]), sub(v[2], v[0]));
    V3 fNormal = nn;
    normalize(fNormal);
    const double planeDist = dot(sub(sC, v[0]), fNormal);
    if (std::fabs(planeDist) > (s.radius + f.halfThick) && !c.isReal()) return false;
    const V3 fC = sub(sC, smul(planeDist, fNormal));
    const V3 outVec[3] = {cross(sub(v[1], v[0]), nn), cross(sub(v[2], v[1]), nn), cross(sub(v[0], v[2]), nn)};
    double ll[3];
    for (int i = 0; i < 3; i++) ll[i] = dot(outVec[i], sub(fC, v[i]));
    const int w = (ll[0] > 0 ? 1 : 0) + (ll[1] > 0 ? 2 : 0) + (ll[2] > 0 ? 4 : 0);
    V3 contPt;
    switch (w) {
      case 0: contPt = fC; break;
      case 1: contPt = closestSegmentPt(fC, v[0], v[1]); break;
      case 2: contPt = closestSegmentPt(fC, v[1], v[2]); break;
      case 4: contPt = closestSegmentPt(fC, v[2], v[0]); break;
      case 3: contPt = v[1]; break;
      case 5: contPt = v[0]; break;
      case 6: contPt = v[2]; break;
      default: throw std::logic_error("Cg2_Facet_Sphere_L6Geom: impossible sphere-facet intersection");
    }
    V3 normal = sub(sC, contPt);
    if (sqn(normal) > pow2(s.radius + f.halfThick) && !c.isReal()) return false;
    const double dist = nrm(normal);
    if (dist != 0) normal = divs(normal, dist);
    else normal = c.geom ? quatRot(c.nodeOri, mk(1, 0, 0)) : mk(1, 0, 0);
    // Facet.cpp:372-376: an edge or vertex contact smoothed towards the neighbouring facet
    static const int edgeNums[7][2] = {{-1, -1}, {0, -1}, {1, -1}, {0, 1}, {2, -1}, {2, 0}, {1, 2}};
    if (w > 0) adjustBoundaryContactGeom(f, v, edgeNums[w], fNormal, outVec, sC, normal, contPt);
    if (f.halfThick > 0) contPt = add(contPt, mul(normal, f.halfThick));
    const double uN = (dist - s.radius) - f.halfThick;
    V3 linVel = zero3();
    if (!std::isnan(f.fakeVel[0])) linVel = ld3(f.fakeVel);
    handleSpheresLike(c, contPt, linVel, zero3(), sC, ld3(s.vel), ld3(s.angVel), normal, contPt, uN,
                      std::max(f.halfThick, s.radius), s.radius);
    return true;
  }
  // ---------------------------------------------------------- phys
  // Cp2_FrictMat_FrictPhys::updateFrictPhys (FrictMat.cpp:20-44)
  void physFrict(Contact& c) const {
    const Material &m1 = mats[par[c.pA].mat], &m2 = mats[par[c.pB].mat];
    double l0 = c.lens[0], l1 = c.lens[1], A = c.contA;
    c.kn = 1 / (1 / (m1.young * A / l0) + 1 / (m2.young * A / l1));
    c.kt = .5 * (m1.ktDivKn + m2.ktDivKn) * c.kn;
    c.tanPhi = std::min(m1.tanPhi, m2.tanPhi);
    c.phys = true;
  }
  // Cp2_HertzMat_HertzPhys::go (Hertz.cpp:34-77)
  void physHertz(Contact& c) const {
    const Material &m1 = mats[par[c.pA].mat], &m2 = mats[par[c.pB].mat];
    double r1 = c.lens[0], r2 = c.lens[1];
    double E1 = m1.young, E2 = m2.young;
    double nu1 = prm.poisson, nu2 = prm.poisson;
    c.K = (4 / 3.) * 1. / ((1 - nu1 * nu1) / E1 + (1 - nu2 * nu2) / E2);
    c.R = 1. / (1. / r1 + 1. / r2);
    double kn0 = c.K * std::sqrt(c.R);
    c.gamma = std::min(m1.surfEnergy, m2.surfEnergy);
    c.alpha = .5 * (m1.alpha + m2.alpha);
    double G1 = E1 / (2 * (1 + nu1)), G2 = E2 / (2 * (1 + nu2));
    double G = .5 * (G1 + G2);
    double nu = .5 * (nu1 + nu2);
    c.kt0 = 2 * std::sqrt(4 * c.R) * G / (2 - nu);
    c.tanPhi = std::min(m1.tanPhi, m2.tanPhi);
    double en = prm.en;
    if (en > 0 && en < 1.) {
      double ma = par[c.pA].mass, mb = par[c.pB].mass;
      double mbar = (ma <= 0 && mb > 0) ? mb : ((ma > 0 && mb <= 0) ? ma : (ma * mb) / (ma + mb));
      double viscAlpha = -std::sqrt(5) * std::log(en) / (std::sqrt(pow2(std::log(en)) + pow2(M_PI)));
      c.alpha_sqrtMK = std::max(0., viscAlpha * std::sqrt(mbar * kn0));
    } else c.alpha_sqrtMK = 0.0;
    c.phys = true;
  }
  // ---------------------------------------------------------- laws
  // Law2_L6Geom_FrictPhys_IdealElPl::go (IdealElPl.cpp:15-125), iniEqlb=false, no rolling
  bool lawIdealElPl(Contact& c) const {
    double uN = c.uN;
    if (uN > 0 && !prm.noBreak) return false;
    c.torque = zero3();
    c.force[0] = c.kn * uN;
    if (prm.noFrict || c.tanPhi == 0.) {
      c.force[1] = 0; c.force[2] = 0;
    } else {
      double vt0 = c.vel[1], vt1 = c.vel[2];
      double dk = prm.dt * c.kt;
      double ft0 = c.force[1] + dk * vt0, ft1 = c.force[2] + dk * vt1;
      double maxFt = std::fabs(c.force[0]) * c.tanPhi;
      if (ft0 * ft0 + ft1 * ft1 > maxFt * maxFt && !prm.noSlip) {
        double FtNorm = std::sqrt(ft0 * ft0 + ft1 * ft1);
        double ratio = maxFt / FtNorm;
        if (FtNorm == 0.) ratio = 0.;
        // IdealElPl.cpp:62-73: dissipation of the slip, while Ft is still the trial value
        if (prm.trackEnergy) c.en[1] = (c.kt != 0. ? maxFt * (FtNorm - maxFt) / c.kt : 0.);
        ft0 = ft0 * ratio; ft1 = ft1 * ratio;
      }
      c.force[1] = ft0; c.force[2] = ft1;
      if (std::isnan(std::max(std::max(c.force[0], c.force[1]), c.force[2])) || std::isnan(c.force[0]) ||
          std::isnan(ft0) || std::isnan(ft1))
        throw std::runtime_error("##" + std::to_string(c.pA) + "+" + std::to_string(c.pB) + " has NaN force!");
    }
    // IdealElPl.cpp:85-92: elastic potential, 0 when it is NaN
    if (prm.trackEnergy) {
      double elast = 0.5 * (pow2(c.force[0]) / c.kn +
                            (c.kt != 0. ? (c.force[1] * c.force[1] + c.force[2] * c.force[2]) / c.kt : 0.));
      if (std::isnan(elast)) elast = 0.;
      c.en[0] = elast;
    }
    // IdealElPl.cpp:93-123: rolling and twisting resistance; the torque was zeroed above, so it
    // is this step's increment, capped (no history)
    if (prm.relRollStiff > 0. && prm.rollTanPhi > 0.) {
      const double charLen = c.lens[0] + c.lens[1];
      if (charLen <= 0) throw std::runtime_error("charLen<=0");
      const double kr = c.kn * charLen;
      if (prm.relTwistStiff > 0) {
        c.torque[0] += prm.dt * prm.relTwistStiff * kr * c.angVel[0];
        const double maxTt = std::fabs(c.force[0] * prm.rollTanPhi * charLen);
        if (std::fabs(c.torque[0]) > maxTt) c.torque[0] = std::copysign(maxTt, c.torque[0]);
      }
      const double dtkr = prm.dt * kr;
      c.torque[1] += dtkr * c.angVel[1];
      c.torque[2] += dtkr * c.angVel[2];
      const double maxTr = std::max(0., std::fabs(c.force[0]) * prm.rollTanPhi * charLen);
      if (c.torque[1] * c.torque[1] + c.torque[2] * c.torque[2] > maxTr * maxTr) {
        const double TrNorm = std::sqrt(c.torque[1] * c.torque[1] + c.torque[2] * c.torque[2]);
        double ratio = maxTr / TrNorm;
        if (TrNorm == 0.) ratio = 0.;
        c.torque[1] *= ratio;
        c.torque[2] *= ratio;
      }
    } else {
      c.torque = zero3();
    }
    return true;
  }
  // Law2_L6Geom_HertzPhys_DMT::go (Hertz.cpp:85-210), DMT/Hertz branch (alpha==0 or gamma==0)
  bool lawHertzDMT(Contact& c) const {
    if (!(c.alpha == 0. || c.gamma == 0.)) throw std::runtime_error("Schwarz branch of Law2_L6Geom_HertzPhys_DMT is not restated");
    c.torque = zero3();
    const double dt = prm.dt;
    const double velN = c.vel[0];
    const double vt0 = c.vel[1], vt1 = c.vel[2];
    double kn0 = c.K * std::sqrt(c.R);
    c.kn = (3 / 2.) * kn0 * std::sqrt(-c.uN);
    if (c.uN > 0) return false;
    double s = std::sqrt(-c.uN);
    double Fne = -kn0 * pown(s, 3);
    double Fna = 4 * M_PI * c.R * c.gamma;
    double eta = (c.alpha_sqrtMK > 0 ? c.alpha_sqrtMK * std::sqrt(std::sqrt(std::max(-c.uN, 0.))) : 0.);
    double Fnv = eta * velN;
    if (c.alpha == 0. && prm.noAttraction && Fne + Fnv > 0) Fnv = -Fne;
    double Fn = Fne + Fna + Fnv;
    c.force[0] = Fn;
    // Hertz.cpp:181: normal viscous dissipation
    if (prm.trackEnergy) c.en[2] = Fnv * velN * dt;
    c.kt = c.kt0 * std::sqrt(c.uN < 0 ? -c.uN : 0);
    double dk = dt * c.kt;
    double ft0 = c.force[1] + dk * vt0, ft1 = c.force[2] + dk * vt1;
    double maxFt = std::fabs(std::min(0., Fn) * c.tanPhi);
    if (ft0 * ft0 + ft1 * ft1 > pow2(maxFt)) {
      double FtNorm = std::sqrt(ft0 * ft0 + ft1 * ft1);
      double ratio = maxFt / FtNorm;
      // Hertz.cpp:191: sliding dissipation
      if (prm.trackEnergy) c.en[1] = (.5 * (FtNorm - maxFt) + maxFt) * (FtNorm - maxFt) / c.kt;
      ft0 = ft0 * ratio; ft1 = ft1 * ratio;
    } else {
      ft0 = ft0 + eta * vt0; ft1 = ft1 + eta * vt1;
      // Hertz.cpp:198: tangential viscous dissipation
      if (prm.trackEnergy) c.en[3] = eta * (vt0 * vt0 + vt1 * vt1) * dt;
    }
    c.force[1] = ft0; c.force[2] = ft1;
    // Hertz.cpp:203-206: elastic potential (in compression)
    if (prm.trackEnergy && c.uN < 0)
      c.en[0] = kn0 * (2 / 5.) * pown(std::sqrt(-c.uN), 5) + 0.5 * (ft0 * ft0 + ft1 * ft1) / c.kt;
    return true;
  }
  // Contact::getForceTorqueBranch (Contact.cpp:24-34); phys.torque is zero for both laws
  void forceBranch(const Contact& c, int pid, V3& F, V3& T) const {
    bool isPA = (pid == c.pA);
    double sign = isPA ? 1. : -1.;
    F = mul(quatRot(c.nodeOri, c.force), sign);
    V3 shift = (!isPA && prm.periodic) ? shiftPos(c.cellDist) : zero3();
    V3 xc = sub(c.nodePos, add(ld3(par[pid].pos), shift));
    // the phys torque (rolling resistance) is added when it is not zero
    if (c.torque[0] == 0 && c.torque[1] == 0 && c.torque[2] == 0) T = cross(xc, F);
    else T = add(cross(xc, F), mul(quatRot(c.nodeOri, c.torque), sign));
  }
  // CompUtils::triangleBarycentrics (CompUtils.cpp:94-112): x - A = u (B - A) + v (C - A), dotted
  // with B - A and C - A gives a 2x2 system, solved with the inverse of its matrix (adjugate over
  // the determinant); returns (1 - u - v, u, v)
  static V3 triangleBarycentrics(const V3& x, const V3& A, const V3& B, const V3& C) {
    const V3 vB = sub(B, A), vC = sub(C, A), vX = sub(x, A);
    const double m00 = dot(vB, vB), m01 = dot(vC, vB), m10 = dot(vB, vC), m11 = dot(vC, vC);
    const double invdet = 1. / (m00 * m11 - m10 * m01);
    const double r0 = dot(vX, vB), r1 = dot(vX, vC);
    const double u = (m11 * invdet) * r0 + (-m01 * invdet) * r1;
    const double v = (-m10 * invdet) * r0 + (m00 * invdet) * r1;
    return mk(1. - u - v, u, v);
  }
  // IntraForce with In2_Facet (Facet.cpp:301-328, bary=true): every facet with contacts spreads
  // each real contact's force over its 3 nodes, weighted by the barycentric coordinates of the
  // contact point projected onto the facet plane; Contact::getForceTorqueBranch (Contact.cpp:
  // 24-34) per node gives F = ori * force * sign and xc = contact point - node position (the
  // facet is always pA, so no periodic shift) and T = ori * torque * sign (rolling
  // resistance; zero otherwise). Node i gets w_i F and xc_i x (w_i F) + w_i T. Contacts are visited in Particle::contacts order (the
  // other id). The facet nodes are massless and outside gravity: the reset at the start of the
  // step (ForceResetter, or Leapfrog.reset of the step before) sets them to zero, so the values
  // kept are this step's sums, which ReadForce reads before Leapfrog resets them.
  void intraForce() {
#pragma omp parallel for schedule(dynamic, 4)
    for (long pid = 0; pid < (long)par.size(); pid++) {
      Particle& f = par[pid];
      if (f.shape != ORC_FACET) continue;
      for (int k = 0; k < 9; k++) f.nodeF[k] = f.nodeT[k] = 0.;
      if (f.contacts.empty()) continue;
      const V3 v[3] = {ld3(f.fv), ld3(f.fv + 3), ld3(f.fv + 6)};
      V3 normal = cross(sub(v[1], v[0]), sub(v[2], v[0]));  // Facet::getNormal (Facet.cpp:74-77)
      normalize(normal);
      for (const auto& oc : f.contacts) {
        const Contact& C = con[oc.second];
        if (!C.isReal()) continue;
        const V3 c = C.nodePos;
        const V3 p = sub(c, smul(dot(sub(c, v[0]), normal), normal));
        const V3 w = triangleBarycentrics(p, v[0], v[1], v[2]);
        const bool isPA = (int)pid == C.pA;
        const V3 Fc = mul(quatRot(C.nodeOri, C.force), isPA ? 1. : -1.);
        const V3 shift = (!isPA && prm.periodic) ? shiftPos(C.cellDist) : zero3();
        // the phys torque (rolling resistance), zero for the other laws
        const bool hasT = !(C.torque[0] == 0 && C.torque[1] == 0 && C.torque[2] == 0);
        const V3 Tc = hasT ? mul(quatRot(C.nodeOri, C.torque), isPA ? 1. : -1.) : zero3();
        for (int i = 0; i < 3; i++) {
          const V3 xc = sub(c, add(v[i], shift));
          const V3 F = mul(Fc, w[i]);
          const V3 T = hasT ? add(cross(xc, F), mul(Tc, w[i])) : cross(xc, F);
          for (int k = 0; k < 3; k++) { f.nodeF[3 * i + k] += F[k]; f.nodeT[3 * i + k] += T[k]; }
        }
      }
    }
  }
  // ContactLoop::run (ContactLoop.cpp:93-268)
  void contactLoop() {
    const bool det = prm.deterministic != 0;
    int TH = 1;
#ifdef _OPENMP
    TH = omp_get_max_threads();
#endif
    std::vector<std::vector<std::pair<int, int>>> pendT(TH);
    std::vector<std::vector<std::pair<int, int>>> errT(TH);
    std::vector<std::string> errMsg(TH);
    if (!det && locks.size() != par.size()) locks = std::vector<std::mutex>(par.size());
    long size = (long)con.size();
    if (prm.trackEnergy)
      for (auto& C : con) C.en[0] = C.en[1] = C.en[2] = C.en[3] = 0.;
#pragma omp parallel for schedule(guided)
    for (long i = 0; i < size; i++) {
      int t = 0;
#ifdef _OPENMP
      t = omp_get_thread_num();
#endif
      Contact& C = con[i];
      try {
        if (!C.isReal() && isFresh(C)) {
          if (par[C.pA].shape == ORC_SPHERE && par[C.pB].shape == ORC_SPHERE)
            C.minDist00Sq = pow2(std::fabs(prm.distFactor * (par[C.pA].radius + par[C.pB].radius)));
          else C.minDist00Sq = -1;
        }
        V3 shift2 = shiftPos(C.cellDist);
        if (!C.isReal() && !isFresh(C) && C.minDist00Sq > 0 &&
            sqn(sub(ld3(par[C.pA].pos), add(ld3(par[C.pB].pos), shift2))) > C.minDist00Sq)
          continue;
        // geometry functor by shape pair (pB is always the sphere); other pairs have no functor
        if (par[C.pB].shape != ORC_SPHERE) continue;
        bool geomOk;
        if (par[C.pA].shape == ORC_WALL) geomOk = geomWallSphere(C, shift2);
        else if (par[C.pA].shape == ORC_FACET) geomOk = geomFacetSphere(C, shift2);
        else geomOk = geomSphereSphere(C, shift2);
        if (!geomOk) continue;
        // ContactLoop.cpp:179-180: phys at creation, and on existing contacts by updatePhys
        if (!C.phys) C.stepCreated = step;
        if (!C.phys || prm.updatePhys > 0) {
          if (prm.physKind == ORC_PHYS_HERTZ) physHertz(C);
          else physFrict(C);
        }
        bool keep = prm.lawKind == ORC_LAW_HERTZDMT ? lawHertzDMT(C) : lawIdealElPl(C);
        if (!keep) {
          // ContactContainer::requestRemoval -> Contact::reset
          C.geom = false; C.phys = false;
          C.force = zero3(); C.torque = zero3();
          pendT[t].push_back({C.pA, C.pB});
        }
        if (prm.applyForces && C.isReal() && !det) {
          for (int pid : {C.pA, C.pB}) {
            // applyForceUninodal: multinodal particles (facets) get no force here
            if (par[pid].shape == ORC_FACET) continue;
            V3 F, T;
            forceBranch(C, pid, F, T);
            std::lock_guard<std::mutex> lk(locks[pid]);
            Particle& p = par[pid];
            for (int k = 0; k < 3; k++) { p.force[k] += F[k]; p.torque[k] += T[k]; }
          }
        }
      } catch (std::exception& e) {
        errMsg[t] = e.what();
      }
    }
    for (auto& m : errMsg)
      if (!m.empty()) throw std::runtime_error(m);
    // ContactLoop.cpp:265-266
    if (prm.updatePhys == 2) prm.updatePhys = 0;
    // ContactLoop.cpp:117,123,216-229,243-244: stress = sum of F branch^T over the real contacts,
    // over the cell volume (periodic cells only); the gradV work is zero (gradV = 0)
    for (double& x : clStress) x = 0.;
    if (prm.evalStress && prm.periodic) {
      for (const auto& C : con) {
        if (!C.isReal()) continue;
        if (par[C.pA].shape == ORC_FACET || par[C.pB].shape == ORC_FACET)
          throw std::runtime_error("ContactLoop.evalStress: multi-nodal particles in contact");
        const V3 b = sub(add(ld3(par[C.pB].pos), shiftPos(C.cellDist)), ld3(par[C.pA].pos));
        const V3 F = quatRot(C.nodeOri, C.force);
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) clStress[3 * i + j] += F[i] * b[j];
      }
      const double V = prm.cell[0] * prm.cell[1] * prm.cell[2];
      for (double& x : clStress) x /= V;
    }
    for (auto& v : pendT)
      for (auto& pr : v) pending.push_back(pr);
    if (prm.trackEnergy) {
      // EnergyTracker: elast (resettable) and the dissipations (increments) are each one sum over
      // the contacts, in ContactContainer order, added to Scene.energy
      double e[4] = {0, 0, 0, 0};
      for (const auto& C : con)
        for (int q = 0; q < 4; q++) e[q] += C.en[q];
      energy[ORC_E_ELAST] += e[0];
      energy[ORC_E_PLAST] += e[1];
      energy[ORC_E_VISCN] += e[2];
      energy[ORC_E_VISCT] += e[3];
    }
    if (det && prm.applyForces) {
      // canonical order: each particle sums its real contacts in increasing order of the other id
#pragma omp parallel for schedule(guided)
      for (long pid = 0; pid < (long)par.size(); pid++) {
        Particle& p = par[pid];
        if (p.shape == ORC_FACET) continue;
        for (const auto& oc : p.contacts) {
          const Contact& C = con[oc.second];
          if (!C.isReal()) continue;
          V3 F, T;
          forceBranch(C, (int)pid, F, T);
          for (int k = 0; k < 3; k++) { p.force[k] += F[k]; p.torque[k] += T[k]; }
        }
      }
    }
  }
  // ---------------------------------------------------------- integration
  void resetForce(Particle& p) const {
    bool hasGravity = !(prm.gravity[0] == 0 && prm.gravity[1] == 0 && prm.gravity[2] == 0);
    for (int k = 0; k < 3; k++) {
      p.force[k] = (hasGravity && !(p.flags & GRAVITY_SKIP)) ? p.mass * prm.gravity[k] : 0.;
      p.torque[k] = 0.;
    }
  }
  // ForceResetter::run (Leapfrog.cpp:15-26)
  void forceResetter() {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)par.size(); i++)
      if (par[i].integrated) resetForce(par[i]);
  }
  // Leapfrog::run, spherical integrator (Leapfrog.cpp:104-297)
  void leapfrog() {
    const double dt = prm.dt, damping = prm.damping;
    // Leapfrog.cpp:132 (gradV is zero here, so a periodic scene is not tracked at step 0)
    const bool track = prm.trackEnergy && (!prm.periodic || step > 0);
    const bool hasGravity = !(prm.gravity[0] == 0 && prm.gravity[1] == 0 && prm.gravity[2] == 0);
#pragma omp parallel for schedule(guided)
    for (long i = 0; i < (long)par.size(); i++) {
      Particle& p = par[i];
      if (!p.integrated) continue;
      bool damp = (damping != 0. && !(p.flags & DAMPING_SKIP));
      const uint32_t bl = p.flags & DOF_ALL;
      p.en[0] = p.en[1] = p.en[2] = 0.;
      if (track) {
        // Leapfrog::doDampingDissipation (Leapfrog.cpp:53-65), with the values before the update
        if (damp && !(p.flags & ENERGY_SKIP)) {
          double fv = (std::fabs(p.vel[0]) * std::fabs(p.force[0]) + std::fabs(p.vel[1]) * std::fabs(p.force[1])) +
                      std::fabs(p.vel[2]) * std::fabs(p.force[2]);
          double tw = (std::fabs(p.angVel[0]) * std::fabs(p.torque[0]) + std::fabs(p.angVel[1]) * std::fabs(p.torque[1])) +
                      std::fabs(p.angVel[2]) * std::fabs(p.torque[2]);
          p.en[2] = fv * damping * dt + tw * damping * dt;
        }
        // Leapfrog::doGravityWork (Leapfrog.cpp:67-74)
        if (hasGravity && !(p.flags & GRAVITY_SKIP)) {
          double gr = 0;
          if (bl == 0)
            gr = -((prm.gravity[0] * p.vel[0] + prm.gravity[1] * p.vel[1]) + prm.gravity[2] * p.vel[2]) * p.mass * dt;
          else
            for (int k = 0; k < 3; k++)
              if (!(bl & (1u << k))) gr -= prm.gravity[k] * p.vel[k] * p.mass * dt;
          p.en[1] = gr;
        }
      }
      // velocities before the update and accelerations, for the kinetic energy
      double pv[3] = {0, 0, 0}, pw[3] = {0, 0, 0}, lacc[3] = {0, 0, 0}, aacc[3] = {0, 0, 0};
      if (bl != DOF_ALL) {
        for (int k = 0; k < 3; k++) { pv[k] = p.vel[k]; pw[k] = p.angVel[k]; }
        double la[3], aa[3] = {0, 0, 0};
        if (bl == 0) for (int k = 0; k < 3; k++) la[k] = p.force[k] / p.mass;
        else for (int k = 0; k < 3; k++) { la[k] = 0; if (!(bl & (1u << k))) la[k] += p.force[k] / p.mass; }
        if (damp)
          for (int k = 0; k < 3; k++) la[k] *= 1 - damping * signum(p.force[k] * (p.vel[k] + 0.5 * dt * la[k]));
        for (int k = 0; k < 3; k++) p.vel[k] += dt * la[k];
        if (!(p.inertia[0] == 0 && p.inertia[1] == 0 && p.inertia[2] == 0)) {
          if (bl == 0) for (int k = 0; k < 3; k++) aa[k] = p.torque[k] / p.inertia[k];
          else for (int k = 0; k < 3; k++) { aa[k] = 0; if (!(bl & (8u << k))) aa[k] += p.torque[k] / p.inertia[k]; }
          if (damp)
            for (int k = 0; k < 3; k++) aa[k] *= 1 - damping * signum(p.torque[k] * (p.angVel[k] + 0.5 * dt * aa[k]));
          for (int k = 0; k < 3; k++) p.angVel[k] += dt * aa[k];
        }
        for (int k = 0; k < 3; k++) { lacc[k] = la[k]; aacc[k] = aa[k]; }
      }
      // Leapfrog::doKineticEnergy (Leapfrog.cpp:76-98), spherical: on-step velocities
      if (track && !(p.flags & ENERGY_SKIP)) {
        double cv[3], cw[3];
        for (int k = 0; k < 3; k++) { cv[k] = pv[k] + .5 * dt * lacc[k]; cw[k] = pw[k] + .5 * dt * aacc[k]; }
        double Et = .5 * p.mass * ((cv[0] * cv[0] + cv[1] * cv[1]) + cv[2] * cv[2]);
        double Er = 0.5 * ((cw[0] * (p.inertia[0] * cw[0]) + cw[1] * (p.inertia[1] * cw[1])) + cw[2] * (p.inertia[2] * cw[2]));
        if (std::isnan(Er) || std::isinf(std::max(std::max(p.inertia[0], p.inertia[1]), p.inertia[2]))) Er = 0;
        p.en[0] = Et + Er;
      }
      for (int k = 0; k < 3; k++) p.pos[k] += p.vel[k] * dt;
      // leapfrogSphericalRotate
      V3 axis = ld3(p.angVel);
      Q4 ori{p.ori[0], p.ori[1], p.ori[2], p.ori[3]};
      if (!(axis[0] == 0 && axis[1] == 0 && axis[2] == 0)) {
        double angle = nrm(axis);
        axis = divs(axis, angle);
        double ha = 0.5 * (angle * dt);
        double sh = std::sin(ha);
        Q4 q{std::cos(ha), sh * axis[0], sh * axis[1], sh * axis[2]};
        ori = quatMul(q, ori);
      }
      double qn = std::sqrt(((ori.w * ori.w + ori.x * ori.x) + ori.y * ori.y) + ori.z * ori.z);
      p.ori[0] = ori.w / qn; p.ori[1] = ori.x / qn; p.ori[2] = ori.y / qn; p.ori[3] = ori.z / qn;
      if (prm.reset) resetForce(p);
    }
    if (track) {
      // kinetic (resettable), grav and nonviscDamp (increments), each one sum in node order
      double e[3] = {0, 0, 0};
      for (const auto& p : par)
        if (p.integrated)
          for (int q = 0; q < 3; q++) e[q] += p.en[q];
      energy[ORC_E_KINETIC] += e[0];
      energy[ORC_E_GRAV] += e[1];
      energy[ORC_E_NONVISCDAMP] += e[2];
    }
  }
  // ---------------------------------------------------------- DynDt
  double nodalCritDtSq(int pid) const {
    const Particle& p = par[pid];
    if ((p.flags & DOF_ALL) == DOF_ALL) return Inf;
    V3 kt = zero3(), kr = zero3();
    for (const auto& oc : p.contacts) {
      const Contact& C = con[oc.second];
      if (!C.isReal()) continue;
      int ix = (C.pA == pid) ? 0 : 1;
      V3 n = quatRot(C.nodeOri, mk(1, 0, 0));
      V3 n2 = mk(n[0] * n[0], n[1] * n[1], n[2] * n[2]);
      double kn = C.kn, ktt = C.kt;
      for (int k = 0; k < 3; k++) kt[k] = kt[k] + (n2[k] * (kn - ktt) + ktt);
      double lk = pow2(C.lens[ix]) * ktt;
      V3 sm = mk(n2[1] + n2[2], n2[2] + n2[0], n2[0] + n2[1]);
      for (int k = 0; k < 3; k++) kr[k] = kr[k] + lk * sm[k];
    }
    double ret = Inf;
    for (int k = 0; k < 3; k++)
      if (kt[k] != 0 && p.mass > 0. && !(p.flags & (1u << k))) ret = std::min(ret, p.mass / std::fabs(kt[k]));
    for (int k = 0; k < 3; k++)
      if (kr[k] != 0 && p.inertia[k] > 0. && !(p.flags & (8u << k))) ret = std::min(ret, p.inertia[k] / std::fabs(kr[k]));
    return 2 * ret;
  }
  double critDtStiffness() const {
    double ret = Inf;
    for (int i = 0; i < (int)par.size(); i++)
      if (par[i].integrated) ret = std::min(ret, nodalCritDtSq(i));
    return std::sqrt(ret);
  }
  bool dynDtActivated() {
    long per = prm.dynDtPeriod;
    if (per <= 0) return false;
    bool initNow = dynStepLast < 0;
    if (initNow) dynStepLast = step;
    if ((step - dynStepLast >= per) || (step % per == 0)) {
      dynStepPrev = dynStepLast; dynStepLast = step;
      return true;
    }
    return false;
  }
  void dynDtRun() {
    double crDt = critDtStiffness();
    if (std::isinf(crDt)) return;
    long nSteps = step - dynStepPrev;
    double maxNewDt = prm.dt * pown(1. + prm.maxRelInc, (unsigned)nSteps);
    double nd = std::min(crDt * prm.dtSafety, maxNewDt);
    if (nd != prm.dt) nextDt = nd;
  }
  // ---------------------------------------------------------- step driver (Scene::doOneStep)
  void doOneStep(const int* eng, int ne) {
    // Scene.cpp:416: the resettable terms start every step at zero
    if (prm.trackEnergy) energy[ORC_E_KINETIC] = energy[ORC_E_ELAST] = 0.;
    for (int e = 0; e < ne; e++) {
      switch (eng[e]) {
        case ORC_E_FORCERESETTER: forceResetter(); break;
        case ORC_E_COLLIDER: colliderRun(); break;
        case ORC_E_CONTACTLOOP: contactLoop(); break;
        case ORC_E_LEAPFROG: leapfrog(); break;
        case ORC_E_DYNDT: if (dynDtActivated()) dynDtRun(); break;
        case ORC_E_INTRAFORCE: intraForce(); break;
        default: throw std::runtime_error("unknown engine code " + std::to_string(eng[e]));
      }
    }
    step++;
    time += prm.dt;
    if (!std::isnan(nextDt)) { prm.dt = nextDt; nextDt = NaN; }
  }
};

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

void orc_default_params(orc_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->dt = 1e-5;
  p->dtSafety = .9;
  p->trsfRenorm = 100;
  p->deterministic = 1;
  p->distFactor = -1;
  p->loneMask = 2;
  p->cell[0] = p->cell[1] = p->cell[2] = 1;
  p->verletDist = -.05;
  p->physKind = ORC_PHYS_FRICT;
  p->lawKind = ORC_LAW_IDEALELPL;
  p->applyForces = 1;
  p->poisson = .2;
  p->en = NaN;
  p->noAttraction = 1;
  p->damping = .2;
  p->reset = 1;
  p->dynDtPeriod = 100;
  p->maxRelInc = 1e-4;
}

orc_scene* orc_create(const orc_params* p) {
  orc_scene* s = new orc_scene();
  if (p) s->prm = *p;
  else orc_default_params(&s->prm);
  return s;
}
void orc_destroy(orc_scene* s) { delete s; }
int orc_set_params(orc_scene* s, const orc_params* p) {
  s->prm = *p;
  return 0;
}

int orc_add_material(orc_scene* s, int kind, double density, double young, double ktDivKn, double tanPhi,
                     double surfEnergy, double alpha) {
  s->mats.push_back(Material{kind, density, young, ktDivKn, tanPhi, surfEnergy, alpha});
  return (int)s->mats.size() - 1;
}

int orc_set_material(orc_scene* s, int m, int kind, double density, double young, double ktDivKn, double tanPhi,
                     double surfEnergy, double alpha) {
  if (m < 0 || m >= (int)s->mats.size()) return -1;
  s->mats[m] = Material{kind, density, young, ktDivKn, tanPhi, surfEnergy, alpha};
  return 0;
}

int orc_get_params(const orc_scene* s, orc_params* out) {
  *out = s->prm;
  return 0;
}

// woo.utils.sphere (utils.py:153) + Sphere::lumpMassInertia (Sphere.cpp:26-33)
int orc_add_sphere(orc_scene* s, const double pos[3], double radius, int mat, int fixed, uint32_t mask) {
  Particle p;
  p.shape = ORC_SPHERE;
  p.radius = radius;
  p.mat = mat;
  p.mask = mask;
  for (int k = 0; k < 3; k++) p.pos[k] = pos[k];
  double m = (4 / 3.) * M_PI * (radius * radius * radius) * s->mats[mat].density;
  p.mass = m;
  for (int k = 0; k < 3; k++) p.inertia[k] = 0. + 1. * (2. / 5.) * m * (radius * radius);
  p.flags = fixed ? DOF_ALL : 0;
  p.integrated = !fixed;
  s->par.push_back(p);
  return (int)s->par.size() - 1;
}

int orc_set_flags(orc_scene* s, int64_t i, uint32_t flags) {
  if (i < 0 || i >= (int64_t)s->par.size()) { g_err = "particle id out of range"; return -1; }
  s->par[i].flags = flags;
  return 0;
}

int orc_add_spheres(orc_scene* s, int64_t n, const double* pos, const double* radius, int mat, uint32_t mask) {
  int first = (int)s->par.size();
  s->par.reserve(s->par.size() + n);
  for (int64_t i = 0; i < n; i++) orc_add_sphere(s, pos + 3 * i, radius[i], mat, 0, mask);
  return first;
}

// woo.utils.wall (utils.py:245): fixed, zero mass
int orc_add_wall(orc_scene* s, const double pos[3], int axis, int sense, int mat, uint32_t mask) {
  Particle p;
  p.shape = ORC_WALL;
  p.axis = axis;
  p.sense = sense;
  p.mat = mat;
  p.mask = mask;
  for (int k = 0; k < 3; k++) p.pos[k] = pos[k];
  p.mass = 0;
  p.flags = DOF_ALL;
  p.integrated = false;
  s->par.push_back(p);
  return (int)s->par.size() - 1;
}

// woo.utils.facet (utils.py): fixed; Facet::lumpMassInertia gives nothing with halfThick=0 and
// the nodes are fixed, so the facet has no mass of its own here
int orc_add_facet(orc_scene* s, const double verts[9], const double fakeVel[3], double halfThick, int mat,
                  uint32_t mask) {
  Particle p;
  p.shape = ORC_FACET;
  p.mat = mat;
  p.mask = mask;
  for (int k = 0; k < 9; k++) p.fv[k] = verts[k];
  for (int k = 0; k < 3; k++) p.fakeVel[k] = fakeVel ? fakeVel[k] : 0.;
  p.halfThick = halfThick;
  for (int k = 0; k < 3; k++) p.pos[k] = verts[k];
  p.mass = 0;
  p.flags = DOF_ALL;
  p.integrated = false;
  s->par.push_back(p);
  return (int)s->par.size() - 1;
}

int orc_set_facet_n21lim(orc_scene* s, int64_t i, const double n21lim[3]) {
  if (i < 0 || i >= (int64_t)s->par.size() || s->par[i].shape != ORC_FACET) return -1;
  for (int k = 0; k < 3; k++) s->par[i].n21lim[k] = n21lim[k];
  return 0;
}

// Facet::computeNeighborAngles (Facet.cpp:180-216) for every facet; a shared node is a vertex
// at the same coordinates. Edge e0 (nodes e0, e0+1) with exactly one neighbouring facet on
// both nodes gets n1 . n2, n2 turned to our orientation (flipped when the neighbour runs the
// edge in the same direction); otherwise NaN.
int orc_facet_neighbor_angles(orc_scene* s) {
  auto same = [](const double* a, const double* b) { return a[0] == b[0] && a[1] == b[1] && a[2] == b[2]; };
  auto normalOf = [](const Particle& f) {
    V3 n = cross(sub(ld3(f.fv + 3), ld3(f.fv)), sub(ld3(f.fv + 6), ld3(f.fv)));
    normalize(n);
    return n;
  };
  const int64_t np = (int64_t)s->par.size();
  for (int64_t i = 0; i < np; i++) {
    Particle& f = s->par[i];
    if (f.shape != ORC_FACET) continue;
    for (int e0 = 0; e0 < 3; e0++) {
      const int e1 = (e0 + 1) % 3;
      int cnt = 0;
      int64_t f2 = -1;
      for (int64_t j = 0; j < np; j++) {
        const Particle& g = s->par[j];
        if (j == i || g.shape != ORC_FACET) continue;
        bool h0 = false, h1 = false;
        for (int q = 0; q < 3; q++) {
          h0 = h0 || same(g.fv + 3 * q, f.fv + 3 * e0);
          h1 = h1 || same(g.fv + 3 * q, f.fv + 3 * e1);
        }
        if (h0 && h1) {
          f2 = j;
          cnt++;
        }
      }
      if (cnt != 1) {
        f.n21lim[e0] = NaN;
        continue;
      }
      const Particle& g = s->par[f2];
      bool inv = false;
      for (int q = 0; q < 3; q++) {
        if (!same(g.fv + 3 * q, f.fv + 3 * e0)) continue;
        inv = same(g.fv + 3 * ((q + 1) % 3), f.fv + 3 * e1);
        break;
      }
      const V3 n2 = mul(normalOf(g), inv ? -1. : 1.);
      f.n21lim[e0] = dot(normalOf(f), n2);
    }
  }
  return 0;
}

int orc_facet_n21lim(const orc_scene* s, int64_t i, double out[3]) {
  if (i < 0 || i >= (int64_t)s->par.size() || s->par[i].shape != ORC_FACET) return -1;
  for (int k = 0; k < 3; k++) out[k] = s->par[i].n21lim[k];
  return 0;
}

int orc_num_particles(const orc_scene* s) { return (int)s->par.size(); }

int orc_get(const orc_scene* s, int field, double* out) {
  size_t n = s->par.size();
  for (size_t i = 0; i < n; i++) {
    const Particle& p = s->par[i];
    switch (field) {
      case ORC_F_POS: for (int k = 0; k < 3; k++) out[3 * i + k] = p.pos[k]; break;
      case ORC_F_VEL: for (int k = 0; k < 3; k++) out[3 * i + k] = p.vel[k]; break;
      case ORC_F_ANGVEL: for (int k = 0; k < 3; k++) out[3 * i + k] = p.angVel[k]; break;
      case ORC_F_ORI: for (int k = 0; k < 4; k++) out[4 * i + k] = p.ori[k]; break;
      case ORC_F_FORCE: for (int k = 0; k < 3; k++) out[3 * i + k] = p.force[k]; break;
      case ORC_F_TORQUE: for (int k = 0; k < 3; k++) out[3 * i + k] = p.torque[k]; break;
      case ORC_F_MASS: out[i] = p.mass; break;
      case ORC_F_INERTIA: for (int k = 0; k < 3; k++) out[3 * i + k] = p.inertia[k]; break;
      default: g_err = "bad field"; return -1;
    }
  }
  return 0;
}

int orc_set(orc_scene* s, int field, const double* in) {
  size_t n = s->par.size();
  for (size_t i = 0; i < n; i++) {
    Particle& p = s->par[i];
    switch (field) {
      case ORC_F_POS: for (int k = 0; k < 3; k++) p.pos[k] = in[3 * i + k]; break;
      case ORC_F_VEL:
        for (int k = 0; k < 3; k++) p.vel[k] = in[3 * i + k];
        break;
      case ORC_F_ANGVEL: for (int k = 0; k < 3; k++) p.angVel[k] = in[3 * i + k]; break;
      case ORC_F_ORI: for (int k = 0; k < 4; k++) p.ori[k] = in[4 * i + k]; break;
      case ORC_F_FORCE: for (int k = 0; k < 3; k++) p.force[k] = in[3 * i + k]; break;
      case ORC_F_TORQUE: for (int k = 0; k < 3; k++) p.torque[k] = in[3 * i + k]; break;
      case ORC_F_MASS: p.mass = in[i]; break;
      case ORC_F_INERTIA: for (int k = 0; k < 3; k++) p.inertia[k] = in[3 * i + k]; break;
      default: g_err = "bad field"; return -1;
    }
  }
  return 0;
}

int orc_run(orc_scene* s, const int* engines, int nEngines, int nSteps) {
  try {
#ifdef _OPENMP
    if (s->prm.nThreads > 0) omp_set_num_threads(s->prm.nThreads);
#endif
    for (int i = 0; i < nSteps; i++) s->doOneStep(engines, nEngines);
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

int orc_facet_nodes(const orc_scene* s, int64_t i, double out[18]) {
  if (i < 0 || i >= (int64_t)s->par.size() || s->par[i].shape != ORC_FACET) {
    g_err = "not a facet";
    return -1;
  }
  for (int k = 0; k < 9; k++) { out[k] = s->par[i].nodeF[k]; out[9 + k] = s->par[i].nodeT[k]; }
  return 0;
}

int64_t orc_step(const orc_scene* s) { return s->step; }
double orc_dt(const orc_scene* s) { return s->prm.dt; }
double orc_verlet_dist(const orc_scene* s) { return s->verletDist; }
int orc_num_refresh(const orc_scene* s) { return s->nRefresh; }

static bool reportContact(const orc_scene* s, const Contact& c, int which) {
  if (which == 1) return c.isReal();
  return c.isReal() || s->contactAabbOverlap(c);
}

int64_t orc_num_contacts(const orc_scene* s, int which) {
  int64_t n = 0;
  for (const auto& c : s->con) n += reportContact(s, c, which);
  return n;
}

int64_t orc_get_contacts(const orc_scene* s, int which, int32_t* ids, int32_t* cellDist, double* data) {
  std::vector<const Contact*> v;
  for (const auto& c : s->con)
    if (reportContact(s, c, which)) v.push_back(&c);
  std::sort(v.begin(), v.end(), [](const Contact* a, const Contact* b) {
    return pairKey(a->pA, a->pB) < pairKey(b->pA, b->pB);
  });
  for (size_t i = 0; i < v.size(); i++) {
    const Contact& c = *v[i];
    if (ids) { ids[2 * i] = c.pA; ids[2 * i + 1] = c.pB; }
    if (cellDist) for (int k = 0; k < 3; k++) cellDist[3 * i + k] = c.cellDist[k];
    if (data) {
      double* d = data + 16 * i;
      bool r = c.isReal();
      for (int k = 0; k < 3; k++) {
        d[k] = r ? c.trsf[0][k] : NaN;
        d[3 + k] = r ? c.trsf[1][k] : NaN;
        d[6 + k] = r ? c.nodePos[k] : NaN;
        d[9 + k] = r ? c.force[k] : NaN;
      }
      d[12] = r ? c.uN : NaN;
      d[13] = r ? c.kn : NaN;
      d[14] = r ? c.kt : NaN;
      d[15] = r ? 1. : 0.;
    }
  }
  return (int64_t)v.size();
}

// DemFuncs::coordNumber (Funcs.cpp:84-102), all particles, infinite box
double orc_coord_number(const orc_scene* s, int skipFree) {
  long C2 = 0, N0 = 0, N1 = 0, N = 0;
  for (const auto& p : s->par) {
    int n = 0;
    for (const auto& oc : p.contacts) n += s->con[oc.second].isReal();
    if (n == 0) N0++;
    else if (n == 1) N1++;
    N++;
    C2 += n;
  }
  if (skipFree) return (C2 - N1) * 1. / (N - N0 - N1);
  return C2 * 1. / N;
}

// DemFuncs::porosity (Funcs.cpp:104-114), returns the solid fraction Vs/V
double orc_solid_fraction(const orc_scene* s, const double boxMin[3], const double boxMax[3]) {
  double Vs = 0;
  double V = ((boxMax[0] - boxMin[0]) * (boxMax[1] - boxMin[1])) * (boxMax[2] - boxMin[2]);
  for (const auto& p : s->par) {
    if (p.shape != ORC_SPHERE) continue;
    bool in = true;
    for (int k = 0; k < 3; k++) in = in && p.pos[k] >= boxMin[k] && p.pos[k] <= boxMax[k];
    if (!in) continue;
    Vs += (4 / 3.) * M_PI * (p.radius * p.radius * p.radius);
  }
  return Vs / V;
}

// sum over integrated nodes of 1/2 m v^2 + 1/2 w.(I w)
void orc_energy(const orc_scene* s, double out[7]) {
  for (int q = 0; q < ORC_E_N; q++) out[q] = s->energy[q];
}

double orc_kinetic_energy(const orc_scene* s) {
  double E = 0;
  for (const auto& p : s->par) {
    if (!p.integrated) continue;
    V3 v = ld3(p.vel), w = ld3(p.angVel);
    double Et = .5 * p.mass * sqn(v);
    double Er = 0.5 * dot(w, mk(p.inertia[0] * w[0], p.inertia[1] * w[1], p.inertia[2] * w[2]));
    E += Et + Er;
  }
  return E;
}

// DemFuncs::stressStiffness, stress part (Funcs.cpp:183-229): over the real contacts in
// ContactContainer order, d0 = |posA - shift - posB| with the node positions,
// n = ori*e_x, F = ori*force, sigma_ij += d0 (fN n_i n_j + (fT_i n_j + fT_j n_i)/2),
// divided by the volume (the periodic cell's when volume <= 0)
int orc_cl_stress(const orc_scene* s, double out[9]) {
  for (int k = 0; k < 9; k++) out[k] = s->clStress[k];
  return 0;
}

int orc_stress(const orc_scene* s, double volume, double out[9]) {
  double S[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (const auto& C : s->con) {
    if (!C.isReal()) continue;
    const auto& A = s->par[C.pA];
    const auto& B = s->par[C.pB];
    V3 posA = ld3(A.pos), posB = ld3(B.pos);
    if (s->prm.periodic)
      posA = sub(posA, mk(s->prm.cell[0] * C.cellDist[0], s->prm.cell[1] * C.cellDist[1], s->prm.cell[2] * C.cellDist[2]));
    // multinodal (skipMultinodal=false): the contact point stands for the facet
    if (A.shape == ORC_FACET) posA = C.nodePos;
    const double d0 = std::sqrt(sqn(sub(posA, posB)));
    const V3 n = quatRot(C.nodeOri, mk(1, 0, 0));
    const V3 F = quatRot(C.nodeOri, C.force);
    const double fN = dot(F, n);
    const V3 fT = sub(F, mul(n, fN));
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) S[i][j] += d0 * (fN * n[i] * n[j] + .5 * (fT[i] * n[j] + fT[j] * n[i]));
  }
  if (volume <= 0) {
    if (!s->prm.periodic) {
      g_err = "a positive volume must be given for aperiodic scenes";
      return -1;
    }
    volume = s->prm.cell[0] * s->prm.cell[1] * s->prm.cell[2];
  }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) out[3 * i + j] = S[i][j] / volume;
  return 0;
}

// DemFuncs::pWaveDt (Funcs.cpp:151-175)
double orc_pwave_dt(const orc_scene* s) {
  double dt = Inf;
  for (const auto& p : s->par) {
    if (p.shape != ORC_SPHERE) continue;
    const Material& m = s->mats[p.mat];
    dt = std::min(dt, p.radius / (1. * std::sqrt(m.young / m.density)));
  }
  return dt;
}

double orc_crit_dt(const orc_scene* s) { return s->critDtStiffness(); }

int64_t orc_brute_aabb_pairs(const orc_scene* s, int32_t* ids, int64_t cap) {
  int n = (int)s->par.size();
  std::vector<std::pair<int, int>> out;
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++) {
      if (!s->par[i].hasBound || !s->par[j].hasBound) continue;
      if (!s->mayCollide(i, j)) continue;
      int a, b, cd[3];
      s->canonicalRoles(i, j, a, b);
      s->nearestImage(a, b, cd);
      if (s->aabbOverlapShift(a, b, cd)) out.push_back({a, b});
    }
  std::sort(out.begin(), out.end(), [](const std::pair<int, int>& x, const std::pair<int, int>& y) {
    return pairKey(x.first, x.second) < pairKey(y.first, y.second);
  });
  if (ids)
    for (int64_t i = 0; i < (int64_t)out.size() && i < cap; i++) { ids[2 * i] = out[i].first; ids[2 * i + 1] = out[i].second; }
  return (int64_t)out.size();
}

}  // extern "C"
