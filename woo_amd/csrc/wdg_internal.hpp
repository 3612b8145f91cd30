// Internal device state of the DEM step engine (not part of the C ABI).
//
// Particles are stored in "storage order": walls first (user-id order), then
// spheres sorted by broadphase cell, re-sorted at every broadphase refresh so
// that neighbours in space are neighbours in memory. uid[] / sto[] map storage
// index <-> user id (Particle.id); the C ABI speaks user ids only.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/woo_amd.h"

namespace wdg {

// ------------------------------------------------------------------ errors
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
#define WDG_HIP(call)                                                                                   \
  do {                                                                                                  \
    hipError_t e_ = (call);                                                                             \
    if (e_ != hipSuccess)                                                                               \
      throw ::wdg::Error(WDG_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_) + " (" __FILE__ \
                                        ":" + std::to_string(__LINE__) + ")");                          \
  } while (0)
#define WDG_LAUNCH_CHECK() WDG_HIP(hipGetLastError())

// ------------------------------------------------------------------ particle flag bits
// low bits are DemData.flags (Particle.hpp:196-199); high bits are engine-private
enum : uint32_t {
  F_DOF_ALL = 63u,
  F_GRAVITY_SKIP = 512u,
  F_DAMPING_SKIP = 2048u,
  F_INTEGRATED = 1u << 16,
  F_SHAPE_SHIFT = 17,   // 2 bits: 0 sphere, 1 wall
  F_AXIS_SHIFT = 19,    // 2 bits: wall axis
  F_SENSE_SHIFT = 21,   // 2 bits: wall sense + 1
};
__host__ __device__ inline int flag_shape(uint32_t f) { return (f >> F_SHAPE_SHIFT) & 3; }
__host__ __device__ inline int flag_axis(uint32_t f) { return (f >> F_AXIS_SHIFT) & 3; }
__host__ __device__ inline int flag_sense(uint32_t f) { return (int)((f >> F_SENSE_SHIFT) & 3) - 1; }

// ------------------------------------------------------------------ contact info word
enum : uint32_t {
  CI_TYPE_MASK = 3u,   // 0 sphere-sphere, 1 wall-sphere
  CI_SWAP = 4u,        // geometric first particle (pA) is B (the larger storage index)
  CI_NEAR = 8u,        // close enough at the last refresh to touch before the next one
};
enum { CT_SS = 0, CT_WS = 1 };
constexpr uint32_t ADJ_PA = 0x80000000u;  // adjacency entry: this particle is pA of the slot

// periodic image offsets packed 3 x 21 bits (two's complement)
__host__ __device__ inline uint64_t pack_cd(int a, int b, int c) {
  const uint64_t M = (1ull << 21) - 1;
  return (uint64_t(uint32_t(a)) & M) | ((uint64_t(uint32_t(b)) & M) << 21) | ((uint64_t(uint32_t(c)) & M) << 42);
}
__host__ __device__ inline int unpack_cd(uint64_t v, int k) {
  int64_t f = (int64_t)((v >> (21 * k)) & ((1ull << 21) - 1));
  if (f & (1ll << 20)) f -= (1ll << 21);
  return (int)f;
}

// device error bits
enum : uint32_t { ERR_NAN_FORCE = 1u, ERR_DOUBLE_IMAGE = 2u, ERR_CD_RANGE = 4u };

// ------------------------------------------------------------------ device buffers
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t cap = 0;
  void ensure(size_t n) {
    if (n <= cap && p) return;
    size_t c = n < 64 ? 64 : n + n / 2;
    if (p) WDG_HIP(hipFree(p));
    p = nullptr;
    WDG_HIP(hipMalloc(&p, c * sizeof(T)));
    cap = c;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  void swap(DBuf& o) {
    std::swap(p, o.p);
    std::swap(cap, o.cap);
  }
};

// one contact set: slots sorted by (lo, hi) storage index; SoA
struct Contacts {
  int64_t n = 0;
  DBuf<uint32_t> A, B;        // lo / hi storage index
  DBuf<uint32_t> info;        // CI_* bits
  DBuf<uint64_t> dist;        // packed cellDist of pB relative to pA (periodic only)
  DBuf<uint32_t> real;        // 1 if geom+phys exist (history valid)
  DBuf<double> hN[3], hY[3];  // L6Geom.trsf columns 0 and 1
  DBuf<double> hP[3];         // contact point (L6Geom.node.pos)
  DBuf<double> hF[3];         // CPhys.force, local frame
  DBuf<double> huN;           // L6Geom.uN
  DBuf<double> F[3];          // global force on pA, written by the contact kernel for the gather
  // per-particle adjacency: entries slot|ADJ_PA grouped by storage index, each group in
  // increasing order of the other particle's user id (the canonical summation order)
  DBuf<uint32_t> adj, adjOff;
  void ensure(int64_t nc, int64_t npar, bool periodic) {
    size_t c = (size_t)(nc > 0 ? nc : 1);
    A.ensure(c); B.ensure(c); info.ensure(c); real.ensure(c);
    if (periodic) dist.ensure(c);
    for (int k = 0; k < 3; k++) { hN[k].ensure(c); hY[k].ensure(c); hP[k].ensure(c); hF[k].ensure(c); F[k].ensure(c); }
    huN.ensure(c);
    adj.ensure(2 * c);
    adjOff.ensure((size_t)npar + 1);
  }
  void release() {
    A.release(); B.release(); info.release(); dist.release(); real.release();
    for (int k = 0; k < 3; k++) { hN[k].release(); hY[k].release(); hP[k].release(); hF[k].release(); F[k].release(); }
    huN.release(); adj.release(); adjOff.release();
  }
};

constexpr int NPD = 24;  // permuted double arrays per particle
constexpr int NPU = 4;   // permuted u32 arrays per particle

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  wdg_field field{};
  std::vector<wdg_material> mats;
  DBuf<wdg_material> dMats;

  // particles, storage order
  int64_t n = 0;
  int64_t nW = 0;       // walls occupy storage [0, nW)
  double rmin = INFINITY, rmax = 0;
  DBuf<double> pos[3], vel[3], angVel[3], ori[4], force[3], torque[3], inertia[3], lastPos[3];
  DBuf<double> mass, radius;
  DBuf<uint32_t> flags, mask, uid, sto;
  DBuf<int32_t> mat;
  DBuf<double> spareD[NPD];
  DBuf<uint32_t> spareU[NPU];

  // contacts, double-buffered across refreshes
  Contacts C[2];
  int cur = 0;
  Contacts& con() { return C[cur]; }

  // broadphase scratch
  DBuf<uint32_t> cellKey, sortIdx, cellKeyTmp, sortIdxTmp, perm, invPerm;
  DBuf<uint32_t> cellStart, cellEnd;
  DBuf<uint32_t> scount, soff;
  DBuf<uint64_t> pairKey, pairKeyTmp, pairDist;
  DBuf<uint32_t> pairInfo, pairVal, pairValTmp;
  DBuf<uint32_t> runFlag, runOff;
  DBuf<uint64_t> adjKey, adjKeyTmp;
  DBuf<uint32_t> adjValTmp, adjPart;
  DBuf<uint32_t> sortHist, scanAux;
  DBuf<double> red;       // reductions
  DBuf<double> xfer;      // user-id-ordered staging for upload/download
  DBuf<uint32_t> dflags;  // [0] dirty, [1] error bits, [2..] scratch counters
  uint32_t* hflags = nullptr;  // pinned host mirror

  // grid of the last refresh (for the on-demand potential-set report)
  struct GridState {
    double org[3], h[3], L[3];
    int nc[3];
    int periodic;
    bool brute;
    bool valid = false;
  } grid;

  // engine state
  bool boundsValid = false;
  bool dirtyKnown = false;     // dflags[0] reflects the current positions
  bool forcesPending = false;  // contact forces not yet gathered into force/torque
  double verletDist = NAN;
  int64_t nRefresh = 0, nCollide = 0;
  int64_t dynStepLast = -1, dynStepPrev = -1;
  // contact model of the last contact loop (DynDt and contact export need it)
  int lastPhysKind = WDG_PHYS_FRICT;
  double lastPoisson = .2, lastEn = NAN, lastViscAlpha = 0;
};

// ------------------------------------------------------------------ internal API
void scan_exclusive(Ctx& c, const uint32_t* in, uint32_t* out, int64_t n, uint32_t* dTotal);
template <class K>
void radix_sort_pairs(Ctx& c, K* keys, uint32_t* vals, K* keysTmp, uint32_t* valsTmp, int64_t n, int bits);
uint32_t read_u32(Ctx& c, const uint32_t* d);

void collide(Ctx& c, const wdg_collide_params* p);
// potential pairs (enlarged AABBs of the last refresh overlapping) as user-id pairs (pA,pB) + cellDist
void potential_pairs(Ctx& c, std::vector<uint32_t>& ida, std::vector<uint32_t>& idb, std::vector<uint64_t>& cd);
void contact_loop(Ctx& c, const wdg_contact_params* p);
void leapfrog(Ctx& c, const wdg_leapfrog_params* p);
void force_reset(Ctx& c);
void flush_forces(Ctx& c);
double crit_dt(Ctx& c);
void observables(Ctx& c, const double lo[3], const double hi[3], double out[3]);
int64_t count_real(Ctx& c);
void check_errors(Ctx& c);
void relax(Ctx& c, int maxIter, double inflate, double tol, double* maxOverlap, int* iters);
void drop_contacts(Ctx& c);

inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }
inline int bits_for(uint64_t maxval) {
  int b = 1;
  while (b < 64 && (maxval >> b) != 0) b++;
  return b;
}

}  // namespace wdg
