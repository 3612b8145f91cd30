// Internal device state of the DEM step engine (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
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
  CI_SWAP = 4u,        // geometric first particle (pA) is the larger id
  CI_AABB = 8u,        // enlarged AABBs overlapped at the last refresh
};
enum { CT_SS = 0, CT_WS = 1 };

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
constexpr int CD_LIMIT = (1 << 20) - 1;

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
  // grow keeping the first `keep` elements
  void grow(size_t n, size_t keep, hipStream_t s) {
    if (n <= cap && p) return;
    size_t c = n < 64 ? 64 : n + n / 2;
    T* q = nullptr;
    WDG_HIP(hipMalloc(&q, c * sizeof(T)));
    if (p && keep) WDG_HIP(hipMemcpyAsync(q, p, keep * sizeof(T), hipMemcpyDeviceToDevice, s));
    WDG_HIP(hipStreamSynchronize(s));
    if (p) WDG_HIP(hipFree(p));
    p = q;
    cap = c;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// one contact set: slots sorted by (min id, max id); SoA
struct Contacts {
  int64_t n = 0;
  DBuf<uint32_t> A, B;        // min id, max id
  DBuf<uint32_t> info;        // CI_* bits
  DBuf<uint64_t> dist;        // packed cellDist of pB relative to pA (periodic only)
  DBuf<uint32_t> real;        // 1 if geom+phys exist (history valid)
  DBuf<double> hN[3], hY[3];  // L6Geom.trsf columns 0 and 1
  DBuf<double> hP[3];         // contact point (L6Geom.node.pos)
  DBuf<double> hF[3];         // CPhys.force, local frame
  DBuf<double> huN;           // L6Geom.uN
  DBuf<double> F[3];          // global force on pA, written by the contact kernel for the gather
  DBuf<uint32_t> offA, offB;  // CSR over min id (slots) and over max id (into idxB), size n+1
  DBuf<uint32_t> idxB;        // slots sorted by (max id, min id)
  void ensure(int64_t nc, int64_t npar, bool periodic) {
    size_t c = (size_t)(nc > 0 ? nc : 1);
    A.ensure(c); B.ensure(c); info.ensure(c); real.ensure(c);
    if (periodic) dist.ensure(c);
    for (int k = 0; k < 3; k++) { hN[k].ensure(c); hY[k].ensure(c); hP[k].ensure(c); hF[k].ensure(c); F[k].ensure(c); }
    huN.ensure(c);
    idxB.ensure(c);
    offA.ensure((size_t)npar + 1); offB.ensure((size_t)npar + 1);
  }
  void release() {
    A.release(); B.release(); info.release(); dist.release(); real.release();
    for (int k = 0; k < 3; k++) { hN[k].release(); hY[k].release(); hP[k].release(); hF[k].release(); F[k].release(); }
    huN.release(); offA.release(); offB.release(); idxB.release();
  }
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  wdg_field field{};
  std::vector<wdg_material> mats;
  DBuf<wdg_material> dMats;

  // particles (SoA, one node each)
  int64_t n = 0;
  int64_t nSph = 0;
  std::vector<int32_t> wallIds;
  DBuf<int32_t> dWalls;
  double rmin = INFINITY, rmax = 0;
  DBuf<double> pos[3], vel[3], angVel[3], ori[4], force[3], torque[3], inertia[3], lastPos[3];
  DBuf<double> mass, radius;
  DBuf<uint32_t> flags, mask;
  DBuf<int32_t> mat;

  // contacts, double-buffered across refreshes
  Contacts C[2];
  int cur = 0;
  Contacts& con() { return C[cur]; }

  // broadphase scratch
  DBuf<uint32_t> cellKey, sortIdx, cellKeyTmp, sortIdxTmp;
  DBuf<uint32_t> cellStart, cellEnd;
  DBuf<double> sx, sy, sz, shs;  // sorted refresh positions and AABB half sizes
  DBuf<uint32_t> smask, scount, soff;
  DBuf<uint64_t> pairKey, pairKeyTmp, pairDist;
  DBuf<uint32_t> pairInfo, pairVal, pairValTmp;
  DBuf<uint32_t> runFlag, runOff;
  DBuf<uint32_t> keyB, keyBTmp, valB, valBTmp;
  DBuf<uint32_t> sortHist, scanAux;
  DBuf<double> red;  // reductions
  DBuf<uint32_t> dflags;  // [0] dirty, [1] error bits, [2..] scratch counters
  uint32_t* hflags = nullptr;  // pinned host mirror

  // engine state
  bool boundsValid = false;
  bool dirtyKnown = false;     // dflags[0] reflects the current positions
  bool forcesPending = false;  // contact forces not yet gathered into force/torque
  double verletDist = NAN;
  int64_t nRefresh = 0, nCollide = 0;
  bool lastBrute = false;
  int gridCells[3] = {0, 0, 0};
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
void contact_loop(Ctx& c, const wdg_contact_params* p);
void leapfrog(Ctx& c, const wdg_leapfrog_params* p);
void force_reset(Ctx& c);
void flush_forces(Ctx& c);
double crit_dt(Ctx& c);
void observables(Ctx& c, const double lo[3], const double hi[3], double out[3]);
int64_t count_real(Ctx& c);
void check_errors(Ctx& c);
void relax(Ctx& c, int maxIter, double inflate, double tol, double* maxOverlap, int* iters);

inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }
inline int bits_for(uint64_t maxval) {
  int b = 1;
  while (b < 64 && (maxval >> b) != 0) b++;
  return b;
}

}  // namespace wdg
