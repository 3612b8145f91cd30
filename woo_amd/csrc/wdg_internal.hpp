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

// a slot's contact is not real (per-slot and per-entry state indices)
constexpr uint32_t NO_STATE = 0xFFFFFFFFu;
// the entry of the contact's pA (Contacts::eidx); slot indices stay below it
constexpr uint32_t EIDX_PA = 0x80000000u;
// dflags word: set when Contacts::eidx was reset, so that the next filter writes the entries of
// every real slot (else only of the slots whose real state changed); cleared by the force pass
constexpr int EIDX_FULL = 48;

// ------------------------------------------------------------------ particle flag bits
// low bits are DemData.flags (Particle.hpp:196-199); high bits are engine-private
enum : uint32_t {
  F_DOF_ALL = 63u,
  F_ENERGY_SKIP = 256u,
  F_GRAVITY_SKIP = 512u,
  F_DAMPING_SKIP = 2048u,
  F_INTEGRATED = 1u << 16,
  F_SHAPE_SHIFT = 17,   // 2 bits: 0 sphere, 1 wall, 2 facet
  F_AXIS_SHIFT = 19,    // 2 bits: wall axis
  F_SENSE_SHIFT = 21,   // 2 bits: wall sense + 1
};
__host__ __device__ inline int flag_shape(uint32_t f) { return (f >> F_SHAPE_SHIFT) & 3; }
__host__ __device__ inline int flag_axis(uint32_t f) { return (f >> F_AXIS_SHIFT) & 3; }
__host__ __device__ inline int flag_sense(uint32_t f) { return (int)((f >> F_SENSE_SHIFT) & 3) - 1; }

// ------------------------------------------------------------------ contact info word
enum : uint32_t {
  CI_TYPE_MASK = 3u,   // 0 sphere-sphere, 1 wall-sphere, 2 facet-sphere
  CI_SWAP = 4u,        // geometric first particle (pA) is B (the larger storage index)
  CI_NEAR = 8u,        // close enough at the last refresh to touch before the next one
};
enum { CT_SS = 0, CT_WS = 1, CT_FS = 2 };
// per boundary body (storage [0, nW)): facet vertices v0 v1 v2, fakeVel, halfThick, n21lim
constexpr int FGEO = 16;
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

// dflags words: [0] dirty flag of a refresh / k_dirty, [1] error bits, [3..5] scan totals,
// [DIRTY_PP + 0/1] dirty flags set by alternate leapfrog launches (each launch clears the other)
constexpr int DIRTY_PP = 12;

// device error bits
enum : uint32_t {
  ERR_NAN_FORCE = 1u, ERR_DOUBLE_IMAGE = 2u, ERR_CD_RANGE = 4u, ERR_ADJ = 8u, ERR_FACET = 16u, ERR_SNAP_IMAGE = 32u
};

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

// slot list of one broadphase refresh: candidate pairs sorted by (lo, hi) storage index, SoA
struct Contacts {
  int64_t n = 0;
  DBuf<uint32_t> A, B;        // lo / hi storage index
  DBuf<uint32_t> info;        // CI_* bits
  DBuf<uint64_t> dist;        // packed cellDist of pB relative to pA (periodic only)
  DBuf<uint32_t> carry;       // at a refresh: 1 + old slot of a carried real contact, else 0
  // per-particle adjacency: entries slot|ADJ_PA grouped by storage index, each group in
  // increasing order of the other particle's user id (the canonical summation order)
  DBuf<uint32_t> adj, adjOff;
  // per adjacency entry: its slot (| EIDX_PA for pA's entry) when the slot is real, else
  // NO_STATE, kept by the contact loop's filter; and per slot the positions of its two entries
  // (pA's, then pB's), written at the refresh
  DBuf<uint32_t> eidx, entPos;
  void ensure(int64_t nc, int64_t npar, bool periodic) {
    size_t c = (size_t)(nc > 0 ? nc : 1);
    A.ensure(c); B.ensure(c); info.ensure(c); carry.ensure(c);
    if (periodic) dist.ensure(c);
    adj.ensure(2 * c);
    eidx.ensure(2 * c);
    entPos.ensure(2 * c);
    adjOff.ensure((size_t)npar + 1);
  }
  void release() {
    A.release(); B.release(); info.release(); dist.release(); carry.release(); adj.release(); adjOff.release();
    eidx.release(); entPos.release();
  }
};

// Real contacts of the current slot list (the Contact objects that have geom and phys).
// Slot s is real when bit s is set. Slots form groups of GROUP; the state of the real
// contacts of a group is packed densely in the group's range, in slot order: the local
// index of s is the number of real slots of the group before s. Reads and writes of the
// state therefore stream, and no global prefix sum is needed to find an index.
constexpr int TILE = 256;     // slots per prefix-count tile (4 words)
constexpr int GROUP = 2048;   // slots per group (8 tiles)
constexpr int GROUP_BS = 1024;  // threads of a group kernel
inline int64_t ntiles(int64_t nslots) { return (nslots + TILE - 1) / TILE; }
inline int64_t ngroups(int64_t nslots) { return (nslots + GROUP - 1) / GROUP; }
constexpr int GP = 8;         // doubles per gather record (FP64 contact form)
constexpr int GP32 = 12;      // floats per gather record (FP32 contact form)
struct Reals {
  DBuf<unsigned long long> bits;  // 64 words per group
  DBuf<uint32_t> tpre;            // per tile: real slots of its group before the tile
  DBuf<uint32_t> gcnt;            // per group: real slots
  DBuf<double> hN[3], hY[3];      // L6Geom.trsf columns 0 and 1
  DBuf<double> hF[3];             // CPhys.force, local frame
  DBuf<double> huN;               // L6Geom.uN
  // gather record per contact: global force on pA (3), contact point (3), cellDist bits, unused
  DBuf<double> gp;
  // material table (epoch) each real contact takes its phys from, while more than one is kept
  DBuf<uint8_t> hE;
  // global phys torque per real contact (rolling resistance; allocated when it is used)
  DBuf<double> gtq;
  // state index of every real slot, written by the contact loop's filter (stale for the other
  // slots); valid only for a set that filter produced
  DBuf<uint32_t> sidx;
  bool sidxValid = false;
  // room for every slot of a list of nslots being real
  void ensure(int64_t nslots) {
    const size_t g = (size_t)ngroups(nslots);
    const size_t c = g > 0 ? g * GROUP : 1;
    bits.ensure(g * (GROUP / 64) + 1);
    tpre.ensure(g * (GROUP / TILE) + 1);
    gcnt.ensure(g + 1);
    for (int k = 0; k < 3; k++) { hN[k].ensure(c); hY[k].ensure(c); hF[k].ensure(c); }
    huN.ensure(c);
    gp.ensure(GP * c);
    hE.ensure(c);
    sidx.ensure(c);
  }
  void ensureRoll(int64_t nslots) {
    const size_t g = (size_t)ngroups(nslots);
    gtq.ensure(3 * (g > 0 ? g * GROUP : 1));
  }
  void release() {
    bits.release(); tpre.release(); gcnt.release();
    for (int k = 0; k < 3; k++) { hN[k].release(); hY[k].release(); hF[k].release(); }
    huN.release(); gp.release(); sidx.release(); hE.release(); gtq.release();
    sidxValid = false;
  }
};
struct RealView {
  const unsigned long long* bits;
  const uint32_t* tpre;
};
// The used part of a group's range starts at a rotated offset (golden-ratio skew), so
// that the densely used regions of consecutive groups spread over all HBM channels
// instead of all starting at the same power-of-two alignment.
constexpr uint32_t GROUP_SKEW = 1265;
__host__ __device__ inline uint32_t state_index(uint32_t g, uint32_t local) {
  return g * (uint32_t)GROUP + ((local + g * GROUP_SKEW) & (uint32_t)(GROUP - 1));
}
// element i of a contact-state array: double in the FP64 form, float (first half) in FP32
template <int PREC>
__device__ __forceinline__ double ldr(const double* p, uint32_t i) {
  if (PREC == 32) return (double)reinterpret_cast<const float*>(p)[i];
  return p[i];
}
__device__ __forceinline__ bool rv_real(const RealView& v, uint32_t s) { return (v.bits[s >> 6] >> (s & 63)) & 1ull; }
// state index of real slot s
__device__ __forceinline__ uint32_t rv_index(const RealView& v, uint32_t s) {
  const uint32_t w = s >> 6;
  uint32_t r = v.tpre[s >> 8];
  for (uint32_t q = (s >> 8) << 2; q < w; q++) r += (uint32_t)__popcll(v.bits[q]);
  r += (uint32_t)__popcll(v.bits[w] & ((1ull << (s & 63)) - 1ull));
  return state_index(s / (uint32_t)GROUP, r);
}

// A wavefront's LDS writes become visible to its other lanes (no workgroup barrier): used
// where each wavefront reads and writes only its own LDS region.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int NPD = 24;  // permuted double arrays per particle
constexpr int NPU = 5;   // permuted u32 arrays per particle

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  wdg_field field{};
  std::vector<wdg_material> mats;  // the current material table
  // ContactLoop.updatePhys 'never' keeps each contact's phys from the materials it was created
  // with: a change of materials while contacts exist appends the new table (an epoch) after the
  // old ones in dMats, and every real contact carries the epoch it takes its phys from
  DBuf<wdg_material> dMats;
  std::vector<wdg_material> epTab;  // host copy of dMats: matEpochs tables of mats.size() each
  int matEpochs = 1;
  int64_t matGen = 0;
  // the last contact loop had rolling resistance: the real contacts' records have a torque
  bool rollActive = false;
  // ContactLoop.evalStress; clStressOn: the last contact loop evaluated it (clStress holds
  // the 9 components, else the stress is zero); clPart: its per-group partial sums
  bool evalStress = false, clStressOn = false;
  // a single dense cell grid orders the spheres by the Morton code of their cell
  bool cellMorton = true;
  // energy tracking of batched scenes: the terms per particle (pe: kinetic, grav, nonviscDamp
  // of the last leapfrog; pc: elast, plast, viscN, viscT of the last contact loop, each contact
  // counted at its pB), per contact (enC), and the particles of each scene in storage order
  DBuf<double> pe, pc, enC, energyS;  // energyS: WDG_EN_N terms per scene
  DBuf<double> enSPart;                // partial sums of the parts of each scene
  int64_t energySScenes = -1;          // the scene count energyS holds terms for
  DBuf<uint32_t> sceneList, sceneListTmp, sceneKey, sceneKeyTmp, sceneOff;
  int64_t sceneListGen = -1;
  DBuf<double> clStress, clPart;  // counts the changes of the epochs' numbering
  int updatePhys = 0;  // 0 never, 1 always, 2 once (then never)

  // particles, storage order
  int64_t n = 0;
  int64_t nW = 0;       // boundary bodies (walls, facets) occupy storage [0, nW)
  int64_t nFacet = 0;
  DBuf<double> fgeo;    // FGEO doubles per boundary body (facet geometry)
  // per boundary body: force on its 3 nodes (9), then torque (9), from the last IntraForce
  // (facets only)
  DBuf<double> fnode;
  // boundary bodies whose contact forces are gathered (walls: storage indices, facets left
  // out), the parts each one's list is cut into (from the largest list at the last refresh),
  // and their partial sums (up to 64 parts of 6 doubles per body)
  DBuf<uint32_t> gbody;
  std::vector<uint32_t> gbodyHost;
  int64_t nGBody = 0;
  int bndParts = 1;
  DBuf<double> bndPart;
  std::vector<uint32_t> bodyUid;  // user id of each boundary body
  DBuf<double> bnd;     // enlarged AABB of each boundary body at the last refresh (min3, max3)
  double rmin = INFINITY, rmax = 0;
  DBuf<double> pos[3], vel[3], angVel[3], ori[4], force[3], torque[3], inertia[3], lastPos[3];
  DBuf<double> mass, radius;
  DBuf<uint32_t> flags, mask, uid, sto;
  DBuf<uint32_t> scene;  // scene of each particle (batched independent scenes)
  int32_t nScenes = 0;   // 0: one scene; else dt per scene in dtScene
  int64_t maxScene = 0;  // largest scene index uploaded
  std::vector<uint32_t> bodyScene;        // scene of each boundary body (host)
  DBuf<uint32_t> sceneBodyOff, sceneBody;  // boundary bodies by scene (batched scenes)
  DBuf<double> dtScene;
  std::vector<double> dtSceneHost;  // the same, on the host (DynDt updates both)
  // per-scene Verlet distances of the last refresh (tuning.skinSteps > 0), and scratch for the
  // scenes' largest speeds
  DBuf<double> vdScene;
  DBuf<unsigned long long> sceneVmax;
  bool vdSceneActive = false;
  DBuf<unsigned long long> sceneBox;  // per scene: bounds (ordered bits) and sphere count
  DBuf<unsigned char> sgrid;          // per scene: its cell grid (batched aperiodic scenes)
  double refreshSec = 0;  // host time spent in refreshes so far
  int pairStage = 24;     // pairs staged per sphere by the refresh's pair pass
  int nbStage = 48;       // neighbours staged per sphere by it
  // graph runner: device words [0] stop, [1] steps executed, [2] renorm flag of the step,
  // [3..4] first step number (int64), [5] trsfRenorm
  DBuf<uint32_t> gstate;
  wdg_tuning tune{2, 8, 0};  // until wdg_tune / wdg_set_tuning: 2 force workgroups per group of slots
  int64_t gen = 0;  // bumped whenever buffers a captured graph uses may change
  bool graphCapture = false;  // kernels being launched are captured for replay
  struct GraphCache {
    hipGraphExec_t exec = nullptr;
    int64_t gen = -1;
    int K = 0, cfBlocks = 0, rcur = -1, parity = -1;
    bool track = false;  // captured with energy tracking
    bool intra = false;  // captured with IntraForce
    int epochs = 1, updAll = 0;  // material epochs and updatePhys of the capture
    bool evalStress = false;
    double dt = 0, damping = 0;
    wdg_contact_params cp{};
  } graph;
  DBuf<int32_t> mat;
  DBuf<double> spareD[NPD];
  DBuf<uint32_t> spareU[NPU];

  // slot lists, double-buffered across refreshes
  Contacts C[2];
  int cur = 0;
  Contacts& con() { return C[cur]; }
  // real contacts of con(), double-buffered across contact loops and refreshes
  Reals R[2];
  int rcur = 0;
  Reals& reals() { return R[rcur]; }
  RealView rview(int k) const { return RealView{R[k].bits.p, R[k].tpre.p}; }
  RealView rview() const { return rview(rcur); }

  // broadphase scratch
  DBuf<uint32_t> cellKey, sortIdx, cellKeyTmp, sortIdxTmp, perm, invPerm;
  DBuf<uint32_t> cellStart;  // per cell key (and one past the last): first sphere of that cell or later
  DBuf<uint32_t> scount, soff;          // pairs per generating sphere, their offsets
  DBuf<uint32_t> stPartner, stInfo;     // staged pairs, a fixed number per sphere
  DBuf<uint64_t> stDist;
  DBuf<uint32_t> stNb, sdeg;            // staged neighbours per sphere, their number
  DBuf<uint32_t> pairMax;               // per pair-generation workgroup: largest pair / neighbour count
  DBuf<float4> posf;                    // float candidate records of the last refresh
  DBuf<uint64_t> adjEnt;                // adjacency entries with their sort key
  DBuf<uint32_t> adjDeg;
  // adjacency of high-degree particles (boundary bodies), sorted by one radix sort
  DBuf<uint32_t> bigOff, bigRank, bigVal, bigValTmp, bigDest, bigList;
  DBuf<uint64_t> bigKey, bigKeyTmp;
  DBuf<uint32_t> sortHist, scanAux;
  DBuf<uint32_t> act;     // slot of each state index (contact loop), per-slot flags (relaxation)
  DBuf<uint4> desc;       // contact loop: the filter's record of each real slot, by state index
  DBuf<uint64_t> descDist;
  // in-memory snapshots of the kinematic state (pos, vel, angVel, ori; 13 doubles per
  // particle, by user id), valid while the particle set they were taken of is loaded
  static constexpr int NSNAP = 4;
  DBuf<double> snap[NSNAP];
  int64_t snapN[NSNAP] = {-1, -1, -1, -1};
  uint64_t snapUpload[NSNAP] = {0, 0, 0, 0};
  // contact history of each snapshot (FP64 contact form; none in the FP32 form): the keys
  // (min uid << 32 | max uid) of the slot list's slots, sorted, real contacts first (the others
  // hold n << 32), with the slot each had; per slot its record in the wdg_contact_list layout
  // and its packed cellDist; the number of real contacts and the Verlet distance of that list
  DBuf<uint64_t> snapKey[NSNAP], snapCd[NSNAP];
  DBuf<uint32_t> snapSlot[NSNAP];
  DBuf<double> snapRec[NSNAP];
  int64_t snapReal[NSNAP] = {0, 0, 0, 0};
  double snapVerlet[NSNAP] = {NAN, NAN, NAN, NAN};
  // each saved contact's material epoch, valid while matGen is the saved one
  DBuf<uint8_t> snapEp[NSNAP];
  int64_t snapMatGen[NSNAP] = {-1, -1, -1, -1};
  DBuf<uint32_t> snapMatch;  // per slot of a restored list: 1 + its slot in the snapshot, or 0
  uint64_t uploads = 0;   // particle uploads so far
  uint64_t serial = 0;    // unique per context of the process
  // Scene.energy (EnergyTracker) while trackEnergy is set: the terms in WDG_EN_* order, and
  // per-workgroup partial sums of the kernel that last produced them
  bool trackEnergy = false;
  int64_t stepNow = 0;    // Scene.step of the step being run, from wdg_run or the last contact
                          // loop (Leapfrog does not track a periodic scene's step 0)
  DBuf<double> energy;
  DBuf<double> enPart;
  DBuf<double> red;       // reductions
  DBuf<double> xfer;      // user-id-ordered staging for upload/download
  DBuf<uint32_t> dflags;  // [0] dirty, [1] error bits, [2..] scratch counters
  uint32_t* hflags = nullptr;  // pinned host mirror

  // grid of the last refresh (for the on-demand potential-set report)
  struct GridState {
    double org[3], h[3], L[3];
    int nc[3];
    int periodic;
    int hashed;       // cell keys are hashed into hmask + 1 buckets (sparse aperiodic scenes)
    uint32_t hmask;
    int morton;       // dense grid with Morton-coded cell keys
    const void* sg;   // batched aperiodic scenes: the grid of every scene (in sgrid), else null
    bool brute;
    bool valid = false;
  } grid;

  // engine state
  bool boundsValid = false;
  bool dirtyKnown = false;     // dflags[dirtyIdx] reflects the current positions
  int dirtyIdx = 0;            // 0 (refresh, k_dirty) or DIRTY_PP + parity (leapfrog)
  int lfParity = 0;            // flag word the next leapfrog launch sets
  bool forcesPending = false;  // contact forces not yet gathered into force/torque
  // force/torque of every integrated particle hold the reset value (m*gravity unless
  // gravity-skipped, zero torque), so the integrator computes them instead of loading
  bool forceClean = false;
  // the slot list's per-entry real slots (Contacts::eidx) describe the current real set;
  // eidxClean: they all hold NO_STATE (after a refresh)
  bool eidxValid = false, eidxClean = false;
  double verletDist = NAN;
  int64_t nRefresh = 0, nCollide = 0;
  // graph runner: the step of the last refresh and the running mean of the refresh interval
  int64_t lastRefreshStep = -1;
  double refreshEvery = 0;
  // hflags[0 .. DIRTY_PP + 1] mirror dflags after a graph batch (the next collider call uses them)
  bool hflagsFresh = false;
  int64_t dynStepLast = -1, dynStepPrev = -1;
  // contact form: 64 (plain FP64) or 32 (contact geometry and law in FP32)
  int precision = 64;
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
// Law2_L6Geom_FrictPhys_IdealElPl with rolling and twisting resistance (IdealElPl.cpp:94)
inline bool rolling(const wdg_contact_params& p) {
  return p.lawKind == WDG_LAW_IDEALELPL && p.relRollStiff > 0. && p.rollTanPhi > 0.;
}
void leapfrog(Ctx& c, const wdg_leapfrog_params* p);
void force_reset(Ctx& c);
void flush_forces(Ctx& c);
// IntraForce([In2_Facet]): the contact forces of each facet spread to its nodes (fnode)
void intra_force(Ctx& c);
void run_graph(Ctx& c, const wdg_program* prog, int64_t nSteps, int64_t* step, double* dt);
uint32_t read_u32(Ctx& c, const uint32_t* d);
int64_t tune(Ctx& c, const wdg_program* prog, int64_t* step, double* dt);
double crit_dt(Ctx& c);
// DynDt's critical timestep of every scene of a batched context (inf: no stiffness)
void crit_dt_scenes(Ctx& c, std::vector<double>& out);
void observables(Ctx& c, const double lo[3], const double hi[3], double out[3]);
int64_t count_real(Ctx& c);
void stress(Ctx& c, double volume, double out[9]);
double kinetic_energy(Ctx& c);
void scratch_setup(int device);
void kinetic_energy_noalloc(Ctx& c, double* dOut);
// bits + tile prefix of R[dst] from a per-slot flag array (nonzero = real)
void reals_from_flags(Ctx& c, const uint32_t* flag, int64_t nslots, int dst);
void check_errors(Ctx& c);
void relax(Ctx& c, int maxIter, double inflate, double tol, double* maxOverlap, int* iters);
void drop_contacts(Ctx& c);
// a new material table (ContactLoop.updatePhys: the old tables the real contacts still refer to
// are kept as earlier epochs)
void set_materials(Ctx& c, const wdg_material* mats, int n);
void snapshot_save(Ctx& c, int slot);
// the contact history part of a snapshot: kept on the device (FP64 form), and restored after a
// full refresh at the snapshot's positions
void snapshot_contacts_save(Ctx& c, int slot);
void snapshot_contacts_load(Ctx& c, int slot);
// the real set of the current slot list emptied (contact history restarts; the slots stay)
void clear_history(Ctx& c);
void snapshot_load(Ctx& c, int slot);
double snapshot_distance(Ctx& c, int slot);
// every per-entry slot of the current slot list to NO_STATE (and EIDX_FULL set)
void eidx_clear(Ctx& c);
void energy_enable(Ctx& c, bool on);
void energy_clear(Ctx& c);
void energy_step_begin(Ctx& c);
// energy[term] = 0 (graph-replay safe)
void energy_zero(Ctx& c, int term);
// adds the sums over nBlocks workgroup partials of nTerms terms to energy[target[q]] (bit q of
// setMask: stores it instead)
void energy_add(Ctx& c, int64_t nBlocks, int nTerms, const int* target, int setMask);
void energy_read(Ctx& c, double out[7]);
// energy tracking per scene (a batched context of several scenes)
inline bool energy_per_scene(const Ctx& c) { return c.trackEnergy && c.nScenes > 1; }
// sizes the per-particle and per-contact buffers for nc slots and lists each scene's particles
// in storage order (again after a change of order); no allocation when called again unchanged
void energy_scenes_prepare(Ctx& c, int64_t nc);
// after the contact loop: each contact's four terms at its pB, then summed per scene
void energy_scenes_contacts(Ctx& c);
// after the leapfrog: its three terms per particle summed per scene
void energy_scenes_particles(Ctx& c);
void energy_scenes_read(Ctx& c, std::vector<double>& out);

// the sum over a workgroup of BS threads of NT values per thread, stored by thread 0..NT-1 at
// out[0..NT-1]; a fixed order (butterfly in the wavefront, then wavefronts in order), so the
// result does not vary from run to run. Every thread of the workgroup must call it.
template <int NT, int BS>
__device__ __forceinline__ void block_sum_store(double (&v)[NT], double* out) {
  __shared__ double sh[BS / 64][NT];
#pragma unroll
  for (int q = 0; q < NT; q++)
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NT; q++) sh[w][q] = v[q];
  __syncthreads();
  if (threadIdx.x < NT) {
    double s = 0.;
    for (int k = 0; k < BS / 64; k++) s += sh[k][threadIdx.x];
    out[threadIdx.x] = s;
  }
}

inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Workgroups are handed to the 8 XCDs round-robin (workgroup b runs on XCD b % 8), and
// each XCD has its own L2. This maps the hardware workgroup index to a logical one so
// that each XCD receives one contiguous run of logical workgroups: neighbouring work
// (neighbouring particles and slots) shares an L2, and work that is uneven along the
// logical index is spread evenly over the XCDs.
constexpr uint32_t NXCD = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t x = b % NXCD, i = b / NXCD, q = nb / NXCD, r = nb % NXCD;
  return x * q + (x < r ? x : r) + i;
}
inline int bits_for(uint64_t maxval) {
  int b = 1;
  while (b < 64 && (maxval >> b) != 0) b++;
  return b;
}

// the engine state behind a C-ABI handle
Ctx& ctx_of(wdg_ctx* h);
}  // namespace wdg
