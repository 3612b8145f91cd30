// Broadphase: Verlet-skin refresh test, spatial re-sort of the particle
// storage, cell-list pair generation over radix-sorted cell keys, and the merge
// that carries contact history across refreshes. Replaces
// InsertionSortCollider::run (InsertionSortCollider.cpp:386-613).
//
// Pair-set semantics kept from the reference:
//  * AABB of a sphere = centre -/+ (r*distFactor') -/+ verletDist, of a wall =
//    infinite except along its axis (Sphere.cpp:52-82, Wall.cpp:25-38,
//    Collision.cpp:125-162);
//  * refresh when any particle moved more than verletDist since the last one
//    (Collision.cpp:80-123);
//  * potential set after a refresh = {pairs whose enlarged AABBs overlap and that
//    mayCollide} U {real contacts}; a broken contact whose AABBs did not overlap
//    is dropped (InsertionSortCollider.hpp:213-218).
// The step kernels only need the pairs that can touch before the next refresh:
// the slots hold the AABB pairs whose centres are within r1+r2+2*verletDist
// (each particle moves at most verletDist until then), plus carried real
// contacts. The full potential set is regenerated on request (potential_pairs).
//
// Refresh: spheres are re-sorted by cell (storage order), the particle arrays
// permuted, pairs generated per sphere over the 27 neighbour cells, then one
// stable radix sort of [new pairs ; old real slots] makes equal keys adjacent
// so that the compaction carries history; finally the per-particle adjacency is
// sorted by (particle, other user id), the canonical summation order.
#include <cstring>

#include "wdg_internal.hpp"
#include "wdg_math.hpp"

namespace wdg {

struct Grid {
  double org[3];
  double h[3];
  int nc[3];
  int periodic;
  double L[3];
};

struct AabbParams {
  double vd;      // verletDist (absolute)
  double absDf;   // |distFactor|
  uint32_t loneMask;
  int periodic;
  double L[3];
};

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ bool may_collide(uint32_t ma, uint32_t mb, uint32_t lone) {
  return (ma & mb) != 0u && (ma & mb & lone) == 0u;
}
// sphere AABB bounds along one axis, same operations as updateAabb
__device__ __forceinline__ void sph_bounds(double p, double hs, double vd, double& lo, double& hi) {
  lo = p - hs;
  hi = p + hs;
  if (vd > 0) {
    hi = hi + vd;
    lo = lo - vd;
  }
}
// nearest periodic image of B relative to A, from refresh positions
__device__ __forceinline__ int near_img(double pa, double pb, double L) {
  double d = (pb - pa) / L;
  return -(int)floor(d + 0.5);
}
// AABB overlap of spheres a (pA) and b (pB) with b shifted by cd; counts overlapping images when periodic
__device__ __forceinline__ bool sph_overlap(const double pa[3], double hsa, const double pb[3], double hsb,
                                            const AabbParams& P, int cd[3], int& nimg) {
  bool ok = true;
  nimg = 1;
  for (int k = 0; k < 3; k++) {
    double alo, ahi, blo, bhi;
    sph_bounds(pa[k], hsa, P.vd, alo, ahi);
    sph_bounds(pb[k], hsb, P.vd, blo, bhi);
    if (P.periodic) {
      cd[k] = near_img(pa[k], pb[k], P.L[k]);
      double s = P.L[k] * cd[k];
      if (!(alo <= bhi + s && ahi >= blo + s)) ok = false;
      int cnt = 0;
      for (int dd = -1; dd <= 1; dd++) {
        double s2 = P.L[k] * (cd[k] + dd);
        if (alo <= bhi + s2 && ahi >= blo + s2) cnt++;
      }
      nimg *= cnt;
    } else {
      cd[k] = 0;
      if (!(alo <= bhi + 0. && ahi >= blo + 0.)) ok = false;
    }
  }
  return ok;
}

// ------------------------------------------------------------------ refresh test
__global__ void k_dirty(int64_t n, const double* __restrict__ px, const double* __restrict__ py,
                        const double* __restrict__ pz, const double* __restrict__ lx, const double* __restrict__ ly,
                        const double* __restrict__ lz, double maxD2, uint32_t* dflags) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool dirty = false;
  if (i < n) {
    d3 d = sub3(mk3(lx[i], ly[i], lz[i]), mk3(px[i], py[i], pz[i]));
    double d2 = 0;
    d2 = fmax(d2, sqn3(d));
    dirty = d2 + 0. > maxD2;
  }
  if (__ballot(dirty) != 0ull && (threadIdx.x & 63) == 0) atomicOr(&dflags[0], 1u);
}

// order-preserving map of doubles to unsigned for atomic min/max
__device__ __forceinline__ unsigned long long ord_bits(double x) {
  unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | (1ull << 63));
}
static inline double ord_to_double(unsigned long long b) {
  unsigned long long r = (b >> 63) ? (b & ~(1ull << 63)) : ~b;
  double d;
  std::memcpy(&d, &r, 8);
  return d;
}

__global__ void k_bbox(int64_t s0, int64_t n, const double* __restrict__ px, const double* __restrict__ py,
                       const double* __restrict__ pz, unsigned long long* out) {
  int64_t s = s0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long mn[3] = {~0ull, ~0ull, ~0ull}, mx[3] = {0, 0, 0};
  if (s < n) {
    double p[3] = {px[s], py[s], pz[s]};
    for (int k = 0; k < 3; k++) mn[k] = mx[k] = ord_bits(p[k]);
  }
  for (int k = 0; k < 3; k++) {
    for (int off = 32; off > 0; off >>= 1) {
      unsigned long long a = __shfl_xor(mn[k], off);
      unsigned long long b = __shfl_xor(mx[k], off);
      mn[k] = a < mn[k] ? a : mn[k];
      mx[k] = b > mx[k] ? b : mx[k];
    }
  }
  if ((threadIdx.x & 63) == 0) {
    for (int k = 0; k < 3; k++) {
      atomicMin(&out[k], mn[k]);
      atomicMax(&out[3 + k], mx[k]);
    }
  }
}

__device__ __forceinline__ uint32_t cell_key(const double p[3], const Grid& g) {
  int c[3];
  for (int k = 0; k < 3; k++) {
    double x = p[k] - g.org[k];
    if (g.periodic) x = x - floor(x / g.L[k]) * g.L[k];
    int ci = (int)floor(x / g.h[k]);
    c[k] = min(max(ci, 0), g.nc[k] - 1);
  }
  return (uint32_t)(((int64_t)c[2] * g.nc[1] + c[1]) * g.nc[0] + c[0]);
}

// cell key of every sphere (storage [s0,n)); value = current storage index
__global__ void k_keys(int64_t s0, int64_t n, const double* __restrict__ px, const double* __restrict__ py,
                       const double* __restrict__ pz, Grid g, uint32_t* key, uint32_t* idx) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t s = s0 + t;
  if (s >= n) return;
  double p[3] = {px[s], py[s], pz[s]};
  key[t] = cell_key(p, g);
  idx[t] = (uint32_t)s;
}

// perm[new storage] = old storage; walls keep their place
__global__ void k_build_perm(int64_t nW, int64_t n, const uint32_t* __restrict__ sorted, uint32_t* perm,
                             uint32_t* invPerm) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t o = i < nW ? (uint32_t)i : sorted[i - nW];
  perm[i] = o;
  invPerm[o] = (uint32_t)i;
}

struct PermArgs {
  int64_t n;
  const uint32_t* perm;
  const double* sd[NPD];
  double* dd[NPD];
  const uint32_t* su[NPU];
  uint32_t* du[NPU];
};
__global__ __launch_bounds__(256) void k_permute(PermArgs a) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t o = a.perm[i];
#pragma unroll
  for (int k = 0; k < NPD; k++) a.dd[k][i] = a.sd[k][o];
#pragma unroll
  for (int k = 0; k < NPU; k++) a.du[k][i] = a.su[k][o];
}

__global__ void k_sto(int64_t n, const uint32_t* __restrict__ uid, uint32_t* sto) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) sto[uid[i]] = (uint32_t)i;
}

__global__ void k_last_pos(int64_t n, const double* __restrict__ px, const double* __restrict__ py,
                           const double* __restrict__ pz, double* lx, double* ly, double* lz) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  lx[i] = px[i];
  ly[i] = py[i];
  lz[i] = pz[i];
}

__global__ void k_cell_ranges(int64_t nSph, int64_t s0, const uint32_t* __restrict__ key, uint32_t* cstart,
                              uint32_t* cend) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nSph) return;
  uint32_t k = key[t];
  if (t == 0 || key[t - 1] != k) cstart[k] = (uint32_t)(s0 + t);
  if (t == nSph - 1 || key[t + 1] != k) cend[k] = (uint32_t)(s0 + t) + 1;
}

enum { MODE_NEAR = 0, MODE_AABB = 1 };

struct PairArgs {
  int64_t nW, n;
  const uint32_t* key;  // cell key of storage nW+t
  const double *px, *py, *pz;  // refresh positions (lastPos), storage order
  const double* rad;
  const uint32_t *mask, *uid, *flags;
  const uint32_t *cstart, *cend;
  Grid g;
  AabbParams P;
  int brute;
  int idBits;
  // outputs
  uint32_t* count;
  const uint32_t* off;
  uint64_t* pkey;    // MODE_NEAR: (lo<<b)|hi storage ; MODE_AABB: (uidA<<32)|uidB
  uint32_t* pinfo;
  uint64_t* pdist;
  uint32_t* dflags;
};

template <bool WRITE, int MODE>
__global__ __launch_bounds__(256) void k_pairs(PairArgs a) {
  const int64_t s = a.nW + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.n) return;
  const double ps[3] = {a.px[s], a.py[s], a.pz[s]};
  const double rs = a.rad[s];
  const double hs = a.P.absDf * rs;
  const uint32_t ms = a.mask[s], us = a.uid[s];
  uint32_t cnt = 0;
  uint32_t o = WRITE ? a.off[s - a.nW] : 0;
  auto consider = [&](int64_t j) {
    if (j <= s) return;
    if (!may_collide(ms, a.mask[j], a.P.loneMask)) return;
    const double pj[3] = {a.px[j], a.py[j], a.pz[j]};
    const double rj = a.rad[j];
    const uint32_t uj = a.uid[j];
    // canonical roles: the smaller user id is pA
    const bool sA = us < uj;
    const double* pa = sA ? ps : pj;
    const double* pb = sA ? pj : ps;
    const double ra = sA ? rs : rj, rb = sA ? rj : rs;
    int cd[3], nimg;
    if (!sph_overlap(pa, a.P.absDf * ra, pb, a.P.absDf * rb, a.P, cd, nimg)) return;
    if (a.P.periodic && nimg > 1) {
      atomicOr(&a.dflags[1], ERR_DOUBLE_IMAGE);
      return;
    }
    if (MODE == MODE_NEAR) {
      // centres within reach before the next refresh
      d3 d = mk3(pb[0] + a.P.L[0] * cd[0] * a.P.periodic - pa[0], pb[1] + a.P.L[1] * cd[1] * a.P.periodic - pa[1],
                 pb[2] + a.P.L[2] * cd[2] * a.P.periodic - pa[2]);
      double reach = a.P.absDf * (ra + rb) + 2 * a.P.vd;
      if (sqn3(d) > reach * reach * (1 + 1e-9)) return;
    }
    if (WRITE) {
      if (MODE == MODE_NEAR) {
        a.pkey[o + cnt] = ((uint64_t)s << a.idBits) | (uint64_t)j;
        a.pinfo[o + cnt] = (uint32_t)CT_SS | CI_NEAR | (sA ? 0u : CI_SWAP);
      } else {
        a.pkey[o + cnt] = sA ? (((uint64_t)us << 32) | uj) : (((uint64_t)uj << 32) | us);
      }
      if (a.P.periodic) a.pdist[o + cnt] = pack_cd(cd[0], cd[1], cd[2]);
    }
    cnt++;
  };
  if (a.brute) {
    for (int64_t j = s + 1; j < a.n; j++) consider(j);
  } else {
    uint32_t k = a.key[s - a.nW];
    int c[3];
    c[0] = (int)(k % (uint32_t)a.g.nc[0]);
    c[1] = (int)((k / (uint32_t)a.g.nc[0]) % (uint32_t)a.g.nc[1]);
    c[2] = (int)(k / ((uint32_t)a.g.nc[0] * (uint32_t)a.g.nc[1]));
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          int q[3] = {c[0] + dx, c[1] + dy, c[2] + dz};
          bool skip = false;
          for (int m = 0; m < 3; m++) {
            if (q[m] < 0 || q[m] >= a.g.nc[m]) {
              if (!a.g.periodic) skip = true;
              q[m] = (q[m] + a.g.nc[m]) % a.g.nc[m];
            }
          }
          if (skip) continue;
          uint32_t qk = (uint32_t)(((int64_t)q[2] * a.g.nc[1] + q[1]) * a.g.nc[0] + q[0]);
          uint32_t t0 = a.cstart[qk], t1 = a.cend[qk];
          for (uint32_t t = t0; t < t1; t++) consider(t);
        }
  }
  // walls (storage [0,nW)) against every sphere; the wall is pA
  for (int64_t w = 0; w < a.nW; w++) {
    if (!may_collide(a.mask[w], ms, a.P.loneMask)) continue;
    uint32_t fl = a.flags[w];
    int ax = flag_axis(fl);
    double wp = ax == 0 ? a.px[w] : (ax == 1 ? a.py[w] : a.pz[w]);
    double wlo = wp, whi = wp;
    if (a.P.vd > 0) {
      whi = whi + a.P.vd;
      wlo = wlo - a.P.vd;
    }
    double slo, shi;
    sph_bounds(ps[ax], hs, a.P.vd, slo, shi);
    if (!(wlo <= shi + 0. && whi >= slo + 0.)) continue;
    if (WRITE) {
      if (MODE == MODE_NEAR) {
        a.pkey[o + cnt] = ((uint64_t)w << a.idBits) | (uint64_t)s;
        a.pinfo[o + cnt] = (uint32_t)CT_WS | CI_NEAR;
      } else {
        a.pkey[o + cnt] = ((uint64_t)a.uid[w] << 32) | us;
      }
      if (a.P.periodic) a.pdist[o + cnt] = pack_cd(0, 0, 0);
    }
    cnt++;
  }
  if (!WRITE) a.count[s - a.nW] = cnt;
}

// old real slots appended behind the new pairs, re-keyed in the new storage order;
// non-real ones get the sentinel key
__global__ void k_append_old(int64_t nOld, int64_t nPairs, const uint32_t* __restrict__ A,
                             const uint32_t* __restrict__ B, RealView rv,
                             const uint32_t* __restrict__ invPerm, int idBits, uint64_t sentinel, uint64_t* pkey) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nOld) return;
  uint64_t k = sentinel;
  if (rv_real(rv, (uint32_t)c)) {
    uint32_t a = invPerm[A[c]], b = invPerm[B[c]];
    uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    k = ((uint64_t)lo << idBits) | (uint64_t)hi;
  }
  pkey[nPairs + c] = k;
}

__global__ void k_iota(int64_t n, uint32_t* v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (uint32_t)i;
}

__global__ void k_run_flags(int64_t n, const uint64_t* __restrict__ key, uint64_t sentinel, uint32_t* flag) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  uint64_t k = key[s];
  flag[s] = (k != sentinel && (s == 0 || key[s - 1] != k)) ? 1u : 0u;
}

struct CompactArgs {
  int64_t n;       // sorted entries
  int64_t nPairs;  // entries below this value index are new pairs
  const uint64_t* key;
  const uint32_t* val;
  const uint32_t* flag;
  const uint32_t* off;
  const uint32_t* pinfo;
  const uint64_t* pdist;
  const uint32_t* invPerm;
  int idBits;
  int periodic;
  // old set
  const uint32_t *oA, *oB, *oInfo;
  const uint64_t* oDist;
  // new set
  uint32_t *A, *B, *info, *carry;
  uint64_t* dist;
};

__global__ void k_compact(CompactArgs a) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.n || !a.flag[s]) return;
  const uint64_t k = a.key[s];
  const uint32_t o = a.off[s];
  int64_t e0 = a.val[s];
  int64_t e1 = (s + 1 < a.n && a.key[s + 1] == k) ? (int64_t)a.val[s + 1] : -1;
  int64_t bp = -1, old = -1;
  if (e0 < a.nPairs) bp = e0; else old = e0 - a.nPairs;
  if (e1 >= 0) {
    if (e1 < a.nPairs) bp = e1; else old = e1 - a.nPairs;
  }
  const uint32_t lo = (uint32_t)(k >> a.idBits), hi = (uint32_t)(k & ((1ull << a.idBits) - 1));
  a.A[o] = lo;
  a.B[o] = hi;
  uint32_t inf;
  uint64_t dist = 0;
  if (bp >= 0) {
    inf = a.pinfo[bp];
    if (a.periodic) dist = a.pdist[bp];
  } else {
    // roles are fixed by user ids: find where the old pA went
    const uint32_t oi = a.oInfo[old];
    const uint32_t oldPA = (oi & CI_SWAP) ? a.oB[old] : a.oA[old];
    const uint32_t newPA = a.invPerm[oldPA];
    inf = (oi & CI_TYPE_MASK) | (newPA == hi ? CI_SWAP : 0u);
  }
  if (old >= 0 && a.periodic) dist = a.oDist[old];
  a.info[o] = inf;
  if (a.periodic) a.dist[o] = dist;
  a.carry[o] = old >= 0 ? (uint32_t)old + 1u : 0u;
}

// contact state of the carried real contacts: old rank -> new rank
struct CarryArgs {
  int64_t n;
  const uint32_t* carry;
  RealView ov, nv;
  const double* o[10];
  double* d[10];
  const double* ogp;
  double* gp;
};
__global__ void k_carry_state(CarryArgs a) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.n) return;
  const uint32_t cy = a.carry[s];
  if (!cy) return;
  const uint32_t ko = rv_index(a.ov, cy - 1u), kn = rv_index(a.nv, (uint32_t)s);
#pragma unroll
  for (int q = 0; q < 10; q++) a.d[q][kn] = a.o[q][ko];
  const double2* src = reinterpret_cast<const double2*>(a.ogp + (int64_t)GP * ko);
  double2* dst = reinterpret_cast<double2*>(a.gp + (int64_t)GP * kn);
#pragma unroll
  for (int q = 0; q < GP / 2; q++) dst[q] = src[q];
}

// two adjacency entries per slot, keyed (particle storage, other user id)
__global__ void k_adj_keys(int64_t nc, const uint32_t* __restrict__ A, const uint32_t* __restrict__ B,
                           const uint32_t* __restrict__ info, const uint32_t* __restrict__ uid, int idBits,
                           uint64_t* key, uint32_t* val) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const uint32_t lo = A[c], hi = B[c];
  const bool swap = (info[c] & CI_SWAP) != 0u;
  key[2 * c] = ((uint64_t)lo << idBits) | (uint64_t)uid[hi];
  val[2 * c] = (uint32_t)c | (swap ? 0u : ADJ_PA);
  key[2 * c + 1] = ((uint64_t)hi << idBits) | (uint64_t)uid[lo];
  val[2 * c + 1] = (uint32_t)c | (swap ? ADJ_PA : 0u);
}

__global__ void k_key_hi(int64_t n, const uint64_t* __restrict__ key, int idBits, uint32_t* part) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) part[i] = (uint32_t)(key[i] >> idBits);
}

// CSR offsets from a sorted id column: off[id] = first index with that id
__global__ void k_offsets(int64_t nc, int64_t npar, const uint32_t* __restrict__ ids, uint32_t* off) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > nc) return;
  int64_t prev = t == 0 ? -1 : (int64_t)ids[t - 1];
  int64_t cur = t == nc ? npar : (int64_t)ids[t];
  for (int64_t id = prev + 1; id <= cur; id++) off[id] = (uint32_t)t;
}

// ------------------------------------------------------------------ host side
static void check_device_errors(Ctx& c, uint32_t err) {
  if (err & ERR_DOUBLE_IMAGE)
    throw Error(WDG_E_GEOM, "InsertionSortCollider: a pair overlaps through two periodic images "
                            "(particle larger than half of the cell?)");
  if (err & ERR_NAN_FORCE) throw Error(WDG_E_NUMERIC, "NaN force in contact");
  if (err & ERR_CD_RANGE) throw Error(WDG_E_GEOM, "periodic cell distance out of range");
}

void check_errors(Ctx& c) {
  uint32_t err = read_u32(c, c.dflags.p + 1);
  if (err) {
    WDG_HIP(hipMemsetAsync(c.dflags.p + 1, 0, sizeof(uint32_t), c.stream));
    check_device_errors(c, err);
  }
}

void drop_contacts(Ctx& c) {
  for (auto& s : c.C) {
    s.n = 0;
    s.ensure(0, c.n, c.field.periodic != 0);
    WDG_HIP(hipMemsetAsync(s.adjOff.p, 0, (c.n + 1) * sizeof(uint32_t), c.stream));
  }
  for (auto& r : c.R) {
    r.ensure(0);
    WDG_HIP(hipMemsetAsync(r.tpre.p, 0, sizeof(uint32_t), c.stream));
  }
  c.boundsValid = false;
  c.dirtyKnown = false;
  c.forcesPending = false;
  c.grid.valid = false;
}

static PairArgs pair_args(Ctx& c, const Grid& g, const AabbParams& P, int idBits) {
  PairArgs pa{};
  pa.nW = c.nW; pa.n = c.n; pa.key = c.cellKey.p;
  pa.px = c.lastPos[0].p; pa.py = c.lastPos[1].p; pa.pz = c.lastPos[2].p; pa.rad = c.radius.p;
  pa.mask = c.mask.p; pa.uid = c.uid.p; pa.flags = c.flags.p;
  pa.cstart = c.cellStart.p; pa.cend = c.cellEnd.p; pa.g = g; pa.P = P; pa.brute = c.grid.brute;
  pa.idBits = idBits; pa.count = c.scount.p; pa.off = c.soff.p; pa.dflags = c.dflags.p;
  return pa;
}

static void refresh(Ctx& c) {
  const int64_t n = c.n, nW = c.nW, nSph = n - nW;
  const double vd = c.verletDist;
  const bool periodic = c.field.periodic != 0;
  c.nRefresh++;

  // grid: cell edge >= largest enlarged-AABB extent
  Grid g{};
  g.periodic = periodic;
  double absDf = std::fabs(c.field.distFactor);
  double ext = 2. * (absDf * c.rmax + vd);
  ext = ext * (1 + 1e-9) + 1e-300;
  bool brute = nSph < 512;
  if (periodic) {
    for (int k = 0; k < 3; k++) {
      g.L[k] = c.field.cell[k];
      g.nc[k] = (int)std::floor(g.L[k] / ext);
      if (g.nc[k] < 3) brute = true;
      g.nc[k] = std::max(g.nc[k], 1);
      g.h[k] = g.L[k] / g.nc[k];
      g.org[k] = 0;
    }
  } else if (nSph > 0) {
    unsigned long long init[6] = {~0ull, ~0ull, ~0ull, 0, 0, 0};
    unsigned long long* dbb = reinterpret_cast<unsigned long long*>(c.red.p);
    WDG_HIP(hipMemcpyAsync(dbb, init, sizeof(init), hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_bbox, dim3(nblocks(nSph, 256)), dim3(256), 0, c.stream, nW, n, c.pos[0].p, c.pos[1].p,
                       c.pos[2].p, dbb);
    WDG_LAUNCH_CHECK();
    unsigned long long hb[6];
    WDG_HIP(hipMemcpyAsync(hb, dbb, sizeof(hb), hipMemcpyDeviceToHost, c.stream));
    WDG_HIP(hipStreamSynchronize(c.stream));
    double lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
      lo[k] = ord_to_double(hb[k]);
      hi[k] = ord_to_double(hb[3 + k]);
    }
    double h = ext;
    for (int it = 0; it < 200; it++) {
      double cells = 1;
      for (int k = 0; k < 3; k++) {
        g.h[k] = h;
        g.nc[k] = (int)std::min(1e6, std::floor((hi[k] - lo[k]) / h) + 1);
        cells *= g.nc[k];
      }
      if (cells <= std::max(8.0 * nSph, 4096.0) && cells < 2e9) break;
      h *= 1.25;
    }
    for (int k = 0; k < 3; k++) g.org[k] = lo[k];
  }
  int64_t ncells = (int64_t)g.nc[0] * g.nc[1] * g.nc[2];

  AabbParams P{};
  P.vd = vd;
  P.absDf = absDf;
  P.loneMask = c.field.loneMask;
  P.periodic = periodic;
  for (int k = 0; k < 3; k++) P.L[k] = c.field.cell[k];

  // 1. spatial order of the spheres and the permutation of all particle arrays
  c.cellKey.ensure(nSph + 1); c.sortIdx.ensure(nSph + 1);
  c.cellKeyTmp.ensure(nSph + 1); c.sortIdxTmp.ensure(nSph + 1);
  c.perm.ensure(n + 1); c.invPerm.ensure(n + 1);
  c.scount.ensure(nSph + 1); c.soff.ensure(nSph + 1);
  if (nSph > 0) {
    Grid gk = g;
    if (brute) { for (int k = 0; k < 3; k++) { gk.nc[k] = 1; gk.h[k] = 1e300; gk.org[k] = 0; } gk.periodic = 0; }
    hipLaunchKernelGGL(k_keys, dim3(nblocks(nSph, 256)), dim3(256), 0, c.stream, nW, n, c.pos[0].p, c.pos[1].p,
                       c.pos[2].p, gk, c.cellKey.p, c.sortIdx.p);
    WDG_LAUNCH_CHECK();
    if (!brute)
      radix_sort_pairs<uint32_t>(c, c.cellKey.p, c.sortIdx.p, c.cellKeyTmp.p, c.sortIdxTmp.p, nSph,
                                 bits_for((uint64_t)(ncells > 1 ? ncells - 1 : 1)));
  }
  hipLaunchKernelGGL(k_build_perm, dim3(nblocks(n, 256)), dim3(256), 0, c.stream, nW, n, c.sortIdx.p, c.perm.p,
                     c.invPerm.p);
  WDG_LAUNCH_CHECK();
  {
    PermArgs pa{};
    pa.n = n;
    pa.perm = c.perm.p;
    DBuf<double>* d[NPD] = {&c.pos[0], &c.pos[1], &c.pos[2], &c.vel[0], &c.vel[1], &c.vel[2], &c.angVel[0],
                            &c.angVel[1], &c.angVel[2], &c.ori[0], &c.ori[1], &c.ori[2], &c.ori[3], &c.force[0],
                            &c.force[1], &c.force[2], &c.torque[0], &c.torque[1], &c.torque[2], &c.inertia[0],
                            &c.inertia[1], &c.inertia[2], &c.mass, &c.radius};
    DBuf<uint32_t>* u[NPU] = {&c.flags, &c.mask, &c.uid, reinterpret_cast<DBuf<uint32_t>*>(&c.mat)};
    for (int k = 0; k < NPD; k++) {
      c.spareD[k].ensure(n + 1);
      pa.sd[k] = d[k]->p;
      pa.dd[k] = c.spareD[k].p;
    }
    for (int k = 0; k < NPU; k++) {
      c.spareU[k].ensure(n + 1);
      pa.su[k] = u[k]->p;
      pa.du[k] = c.spareU[k].p;
    }
    hipLaunchKernelGGL(k_permute, dim3(nblocks(n, 256)), dim3(256), 0, c.stream, pa);
    WDG_LAUNCH_CHECK();
    for (int k = 0; k < NPD; k++) d[k]->swap(c.spareD[k]);
    for (int k = 0; k < NPU; k++) u[k]->swap(c.spareU[k]);
  }
  hipLaunchKernelGGL(k_sto, dim3(nblocks(n, 256)), dim3(256), 0, c.stream, n, c.uid.p, c.sto.p);
  WDG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_last_pos, dim3(nblocks(n, 256)), dim3(256), 0, c.stream, n, c.pos[0].p, c.pos[1].p, c.pos[2].p,
                     c.lastPos[0].p, c.lastPos[1].p, c.lastPos[2].p);
  WDG_LAUNCH_CHECK();
  if (nSph > 0 && !brute) {
    c.cellStart.ensure(ncells);
    c.cellEnd.ensure(ncells);
    WDG_HIP(hipMemsetAsync(c.cellStart.p, 0, ncells * sizeof(uint32_t), c.stream));
    WDG_HIP(hipMemsetAsync(c.cellEnd.p, 0, ncells * sizeof(uint32_t), c.stream));
    hipLaunchKernelGGL(k_cell_ranges, dim3(nblocks(nSph, 256)), dim3(256), 0, c.stream, nSph, nW, c.cellKey.p,
                       c.cellStart.p, c.cellEnd.p);
    WDG_LAUNCH_CHECK();
  }
  for (int k = 0; k < 3; k++) {
    c.grid.org[k] = g.org[k]; c.grid.h[k] = g.h[k]; c.grid.L[k] = g.L[k]; c.grid.nc[k] = g.nc[k];
  }
  c.grid.periodic = periodic;
  c.grid.brute = brute;
  c.grid.valid = true;

  // 2. pairs that can touch before the next refresh
  const int idBits = bits_for((uint64_t)(n > 1 ? n - 1 : 1));
  const uint64_t sentinel = (2 * idBits >= 64) ? ~0ull : ((1ull << (2 * idBits)) - 1);
  PairArgs pa = pair_args(c, g, P, idBits);
  int64_t nPairs = 0;
  if (nSph > 0) {
    hipLaunchKernelGGL((k_pairs<false, MODE_NEAR>), dim3(nblocks(nSph, 256)), dim3(256), 0, c.stream, pa);
    WDG_LAUNCH_CHECK();
    scan_exclusive(c, c.scount.p, c.soff.p, nSph, c.dflags.p + 4);
    nPairs = read_u32(c, c.dflags.p + 4);
  }
  Contacts& O = c.con();
  const int64_t nOld = O.n;
  const int64_t tot = nPairs + nOld;
  c.pairKey.ensure(tot + 1); c.pairKeyTmp.ensure(tot + 1); c.pairVal.ensure(tot + 1); c.pairValTmp.ensure(tot + 1);
  c.pairInfo.ensure(nPairs + 1); c.runFlag.ensure(tot + 1); c.runOff.ensure(tot + 1);
  if (periodic) c.pairDist.ensure(nPairs + 1);
  if (nPairs > 0) {
    pa.pkey = c.pairKey.p; pa.pinfo = c.pairInfo.p; pa.pdist = c.pairDist.p;
    hipLaunchKernelGGL((k_pairs<true, MODE_NEAR>), dim3(nblocks(nSph, 256)), dim3(256), 0, c.stream, pa);
    WDG_LAUNCH_CHECK();
  }
  // 3. merge with the old real contacts (history)
  if (nOld > 0) {
    hipLaunchKernelGGL(k_append_old, dim3(nblocks(nOld, 256)), dim3(256), 0, c.stream, nOld, nPairs, O.A.p, O.B.p,
                       c.rview(), c.invPerm.p, idBits, sentinel, c.pairKey.p);
    WDG_LAUNCH_CHECK();
  }
  int64_t nNew = 0;
  if (tot > 0) {
    hipLaunchKernelGGL(k_iota, dim3(nblocks(tot, 256)), dim3(256), 0, c.stream, tot, c.pairVal.p);
    WDG_LAUNCH_CHECK();
    radix_sort_pairs<uint64_t>(c, c.pairKey.p, c.pairVal.p, c.pairKeyTmp.p, c.pairValTmp.p, tot, 2 * idBits);
    hipLaunchKernelGGL(k_run_flags, dim3(nblocks(tot, 256)), dim3(256), 0, c.stream, tot, c.pairKey.p, sentinel,
                       c.runFlag.p);
    WDG_LAUNCH_CHECK();
    scan_exclusive(c, c.runFlag.p, c.runOff.p, tot, c.dflags.p + 5);
    nNew = read_u32(c, c.dflags.p + 5);
  }
  Contacts& N = c.C[1 - c.cur];
  N.ensure(nNew, n, periodic);
  N.n = nNew;
  if (nNew > 0) {
    CompactArgs ca{};
    ca.n = tot; ca.nPairs = nPairs; ca.key = c.pairKey.p; ca.val = c.pairVal.p; ca.flag = c.runFlag.p;
    ca.off = c.runOff.p; ca.pinfo = c.pairInfo.p; ca.pdist = c.pairDist.p; ca.invPerm = c.invPerm.p;
    ca.idBits = idBits; ca.periodic = periodic;
    ca.oA = O.A.p; ca.oB = O.B.p; ca.oInfo = O.info.p; ca.oDist = O.dist.p;
    ca.A = N.A.p; ca.B = N.B.p; ca.info = N.info.p; ca.carry = N.carry.p; ca.dist = N.dist.p;
    hipLaunchKernelGGL(k_compact, dim3(nblocks(tot, 256)), dim3(256), 0, c.stream, ca);
    WDG_LAUNCH_CHECK();
  }
  // real set of the new list: the carried contacts, with their state moved to the new ranks
  const int rNew = 1 - c.rcur;
  reals_from_flags(c, N.carry.p, nNew, rNew);
  if (nNew > 0) {
    Reals &RO = c.R[c.rcur], &RN = c.R[rNew];
    CarryArgs ya{};
    ya.n = nNew; ya.carry = N.carry.p; ya.ov = c.rview(c.rcur); ya.nv = c.rview(rNew);
    DBuf<double>* so[10] = {&RO.hN[0], &RO.hN[1], &RO.hN[2], &RO.hY[0], &RO.hY[1],
                            &RO.hY[2], &RO.hF[0], &RO.hF[1], &RO.hF[2], &RO.huN};
    DBuf<double>* sn[10] = {&RN.hN[0], &RN.hN[1], &RN.hN[2], &RN.hY[0], &RN.hY[1],
                            &RN.hY[2], &RN.hF[0], &RN.hF[1], &RN.hF[2], &RN.huN};
    for (int q = 0; q < 10; q++) { ya.o[q] = so[q]->p; ya.d[q] = sn[q]->p; }
    ya.ogp = RO.gp.p; ya.gp = RN.gp.p;
    hipLaunchKernelGGL(k_carry_state, dim3(nblocks(nNew, 256)), dim3(256), 0, c.stream, ya);
    WDG_LAUNCH_CHECK();
  }
  // 4. per-particle adjacency in canonical order (particle storage, other user id)
  const int64_t nAdj = 2 * nNew;
  if (nAdj > 0) {
    c.adjKey.ensure(nAdj); c.adjKeyTmp.ensure(nAdj); c.adjValTmp.ensure(nAdj); c.adjPart.ensure(nAdj);
    hipLaunchKernelGGL(k_adj_keys, dim3(nblocks(nNew, 256)), dim3(256), 0, c.stream, nNew, N.A.p, N.B.p, N.info.p,
                       c.uid.p, idBits, c.adjKey.p, N.adj.p);
    WDG_LAUNCH_CHECK();
    radix_sort_pairs<uint64_t>(c, c.adjKey.p, N.adj.p, c.adjKeyTmp.p, c.adjValTmp.p, nAdj, 2 * idBits);
    hipLaunchKernelGGL(k_key_hi, dim3(nblocks(nAdj, 256)), dim3(256), 0, c.stream, nAdj, c.adjKey.p, idBits,
                       c.adjPart.p);
    WDG_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_offsets, dim3(nblocks(nAdj + 1, 256)), dim3(256), 0, c.stream, nAdj, n,
                     nAdj > 0 ? c.adjPart.p : N.A.p, N.adjOff.p);
  WDG_LAUNCH_CHECK();
  c.cur = 1 - c.cur;
  c.rcur = rNew;
  c.boundsValid = true;
  WDG_HIP(hipMemsetAsync(c.dflags.p, 0, sizeof(uint32_t), c.stream));
  uint32_t err = read_u32(c, c.dflags.p + 1);
  c.dirtyKnown = true;
  c.dirtyIdx = 0;
  if (err) {
    WDG_HIP(hipMemsetAsync(c.dflags.p + 1, 0, sizeof(uint32_t), c.stream));
    check_device_errors(c, err);
  }
}

void collide(Ctx& c, const wdg_collide_params* p) {
  if (c.n == 0) return;
  c.nCollide++;
  if (c.forcesPending) flush_forces(c);
  // Collision.cpp:164-184 setVerletDist: resolved once
  if (std::isnan(c.verletDist)) {
    double vd = p->verletDist;
    if (vd < 0) vd = std::isinf(c.rmin) ? 0. : std::fabs(vd) * c.rmin;
    c.verletDist = vd;
  }
  bool full = !c.boundsValid || p->forceInit;
  if (!full && c.verletDist == 0) full = true;
  if (!full) {
    if (!c.dirtyKnown) {
      WDG_HIP(hipMemsetAsync(c.dflags.p, 0, sizeof(uint32_t), c.stream));
      hipLaunchKernelGGL(k_dirty, dim3(nblocks(c.n, 256)), dim3(256), 0, c.stream, c.n, c.pos[0].p, c.pos[1].p,
                         c.pos[2].p, c.lastPos[0].p, c.lastPos[1].p, c.lastPos[2].p, c.verletDist * c.verletDist,
                         c.dflags.p);
      WDG_LAUNCH_CHECK();
      c.dirtyKnown = true;
      c.dirtyIdx = 0;
    }
    WDG_HIP(hipMemcpyAsync(c.hflags, c.dflags.p, (DIRTY_PP + 2) * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           c.stream));
    WDG_HIP(hipStreamSynchronize(c.stream));
    if (c.hflags[1]) {
      uint32_t err = c.hflags[1];
      WDG_HIP(hipMemsetAsync(c.dflags.p + 1, 0, sizeof(uint32_t), c.stream));
      check_device_errors(c, err);
    }
    if (c.hflags[c.dirtyIdx]) full = true;
  }
  if (full) refresh(c);
}

// potential set of the last refresh: enlarged-AABB pairs from the refresh positions
void potential_pairs(Ctx& c, std::vector<uint32_t>& ida, std::vector<uint32_t>& idb, std::vector<uint64_t>& cd) {
  ida.clear(); idb.clear(); cd.clear();
  const int64_t nSph = c.n - c.nW;
  if (!c.grid.valid || nSph <= 0) return;
  Grid g{};
  for (int k = 0; k < 3; k++) {
    g.org[k] = c.grid.org[k]; g.h[k] = c.grid.h[k]; g.L[k] = c.grid.L[k]; g.nc[k] = c.grid.nc[k];
  }
  g.periodic = c.grid.periodic;
  AabbParams P{};
  P.vd = c.verletDist;
  P.absDf = std::fabs(c.field.distFactor);
  P.loneMask = c.field.loneMask;
  P.periodic = c.field.periodic;
  for (int k = 0; k < 3; k++) P.L[k] = c.field.cell[k];
  PairArgs pa = pair_args(c, g, P, 32);
  hipLaunchKernelGGL((k_pairs<false, MODE_AABB>), dim3(nblocks(nSph, 256)), dim3(256), 0, c.stream, pa);
  WDG_LAUNCH_CHECK();
  scan_exclusive(c, c.scount.p, c.soff.p, nSph, c.dflags.p + 4);
  int64_t np = read_u32(c, c.dflags.p + 4);
  if (np == 0) return;
  DBuf<uint64_t> k, d;
  k.ensure(np);
  d.ensure(np);
  pa.pkey = k.p; pa.pdist = d.p;
  hipLaunchKernelGGL((k_pairs<true, MODE_AABB>), dim3(nblocks(nSph, 256)), dim3(256), 0, c.stream, pa);
  WDG_LAUNCH_CHECK();
  std::vector<uint64_t> hk((size_t)np), hd;
  WDG_HIP(hipMemcpyAsync(hk.data(), k.p, np * 8, hipMemcpyDeviceToHost, c.stream));
  if (c.field.periodic) {
    hd.resize((size_t)np);
    WDG_HIP(hipMemcpyAsync(hd.data(), d.p, np * 8, hipMemcpyDeviceToHost, c.stream));
  }
  WDG_HIP(hipStreamSynchronize(c.stream));
  k.release();
  d.release();
  for (int64_t i = 0; i < np; i++) {
    ida.push_back((uint32_t)(hk[i] >> 32));
    idb.push_back((uint32_t)(hk[i] & 0xFFFFFFFFu));
    cd.push_back(c.field.periodic ? hd[i] : 0);
  }
}

}  // namespace wdg
