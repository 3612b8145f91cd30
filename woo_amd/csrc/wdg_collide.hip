// Broadphase: Verlet-skin refresh test, cell-list pair generation over
// radix-sorted cell keys, and the merge that carries contact history across
// refreshes. Replaces InsertionSortCollider::run (InsertionSortCollider.cpp:386-613).
//
// Pair-set semantics kept from the reference:
//  * AABB of a sphere = centre -/+ (r*distFactor') -/+ verletDist, of a wall =
//    infinite except along its axis (Sphere.cpp:52-82, Wall.cpp:25-38,
//    Collision.cpp:125-162);
//  * refresh when any particle moved more than verletDist since the last one
//    (Collision.cpp:80-123);
//  * after a refresh the contact set is {pairs whose enlarged AABBs overlap and
//    that mayCollide} U {real contacts}; a broken contact whose AABBs did not
//    overlap is dropped (InsertionSortCollider.hpp:213-218).
// Contact slots are sorted by (min id, max id); history of real contacts is
// carried to the new slots by one stable sort of [new pairs ; old real slots],
// so equal keys become adjacent and the compaction joins them.
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
  double dfBox;   // (distFactor>0 ? distFactor : 1)
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
__host__ inline double ord_to_double(unsigned long long b) {
  unsigned long long r = (b >> 63) ? (b & ~(1ull << 63)) : ~b;
  double d;
  std::memcpy(&d, &r, 8);
  return d;
}

__global__ void k_bbox(int64_t nSph, const int32_t* __restrict__ sphIds, const double* __restrict__ px,
                       const double* __restrict__ py, const double* __restrict__ pz, unsigned long long* out) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long mn[3] = {~0ull, ~0ull, ~0ull}, mx[3] = {0, 0, 0};
  if (s < nSph) {
    int i = sphIds[s];
    double p[3] = {px[i], py[i], pz[i]};
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

// cell key per sphere; refresh positions are recorded for every particle (lastPos)
__global__ void k_keys(int64_t nSph, const int32_t* __restrict__ sphIds, const double* __restrict__ px,
                       const double* __restrict__ py, const double* __restrict__ pz, Grid g, uint32_t* key,
                       uint32_t* idx) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nSph) return;
  int i = sphIds[s];
  double p[3] = {px[i], py[i], pz[i]};
  int c[3];
  for (int k = 0; k < 3; k++) {
    double x = p[k] - g.org[k];
    if (g.periodic) x = x - floor(x / g.L[k]) * g.L[k];
    int ci = (int)floor(x / g.h[k]);
    c[k] = min(max(ci, 0), g.nc[k] - 1);
  }
  key[s] = (uint32_t)(((int64_t)c[2] * g.nc[1] + c[1]) * g.nc[0] + c[0]);
  idx[s] = (uint32_t)i;
}

__global__ void k_last_pos(int64_t n, const double* __restrict__ px, const double* __restrict__ py,
                           const double* __restrict__ pz, double* lx, double* ly, double* lz) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  lx[i] = px[i];
  ly[i] = py[i];
  lz[i] = pz[i];
}

__global__ void k_gather_sorted(int64_t nSph, const uint32_t* __restrict__ idx, const double* __restrict__ px,
                                const double* __restrict__ py, const double* __restrict__ pz,
                                const double* __restrict__ rad, const uint32_t* __restrict__ mask, double dfBox,
                                double* sx, double* sy, double* sz, double* shs, uint32_t* smask) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nSph) return;
  uint32_t i = idx[s];
  sx[s] = px[i];
  sy[s] = py[i];
  sz[s] = pz[i];
  shs[s] = dfBox * rad[i];
  smask[s] = mask[i];
}

__global__ void k_cell_ranges(int64_t nSph, const uint32_t* __restrict__ key, uint32_t* cstart, uint32_t* cend) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nSph) return;
  uint32_t k = key[s];
  if (s == 0 || key[s - 1] != k) cstart[k] = (uint32_t)s;
  if (s == nSph - 1 || key[s + 1] != k) cend[k] = (uint32_t)s + 1;
}

struct PairArgs {
  int64_t nSph;
  const uint32_t* key;
  const uint32_t* idx;
  const double *sx, *sy, *sz, *shs;
  const uint32_t* smask;
  const uint32_t *cstart, *cend;
  Grid g;
  AabbParams P;
  int brute;
  // walls
  int nWalls;
  const int32_t* walls;
  const double *px, *py, *pz;
  const uint32_t* flags;
  const uint32_t* mask;
  int idBits;
  // outputs
  uint32_t* count;
  const uint32_t* off;
  uint64_t* pkey;
  uint32_t* pinfo;
  uint64_t* pdist;
  uint32_t* dflags;
};

template <bool WRITE>
__global__ __launch_bounds__(256) void k_pairs(PairArgs a) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.nSph) return;
  const uint32_t i = a.idx[s];
  const double pi[3] = {a.sx[s], a.sy[s], a.sz[s]};
  const double hi = a.shs[s];
  const uint32_t mi = a.smask[s];
  uint32_t cnt = 0;
  uint32_t o = WRITE ? a.off[s] : 0;
  auto consider = [&](int64_t t) {
    uint32_t j = a.idx[t];
    if (j <= i) return;
    if (!may_collide(mi, a.smask[t], a.P.loneMask)) return;
    const double pj[3] = {a.sx[t], a.sy[t], a.sz[t]};
    int cd[3], nimg;
    if (!sph_overlap(pi, hi, pj, a.shs[t], a.P, cd, nimg)) return;
    if (a.P.periodic && nimg > 1) {
      atomicOr(&a.dflags[1], ERR_DOUBLE_IMAGE);
      return;
    }
    if (WRITE) {
      a.pkey[o + cnt] = ((uint64_t)i << a.idBits) | (uint64_t)j;
      a.pinfo[o + cnt] = (uint32_t)CT_SS | CI_AABB;
      if (a.P.periodic) a.pdist[o + cnt] = pack_cd(cd[0], cd[1], cd[2]);
    }
    cnt++;
  };
  if (a.brute) {
    for (int64_t t = 0; t < a.nSph; t++) consider(t);
  } else {
    uint32_t k = a.key[s];
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
  // walls: every sphere is tested against every wall (walls are few)
  for (int w = 0; w < a.nWalls; w++) {
    uint32_t wid = (uint32_t)a.walls[w];
    if (!may_collide(a.mask[wid], mi, a.P.loneMask)) continue;
    uint32_t fl = a.flags[wid];
    int ax = flag_axis(fl);
    double wp = ax == 0 ? a.px[wid] : (ax == 1 ? a.py[wid] : a.pz[wid]);
    double wlo = wp, whi = wp;
    if (a.P.vd > 0) {
      whi = whi + a.P.vd;
      wlo = wlo - a.P.vd;
    }
    double slo, shi;
    sph_bounds(pi[ax], hi, a.P.vd, slo, shi);
    if (!(wlo <= shi + 0. && whi >= slo + 0.)) continue;
    if (WRITE) {
      uint32_t lo = wid < i ? wid : i, hi2 = wid < i ? i : wid;
      a.pkey[o + cnt] = ((uint64_t)lo << a.idBits) | (uint64_t)hi2;
      a.pinfo[o + cnt] = (uint32_t)CT_WS | CI_AABB | (wid > i ? CI_SWAP : 0u);
      if (a.P.periodic) a.pdist[o + cnt] = pack_cd(0, 0, 0);
    }
    cnt++;
  }
  if (!WRITE) a.count[s] = cnt;
}

// old real slots appended behind the new pairs; non-real ones get the sentinel key
__global__ void k_append_old(int64_t nOld, int64_t nPairs, const uint32_t* __restrict__ A,
                             const uint32_t* __restrict__ B, const uint32_t* __restrict__ real, int idBits,
                             uint64_t sentinel, uint64_t* pkey) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nOld) return;
  pkey[nPairs + c] = real[c] ? (((uint64_t)A[c] << idBits) | (uint64_t)B[c]) : sentinel;
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
  int idBits;
  int periodic;
  // old set
  const uint32_t* oInfo;
  const uint64_t* oDist;
  const double *oN0, *oN1, *oN2, *oY0, *oY1, *oY2, *oP0, *oP1, *oP2, *oF0, *oF1, *oF2, *ouN;
  // new set
  uint32_t *A, *B, *info, *real;
  uint64_t* dist;
  double *N0, *N1, *N2, *Y0, *Y1, *Y2, *P0, *P1, *P2, *F0, *F1, *F2, *uN;
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
  a.A[o] = (uint32_t)(k >> a.idBits);
  a.B[o] = (uint32_t)(k & ((1ull << a.idBits) - 1));
  uint32_t inf;
  uint64_t dist = 0;
  if (bp >= 0) {
    inf = a.pinfo[bp];
    if (a.periodic) dist = a.pdist[bp];
  } else {
    inf = a.oInfo[old] & ~CI_AABB;
  }
  if (old >= 0 && a.periodic) dist = a.oDist[old];
  a.info[o] = inf;
  if (a.periodic) a.dist[o] = dist;
  if (old >= 0) {
    a.real[o] = 1u;
    a.N0[o] = a.oN0[old]; a.N1[o] = a.oN1[old]; a.N2[o] = a.oN2[old];
    a.Y0[o] = a.oY0[old]; a.Y1[o] = a.oY1[old]; a.Y2[o] = a.oY2[old];
    a.P0[o] = a.oP0[old]; a.P1[o] = a.oP1[old]; a.P2[o] = a.oP2[old];
    a.F0[o] = a.oF0[old]; a.F1[o] = a.oF1[old]; a.F2[o] = a.oF2[old];
    a.uN[o] = a.ouN[old];
  } else {
    a.real[o] = 0u;
  }
}

// CSR offsets from a sorted id column: off[id] = first slot with that id
__global__ void k_offsets(int64_t nc, int64_t npar, const uint32_t* __restrict__ ids, uint32_t* off) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > nc) return;
  int64_t prev = t == 0 ? -1 : (int64_t)ids[t - 1];
  int64_t cur = t == nc ? npar : (int64_t)ids[t];
  for (int64_t id = prev + 1; id <= cur; id++) off[id] = (uint32_t)t;
}

__global__ void k_copy_u32(int64_t n, const uint32_t* __restrict__ in, uint32_t* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

// ------------------------------------------------------------------ host side
static void check_device_errors(Ctx& c, uint32_t err) {
  if (err & ERR_DOUBLE_IMAGE)
    throw Error(WDG_E_GEOM, "InsertionSortCollider: a pair overlaps through two periodic images "
                            "(particle larger than half of the cell?)");
  if (err & ERR_NAN_FORCE) throw Error(WDG_E_NUMERIC, "NaN force in contact");
  if (err & ERR_CD_RANGE) throw Error(WDG_E_GEOM, "periodic cell distance out of range");
}

static void refresh(Ctx& c) {
  const int64_t n = c.n, nSph = c.nSph;
  const double vd = c.verletDist;
  const double dfBox = c.field.distFactor > 0 ? c.field.distFactor : 1.;
  const bool periodic = c.field.periodic != 0;
  c.nRefresh++;

  // refresh positions
  hipLaunchKernelGGL(k_last_pos, dim3(nblocks(n, 256)), dim3(256), 0, c.stream, n, c.pos[0].p, c.pos[1].p, c.pos[2].p,
                     c.lastPos[0].p, c.lastPos[1].p, c.lastPos[2].p);
  WDG_LAUNCH_CHECK();

  // grid: cell edge >= largest enlarged-AABB extent
  Grid g{};
  g.periodic = periodic;
  double ext = 2. * (dfBox * c.rmax + vd);
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
    c.dflags.ensure(64);
    unsigned long long init[6] = {~0ull, ~0ull, ~0ull, 0, 0, 0};
    unsigned long long* dbb = reinterpret_cast<unsigned long long*>(c.red.p);
    WDG_HIP(hipMemcpyAsync(dbb, init, sizeof(init), hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_bbox, dim3(nblocks(nSph, 256)), dim3(256), 0, c.stream, nSph, c.dWalls.p + c.wallIds.size(),
                       c.pos[0].p, c.pos[1].p, c.pos[2].p, dbb);
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
  c.lastBrute = brute;
  for (int k = 0; k < 3; k++) c.gridCells[k] = brute ? 0 : g.nc[k];

  AabbParams P{};
  P.vd = vd;
  P.dfBox = dfBox;
  P.loneMask = c.field.loneMask;
  P.periodic = periodic;
  for (int k = 0; k < 3; k++) P.L[k] = c.field.cell[k];

  // sorted sphere order and cell ranges
  c.cellKey.ensure(nSph + 1); c.sortIdx.ensure(nSph + 1);
  c.cellKeyTmp.ensure(nSph + 1); c.sortIdxTmp.ensure(nSph + 1);
  c.sx.ensure(nSph + 1); c.sy.ensure(nSph + 1); c.sz.ensure(nSph + 1); c.shs.ensure(nSph + 1);
  c.smask.ensure(nSph + 1); c.scount.ensure(nSph + 1); c.soff.ensure(nSph + 1);
  const int32_t* sphIds = c.dWalls.p + c.wallIds.size();
  if (nSph > 0) {
    Grid gk = g;
    if (brute) { for (int k = 0; k < 3; k++) { gk.nc[k] = 1; gk.h[k] = 1e300; gk.org[k] = 0; } gk.periodic = 0; }
    hipLaunchKernelGGL(k_keys, dim3(nblocks(nSph, 256)), dim3(256), 0, c.stream, nSph, sphIds, c.pos[0].p, c.pos[1].p,
                       c.pos[2].p, gk, c.cellKey.p, c.sortIdx.p);
    WDG_LAUNCH_CHECK();
    if (!brute) {
      radix_sort_pairs<uint32_t>(c, c.cellKey.p, c.sortIdx.p, c.cellKeyTmp.p, c.sortIdxTmp.p, nSph,
                                 bits_for((uint64_t)(ncells > 1 ? ncells - 1 : 1)));
      c.cellStart.ensure(ncells);
      c.cellEnd.ensure(ncells);
      WDG_HIP(hipMemsetAsync(c.cellStart.p, 0, ncells * sizeof(uint32_t), c.stream));
      WDG_HIP(hipMemsetAsync(c.cellEnd.p, 0, ncells * sizeof(uint32_t), c.stream));
      hipLaunchKernelGGL(k_cell_ranges, dim3(nblocks(nSph, 256)), dim3(256), 0, c.stream, nSph, c.cellKey.p,
                         c.cellStart.p, c.cellEnd.p);
      WDG_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_gather_sorted, dim3(nblocks(nSph, 256)), dim3(256), 0, c.stream, nSph, c.sortIdx.p,
                       c.pos[0].p, c.pos[1].p, c.pos[2].p, c.radius.p, c.mask.p, dfBox, c.sx.p, c.sy.p, c.sz.p,
                       c.shs.p, c.smask.p);
    WDG_LAUNCH_CHECK();
  }

  const int idBits = bits_for((uint64_t)(n > 1 ? n - 1 : 1));
  const uint64_t sentinel = (2 * idBits >= 64) ? ~0ull : ((1ull << (2 * idBits)) - 1);

  PairArgs pa{};
  pa.nSph = nSph; pa.key = c.cellKey.p; pa.idx = c.sortIdx.p;
  pa.sx = c.sx.p; pa.sy = c.sy.p; pa.sz = c.sz.p; pa.shs = c.shs.p; pa.smask = c.smask.p;
  pa.cstart = c.cellStart.p; pa.cend = c.cellEnd.p; pa.g = g; pa.P = P; pa.brute = brute;
  pa.nWalls = (int)c.wallIds.size(); pa.walls = c.dWalls.p;
  pa.px = c.pos[0].p; pa.py = c.pos[1].p; pa.pz = c.pos[2].p; pa.flags = c.flags.p; pa.mask = c.mask.p;
  pa.idBits = idBits; pa.count = c.scount.p; pa.off = c.soff.p; pa.dflags = c.dflags.p;
  int64_t nPairs = 0;
  if (nSph > 0) {
    hipLaunchKernelGGL(k_pairs<false>, dim3(nblocks(nSph, 256)), dim3(256), 0, c.stream, pa);
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
    hipLaunchKernelGGL(k_pairs<true>, dim3(nblocks(nSph, 256)), dim3(256), 0, c.stream, pa);
    WDG_LAUNCH_CHECK();
  }
  if (nOld > 0) {
    hipLaunchKernelGGL(k_append_old, dim3(nblocks(nOld, 256)), dim3(256), 0, c.stream, nOld, nPairs, O.A.p, O.B.p,
                       O.real.p, idBits, sentinel, c.pairKey.p);
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
    ca.off = c.runOff.p; ca.pinfo = c.pairInfo.p; ca.pdist = c.pairDist.p; ca.idBits = idBits; ca.periodic = periodic;
    ca.oInfo = O.info.p; ca.oDist = O.dist.p;
    ca.oN0 = O.hN[0].p; ca.oN1 = O.hN[1].p; ca.oN2 = O.hN[2].p;
    ca.oY0 = O.hY[0].p; ca.oY1 = O.hY[1].p; ca.oY2 = O.hY[2].p;
    ca.oP0 = O.hP[0].p; ca.oP1 = O.hP[1].p; ca.oP2 = O.hP[2].p;
    ca.oF0 = O.hF[0].p; ca.oF1 = O.hF[1].p; ca.oF2 = O.hF[2].p; ca.ouN = O.huN.p;
    ca.A = N.A.p; ca.B = N.B.p; ca.info = N.info.p; ca.real = N.real.p; ca.dist = N.dist.p;
    ca.N0 = N.hN[0].p; ca.N1 = N.hN[1].p; ca.N2 = N.hN[2].p;
    ca.Y0 = N.hY[0].p; ca.Y1 = N.hY[1].p; ca.Y2 = N.hY[2].p;
    ca.P0 = N.hP[0].p; ca.P1 = N.hP[1].p; ca.P2 = N.hP[2].p;
    ca.F0 = N.hF[0].p; ca.F1 = N.hF[1].p; ca.F2 = N.hF[2].p; ca.uN = N.huN.p;
    hipLaunchKernelGGL(k_compact, dim3(nblocks(tot, 256)), dim3(256), 0, c.stream, ca);
    WDG_LAUNCH_CHECK();
  }
  // CSR over min id, and the (max id, min id)-ordered transpose for the gather
  hipLaunchKernelGGL(k_offsets, dim3(nblocks(nNew + 1, 256)), dim3(256), 0, c.stream, nNew, n, N.A.p, N.offA.p);
  WDG_LAUNCH_CHECK();
  if (nNew > 0) {
    c.keyB.ensure(nNew); c.keyBTmp.ensure(nNew); c.valBTmp.ensure(nNew);
    hipLaunchKernelGGL(k_copy_u32, dim3(nblocks(nNew, 256)), dim3(256), 0, c.stream, nNew, N.B.p, c.keyB.p);
    WDG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_iota, dim3(nblocks(nNew, 256)), dim3(256), 0, c.stream, nNew, N.idxB.p);
    WDG_LAUNCH_CHECK();
    radix_sort_pairs<uint32_t>(c, c.keyB.p, N.idxB.p, c.keyBTmp.p, c.valBTmp.p, nNew, idBits);
  }
  hipLaunchKernelGGL(k_offsets, dim3(nblocks(nNew + 1, 256)), dim3(256), 0, c.stream, nNew, n,
                     nNew > 0 ? c.keyB.p : N.A.p, N.offB.p);
  WDG_LAUNCH_CHECK();
  c.cur = 1 - c.cur;
  c.boundsValid = true;
  // clean after the refresh; device error bits are checked now
  WDG_HIP(hipMemsetAsync(c.dflags.p, 0, sizeof(uint32_t), c.stream));
  uint32_t err = read_u32(c, c.dflags.p + 1);
  c.dirtyKnown = true;
  if (err) {
    WDG_HIP(hipMemsetAsync(c.dflags.p + 1, 0, sizeof(uint32_t), c.stream));
    check_device_errors(c, err);
  }
}

void check_errors(Ctx& c) {
  uint32_t err = read_u32(c, c.dflags.p + 1);
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
    }
    WDG_HIP(hipMemcpyAsync(c.hflags, c.dflags.p, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
    WDG_HIP(hipStreamSynchronize(c.stream));
    if (c.hflags[1]) {
      uint32_t err = c.hflags[1];
      WDG_HIP(hipMemsetAsync(c.dflags.p + 1, 0, sizeof(uint32_t), c.stream));
      check_device_errors(c, err);
    }
    if (c.hflags[0]) full = true;
  }
  if (full) refresh(c);
}

}  // namespace wdg
