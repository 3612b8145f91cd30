// extern "C" entry points of the DEM step engine (see include/woo_amd.h).
// The ABI speaks user ids (Particle.id) and user-ordered arrays; the device keeps
// particles in storage order (walls, then spheres sorted by cell), mapped by uid/sto.
#include <atomic>
#include <unordered_map>
#include <algorithm>
#include <cstring>
#include <numeric>

#include "wdg_internal.hpp"
#include "wdg_math.hpp"
#include "wdg_physics.hpp"

struct wdg_ctx {
  wdg::Ctx c;
};
wdg::Ctx& wdg::ctx_of(wdg_ctx* h) { return h->c; }

namespace wdg {

// out[u] = in[sto[u]]
__global__ void k_gather_uid(int64_t n, const uint32_t* __restrict__ sto, const double* __restrict__ in, double* out) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u < n) out[u] = in[sto[u]];
}
// out[sto[u]] = in[u]
__global__ void k_scatter_uid(int64_t n, const uint32_t* __restrict__ sto, const double* __restrict__ in, double* out) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u < n) out[sto[u]] = in[u];
}

// one record of 16 doubles per slot, for inspection and parity tests
struct ExportArgs {
  int64_t nc;
  const uint32_t *A, *B, *info;
  RealView rv;
  const double *N0, *N1, *N2, *Y0, *Y1, *Y2, *F0, *F1, *F2, *uN, *gp;
  const double *rad, *mass;
  const double *px, *py, *pz;
  const int32_t* mat;
  const wdg_material* mats;
  const uint8_t* hE;  // each real contact's material epoch (null: one table)
  int nMats;
  const double* fgeo;
  int hertz;
  double poisson, en, viscAlpha;
  double* out;
  uint8_t* outE;  // the epochs per slot (snapshots), or null
};

template <int PREC>
__global__ void k_export(ExportArgs a) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.nc) return;
  double* o = a.out + 16 * s;
  if (!rv_real(a.rv, (uint32_t)s)) {
    for (int k = 0; k < 15; k++) o[k] = NAN;
    o[15] = 0.;
    return;
  }
  const uint32_t r = rv_index(a.rv, (uint32_t)s);
  const uint32_t inf = a.info[s];
  const bool swap = (inf & CI_SWAP) != 0u;
  const uint32_t pA = swap ? a.B[s] : a.A[s], pB = swap ? a.A[s] : a.B[s];
  double r1, r2;
  contact_radii(inf, pA, pB, a.rad, a.fgeo, r1, r2);
  double l0, l1, A;
  lens_area(r1, r2, l0, l1, A);
  const uint32_t ep = a.hE ? (uint32_t)a.hE[r] : 0u;
  if (a.outE) a.outE[s] = (uint8_t)ep;
  const uint32_t eb = ep * (uint32_t)a.nMats;
  const wdg_material mA = a.mats[eb + a.mat[pA]], mB = a.mats[eb + a.mat[pB]];
  PhysVals ph;
  double kn, kt;
  const double uN = ldr<PREC>(a.uN, r);
  if (a.hertz) {
    phys_hertz(mA, mB, l0, l1, a.mass[pA], a.mass[pB], a.poisson, a.en, a.viscAlpha, ph);
    double kn0 = ph.K * sqrt(ph.R);
    kn = (3 / 2.) * kn0 * sqrt(-uN);
    kt = ph.kt0 * sqrt(uN < 0 ? -uN : 0);
  } else {
    phys_frict(mA, mB, l0, l1, A, ph);
    kn = ph.kn;
    kt = ph.kt;
  }
  o[0] = ldr<PREC>(a.N0, r); o[1] = ldr<PREC>(a.N1, r); o[2] = ldr<PREC>(a.N2, r);
  o[3] = ldr<PREC>(a.Y0, r); o[4] = ldr<PREC>(a.Y1, r); o[5] = ldr<PREC>(a.Y2, r);
  if (PREC == 32) {
    // the record holds the contact point relative to pA
    const float* g = reinterpret_cast<const float*>(a.gp) + (int64_t)GP32 * r;
    o[6] = a.px[pA] + (double)g[3]; o[7] = a.py[pA] + (double)g[4]; o[8] = a.pz[pA] + (double)g[5];
  } else {
    const double* g = a.gp + (int64_t)GP * r;
    o[6] = g[3]; o[7] = g[4]; o[8] = g[5];
  }
  o[9] = ldr<PREC>(a.F0, r); o[10] = ldr<PREC>(a.F1, r); o[11] = ldr<PREC>(a.F2, r);
  o[12] = uN; o[13] = kn; o[14] = kt; o[15] = 1.;
}

// the record export of the current slot list into out (16 doubles per slot)
static ExportArgs export_args(Ctx& c, double* out) {
  Contacts& C = c.con();
  Reals& RS = c.reals();
  ExportArgs a{};
  a.nc = C.n; a.A = C.A.p; a.B = C.B.p; a.info = C.info.p; a.rv = c.rview();
  a.N0 = RS.hN[0].p; a.N1 = RS.hN[1].p; a.N2 = RS.hN[2].p;
  a.Y0 = RS.hY[0].p; a.Y1 = RS.hY[1].p; a.Y2 = RS.hY[2].p;
  a.F0 = RS.hF[0].p; a.F1 = RS.hF[1].p; a.F2 = RS.hF[2].p; a.uN = RS.huN.p; a.gp = RS.gp.p;
  a.rad = c.radius.p; a.mass = c.mass.p; a.mat = c.mat.p; a.mats = c.dMats.p; a.fgeo = c.fgeo.p;
  a.nMats = (int)c.mats.size();
  if (c.matEpochs > 1) a.hE = RS.hE.p;
  a.px = c.pos[0].p; a.py = c.pos[1].p; a.pz = c.pos[2].p;
  a.hertz = c.lastPhysKind == WDG_PHYS_HERTZ;
  a.poisson = c.lastPoisson; a.en = c.lastEn; a.viscAlpha = c.lastViscAlpha;
  a.out = out;
  return a;
}

// a checkpoint's real contacts written back into the current real set: per slot, 1 + the
// index of its record (0: not real), records in the wdg_contact_list layout
struct RestoreArgs {
  int64_t nc;
  const uint32_t* rec;
  const double* data;
  RealView rv;
  const uint64_t* dist;  // packed cellDist per slot (periodic cells; else null)
  double *N0, *N1, *N2, *Y0, *Y1, *Y2, *F0, *F1, *F2, *uN, *gp;
  // material epochs (while more than one table is kept): each record's (else all take epoch)
  uint8_t* hE;
  const uint8_t* recE;
  uint32_t epoch;
};

__global__ void k_restore(RestoreArgs a) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.nc) return;
  const uint32_t q = a.rec[s];
  if (q == 0u) return;
  const double* d = a.data + (int64_t)16 * (q - 1u);
  const uint32_t k = rv_index(a.rv, (uint32_t)s);
  if (a.hE) a.hE[k] = a.recE ? a.recE[q - 1u] : (uint8_t)a.epoch;
  const d3 T0 = mk3(d[0], d[1], d[2]), T1 = mk3(d[3], d[4], d[5]);
  a.N0[k] = T0.x; a.N1[k] = T0.y; a.N2[k] = T0.z;
  a.Y0[k] = T1.x; a.Y1[k] = T1.y; a.Y2[k] = T1.z;
  a.F0[k] = d[9]; a.F1[k] = d[10]; a.F2[k] = d[11];
  a.uN[k] = d[12];
  // the gather record as the force pass writes it: global force on pA, contact point, cellDist
  const d3 G = quat_rot(quat_from_cols(T0, T1, cross3(T0, T1)), mk3(d[9], d[10], d[11]));
  double2* g2 = reinterpret_cast<double2*>(a.gp + (int64_t)GP * k);
  g2[0] = make_double2(G.x, G.y);
  g2[1] = make_double2(G.z, d[6]);
  g2[2] = make_double2(d[7], d[8]);
  g2[3] = make_double2(__longlong_as_double((long long)(a.dist ? a.dist[s] : 0ull)), 0.);
}

// Contact history of a checkpoint (the save half is wdg_contact_list(which=1)): a full
// collider refresh at the uploaded positions, then each saved contact's slot is found by its
// particles' user ids and made real with the saved state
void restore_contacts(Ctx& c, int64_t n, const int32_t* ids, const int32_t* cellDist, const double* data,
                      double verletDist) {
  if (c.precision != 64) throw Error(WDG_E_STATE, "restoring contacts needs the FP64 contact form");
  if (c.forcesPending) flush_forces(c);
  drop_contacts(c);
  if (c.n == 0) {
    if (n > 0) throw Error(WDG_E_ARG, "no particles to restore contacts between");
    return;
  }
  wdg_collide_params cp{verletDist, 1};
  collide(c, &cp);
  Contacts& C = c.con();
  const int64_t nc = C.n;
  std::vector<uint32_t> A((size_t)nc), B((size_t)nc), info((size_t)nc), uid((size_t)c.n);
  std::vector<uint64_t> dist;
  WDG_HIP(hipStreamSynchronize(c.stream));
  WDG_HIP(hipMemcpy(uid.data(), c.uid.p, c.n * 4, hipMemcpyDeviceToHost));
  if (nc) {
    WDG_HIP(hipMemcpy(A.data(), C.A.p, nc * 4, hipMemcpyDeviceToHost));
    WDG_HIP(hipMemcpy(B.data(), C.B.p, nc * 4, hipMemcpyDeviceToHost));
    WDG_HIP(hipMemcpy(info.data(), C.info.p, nc * 4, hipMemcpyDeviceToHost));
    if (c.field.periodic) {
      dist.resize((size_t)nc);
      WDG_HIP(hipMemcpy(dist.data(), C.dist.p, nc * 8, hipMemcpyDeviceToHost));
    }
  }
  std::unordered_map<uint64_t, int64_t> slot;
  slot.reserve((size_t)nc);
  for (int64_t s = 0; s < nc; s++) {
    const uint64_t a = uid[A[s]], b = uid[B[s]];
    slot[(std::min(a, b) << 32) | std::max(a, b)] = s;
  }
  std::vector<uint32_t> rec((size_t)std::max<int64_t>(nc, 1), 0u);
  for (int64_t k = 0; k < n; k++) {
    const int64_t ia = ids[2 * k], ib = ids[2 * k + 1];
    const std::string name = "##" + std::to_string(ia) + "+" + std::to_string(ib);
    if (ia < 0 || ib < 0 || ia >= c.n || ib >= c.n || ia == ib) throw Error(WDG_E_ARG, name + ": bad particle ids");
    const uint64_t lo = (uint64_t)std::min(ia, ib), hi = (uint64_t)std::max(ia, ib);
    auto it = slot.find((lo << 32) | hi);
    if (it == slot.end())
      throw Error(WDG_E_ARG, name + ": the pair's bounding boxes do not overlap (positions not those of the checkpoint?)");
    const int64_t s = it->second;
    const bool swap = (info[s] & CI_SWAP) != 0u;
    if ((int64_t)uid[swap ? B[s] : A[s]] != ia) throw Error(WDG_E_ARG, name + ": particles not in the contact's role order");
    if (c.field.periodic && cellDist)
      for (int m = 0; m < 3; m++)
        if (unpack_cd(dist[s], m) != cellDist[3 * k + m]) throw Error(WDG_E_ARG, name + ": cellDist differs from the nearest image");
    if (rec[s]) throw Error(WDG_E_ARG, name + ": listed twice");
    rec[s] = (uint32_t)(k + 1);
  }
  if (nc == 0) return;
  DBuf<uint32_t> drec;
  DBuf<double> ddata;
  drec.ensure((size_t)nc);
  ddata.ensure((size_t)std::max<int64_t>(n, 1) * 16);
  WDG_HIP(hipMemcpy(drec.p, rec.data(), nc * 4, hipMemcpyHostToDevice));
  if (n) WDG_HIP(hipMemcpy(ddata.p, data, n * 16 * sizeof(double), hipMemcpyHostToDevice));
  reals_from_flags(c, drec.p, nc, c.rcur);
  Reals& R = c.reals();
  RestoreArgs a{};
  a.nc = nc; a.rec = drec.p; a.data = ddata.p; a.rv = c.rview();
  a.dist = c.field.periodic ? C.dist.p : nullptr;
  a.N0 = R.hN[0].p; a.N1 = R.hN[1].p; a.N2 = R.hN[2].p;
  a.Y0 = R.hY[0].p; a.Y1 = R.hY[1].p; a.Y2 = R.hY[2].p;
  a.F0 = R.hF[0].p; a.F1 = R.hF[1].p; a.F2 = R.hF[2].p; a.uN = R.huN.p; a.gp = R.gp.p;
  // a checkpoint's contacts take the current materials
  if (c.matEpochs > 1) { a.hE = R.hE.p; a.epoch = (uint32_t)(c.matEpochs - 1); }
  hipLaunchKernelGGL(k_restore, dim3(nblocks(nc, 256)), dim3(256), 0, c.stream, a);
  WDG_LAUNCH_CHECK();
  WDG_HIP(hipStreamSynchronize(c.stream));
  drec.release();
  ddata.release();
  c.forcesPending = false;
  c.rollActive = false;  // the records hold no phys torque; the next contact loop writes it
  c.gen++;
}

// ------------------------------------------------------------------ contact history of snapshots
// Per slot: its key (min uid << 32 | max uid) when real, else nkey (above every real key), the
// slot itself as the value, and its packed cellDist
__global__ void k_snap_keys(int64_t nc, const uint32_t* __restrict__ A, const uint32_t* __restrict__ B, RealView rv,
                            const uint32_t* __restrict__ uid, uint64_t nkey, const uint64_t* __restrict__ dist,
                            uint64_t* key, uint32_t* val, uint64_t* cd) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nc) return;
  uint64_t k = nkey;
  if (rv_real(rv, (uint32_t)s)) {
    const uint64_t a = uid[A[s]], b = uid[B[s]];
    k = (a < b ? a : b) << 32 | (a < b ? b : a);
  }
  key[s] = k;
  val[s] = (uint32_t)s;
  if (dist) cd[s] = dist[s];
}

// Per slot of the refreshed list: 1 + the slot of the same pair among the snapshot's real
// contacts (binary search over the sorted keys), or 0. The pair's periodic image must be the
// saved one.
__global__ void k_snap_match(int64_t nc, const uint32_t* __restrict__ A, const uint32_t* __restrict__ B,
                             const uint32_t* __restrict__ uid, const uint64_t* __restrict__ dist,
                             const uint64_t* __restrict__ key, const uint32_t* __restrict__ slot, int64_t nOld,
                             const uint64_t* __restrict__ oldCd, uint32_t* rec, uint32_t* dflags) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nc) return;
  const uint64_t a = uid[A[s]], b = uid[B[s]];
  const uint64_t k = (a < b ? a : b) << 32 | (a < b ? b : a);
  int64_t lo = 0, hi = nOld;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (key[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  uint32_t r = 0u;
  if (lo < nOld && key[lo] == k) {
    const uint32_t os = slot[lo];
    if (dist && oldCd[os] != dist[s]) atomicOr(&dflags[1], ERR_SNAP_IMAGE);
    r = os + 1u;
  }
  rec[s] = r;
}

void snapshot_contacts_save(Ctx& c, int slot) {
  c.snapReal[slot] = 0;
  Contacts& C = c.con();
  if (c.precision != 64 || C.n == 0 || c.n == 0) return;
  const int64_t nReal = count_real(c);
  if (nReal == 0) return;
  const int64_t nc = C.n;
  const bool per = c.field.periodic != 0;
  c.snapKey[slot].ensure(nc);
  c.snapSlot[slot].ensure(nc);
  c.snapRec[slot].ensure((size_t)nc * 16);
  if (per) c.snapCd[slot].ensure(nc);
  c.bigKeyTmp.ensure(nc);
  c.bigValTmp.ensure(nc);
  hipLaunchKernelGGL(k_snap_keys, dim3(nblocks(nc, 256)), dim3(256), 0, c.stream, nc, (const uint32_t*)C.A.p,
                     (const uint32_t*)C.B.p, c.rview(), (const uint32_t*)c.uid.p, (uint64_t)c.n << 32,
                     per ? (const uint64_t*)C.dist.p : nullptr, c.snapKey[slot].p, c.snapSlot[slot].p,
                     per ? c.snapCd[slot].p : nullptr);
  WDG_LAUNCH_CHECK();
  radix_sort_pairs<uint64_t>(c, c.snapKey[slot].p, c.snapSlot[slot].p, c.bigKeyTmp.p, c.bigValTmp.p, nc,
                             32 + bits_for((uint64_t)c.n));
  ExportArgs ea = export_args(c, c.snapRec[slot].p);
  c.snapMatGen[slot] = -1;
  if (c.matEpochs > 1) {
    c.snapEp[slot].ensure(nc);
    ea.outE = c.snapEp[slot].p;
    c.snapMatGen[slot] = c.matGen;
  }
  hipLaunchKernelGGL(k_export<64>, dim3(nblocks(nc, 256)), dim3(256), 0, c.stream, ea);
  WDG_LAUNCH_CHECK();
  c.snapReal[slot] = nReal;
  c.snapVerlet[slot] = c.verletDist;
}

// after the kinematic state of the snapshot was loaded and the contacts dropped
void snapshot_contacts_load(Ctx& c, int slot) {
  const int64_t nReal = c.snapReal[slot];
  if (nReal == 0 || c.precision != 64) return;
  // the slot list of the snapshot's Verlet distance, rebuilt at its positions: it holds every
  // saved real contact (their particles touch, so their boxes overlap)
  c.verletDist = c.snapVerlet[slot];
  wdg_collide_params cp{c.verletDist, 1};
  collide(c, &cp);
  Contacts& C = c.con();
  const int64_t nc = C.n;
  const bool per = c.field.periodic != 0;
  c.snapMatch.ensure((size_t)std::max<int64_t>(nc, 1));
  if (nc > 0) {
    hipLaunchKernelGGL(k_snap_match, dim3(nblocks(nc, 256)), dim3(256), 0, c.stream, nc, (const uint32_t*)C.A.p,
                       (const uint32_t*)C.B.p, (const uint32_t*)c.uid.p, per ? (const uint64_t*)C.dist.p : nullptr,
                       (const uint64_t*)c.snapKey[slot].p, (const uint32_t*)c.snapSlot[slot].p, nReal,
                       per ? (const uint64_t*)c.snapCd[slot].p : nullptr, c.snapMatch.p, c.dflags.p);
    WDG_LAUNCH_CHECK();
    reals_from_flags(c, c.snapMatch.p, nc, c.rcur);
    Reals& R = c.reals();
    RestoreArgs a{};
    a.nc = nc; a.rec = c.snapMatch.p; a.data = c.snapRec[slot].p; a.rv = c.rview();
    a.dist = per ? C.dist.p : nullptr;
    a.N0 = R.hN[0].p; a.N1 = R.hN[1].p; a.N2 = R.hN[2].p;
    a.Y0 = R.hY[0].p; a.Y1 = R.hY[1].p; a.Y2 = R.hY[2].p;
    a.F0 = R.hF[0].p; a.F1 = R.hF[1].p; a.F2 = R.hF[2].p; a.uN = R.huN.p; a.gp = R.gp.p;
    // the saved epochs while the tables are numbered as at the save, else the current one
    if (c.matEpochs > 1) {
      a.hE = R.hE.p;
      a.epoch = (uint32_t)(c.matEpochs - 1);
      if (c.snapMatGen[slot] == c.matGen) a.recE = c.snapEp[slot].p;
    }
    hipLaunchKernelGGL(k_restore, dim3(nblocks(nc, 256)), dim3(256), 0, c.stream, a);
    WDG_LAUNCH_CHECK();
  }
  check_errors(c);
  const int64_t got = count_real(c);
  if (got != nReal)
    throw Error(WDG_E_STATE, "snapshot: " + std::to_string(nReal - got) +
                                 " saved contacts are not in the slot list rebuilt at the snapshot's positions");
  c.forcesPending = false;
  c.rollActive = false;
  c.gen++;
}

}  // namespace wdg

namespace {
thread_local std::string g_err;

template <class F>
int guard(wdg_ctx* ctx, F&& f) {
  try {
    if (ctx) WDG_HIP(hipSetDevice(ctx->c.device));
    f();
    return WDG_OK;
  } catch (wdg::Error& e) {
    g_err = e.what();
    return e.code;
  } catch (std::exception& e) {
    g_err = e.what();
    return WDG_E_STATE;
  }
}

wdg::DBuf<double>* field_bufs(wdg::Ctx& c, int field, int& w) {
  switch (field) {
    case WDG_F_POS: w = 3; return c.pos;
    case WDG_F_VEL: w = 3; return c.vel;
    case WDG_F_ANGVEL: w = 3; return c.angVel;
    case WDG_F_ORI: w = 4; return c.ori;
    case WDG_F_FORCE: w = 3; return c.force;
    case WDG_F_TORQUE: w = 3; return c.torque;
    case WDG_F_MASS: w = 1; return &c.mass;
    case WDG_F_INERTIA: w = 3; return c.inertia;
  }
  throw wdg::Error(WDG_E_ARG, "unknown field " + std::to_string(field));
}

// user-ordered AoS host array -> storage-ordered SoA device arrays
void put_field(wdg::Ctx& c, wdg::DBuf<double>* dst, int w, const double* host) {
  using namespace wdg;
  const int64_t n = c.n;
  if (!n) return;
  c.xfer.ensure(n);
  std::vector<double> col((size_t)n);
  for (int k = 0; k < w; k++) {
    for (int64_t u = 0; u < n; u++) col[u] = host[u * w + k];
    WDG_HIP(hipMemcpyAsync(c.xfer.p, col.data(), n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_scatter_uid, dim3(nblocks(n, 256)), dim3(256), 0, c.stream, n, (const uint32_t*)c.sto.p,
                       (const double*)c.xfer.p, dst[k].p);
    WDG_LAUNCH_CHECK();
    WDG_HIP(hipStreamSynchronize(c.stream));
  }
}
void get_field(wdg::Ctx& c, wdg::DBuf<double>* src, int w, double* host) {
  using namespace wdg;
  const int64_t n = c.n;
  if (!n) return;
  c.xfer.ensure(n);
  std::vector<double> col((size_t)n);
  for (int k = 0; k < w; k++) {
    hipLaunchKernelGGL(k_gather_uid, dim3(nblocks(n, 256)), dim3(256), 0, c.stream, n, (const uint32_t*)c.sto.p,
                       (const double*)src[k].p, c.xfer.p);
    WDG_LAUNCH_CHECK();
    WDG_HIP(hipMemcpyAsync(col.data(), c.xfer.p, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    WDG_HIP(hipStreamSynchronize(c.stream));
    for (int64_t u = 0; u < n; u++) host[u * w + k] = col[u];
  }
}

struct Record {
  uint32_t a, b;  // user ids of pA, pB
  int cd[3];
  int64_t slot;   // -1: potential only
};
}  // namespace

using namespace wdg;

static std::vector<Record> collect_contacts(Ctx& c, int which, std::vector<double>* rec) {
  Contacts& C = c.con();
  const int64_t nc = C.n;
  std::vector<uint32_t> A((size_t)nc), B((size_t)nc), info((size_t)nc), uid((size_t)c.n);
  std::vector<unsigned long long> bits((size_t)(nc + 63) / 64);
  Reals& RS = c.reals();
  std::vector<uint64_t> dist;
  WDG_HIP(hipStreamSynchronize(c.stream));
  if (c.n) WDG_HIP(hipMemcpy(uid.data(), c.uid.p, c.n * 4, hipMemcpyDeviceToHost));
  if (nc) {
    WDG_HIP(hipMemcpy(A.data(), C.A.p, nc * 4, hipMemcpyDeviceToHost));
    WDG_HIP(hipMemcpy(B.data(), C.B.p, nc * 4, hipMemcpyDeviceToHost));
    WDG_HIP(hipMemcpy(info.data(), C.info.p, nc * 4, hipMemcpyDeviceToHost));
    WDG_HIP(hipMemcpy(bits.data(), RS.bits.p, bits.size() * 8, hipMemcpyDeviceToHost));
    if (c.field.periodic) {
      dist.resize((size_t)nc);
      WDG_HIP(hipMemcpy(dist.data(), C.dist.p, nc * 8, hipMemcpyDeviceToHost));
    }
    if (rec) {
      DBuf<double> out;
      out.ensure((size_t)nc * 16);
      ExportArgs a = export_args(c, out.p);
      hipLaunchKernelGGL(c.precision == 32 ? k_export<32> : k_export<64>, dim3(nblocks(nc, 256)), dim3(256), 0,
                         c.stream, a);
      WDG_LAUNCH_CHECK();
      rec->resize((size_t)nc * 16);
      WDG_HIP(hipMemcpyAsync(rec->data(), out.p, rec->size() * 8, hipMemcpyDeviceToHost, c.stream));
      WDG_HIP(hipStreamSynchronize(c.stream));
      out.release();
    }
  }
  std::vector<Record> R;
  auto key = [](uint32_t a, uint32_t b) { return std::make_pair(std::min(a, b), std::max(a, b)); };
  std::vector<std::pair<std::pair<uint32_t, uint32_t>, size_t>> order;
  for (int64_t s = 0; s < nc; s++) {
    if (!((bits[(size_t)s >> 6] >> (s & 63)) & 1ull)) continue;
    bool swap = (info[s] & CI_SWAP) != 0;
    Record r;
    r.a = uid[swap ? B[s] : A[s]];
    r.b = uid[swap ? A[s] : B[s]];
    for (int m = 0; m < 3; m++) r.cd[m] = c.field.periodic ? unpack_cd(dist[s], m) : 0;
    r.slot = s;
    R.push_back(r);
  }
  if (which == 0) {
    // potential set: enlarged AABBs of the last refresh, plus the real contacts
    std::vector<uint32_t> pa, pb;
    std::vector<uint64_t> pcd;
    potential_pairs(c, pa, pb, pcd);
    std::vector<std::pair<uint32_t, uint32_t>> have;
    for (auto& r : R) have.push_back(key(r.a, r.b));
    std::sort(have.begin(), have.end());
    for (size_t i = 0; i < pa.size(); i++) {
      if (std::binary_search(have.begin(), have.end(), key(pa[i], pb[i]))) continue;
      Record r;
      r.a = pa[i];
      r.b = pb[i];
      for (int m = 0; m < 3; m++) r.cd[m] = c.field.periodic ? unpack_cd(pcd[i], m) : 0;
      r.slot = -1;
      R.push_back(r);
    }
  }
  std::sort(R.begin(), R.end(), [&](const Record& x, const Record& y) { return key(x.a, x.b) < key(y.a, y.b); });
  return R;
}

extern "C" {

const char* wdg_last_error(void) { return g_err.c_str(); }

int wdg_ctx_create(int device, wdg_ctx** out) {
  if (!out) { g_err = "out is NULL"; return WDG_E_ARG; }
  *out = nullptr;
  wdg_ctx* ctx = new wdg_ctx();
  ctx->c.device = device;
  static std::atomic<uint64_t> serials{0};
  ctx->c.serial = ++serials;
  int rc = guard(ctx, [&] {
    int ndev = 0;
    WDG_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) throw Error(WDG_E_ARG, "no HIP device " + std::to_string(device));
    WDG_HIP(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
    ctx->c.dflags.ensure(64);
    WDG_HIP(hipMemset(ctx->c.dflags.p, 0, 64 * sizeof(uint32_t)));
    WDG_HIP(hipHostMalloc((void**)&ctx->c.hflags, 64 * sizeof(uint32_t), hipHostMallocDefault));
    ctx->c.red.ensure(64);
    // the device's copy of the no-allocation entry points' scratch array
    scratch_setup(device);
    ctx->c.field.distFactor = -1;
    ctx->c.field.loneMask = 2;
    ctx->c.field.cell[0] = ctx->c.field.cell[1] = ctx->c.field.cell[2] = 1;
  });
  if (rc != WDG_OK) {
    delete ctx;
    return rc;
  }
  *out = ctx;
  return WDG_OK;
}

int wdg_ctx_destroy(wdg_ctx* ctx) {
  if (!ctx) return WDG_OK;
  int rc = guard(ctx, [&] {
    Ctx& c = ctx->c;
    (void)hipStreamSynchronize(c.stream);
    for (auto& s : c.C) s.release();
    for (auto& r : c.R) r.release();
    c.act.release(); c.desc.release(); c.descDist.release(); c.energy.release(); c.enPart.release();
    c.energyS.release(); c.pe.release(); c.pc.release(); c.enC.release(); c.enSPart.release();
    c.energySScenes = -1;
    DBuf<double>* arrs[] = {c.pos, c.vel, c.angVel, c.force, c.torque, c.inertia, c.lastPos};
    for (auto* a : arrs) for (int k = 0; k < 3; k++) a[k].release();
    for (int k = 0; k < 4; k++) c.ori[k].release();
    for (int k = 0; k < NPD; k++) c.spareD[k].release();
    for (int k = 0; k < NPU; k++) c.spareU[k].release();
    c.mass.release(); c.radius.release(); c.flags.release(); c.mask.release(); c.mat.release();
    c.uid.release(); c.sto.release(); c.dMats.release();
    c.cellKey.release(); c.sortIdx.release(); c.cellKeyTmp.release(); c.sortIdxTmp.release();
    c.perm.release(); c.invPerm.release(); c.cellStart.release();
    c.scount.release(); c.soff.release(); c.stPartner.release(); c.stInfo.release(); c.stDist.release();
    c.adjEnt.release(); c.adjDeg.release(); c.stNb.release(); c.sdeg.release(); c.posf.release();
    c.fgeo.release(); c.bnd.release(); c.pairMax.release(); c.scene.release(); c.dtScene.release();
    c.sceneBodyOff.release(); c.sceneBody.release();
    c.fnode.release(); c.gbody.release(); c.bndPart.release(); c.vdScene.release(); c.sceneVmax.release();
    c.sceneBox.release(); c.sgrid.release();
    c.gstate.release();
    if (c.graph.exec) (void)hipGraphExecDestroy(c.graph.exec);
    c.graph.exec = nullptr; c.bigOff.release(); c.bigRank.release(); c.bigVal.release();
    c.bigValTmp.release(); c.bigDest.release(); c.bigKey.release(); c.bigKeyTmp.release(); c.bigList.release();
    for (auto& b : c.snap) b.release();
    for (int k = 0; k < Ctx::NSNAP; k++) {
      c.snapKey[k].release(); c.snapCd[k].release(); c.snapSlot[k].release(); c.snapRec[k].release();
      c.snapEp[k].release();
    }
    c.snapMatch.release();
    c.sortHist.release(); c.scanAux.release(); c.red.release(); c.xfer.release(); c.dflags.release();
    if (c.hflags) (void)hipHostFree(c.hflags);
    if (c.stream) (void)hipStreamDestroy(c.stream);
  });
  delete ctx;
  return rc;
}

int wdg_set_field(wdg_ctx* ctx, const wdg_field* f) {
  if (!ctx || !f) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    if (f->periodic)
      for (int k = 0; k < 3; k++)
        if (!(f->cell[k] > 0)) throw Error(WDG_E_ARG, "periodic cell size must be positive");
    bool changed = f->periodic != c.field.periodic;
    for (int k = 0; k < 3; k++) changed = changed || f->cell[k] != c.field.cell[k];
    changed = changed || f->distFactor != c.field.distFactor || f->loneMask != c.field.loneMask;
    // the reset value of the forces depends on gravity, and a captured step graph holds it by
    // value in its leapfrog arguments
    bool gravChanged = false;
    for (int k = 0; k < 3; k++) gravChanged = gravChanged || f->gravity[k] != c.field.gravity[k];
    if (gravChanged) {
      c.forceClean = false;
      c.gen++;
    }
    c.field = *f;
    if (changed) {
      c.boundsValid = false;
      c.gen++;
    }
  });
}

int wdg_set_materials(wdg_ctx* ctx, const wdg_material* mats, int32_t n) {
  if (!ctx || (n > 0 && !mats) || n < 0) { g_err = "bad materials"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    set_materials(ctx->c, mats, n);
  });
}

static void scene_bodies(Ctx& c);

int wdg_upload_particles(wdg_ctx* ctx, const wdg_particles* p) {
  if (!ctx || !p || p->n < 0 || (p->n > 0 && (!p->shape || !p->pos || !p->mass || !p->inertia))) {
    g_err = "bad particle description";
    return WDG_E_ARG;
  }
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    const int64_t n = p->n;
    if (n >= (int64_t)1 << 31) throw Error(WDG_E_CAPACITY, "too many particles");
    WDG_HIP(hipStreamSynchronize(c.stream));
    // storage order: boundary bodies (walls, facets; user-id order), then spheres (user-id
    // order until the first refresh)
    std::vector<uint32_t> order;
    for (int pass = 0; pass < 2; pass++)
      for (int64_t i = 0; i < n; i++) {
        int shape = p->shape[i];
        if (shape != WDG_SPHERE && shape != WDG_WALL && shape != WDG_FACET)
          throw Error(WDG_E_ARG, "unsupported shape " + std::to_string(shape));
        if ((pass == 0) == (shape != WDG_SPHERE)) order.push_back((uint32_t)i);
      }
    std::vector<uint32_t> flags((size_t)n), mask((size_t)n), uid((size_t)n), sto((size_t)n), scn((size_t)n);
    int64_t maxScene = 0;
    std::vector<int32_t> mat((size_t)n);
    std::vector<double> col[NPD];
    for (auto& v : col) v.resize((size_t)n);
    c.rmin = INFINITY;
    c.rmax = 0;
    int64_t nW = 0, nF = 0;
    std::vector<double> fgeo;
    for (int64_t s = 0; s < n; s++) {
      const int64_t i = order[s];
      const int shape = p->shape[i];
      int m = p->material ? p->material[i] : 0;
      if (m < 0 || m >= (int)c.mats.size()) throw Error(WDG_E_ARG, "material index out of range");
      uint32_t f = (p->flags ? p->flags[i] : 0u) & 0xFFFFu;
      if (p->integrated ? p->integrated[i] : 1) f |= F_INTEGRATED;
      f |= (uint32_t)shape << F_SHAPE_SHIFT;
      double r = 0.;
      if (shape == WDG_WALL) {
        int ax = p->wallAxis ? p->wallAxis[i] : 0, se = p->wallSense ? p->wallSense[i] : 0;
        if (ax < 0 || ax > 2 || se < -1 || se > 1) throw Error(WDG_E_ARG, "bad wall axis/sense");
        f |= (uint32_t)ax << F_AXIS_SHIFT;
        f |= (uint32_t)(se + 1) << F_SENSE_SHIFT;
        nW++;
      } else if (shape == WDG_FACET) {
        if (!p->facetVerts) throw Error(WDG_E_ARG, "facets need facetVerts");
        for (int k = 0; k < 3; k++)
          if ((p->vel && p->vel[3 * i + k] != 0.) || (p->angVel && p->angVel[3 * i + k] != 0.))
            throw Error(WDG_E_ARG, "Facet #" + std::to_string(i) + ": moving facet nodes are not supported (use fakeVel)");
        const double ht = p->halfThick ? p->halfThick[i] : 0.;
        if (!(ht >= 0.)) throw Error(WDG_E_ARG, "Facet #" + std::to_string(i) + ": halfThick must be >= 0");
        // fixed, outside DemField.nodes: the vertices never move
        f = (f & ~(uint32_t)F_INTEGRATED) | F_DOF_ALL;
        fgeo.resize((size_t)FGEO * (nW + 1), 0.);
        double* g = fgeo.data() + (size_t)FGEO * nW;
        for (int k = 0; k < 9; k++) g[k] = p->facetVerts[9 * i + k];
        for (int k = 0; k < 3; k++) g[9 + k] = p->fakeVel ? p->fakeVel[3 * i + k] : 0.;
        g[12] = ht;
        for (int k = 0; k < 3; k++) g[13 + k] = p->facetN21lim ? p->facetN21lim[3 * i + k] : NAN;
        nW++;
        nF++;
      } else {
        r = p->radius ? p->radius[i] : 0.;
        if (!(r > 0)) throw Error(WDG_E_ARG, "Sphere #" + std::to_string(i) + ": radius must be positive");
        c.rmin = std::min(c.rmin, r);
        c.rmax = std::max(c.rmax, r);
      }
      flags[s] = f;
      const int32_t sc = p->scene ? p->scene[i] : 0;
      if (sc < 0) throw Error(WDG_E_ARG, "negative scene index");
      scn[s] = (uint32_t)sc;
      maxScene = std::max<int64_t>(maxScene, sc);
      mask[s] = p->mask ? p->mask[i] : (shape != WDG_SPHERE ? 3u : 5u);
      mat[s] = m;
      uid[s] = (uint32_t)i;
      sto[i] = (uint32_t)s;
      // the permuted array list of the refresh: pos3 vel3 angVel3 ori4 force3 torque3 inertia3 mass radius
      for (int k = 0; k < 3; k++) {
        col[k][s] = p->pos[3 * i + k];
        col[3 + k][s] = p->vel ? p->vel[3 * i + k] : 0.;
        col[6 + k][s] = p->angVel ? p->angVel[3 * i + k] : 0.;
        col[13 + k][s] = p->force ? p->force[3 * i + k] : 0.;
        col[16 + k][s] = p->torque ? p->torque[3 * i + k] : 0.;
        col[19 + k][s] = p->inertia[3 * i + k];
      }
      for (int k = 0; k < 4; k++) col[9 + k][s] = p->ori ? p->ori[4 * i + k] : (k == 0 ? 1. : 0.);
      col[22][s] = p->mass[i];
      col[23][s] = r;
      if (shape == WDG_FACET)
        for (int k = 0; k < 3; k++) col[k][s] = p->facetVerts[9 * i + k];  // the node is vertex 0
    }
    c.n = n;
    c.nW = nW;
    c.nFacet = nF;
    fgeo.resize((size_t)FGEO * (nW + 1), 0.);
    c.fgeo.ensure(fgeo.size());
    WDG_HIP(hipMemcpy(c.fgeo.p, fgeo.data(), fgeo.size() * sizeof(double), hipMemcpyHostToDevice));
    c.bnd.ensure((size_t)6 * (nW + 1));
    c.fnode.ensure((size_t)18 * (nW + 1));
    {
      std::vector<uint32_t> gb;
      for (int64_t b = 0; b < nW; b++)
        if (((flags[b] >> F_SHAPE_SHIFT) & 3u) != (uint32_t)WDG_FACET) gb.push_back((uint32_t)b);
      c.nGBody = (int64_t)gb.size();
      c.gbodyHost = gb;
      c.gbody.ensure(gb.size() + 1);
      if (!gb.empty()) WDG_HIP(hipMemcpy(c.gbody.p, gb.data(), gb.size() * 4, hipMemcpyHostToDevice));
      c.bndPart.ensure((size_t)64 * 6 * (gb.size() + 1));
      c.bndParts = 1;
    }
    WDG_HIP(hipMemset(c.fnode.p, 0, (size_t)18 * (nW + 1) * sizeof(double)));
    c.bodyUid.assign(order.begin(), order.begin() + nW);
    DBuf<double>* d[NPD] = {&c.pos[0], &c.pos[1], &c.pos[2], &c.vel[0], &c.vel[1], &c.vel[2], &c.angVel[0],
                            &c.angVel[1], &c.angVel[2], &c.ori[0], &c.ori[1], &c.ori[2], &c.ori[3], &c.force[0],
                            &c.force[1], &c.force[2], &c.torque[0], &c.torque[1], &c.torque[2], &c.inertia[0],
                            &c.inertia[1], &c.inertia[2], &c.mass, &c.radius};
    for (int k = 0; k < NPD; k++) {
      d[k]->ensure(n + 1);
      if (n) WDG_HIP(hipMemcpy(d[k]->p, col[k].data(), n * sizeof(double), hipMemcpyHostToDevice));
    }
    for (int k = 0; k < 3; k++) {
      c.lastPos[k].ensure(n + 1);
      if (n) WDG_HIP(hipMemcpy(c.lastPos[k].p, col[k].data(), n * sizeof(double), hipMemcpyHostToDevice));
    }
    c.flags.ensure(n + 1); c.mask.ensure(n + 1); c.mat.ensure(n + 1); c.uid.ensure(n + 1); c.sto.ensure(n + 1);
    c.scene.ensure(n + 1);
    c.maxScene = maxScene;
    c.bodyScene.assign(scn.begin(), scn.begin() + nW);
    scene_bodies(c);
    c.gen++;
    if (n) WDG_HIP(hipMemcpy(c.scene.p, scn.data(), n * 4, hipMemcpyHostToDevice));
    if (n) {
      WDG_HIP(hipMemcpy(c.flags.p, flags.data(), n * 4, hipMemcpyHostToDevice));
      WDG_HIP(hipMemcpy(c.mask.p, mask.data(), n * 4, hipMemcpyHostToDevice));
      WDG_HIP(hipMemcpy(c.mat.p, mat.data(), n * 4, hipMemcpyHostToDevice));
      WDG_HIP(hipMemcpy(c.uid.p, uid.data(), n * 4, hipMemcpyHostToDevice));
      WDG_HIP(hipMemcpy(c.sto.p, sto.data(), n * 4, hipMemcpyHostToDevice));
    }
    drop_contacts(c);
    WDG_HIP(hipStreamSynchronize(c.stream));
    c.uploads++;
    c.forceClean = false;
    c.verletDist = NAN;
    c.dynStepLast = c.dynStepPrev = -1;
  });
}

int wdg_upload(wdg_ctx* ctx, int32_t field, const double* host) {
  if (!ctx || !host) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    int w;
    DBuf<double>* b = field_bufs(c, field, w);
    if ((field == WDG_F_FORCE || field == WDG_F_TORQUE) && c.forcesPending) flush_forces(c);
    put_field(c, b, w, host);
    if (field == WDG_F_POS) c.dirtyKnown = false;
    // new forces, or a new mass changes the gravity part of the reset value
    if (field == WDG_F_FORCE || field == WDG_F_TORQUE || field == WDG_F_MASS) c.forceClean = false;
  });
}

int wdg_download(wdg_ctx* ctx, int32_t field, double* host) {
  if (!ctx || !host) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    int w;
    DBuf<double>* b = field_bufs(c, field, w);
    if ((field == WDG_F_FORCE || field == WDG_F_TORQUE) && c.forcesPending) flush_forces(c);
    get_field(c, b, w, host);
  });
}

int wdg_collide(wdg_ctx* ctx, const wdg_collide_params* p) {
  if (!ctx || !p) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] { collide(ctx->c, p); });
}

int wdg_contact_loop(wdg_ctx* ctx, const wdg_contact_params* p) {
  if (!ctx || !p) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] { contact_loop(ctx->c, p); });
}

int wdg_leapfrog(wdg_ctx* ctx, const wdg_leapfrog_params* p) {
  if (!ctx || !p) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] { leapfrog(ctx->c, p); });
}

int wdg_intra_force(wdg_ctx* ctx) {
  if (!ctx) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] { intra_force(ctx->c); });
}

int wdg_facet_nodes(wdg_ctx* ctx, const int64_t* ids, int64_t n, double* out) {
  if (!ctx || n < 0 || (n > 0 && (!ids || !out))) { g_err = "bad arguments"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    std::vector<int64_t> body((size_t)n);
    std::unordered_map<uint32_t, int64_t> byUid;
    for (int64_t b = 0; b < c.nW; b++) byUid[c.bodyUid[b]] = b;
    std::vector<uint32_t> fl((size_t)c.nW);
    if (c.nW) WDG_HIP(hipMemcpyAsync(fl.data(), c.flags.p, c.nW * 4, hipMemcpyDeviceToHost, c.stream));
    std::vector<double> all((size_t)18 * c.nW);
    if (c.nW) WDG_HIP(hipMemcpyAsync(all.data(), c.fnode.p, all.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    WDG_HIP(hipStreamSynchronize(c.stream));
    for (int64_t j = 0; j < n; j++) {
      auto it = ids[j] >= 0 && ids[j] < c.n ? byUid.find((uint32_t)ids[j]) : byUid.end();
      if (it == byUid.end() || flag_shape(fl[it->second]) != WDG_FACET)
        throw Error(WDG_E_ARG, "particle #" + std::to_string(ids[j]) + " is not a facet");
      std::copy(&all[18 * it->second], &all[18 * it->second] + 18, out + 18 * j);
    }
  });
}

int wdg_force_reset(wdg_ctx* ctx) {
  if (!ctx) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] { force_reset(ctx->c); });
}

int wdg_crit_dt(wdg_ctx* ctx, double* critDt) {
  if (!ctx || !critDt) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] { *critDt = crit_dt(ctx->c); });
}

int wdg_contact_count(wdg_ctx* ctx, int32_t which, int64_t* n) {
  if (!ctx || !n) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    if (which == 1) { *n = count_real(ctx->c); return; }
    *n = (int64_t)collect_contacts(ctx->c, which, nullptr).size();
  });
}

int wdg_contact_list(wdg_ctx* ctx, int32_t which, int32_t* ids, int32_t* cellDist, double* data) {
  if (!ctx) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    std::vector<double> rec;
    std::vector<Record> R = collect_contacts(ctx->c, which, data ? &rec : nullptr);
    for (size_t k = 0; k < R.size(); k++) {
      if (ids) { ids[2 * k] = (int32_t)R[k].a; ids[2 * k + 1] = (int32_t)R[k].b; }
      if (cellDist) for (int m = 0; m < 3; m++) cellDist[3 * k + m] = R[k].cd[m];
      if (data) {
        if (R[k].slot >= 0) std::memcpy(data + 16 * k, rec.data() + 16 * R[k].slot, 16 * sizeof(double));
        else { for (int m = 0; m < 15; m++) data[16 * k + m] = NAN; data[16 * k + 15] = 0.; }
      }
    }
  });
}

int wdg_restore_contacts(wdg_ctx* ctx, int64_t n, const int32_t* ids, const int32_t* cellDist, const double* data,
                         double verletDist) {
  if (!ctx || n < 0 || (n > 0 && (!ids || !data))) { g_err = "bad arguments"; return WDG_E_ARG; }
  return guard(ctx, [&] { restore_contacts(ctx->c, n, ids, cellDist, data, verletDist); });
}

int wdg_observables(wdg_ctx* ctx, const double boxMin[3], const double boxMax[3], double out[3]) {
  if (!ctx || !boxMin || !boxMax || !out) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] { observables(ctx->c, boxMin, boxMax, out); });
}

int wdg_set_precision(wdg_ctx* ctx, int32_t bits) {
  if (!ctx || (bits != 32 && bits != 64)) { g_err = "precision must be 32 or 64"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    if (c.precision == bits) return;
    // the contact state is laid out per precision: start the history afresh
    if (c.forcesPending) flush_forces(c);
    drop_contacts(c);
    c.precision = bits;
    c.gen++;
  });
}

// boundary bodies grouped by scene (their storage indices never change)
static void scene_bodies(Ctx& c) {
  const int64_t ns = std::max<int64_t>(c.nScenes, c.maxScene + 1);
  std::vector<uint32_t> off((size_t)ns + 1, 0u), body(c.bodyScene.size());
  for (uint32_t sc : c.bodyScene) off[sc + 1]++;
  for (int64_t k = 0; k < ns; k++) off[k + 1] += off[k];
  std::vector<uint32_t> fill(off.begin(), off.end() - 1);
  for (size_t w = 0; w < c.bodyScene.size(); w++) body[fill[c.bodyScene[w]]++] = (uint32_t)w;
  c.sceneBodyOff.ensure(off.size());
  c.sceneBody.ensure(std::max<size_t>(body.size(), 1));
  WDG_HIP(hipMemcpy(c.sceneBodyOff.p, off.data(), off.size() * 4, hipMemcpyHostToDevice));
  if (!body.empty()) WDG_HIP(hipMemcpy(c.sceneBody.p, body.data(), body.size() * 4, hipMemcpyHostToDevice));
}

int wdg_set_scenes(wdg_ctx* ctx, int32_t nScenes, const double* dt) {
  if (!ctx || nScenes < 0 || (nScenes > 0 && !dt)) { g_err = "bad arguments"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    for (int32_t k = 0; k < nScenes; k++)
      if (!(dt[k] > 0)) throw Error(WDG_E_ARG, "scene " + std::to_string(k) + ": dt must be positive");
    if (c.forcesPending) flush_forces(c);
    WDG_HIP(hipStreamSynchronize(c.stream));
    c.nScenes = nScenes;
    scene_bodies(c);
    c.dtScene.ensure((size_t)std::max(nScenes, 1));
    if (nScenes) WDG_HIP(hipMemcpy(c.dtScene.p, dt, nScenes * sizeof(double), hipMemcpyHostToDevice));
    c.dtSceneHost.assign(dt, dt + nScenes);
    c.gen++;
  });
}

int wdg_run_graph(wdg_ctx* ctx, const wdg_program* prog, int64_t nSteps, int64_t* step, double* dt) {
  if (!ctx || !prog || !step || !dt || nSteps < 0) { g_err = "bad arguments"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    run_graph(ctx->c, prog, nSteps, step, dt);
  });
}

int wdg_tune(wdg_ctx* ctx, const wdg_program* prog, int64_t* step, double* dt, wdg_tuning* out, int64_t* stepsTaken) {
  if (!ctx || !prog || !step || !dt) { g_err = "bad arguments"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    const int64_t n = tune(ctx->c, prog, step, dt);
    if (out) *out = ctx->c.tune;
    if (stepsTaken) *stepsTaken = n;
  });
}

int wdg_get_tuning(wdg_ctx* ctx, wdg_tuning* out) {
  if (!ctx || !out) { g_err = "NULL argument"; return WDG_E_ARG; }
  *out = ctx->c.tune;
  return WDG_OK;
}

int wdg_set_update_phys(wdg_ctx* ctx, int32_t mode) {
  if (!ctx || mode < WDG_UPDATE_PHYS_NEVER || mode > WDG_UPDATE_PHYS_ONCE) {
    g_err = "updatePhys: 0 (never), 1 (always) or 2 (once)";
    return WDG_E_ARG;
  }
  ctx->c.updatePhys = mode;
  return WDG_OK;
}

int wdg_set_eval_stress(wdg_ctx* ctx, int32_t on) {
  if (!ctx) { g_err = "null argument"; return WDG_E_ARG; }
  ctx->c.evalStress = on != 0;
  return WDG_OK;
}

int wdg_contact_loop_stress(wdg_ctx* ctx, double out[9]) {
  if (!ctx || !out) { g_err = "null argument"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    for (int k = 0; k < 9; k++) out[k] = 0.;
    if (!c.clStressOn) return;
    WDG_HIP(hipMemcpyAsync(out, c.clStress.p, 9 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    WDG_HIP(hipStreamSynchronize(c.stream));
  });
}

int wdg_get_update_phys(wdg_ctx* ctx, int32_t* mode) {
  if (!ctx || !mode) { g_err = "null argument"; return WDG_E_ARG; }
  *mode = ctx->c.updatePhys;
  return WDG_OK;
}

int wdg_set_tuning(wdg_ctx* ctx, const wdg_tuning* t) {
  if (!ctx || !t || !(t->cfBlocks == 1 || t->cfBlocks == 2 || t->cfBlocks == 4 || t->cfBlocks == 8) ||
      t->graphSteps < 2 || (t->graphSteps & 1) || t->graphSteps > 256 || t->skinSteps < 0 || t->skinSteps > 1024) {
    g_err = "tuning: cfBlocks in {1,2,4,8}, graphSteps even in [2,256], skinSteps in [0,1024]";
    return WDG_E_ARG;
  }
  // a new skin takes effect at a refresh: the next collider call makes one
  if (t->skinSteps != ctx->c.tune.skinSteps) ctx->c.boundsValid = false;
  ctx->c.tune = *t;
  return WDG_OK;
}

int wdg_stress(wdg_ctx* ctx, double volume, double out[9]) {
  if (!ctx || !out) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] { stress(ctx->c, volume, out); });
}

int wdg_kinetic_energy(wdg_ctx* ctx, double* ek) {
  if (!ctx || !ek) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] { *ek = kinetic_energy(ctx->c); });
}

int wdg_kinetic_energy_noalloc(wdg_ctx* ctx, double* dOut) {
  if (!ctx || !dOut) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] { kinetic_energy_noalloc(ctx->c, dOut); });
}

int wdg_set_energy_tracking(wdg_ctx* ctx, int32_t on) {
  if (!ctx) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] { energy_enable(ctx->c, on != 0); });
}

int wdg_energy(wdg_ctx* ctx, double out[7]) {
  if (!ctx || !out) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] { energy_read(ctx->c, out); });
}

int wdg_energy_set(wdg_ctx* ctx, const double in[7]) {
  if (!ctx || !in) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    if (c.nScenes > 1) throw Error(WDG_E_STATE, "a batched context of several scenes: wdg_energy_scenes_set");
    c.energy.ensure(WDG_EN_N);
    WDG_HIP(hipMemcpyAsync(c.energy.p, in, WDG_EN_N * sizeof(double), hipMemcpyHostToDevice, c.stream));
    WDG_HIP(hipStreamSynchronize(c.stream));
  });
}

int wdg_energy_scenes(wdg_ctx* ctx, double* out) {
  if (!ctx || !out) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    std::vector<double> v;
    energy_scenes_read(ctx->c, v);
    std::memcpy(out, v.data(), v.size() * sizeof(double));
  });
}

int wdg_energy_scenes_set(wdg_ctx* ctx, const double* in) {
  if (!ctx || !in) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    if (c.nScenes <= 1) throw Error(WDG_E_STATE, "not a batched context of several scenes (wdg_set_scenes)");
    c.energyS.ensure((size_t)c.nScenes * WDG_EN_N);
    WDG_HIP(hipMemcpyAsync(c.energyS.p, in, (size_t)c.nScenes * WDG_EN_N * sizeof(double), hipMemcpyHostToDevice,
                           c.stream));
    WDG_HIP(hipStreamSynchronize(c.stream));
    c.energySScenes = c.nScenes;
  });
}

int wdg_energy_clear(wdg_ctx* ctx) {
  if (!ctx) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] { energy_clear(ctx->c); });
}

int wdg_stats(wdg_ctx* ctx, wdg_counters* out) {
  if (!ctx || !out) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    std::memset(out, 0, sizeof(*out));
    out->nParticles = c.n;
    out->nCandidates = c.con().n;
    out->nReal = count_real(c);
    out->nRefresh = c.nRefresh;
    out->nSteps = c.nCollide;
    out->verletDist = c.verletDist;
    out->brute = c.grid.brute;
    for (int k = 0; k < 3; k++) out->gridCells[k] = c.grid.brute ? 0 : c.grid.nc[k];
    out->refreshSeconds = c.refreshSec;
  });
}

// pown: square-and-multiply (Math.hpp:166-175)
static double pown(double x, unsigned n) {
  unsigned d = n >> 1, r = n & 1;
  double x2d = d == 0 ? 1. : pown(x * x, d);
  double xr = r == 0 ? 1. : x;
  return x2d * xr;
}

// Scene::doOneStep (Scene.cpp:366-467) for a fixed engine list
int wdg_run(wdg_ctx* ctx, const wdg_program* prog, int64_t nSteps, int64_t* step, double* dt) {
  if (!ctx || !prog || !step || !dt || prog->nEngines < 0 || prog->nEngines > 8) {
    g_err = "bad arguments";
    return WDG_E_ARG;
  }
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    std::vector<double> crs, nextScene;
    for (int64_t s = 0; s < nSteps; s++) {
      double nextDt = NAN;
      bool sceneDt = false;
      c.stepNow = *step;
      if (c.trackEnergy) energy_step_begin(c);
      for (int e = 0; e < prog->nEngines; e++) {
        switch (prog->engines[e]) {
          case WDG_E_FORCERESETTER: force_reset(c); break;
          case WDG_E_INTRAFORCE: intra_force(c); break;
          case WDG_E_COLLIDER: collide(c, &prog->collide); break;
          case WDG_E_CONTACTLOOP: {
            wdg_contact_params p = prog->contact;
            p.step = *step;
            p.dt = *dt;
            contact_loop(c, &p);
          } break;
          case WDG_E_LEAPFROG: {
            wdg_leapfrog_params p = prog->leapfrog;
            p.dt = *dt;
            leapfrog(c, &p);
          } break;
          case WDG_E_DYNDT: {
            // PeriodicEngine::isActivated (Engine.cpp:170-196) with stepModulo and initRun
            long per = prog->dynDtPeriod;
            if (per <= 0) break;
            if (c.dynStepLast < 0) c.dynStepLast = *step;
            if ((*step - c.dynStepLast >= per) || (*step % per == 0)) {
              c.dynStepPrev = c.dynStepLast;
              c.dynStepLast = *step;
              if (c.nScenes > 0) {
                // batched scenes: DynDt::run of each scene with its own dt (every scene is
                // stepped by this loop, so all share the activation steps)
                crit_dt_scenes(c, crs);
                const int64_t ns = *step - c.dynStepPrev;
                nextScene.assign(crs.size(), NAN);
                for (size_t q = 0; q < crs.size(); q++) {
                  if (std::isinf(crs[q])) continue;
                  const double d0 = c.dtSceneHost[q];
                  const double nd = std::min(crs[q] * prog->dtSafety, d0 * pown(1. + prog->maxRelInc, (unsigned)ns));
                  if (nd != d0) {
                    nextScene[q] = nd;
                    sceneDt = true;
                  }
                }
                break;
              }
              // DynDt::run (DynDt.cpp:90-113)
              double cr = crit_dt(c);
              if (std::isinf(cr)) break;
              int64_t ns = *step - c.dynStepPrev;
              double maxNewDt = *dt * pown(1. + prog->maxRelInc, (unsigned)ns);
              double nd = std::min(cr * prog->dtSafety, maxNewDt);
              if (nd != *dt) nextDt = nd;
            }
          } break;
          default: throw Error(WDG_E_ARG, "unknown engine code " + std::to_string(prog->engines[e]));
        }
      }
      *step += 1;
      if (!std::isnan(nextDt)) *dt = nextDt;
      if (sceneDt) {
        for (size_t q = 0; q < nextScene.size(); q++)
          if (!std::isnan(nextScene[q])) c.dtSceneHost[q] = nextScene[q];
        WDG_HIP(hipMemcpyAsync(c.dtScene.p, c.dtSceneHost.data(), c.dtSceneHost.size() * sizeof(double),
                               hipMemcpyHostToDevice, c.stream));
        WDG_HIP(hipStreamSynchronize(c.stream));
      }
    }
  });
}

int wdg_crit_dt_scenes(wdg_ctx* ctx, double* out) {
  if (!ctx || !out) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    std::vector<double> v;
    crit_dt_scenes(ctx->c, v);
    std::copy(v.begin(), v.end(), out);
  });
}

int wdg_get_scene_dt(wdg_ctx* ctx, double* out) {
  if (!ctx || !out) { g_err = "NULL argument"; return WDG_E_ARG; }
  if (ctx->c.nScenes <= 0) { g_err = "not a batched context (wdg_set_scenes)"; return WDG_E_STATE; }
  std::copy(ctx->c.dtSceneHost.begin(), ctx->c.dtSceneHost.end(), out);
  return WDG_OK;
}

int wdg_relax(wdg_ctx* ctx, int32_t maxIter, double inflate, double tol, double* maxOverlap, int32_t* iters) {
  if (!ctx || !maxOverlap || !iters || maxIter < 0 || !(inflate >= 0)) {
    g_err = "bad arguments";
    return WDG_E_ARG;
  }
  return guard(ctx, [&] {
    int it = 0;
    relax(ctx->c, maxIter, inflate, tol, maxOverlap, &it);
    *iters = it;
  });
}

int wdg_sync(wdg_ctx* ctx) {
  if (!ctx) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    WDG_HIP(hipStreamSynchronize(ctx->c.stream));
    check_errors(ctx->c);
  });
}

int wdg_snapshot(wdg_ctx* ctx, int32_t slot) {
  if (!ctx) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] { snapshot_save(ctx->c, slot); });
}

int wdg_restore(wdg_ctx* ctx, int32_t slot) {
  if (!ctx) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] { snapshot_load(ctx->c, slot); });
}

int wdg_snapshot_distance(wdg_ctx* ctx, int32_t slot, double* maxDist) {
  if (!ctx || !maxDist) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] { *maxDist = snapshot_distance(ctx->c, slot); });
}

void* wdg_stream(wdg_ctx* ctx) { return ctx ? (void*)ctx->c.stream : nullptr; }

}  // extern "C"
