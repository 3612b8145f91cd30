// extern "C" entry points of the DEM step engine (see include/woo_amd.h).
#include <algorithm>
#include <cstring>

#include "wdg_internal.hpp"
#include "wdg_math.hpp"
#include "wdg_physics.hpp"

struct wdg_ctx {
  wdg::Ctx c;
};

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

// upload n*w doubles (AoS, row-major) into w SoA arrays
void upload_soa(wdg::Ctx& c, wdg::DBuf<double>* dst, int w, const double* host, int64_t n, double fill[4]) {
  std::vector<double> col((size_t)n);
  for (int k = 0; k < w; k++) {
    dst[k].ensure((size_t)n + 1);
    for (int64_t i = 0; i < n; i++) col[i] = host ? host[i * w + k] : fill[k];
    if (n) WDG_HIP(hipMemcpy(dst[k].p, col.data(), n * sizeof(double), hipMemcpyHostToDevice));
  }
}
void download_soa(wdg::Ctx& c, wdg::DBuf<double>* src, int w, double* host, int64_t n) {
  std::vector<double> col((size_t)n);
  WDG_HIP(hipStreamSynchronize(c.stream));
  for (int k = 0; k < w; k++) {
    if (n) WDG_HIP(hipMemcpy(col.data(), src[k].p, n * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) host[i * w + k] = col[i];
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
}  // namespace

namespace wdg {

// one record of 16 doubles per slot, for inspection and parity tests
struct ExportArgs {
  int64_t nc;
  const uint32_t *A, *B, *info, *real;
  const double *N0, *N1, *N2, *Y0, *Y1, *Y2, *P0, *P1, *P2, *F0, *F1, *F2, *uN;
  const double *rad, *mass;
  const int32_t* mat;
  const wdg_material* mats;
  int hertz;
  double poisson, en, viscAlpha;
  double* out;
};

__global__ void k_export(ExportArgs a) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.nc) return;
  double* o = a.out + 16 * s;
  const bool real = a.real[s] != 0u;
  if (!real) {
    for (int k = 0; k < 15; k++) o[k] = NAN;
    o[15] = 0.;
    return;
  }
  const uint32_t inf = a.info[s];
  const bool swap = (inf & CI_SWAP) != 0u;
  const uint32_t pA = swap ? a.B[s] : a.A[s], pB = swap ? a.A[s] : a.B[s];
  double r1, r2;
  if ((inf & CI_TYPE_MASK) == CT_SS) { r1 = a.rad[pA]; r2 = a.rad[pB]; }
  else { r1 = -a.rad[pB]; r2 = a.rad[pB]; }
  double l0, l1, A;
  lens_area(r1, r2, l0, l1, A);
  const wdg_material mA = a.mats[a.mat[pA]], mB = a.mats[a.mat[pB]];
  PhysVals ph;
  double kn, kt;
  const double uN = a.uN[s];
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
  o[0] = a.N0[s]; o[1] = a.N1[s]; o[2] = a.N2[s];
  o[3] = a.Y0[s]; o[4] = a.Y1[s]; o[5] = a.Y2[s];
  o[6] = a.P0[s]; o[7] = a.P1[s]; o[8] = a.P2[s];
  o[9] = a.F0[s]; o[10] = a.F1[s]; o[11] = a.F2[s];
  o[12] = uN; o[13] = kn; o[14] = kt; o[15] = 1.;
}

}  // namespace wdg

using namespace wdg;

extern "C" {

const char* wdg_last_error(void) { return g_err.c_str(); }

int wdg_ctx_create(int device, wdg_ctx** out) {
  if (!out) { g_err = "out is NULL"; return WDG_E_ARG; }
  *out = nullptr;
  wdg_ctx* ctx = new wdg_ctx();
  ctx->c.device = device;
  int rc = guard(ctx, [&] {
    int ndev = 0;
    WDG_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) throw Error(WDG_E_ARG, "no HIP device " + std::to_string(device));
    WDG_HIP(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
    ctx->c.dflags.ensure(64);
    WDG_HIP(hipMemset(ctx->c.dflags.p, 0, 64 * sizeof(uint32_t)));
    WDG_HIP(hipHostMalloc((void**)&ctx->c.hflags, 64 * sizeof(uint32_t), hipHostMallocDefault));
    ctx->c.red.ensure(64);
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
    DBuf<double>* arrs[] = {c.pos, c.vel, c.angVel, c.force, c.torque, c.inertia, c.lastPos};
    for (auto* a : arrs) for (int k = 0; k < 3; k++) a[k].release();
    for (int k = 0; k < 4; k++) c.ori[k].release();
    c.mass.release(); c.radius.release(); c.flags.release(); c.mask.release(); c.mat.release();
    c.dMats.release(); c.dWalls.release();
    c.cellKey.release(); c.sortIdx.release(); c.cellKeyTmp.release(); c.sortIdxTmp.release();
    c.cellStart.release(); c.cellEnd.release(); c.sx.release(); c.sy.release(); c.sz.release(); c.shs.release();
    c.smask.release(); c.scount.release(); c.soff.release(); c.pairKey.release(); c.pairKeyTmp.release();
    c.pairDist.release(); c.pairInfo.release(); c.pairVal.release(); c.pairValTmp.release(); c.runFlag.release();
    c.runOff.release(); c.keyB.release(); c.keyBTmp.release(); c.valB.release(); c.valBTmp.release();
    c.sortHist.release(); c.scanAux.release(); c.red.release(); c.dflags.release();
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
    c.field = *f;
    if (changed) c.boundsValid = false;
  });
}

int wdg_set_materials(wdg_ctx* ctx, const wdg_material* mats, int32_t n) {
  if (!ctx || (n > 0 && !mats) || n < 0) { g_err = "bad materials"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    c.mats.assign(mats, mats + n);
    c.dMats.ensure((size_t)std::max(n, 1));
    if (n) WDG_HIP(hipMemcpy(c.dMats.p, mats, n * sizeof(wdg_material), hipMemcpyHostToDevice));
  });
}

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
    std::vector<uint32_t> flags((size_t)n), mask((size_t)n);
    std::vector<int32_t> mat((size_t)n), walls, sph;
    std::vector<double> rad((size_t)n, 0.), mass((size_t)n);
    c.rmin = INFINITY;
    c.rmax = 0;
    for (int64_t i = 0; i < n; i++) {
      int shape = p->shape[i];
      if (shape != WDG_SPHERE && shape != WDG_WALL) throw Error(WDG_E_ARG, "unsupported shape " + std::to_string(shape));
      int m = p->material ? p->material[i] : 0;
      if (m < 0 || m >= (int)c.mats.size()) throw Error(WDG_E_ARG, "material index out of range");
      uint32_t f = (p->flags ? p->flags[i] : 0u) & 0xFFFFu;
      if (p->integrated ? p->integrated[i] : 1) f |= F_INTEGRATED;
      f |= (uint32_t)shape << F_SHAPE_SHIFT;
      if (shape == WDG_WALL) {
        int ax = p->wallAxis ? p->wallAxis[i] : 0, se = p->wallSense ? p->wallSense[i] : 0;
        if (ax < 0 || ax > 2 || se < -1 || se > 1) throw Error(WDG_E_ARG, "bad wall axis/sense");
        f |= (uint32_t)ax << F_AXIS_SHIFT;
        f |= (uint32_t)(se + 1) << F_SENSE_SHIFT;
        walls.push_back((int32_t)i);
      } else {
        double r = p->radius ? p->radius[i] : 0.;
        if (!(r > 0)) throw Error(WDG_E_ARG, "Sphere #" + std::to_string(i) + ": radius must be positive");
        rad[i] = r;
        c.rmin = std::min(c.rmin, r);
        c.rmax = std::max(c.rmax, r);
        sph.push_back((int32_t)i);
      }
      flags[i] = f;
      mask[i] = p->mask ? p->mask[i] : (shape == WDG_WALL ? 3u : 5u);
      mat[i] = m;
      mass[i] = p->mass[i];
    }
    c.n = n;
    c.nSph = (int64_t)sph.size();
    c.wallIds = walls;
    std::vector<int32_t> ids(walls);
    ids.insert(ids.end(), sph.begin(), sph.end());
    c.dWalls.ensure(ids.size() + 1);
    if (!ids.empty()) WDG_HIP(hipMemcpy(c.dWalls.p, ids.data(), ids.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    double zero[4] = {0, 0, 0, 0}, qid[4] = {1, 0, 0, 0};
    upload_soa(c, c.pos, 3, p->pos, n, zero);
    upload_soa(c, c.vel, 3, p->vel, n, zero);
    upload_soa(c, c.angVel, 3, p->angVel, n, zero);
    upload_soa(c, c.ori, 4, p->ori, n, qid);
    upload_soa(c, c.force, 3, p->force, n, zero);
    upload_soa(c, c.torque, 3, p->torque, n, zero);
    upload_soa(c, c.inertia, 3, p->inertia, n, zero);
    upload_soa(c, &c.mass, 1, mass.data(), n, zero);
    upload_soa(c, &c.radius, 1, rad.data(), n, zero);
    upload_soa(c, c.lastPos, 3, p->pos, n, zero);
    c.flags.ensure(n + 1); c.mask.ensure(n + 1); c.mat.ensure(n + 1);
    if (n) {
      WDG_HIP(hipMemcpy(c.flags.p, flags.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
      WDG_HIP(hipMemcpy(c.mask.p, mask.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
      WDG_HIP(hipMemcpy(c.mat.p, mat.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    // contacts are dropped: the next collider call is a full run
    for (auto& s : c.C) {
      s.n = 0;
      s.ensure(0, n, c.field.periodic != 0);
      WDG_HIP(hipMemset(s.offA.p, 0, (n + 1) * sizeof(uint32_t)));
      WDG_HIP(hipMemset(s.offB.p, 0, (n + 1) * sizeof(uint32_t)));
    }
    c.boundsValid = false;
    c.dirtyKnown = false;
    c.forcesPending = false;
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
    WDG_HIP(hipStreamSynchronize(c.stream));
    double zero[4] = {0, 0, 0, 0};
    if (field == WDG_F_MASS) {
      WDG_HIP(hipMemcpy(c.mass.p, host, c.n * sizeof(double), hipMemcpyHostToDevice));
    } else {
      upload_soa(c, b, w, host, c.n, zero);
    }
    if (field == WDG_F_POS) c.dirtyKnown = false;
  });
}

int wdg_download(wdg_ctx* ctx, int32_t field, double* host) {
  if (!ctx || !host) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    int w;
    DBuf<double>* b = field_bufs(c, field, w);
    if ((field == WDG_F_FORCE || field == WDG_F_TORQUE) && c.forcesPending) flush_forces(c);
    download_soa(c, b, w, host, c.n);
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
    Ctx& c = ctx->c;
    Contacts& C = c.con();
    if (which == 1) { *n = count_real(c); return; }
    std::vector<uint32_t> info((size_t)C.n), real((size_t)C.n);
    WDG_HIP(hipStreamSynchronize(c.stream));
    if (C.n) {
      WDG_HIP(hipMemcpy(info.data(), C.info.p, C.n * sizeof(uint32_t), hipMemcpyDeviceToHost));
      WDG_HIP(hipMemcpy(real.data(), C.real.p, C.n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    }
    int64_t k = 0;
    for (int64_t s = 0; s < C.n; s++) k += (real[s] || (info[s] & CI_AABB)) ? 1 : 0;
    *n = k;
  });
}

int wdg_contact_list(wdg_ctx* ctx, int32_t which, int32_t* ids, int32_t* cellDist, double* data) {
  if (!ctx) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    Contacts& C = c.con();
    const int64_t nc = C.n;
    std::vector<uint32_t> A((size_t)nc), B((size_t)nc), info((size_t)nc), real((size_t)nc);
    std::vector<uint64_t> dist;
    std::vector<double> rec;
    WDG_HIP(hipStreamSynchronize(c.stream));
    if (nc) {
      WDG_HIP(hipMemcpy(A.data(), C.A.p, nc * 4, hipMemcpyDeviceToHost));
      WDG_HIP(hipMemcpy(B.data(), C.B.p, nc * 4, hipMemcpyDeviceToHost));
      WDG_HIP(hipMemcpy(info.data(), C.info.p, nc * 4, hipMemcpyDeviceToHost));
      WDG_HIP(hipMemcpy(real.data(), C.real.p, nc * 4, hipMemcpyDeviceToHost));
      if (c.field.periodic) {
        dist.resize((size_t)nc);
        WDG_HIP(hipMemcpy(dist.data(), C.dist.p, nc * 8, hipMemcpyDeviceToHost));
      }
      if (data) {
        DBuf<double> out;
        out.ensure((size_t)nc * 16);
        ExportArgs a{};
        a.nc = nc; a.A = C.A.p; a.B = C.B.p; a.info = C.info.p; a.real = C.real.p;
        a.N0 = C.hN[0].p; a.N1 = C.hN[1].p; a.N2 = C.hN[2].p;
        a.Y0 = C.hY[0].p; a.Y1 = C.hY[1].p; a.Y2 = C.hY[2].p;
        a.P0 = C.hP[0].p; a.P1 = C.hP[1].p; a.P2 = C.hP[2].p;
        a.F0 = C.hF[0].p; a.F1 = C.hF[1].p; a.F2 = C.hF[2].p; a.uN = C.huN.p;
        a.rad = c.radius.p; a.mass = c.mass.p; a.mat = c.mat.p; a.mats = c.dMats.p;
        a.hertz = c.lastPhysKind == WDG_PHYS_HERTZ;
        a.poisson = c.lastPoisson; a.en = c.lastEn; a.viscAlpha = c.lastViscAlpha;
        a.out = out.p;
        hipLaunchKernelGGL(k_export, dim3(nblocks(nc, 256)), dim3(256), 0, c.stream, a);
        WDG_LAUNCH_CHECK();
        rec.resize((size_t)nc * 16);
        WDG_HIP(hipMemcpyAsync(rec.data(), out.p, rec.size() * 8, hipMemcpyDeviceToHost, c.stream));
        WDG_HIP(hipStreamSynchronize(c.stream));
        out.release();
      }
    }
    int64_t k = 0;
    for (int64_t s = 0; s < nc; s++) {
      bool take = which == 1 ? real[s] != 0 : (real[s] || (info[s] & CI_AABB));
      if (!take) continue;
      bool swap = (info[s] & CI_SWAP) != 0;
      if (ids) {
        ids[2 * k] = (int32_t)(swap ? B[s] : A[s]);
        ids[2 * k + 1] = (int32_t)(swap ? A[s] : B[s]);
      }
      if (cellDist)
        for (int m = 0; m < 3; m++) cellDist[3 * k + m] = c.field.periodic ? unpack_cd(dist[s], m) : 0;
      if (data) std::memcpy(data + 16 * k, rec.data() + 16 * s, 16 * sizeof(double));
      k++;
    }
  });
}

int wdg_observables(wdg_ctx* ctx, const double boxMin[3], const double boxMax[3], double out[3]) {
  if (!ctx || !boxMin || !boxMax || !out) { g_err = "NULL argument"; return WDG_E_ARG; }
  return guard(ctx, [&] { observables(ctx->c, boxMin, boxMax, out); });
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
    out->brute = c.lastBrute;
    for (int k = 0; k < 3; k++) out->gridCells[k] = c.gridCells[k];
  });
}

// Scene::doOneStep (Scene.cpp:366-467) for a fixed engine list
int wdg_run(wdg_ctx* ctx, const wdg_program* prog, int64_t nSteps, int64_t* step, double* dt) {
  if (!ctx || !prog || !step || !dt || prog->nEngines < 0 || prog->nEngines > 8) {
    g_err = "bad arguments";
    return WDG_E_ARG;
  }
  return guard(ctx, [&] {
    Ctx& c = ctx->c;
    for (int64_t s = 0; s < nSteps; s++) {
      double nextDt = NAN;
      for (int e = 0; e < prog->nEngines; e++) {
        switch (prog->engines[e]) {
          case WDG_E_FORCERESETTER: force_reset(c); break;
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
              // DynDt::run (DynDt.cpp:90-113)
              double cr = crit_dt(c);
              if (std::isinf(cr)) break;
              int64_t ns = *step - c.dynStepPrev;
              double base = 1. + prog->maxRelInc, pw = 1.;
              // pown: square-and-multiply (Math.hpp:166-175)
              struct P {
                static double pown(double x, unsigned n) {
                  unsigned d = n >> 1, r = n & 1;
                  double x2d = d == 0 ? 1. : pown(x * x, d);
                  double xr = r == 0 ? 1. : x;
                  return x2d * xr;
                }
              };
              pw = P::pown(base, (unsigned)ns);
              double maxNewDt = *dt * pw;
              double nd = std::min(cr * prog->dtSafety, maxNewDt);
              if (nd != *dt) nextDt = nd;
            }
          } break;
          default: throw Error(WDG_E_ARG, "unknown engine code " + std::to_string(prog->engines[e]));
        }
      }
      *step += 1;
      if (!std::isnan(nextDt)) *dt = nextDt;
    }
  });
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

void* wdg_stream(wdg_ctx* ctx) { return ctx ? (void*)ctx->c.stream : nullptr; }

}  // extern "C"
