// Initial-condition generation: overlap removal for dense random packings.
//
// Spheres placed at random at the target solid fraction overlap; each
// iteration moves both members of every overlapping pair apart by half the
// overlap of slightly inflated radii (a Jacobi sweep over the contact slots of
// the broadphase), until the largest overlap falls below a tolerance. This is
// set-up work (the role of SpherePack/randomPeriPack, SpherePack.cpp:140-247,
// pack.py:572) and runs before the step loop; afterwards the contact state is
// dropped so that the simulation starts without history.
#include <cstring>

#include "wdg_internal.hpp"
#include "wdg_math.hpp"

namespace wdg {

struct RelaxArgs {
  int64_t nc;
  const uint32_t *A, *B, *info;
  const uint64_t* dist;
  uint32_t* active;
  double *G0, *G1, *G2;
  const double *px, *py, *pz, *rad;
  double L[3];
  int periodic;
  unsigned long long* maxOv;
};

// push of pB away from pA along the centre line (pA gets the opposite)
__global__ __launch_bounds__(256) void k_relax_pairs(RelaxArgs a) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double ov = 0.;
  if (c < a.nc) {
    a.active[c] = 0u;
    const uint32_t inf = a.info[c];
    if ((inf & CI_TYPE_MASK) == CT_SS) {
      const bool swap = (inf & CI_SWAP) != 0u;
      const uint32_t i = swap ? a.B[c] : a.A[c], j = swap ? a.A[c] : a.B[c];
      d3 s = mk3(0., 0., 0.);
      if (a.periodic) {
        uint64_t cd = a.dist[c];
        s = mk3(a.L[0] * unpack_cd(cd, 0), a.L[1] * unpack_cd(cd, 1), a.L[2] * unpack_cd(cd, 2));
      }
      d3 d = sub3(add3(mk3(a.px[j], a.py[j], a.pz[j]), s), mk3(a.px[i], a.py[i], a.pz[i]));
      double dist = sqrt(sqn3(d));
      ov = (a.rad[i] + a.rad[j]) - dist;
      if (ov > 0 && dist > 0) {
        d3 p = mul3(d, 0.5 * ov / dist);
        a.G0[c] = p.x;
        a.G1[c] = p.y;
        a.G2[c] = p.z;
        a.active[c] = 1u;
      } else {
        ov = 0.;
      }
    }
  }
  unsigned long long b = (unsigned long long)__double_as_longlong(ov);
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(b, off);
    b = o > b ? o : b;
  }
  if ((threadIdx.x & 63) == 0 && b) atomicMax(a.maxOv, b);
}

__global__ __launch_bounds__(256) void k_relax_move(int64_t n, const uint32_t* __restrict__ flags,
                                                     const uint32_t* __restrict__ adj,
                                                     const uint32_t* __restrict__ adjOff,
                                                     const uint32_t* __restrict__ active, const double* __restrict__ G0,
                                                     const double* __restrict__ G1, const double* __restrict__ G2,
                                                     double* px, double* py, double* pz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flag_shape(flags[i]) != WDG_SPHERE || !(flags[i] & F_INTEGRATED)) return;
  d3 m = mk3(0., 0., 0.);
  for (uint32_t k = adjOff[i]; k < adjOff[i + 1]; k++) {
    const uint32_t e = adj[k], s = e & ~ADJ_PA;
    if (!active[s]) continue;
    d3 p = mk3(G0[s], G1[s], G2[s]);
    m = (e & ADJ_PA) ? sub3(m, p) : add3(m, p);
  }
  px[i] += m.x;
  py[i] += m.y;
  pz[i] += m.z;
}

__global__ void k_scale(int64_t n, const double* __restrict__ in, double s, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i] * s;
}

void relax(Ctx& c, int maxIter, double inflate, double tol, double* maxOverlap, int* iters) {
  if (c.n == 0) { *maxOverlap = 0; *iters = 0; return; }
  if (c.forcesPending) flush_forces(c);
  // relax with inflated radii (so that true radii end up overlap-free); the radii are
  // restored bit for bit from a copy (they are permuted along with the particles)
  DBuf<double> rtrue;
  rtrue.ensure(c.n);
  WDG_HIP(hipMemcpyAsync(rtrue.p, c.radius.p, c.n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  hipLaunchKernelGGL(k_scale, dim3(nblocks(c.n, 256)), dim3(256), 0, c.stream, c.n, (const double*)rtrue.p,
                     1 + inflate, c.radius.p);
  WDG_LAUNCH_CHECK();
  // the particle arrays are permuted at every refresh, so the true radii are kept in user-id order
  DBuf<double> ruid;
  ruid.ensure(c.n);
  {
    std::vector<double> h((size_t)c.n);
    std::vector<uint32_t> u((size_t)c.n);
    WDG_HIP(hipMemcpyAsync(h.data(), rtrue.p, c.n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    WDG_HIP(hipMemcpyAsync(u.data(), c.uid.p, c.n * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
    WDG_HIP(hipStreamSynchronize(c.stream));
    std::vector<double> byUid((size_t)c.n);
    for (int64_t s = 0; s < c.n; s++) byUid[u[s]] = h[s];
    WDG_HIP(hipMemcpy(ruid.p, byUid.data(), c.n * sizeof(double), hipMemcpyHostToDevice));
  }
  rtrue.release();
  const double rminSave = c.rmin, rmaxSave = c.rmax;
  c.rmin *= 1 + inflate;
  c.rmax *= 1 + inflate;
  wdg_collide_params cp{};
  // generous skin: fewer rebuilds while particles move by up to half an overlap per sweep
  cp.verletDist = 0.3 * c.rmin;
  cp.forceInit = 1;
  const double saveVd = c.verletDist;
  c.verletDist = NAN;
  unsigned long long* dmax = reinterpret_cast<unsigned long long*>(c.red.p) + 8;
  DBuf<double> push[3];  // per-slot push of pB
  double ov = INFINITY;
  int it = 0;
  for (; it < maxIter; it++) {
    collide(c, &cp);
    cp.forceInit = 0;
    Contacts& C = c.con();
    WDG_HIP(hipMemsetAsync(dmax, 0, sizeof(unsigned long long), c.stream));
    if (C.n > 0) {
      c.act.ensure(C.n);
      for (int k = 0; k < 3; k++) push[k].ensure(C.n);
      RelaxArgs a{};
      a.nc = C.n; a.A = C.A.p; a.B = C.B.p; a.info = C.info.p; a.dist = C.dist.p; a.active = c.act.p;
      a.G0 = push[0].p; a.G1 = push[1].p; a.G2 = push[2].p;
      a.px = c.pos[0].p; a.py = c.pos[1].p; a.pz = c.pos[2].p; a.rad = c.radius.p;
      for (int k = 0; k < 3; k++) a.L[k] = c.field.cell[k];
      a.periodic = c.field.periodic;
      a.maxOv = dmax;
      hipLaunchKernelGGL(k_relax_pairs, dim3(nblocks(C.n, 256)), dim3(256), 0, c.stream, a);
      WDG_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_relax_move, dim3(nblocks(c.n, 256)), dim3(256), 0, c.stream, c.n, c.flags.p, C.adj.p,
                         C.adjOff.p, c.act.p, push[0].p, push[1].p, push[2].p, c.pos[0].p, c.pos[1].p, c.pos[2].p);
      WDG_LAUNCH_CHECK();
      c.dirtyKnown = false;
    }
    if (it % 8 == 7 || it == maxIter - 1) {
      unsigned long long hb;
      WDG_HIP(hipMemcpyAsync(&hb, dmax, sizeof(hb), hipMemcpyDeviceToHost, c.stream));
      WDG_HIP(hipStreamSynchronize(c.stream));
      std::memcpy(&ov, &hb, 8);
      if (ov <= tol) { it++; break; }
    }
  }
  // restore the true radii in the current storage order: radius[s] = ruid[uid[s]]
  {
    std::vector<double> byUid((size_t)c.n), h((size_t)c.n);
    std::vector<uint32_t> u((size_t)c.n);
    WDG_HIP(hipMemcpyAsync(byUid.data(), ruid.p, c.n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    WDG_HIP(hipMemcpyAsync(u.data(), c.uid.p, c.n * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
    WDG_HIP(hipStreamSynchronize(c.stream));
    for (int64_t s = 0; s < c.n; s++) h[s] = byUid[u[s]];
    WDG_HIP(hipMemcpy(c.radius.p, h.data(), c.n * sizeof(double), hipMemcpyHostToDevice));
  }
  ruid.release();
  for (int k = 0; k < 3; k++) push[k].release();
  c.rmin = rminSave;
  c.rmax = rmaxSave;
  drop_contacts(c);
  WDG_HIP(hipStreamSynchronize(c.stream));
  c.verletDist = saveVd;
  c.nRefresh = 0;
  c.nCollide = 0;
  *maxOverlap = ov;
  *iters = it;
}

}  // namespace wdg
