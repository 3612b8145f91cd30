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
  double grow;  // 1 + inflate
  unsigned long long* maxOv;
};

__global__ __launch_bounds__(256) void k_relax_pairs(RelaxArgs a) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double ov = 0.;
  if (c < a.nc) {
    a.active[c] = 0u;
    const uint32_t inf = a.info[c];
    if ((inf & CI_TYPE_MASK) == CT_SS) {
      const uint32_t i = a.A[c], j = a.B[c];
      d3 s = mk3(0., 0., 0.);
      if (a.periodic) {
        uint64_t cd = a.dist[c];
        s = mk3(a.L[0] * unpack_cd(cd, 0), a.L[1] * unpack_cd(cd, 1), a.L[2] * unpack_cd(cd, 2));
      }
      d3 d = sub3(add3(mk3(a.px[j], a.py[j], a.pz[j]), s), mk3(a.px[i], a.py[i], a.pz[i]));
      double dist = sqrt(sqn3(d));
      double rr = (a.rad[i] + a.rad[j]) * a.grow;
      ov = rr - dist;
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
                                                     const uint32_t* __restrict__ offA, const uint32_t* __restrict__ offB,
                                                     const uint32_t* __restrict__ idxB, const uint32_t* __restrict__ active,
                                                     const double* __restrict__ G0, const double* __restrict__ G1,
                                                     const double* __restrict__ G2, double* px, double* py, double* pz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flag_shape(flags[i]) != WDG_SPHERE || !(flags[i] & F_INTEGRATED)) return;
  d3 m = mk3(0., 0., 0.);
  for (uint32_t k = offB[i]; k < offB[i + 1]; k++) {
    uint32_t s = idxB[k];
    if (active[s]) m = add3(m, mk3(G0[s], G1[s], G2[s]));
  }
  for (uint32_t s = offA[i]; s < offA[i + 1]; s++)
    if (active[s]) m = sub3(m, mk3(G0[s], G1[s], G2[s]));
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
  // relax with inflated radii (so that true radii end up overlap-free); originals restored bit for bit
  DBuf<double> rsave;
  rsave.ensure(c.n);
  WDG_HIP(hipMemcpyAsync(rsave.p, c.radius.p, c.n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  hipLaunchKernelGGL(k_scale, dim3(nblocks(c.n, 256)), dim3(256), 0, c.stream, c.n, (const double*)rsave.p,
                     1 + inflate, c.radius.p);
  WDG_LAUNCH_CHECK();
  const double rminSave = c.rmin, rmaxSave = c.rmax;
  c.rmin *= 1 + inflate;
  c.rmax *= 1 + inflate;
  wdg_collide_params cp{};
  // generous skin: fewer rebuilds while particles move by up to half an overlap per sweep
  cp.verletDist = 0.3 * c.rmin;
  cp.forceInit = 1;
  const double saveVd = c.verletDist;
  c.verletDist = NAN;
  inflate = 0;
  unsigned long long* dmax = reinterpret_cast<unsigned long long*>(c.red.p) + 8;
  double ov = INFINITY;
  int it = 0;
  for (; it < maxIter; it++) {
    collide(c, &cp);
    cp.forceInit = 0;
    Contacts& C = c.con();
    WDG_HIP(hipMemsetAsync(dmax, 0, sizeof(unsigned long long), c.stream));
    if (C.n > 0) {
      RelaxArgs a{};
      a.nc = C.n; a.A = C.A.p; a.B = C.B.p; a.info = C.info.p; a.dist = C.dist.p; a.active = C.real.p;
      a.G0 = C.F[0].p; a.G1 = C.F[1].p; a.G2 = C.F[2].p;
      a.px = c.pos[0].p; a.py = c.pos[1].p; a.pz = c.pos[2].p; a.rad = c.radius.p;
      for (int k = 0; k < 3; k++) a.L[k] = c.field.cell[k];
      a.periodic = c.field.periodic;
      a.grow = 1 + inflate;
      a.maxOv = dmax;
      hipLaunchKernelGGL(k_relax_pairs, dim3(nblocks(C.n, 256)), dim3(256), 0, c.stream, a);
      WDG_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_relax_move, dim3(nblocks(c.n, 256)), dim3(256), 0, c.stream, c.n, c.flags.p, C.offA.p,
                         C.offB.p, C.idxB.p, C.real.p, C.F[0].p, C.F[1].p, C.F[2].p, c.pos[0].p, c.pos[1].p,
                         c.pos[2].p);
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
  // drop the relaxation contacts; the simulation starts without history
  WDG_HIP(hipMemcpyAsync(c.radius.p, rsave.p, c.n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  c.rmin = rminSave;
  c.rmax = rmaxSave;
  for (auto& s : c.C) {
    s.n = 0;
    WDG_HIP(hipMemsetAsync(s.offA.p, 0, (c.n + 1) * sizeof(uint32_t), c.stream));
    WDG_HIP(hipMemsetAsync(s.offB.p, 0, (c.n + 1) * sizeof(uint32_t), c.stream));
  }
  WDG_HIP(hipStreamSynchronize(c.stream));
  rsave.release();
  c.boundsValid = false;
  c.dirtyKnown = false;
  c.forcesPending = false;
  c.verletDist = saveVd;
  c.nRefresh = 0;
  c.nCollide = 0;
  *maxOverlap = ov;
  *iters = it;
}

}  // namespace wdg
