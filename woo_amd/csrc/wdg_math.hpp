// Device vector / quaternion math for the DEM kernels (gfx950, FP64).
//
// Evaluation order is fixed so that results are reproducible bit for bit:
//  * 3-sums left to right, (a0+a1)+a2;
//  * cross(a,b) = (a1*b2-a2*b1, a2*b0-a0*b2, a0*b1-a1*b0);
//  * normalize divides each component by sqrt(squaredNorm) when positive;
//  * compiled with -ffp-contract=off, so no product is fused into an add.
// These are the Eigen operations used by the reference (L6Geom.cpp, Contact.cpp,
// Leapfrog.cpp) restated for the device.
#pragma once
#include <hip/hip_runtime.h>

namespace wdg {

struct d3 {
  double x, y, z;
};
struct q4 {
  double w, x, y, z;
};

__device__ __forceinline__ d3 mk3(double a, double b, double c) { return d3{a, b, c}; }
__device__ __forceinline__ d3 add3(d3 a, d3 b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 sub3(d3 a, d3 b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 mul3(d3 a, double s) { return d3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ d3 div3(d3 a, double s) { return d3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ double dot3(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double sqn3(d3 a) { return dot3(a, a); }
__device__ __forceinline__ d3 cross3(d3 a, d3 b) {
  return d3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ d3 normalized3(d3 a) {
  double z = sqn3(a);
  if (z > 0.) {
    double s = sqrt(z);
    return d3{a.x / s, a.y / s, a.z / s};
  }
  return a;
}
__device__ __forceinline__ double comp3(d3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
__device__ __forceinline__ void setcomp3(d3& a, int k, double v) {
  if (k == 0) a.x = v;
  else if (k == 1) a.y = v;
  else a.z = v;
}

// Rotation matrix (given by columns) to quaternion: Shoemake's method (trace
// branch when the trace is positive, otherwise pivot on the largest diagonal).
__device__ __forceinline__ q4 quat_from_cols(d3 c0, d3 c1, d3 c2) {
  // m[r][c]
  double m[3][3] = {{c0.x, c1.x, c2.x}, {c0.y, c1.y, c2.y}, {c0.z, c1.z, c2.z}};
  q4 q;
  double t = m[0][0] + m[1][1] + m[2][2];
  if (t > 0.) {
    t = sqrt(t + 1.0);
    q.w = 0.5 * t;
    t = 0.5 / t;
    q.x = (m[2][1] - m[1][2]) * t;
    q.y = (m[0][2] - m[2][0]) * t;
    q.z = (m[1][0] - m[0][1]) * t;
  } else {
    // largest diagonal element i, then j = i+1, k = i+2 (mod 3); the three cases are
    // written out so that nothing is indexed at run time
    int i = 0;
    if (m[1][1] > m[0][0]) i = 1;
    if (m[2][2] > (i == 0 ? m[0][0] : m[1][1])) i = 2;
    if (i == 0) {
      t = sqrt(m[0][0] - m[1][1] - m[2][2] + 1.0);
      q.x = 0.5 * t;
      t = 0.5 / t;
      q.w = (m[2][1] - m[1][2]) * t;
      q.y = (m[1][0] + m[0][1]) * t;
      q.z = (m[2][0] + m[0][2]) * t;
    } else if (i == 1) {
      t = sqrt(m[1][1] - m[2][2] - m[0][0] + 1.0);
      q.y = 0.5 * t;
      t = 0.5 / t;
      q.w = (m[0][2] - m[2][0]) * t;
      q.z = (m[2][1] + m[1][2]) * t;
      q.x = (m[0][1] + m[1][0]) * t;
    } else {
      t = sqrt(m[2][2] - m[0][0] - m[1][1] + 1.0);
      q.z = 0.5 * t;
      t = 0.5 / t;
      q.w = (m[1][0] - m[0][1]) * t;
      q.x = (m[0][2] + m[2][0]) * t;
      q.y = (m[1][2] + m[2][1]) * t;
    }
  }
  return q;
}

// q*v = v + w*(2 q×v) + q×(2 q×v)
__device__ __forceinline__ d3 quat_rot(q4 q, d3 v) {
  d3 qv = d3{q.x, q.y, q.z};
  d3 uv = cross3(qv, v);
  uv = add3(uv, uv);
  d3 r = add3(v, mul3(uv, q.w));
  return add3(r, cross3(qv, uv));
}

__device__ __forceinline__ q4 quat_mul(q4 a, q4 b) {
  q4 r;
  r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
  r.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
  r.y = a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z;
  r.z = a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x;
  return r;
}

__device__ __forceinline__ double sgn(double f) { return f < 0 ? -1. : (f > 0 ? 1. : 0.); }
__device__ __forceinline__ double p2(double x) { return x * x; }

}  // namespace wdg
