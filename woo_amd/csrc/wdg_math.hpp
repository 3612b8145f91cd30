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
    int i = 0;
    if (m[1][1] > m[0][0]) i = 1;
    if (m[2][2] > m[i][i]) i = 2;
    int j = (i + 1) % 3, k = (j + 1) % 3;
    t = sqrt(m[i][i] - m[j][j] - m[k][k] + 1.0);
    double v[3];
    v[i] = 0.5 * t;
    t = 0.5 / t;
    q.w = (m[k][j] - m[j][k]) * t;
    v[j] = (m[j][i] + m[i][j]) * t;
    v[k] = (m[k][i] + m[i][k]) * t;
    q.x = v[0];
    q.y = v[1];
    q.z = v[2];
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
