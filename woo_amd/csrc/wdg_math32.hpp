// Float vector and quaternion helpers for the FP32 contact form (same formulas as the
// FP64 helpers: 3-sums left to right, Shoemake's matrix->quaternion, q*v = v + w(2 q x v)
// + q x (2 q x v)).
#pragma once
#include <hip/hip_runtime.h>

namespace wdg {

struct f3 {
  float x, y, z;
};
struct fq4 {
  float w, x, y, z;
};
__host__ __device__ inline f3 mkf3(float x, float y, float z) { return f3{x, y, z}; }
__host__ __device__ inline f3 add3(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline f3 sub3(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline f3 mul3(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline f3 div3(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__host__ __device__ inline float dot3(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__host__ __device__ inline float sqn3(f3 a) { return dot3(a, a); }
__host__ __device__ inline f3 cross3(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline f3 normalized3(f3 a) { return div3(a, sqrtf(sqn3(a))); }
__host__ __device__ inline float comp3(f3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
__host__ __device__ inline void setcomp3(f3& a, int k, float v) {
  if (k == 0) a.x = v;
  else if (k == 1) a.y = v;
  else a.z = v;
}

// rotation matrix with the given columns -> unit quaternion (Shoemake), three cases
// written out as in the FP64 version
__host__ __device__ inline fq4 quat_from_cols(f3 c0, f3 c1, f3 c2) {
  const float m00 = c0.x, m10 = c0.y, m20 = c0.z, m01 = c1.x, m11 = c1.y, m21 = c1.z, m02 = c2.x, m12 = c2.y,
              m22 = c2.z;
  fq4 q;
  float t = m00 + m11 + m22;
  if (t > 0.f) {
    t = sqrtf(t + 1.0f);
    q.w = 0.5f * t;
    t = 0.5f / t;
    q.x = (m21 - m12) * t;
    q.y = (m02 - m20) * t;
    q.z = (m10 - m01) * t;
  } else {
    int i = 0;
    if (m11 > m00) i = 1;
    if (m22 > (i == 0 ? m00 : m11)) i = 2;
    if (i == 0) {
      t = sqrtf(m00 - m11 - m22 + 1.0f);
      q.x = 0.5f * t;
      t = 0.5f / t;
      q.w = (m21 - m12) * t;
      q.y = (m10 + m01) * t;
      q.z = (m20 + m02) * t;
    } else if (i == 1) {
      t = sqrtf(m11 - m22 - m00 + 1.0f);
      q.y = 0.5f * t;
      t = 0.5f / t;
      q.w = (m02 - m20) * t;
      q.z = (m21 + m12) * t;
      q.x = (m01 + m10) * t;
    } else {
      t = sqrtf(m22 - m00 - m11 + 1.0f);
      q.z = 0.5f * t;
      t = 0.5f / t;
      q.w = (m10 - m01) * t;
      q.x = (m02 + m20) * t;
      q.y = (m12 + m21) * t;
    }
  }
  return q;
}
__host__ __device__ inline f3 quat_rot(fq4 q, f3 v) {
  const f3 u = f3{q.x, q.y, q.z};
  const f3 t = mul3(cross3(u, v), 2.f);
  return add3(add3(v, mul3(t, q.w)), cross3(u, t));
}

}  // namespace wdg
