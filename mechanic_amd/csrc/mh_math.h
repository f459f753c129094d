// Device vector math with a fixed evaluation order.
//
// Every helper spells out the order of its IEEE-754 binary32 operations, so a
// kernel built with -ffp-contract=off rounds exactly like any other
// implementation that evaluates the same expressions left to right. The forms
// follow the WGSL builtins the reference's kernels call: dot is
// (x*x' + y*y') + z*z', cross is the textbook determinant, and quat_multiply
// and quat_rotate are those of crates/mechanic-gpu/src/kernels/physics.wgsl:46-56.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mh {

struct v3 {
  float x, y, z;
};

__device__ __forceinline__ v3 mk(float x, float y, float z) { return v3{x, y, z}; }
__device__ __forceinline__ v3 xyz(float4 v) { return v3{v.x, v.y, v.z}; }
__device__ __forceinline__ float4 f4(v3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
__device__ __forceinline__ v3 operator+(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ v3 operator-(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ v3 operator-(v3 a) { return v3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ v3 operator*(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ v3 operator*(float s, v3 a) { return v3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ v3 operator/(v3 a, float s) { return v3{a.x / s, a.y / s, a.z / s}; }

__device__ __forceinline__ float dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float dot4(float4 a, float4 b) {
  return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w;
}
__device__ __forceinline__ v3 cross(v3 a, v3 b) {
  return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// Select-based min/max: identical results for signed zeros and NaN on every
// implementation that evaluates the same comparison.
__device__ __forceinline__ float fmx(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float fmn(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ v3 vmin(v3 a, v3 b) { return v3{fmn(a.x, b.x), fmn(a.y, b.y), fmn(a.z, b.z)}; }
__device__ __forceinline__ v3 vmax(v3 a, v3 b) { return v3{fmx(a.x, b.x), fmx(a.y, b.y), fmx(a.z, b.z)}; }
__device__ __forceinline__ v3 vabs(v3 a) { return v3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
__device__ __forceinline__ float length(v3 a) { return sqrtf(dot(a, a)); }
// 1/sqrt with a correctly rounded sqrt and a correctly rounded divide.
__device__ __forceinline__ float rsqrt_cr(float x) { return 1.0f / sqrtf(x); }
__device__ __forceinline__ float signf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

// Quaternions are float4 (x, y, z, w).
__device__ __forceinline__ float4 quat_multiply(float4 a, float4 b) {
  v3 av = xyz(a), bv = xyz(b);
  v3 v = (a.w * bv + b.w * av) + cross(av, bv);
  return f4(v, a.w * b.w - dot(av, bv));
}
__device__ __forceinline__ v3 quat_rotate(float4 q, v3 v) {
  v3 qv = xyz(q);
  v3 t = 2.0f * cross(qv, v);
  return (v + q.w * t) + cross(qv, t);
}
__device__ __forceinline__ float4 quat_conj(float4 q) { return make_float4(-q.x, -q.y, -q.z, q.w); }

// f32 -> i32 after round-half-to-even, saturating like the GPU conversion: NaN
// -> 0, values at or above 2^31 -> INT_MAX, below -2^31 -> INT_MIN. gfx950's
// v_cvt_i32_f32 saturates exactly so (mh_device_round_i32 checks it against
// these rules), in one instruction where the compare-and-select form takes
// eight; the contact solver converts twelve values a contact a pass.
__device__ __forceinline__ int32_t round_to_i32(float x) {
  const float r = __builtin_rintf(x);
  int32_t out;
  asm("v_cvt_i32_f32 %0, %1" : "=v"(out) : "v"(r));
  return out;
}

__device__ __forceinline__ bool finite4(float4 v) {
  const float lim = 3.402823e+38f;
  return (v.x == v.x) && (v.y == v.y) && (v.z == v.z) && (v.w == v.w) && fabsf(v.x) < lim &&
         fabsf(v.y) < lim && fabsf(v.z) < lim && fabsf(v.w) < lim;
}

}  // namespace mh
