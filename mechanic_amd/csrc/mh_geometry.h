// Collider geometry shared by the broadphase, narrowphase and ground kernels.
//
// Each helper evaluates the same expression, in the same order, as the
// reference function it names, so values computed once here and reused are
// bit-identical to recomputing them per call as the reference does.
#pragma once

#include "internal.h"
#include "mh_math.h"

namespace mh {

constexpr float kFltBig = 3.402823e+38f;

struct ColliderPose {
  v3 center;       // collider_center (lbvh.wgsl:73-77, collision.wgsl:164-168)
  float4 rotation; // collider_rotation (collision.wgsl:170-173)
  v3 axes[3];      // quat_rotate(rotation, unit axis)
  v3 half;         // half extents xyz
  float half_w;    // outer radius of a full-cylinder segment
  float center_w;  // segment-centre radius of a full-cylinder segment
  uint32_t body;
  uint32_t role;   // ground-contact role (abi.rs:133)
};

__device__ __forceinline__ ColliderPose collider_pose(const Dev& d, uint32_t c) {
  const float4* row = d.colliders + 4 * (size_t)c;
  float4 lc = row[0], lr = row[1], he = row[2], meta = row[3];
  ColliderPose p;
  p.body = __float_as_uint(meta.x);
  p.role = __float_as_uint(meta.w);
  float4 bp = d.pos[p.body];
  float4 br = d.rot[p.body];
  p.center = xyz(bp) + quat_rotate(br, xyz(lc));
  p.rotation = quat_multiply(br, lr);
  p.axes[0] = quat_rotate(p.rotation, mk(1.0f, 0.0f, 0.0f));
  p.axes[1] = quat_rotate(p.rotation, mk(0.0f, 1.0f, 0.0f));
  p.axes[2] = quat_rotate(p.rotation, mk(0.0f, 0.0f, 1.0f));
  p.half = xyz(he);
  p.half_w = he.w;
  p.center_w = lc.w;
  return p;
}

// The world pose the Morton kernel stored for collider c this tick; the same
// bits collider_pose() computes (cylinder ground data excluded).
__device__ __forceinline__ void store_pose(const Dev& d, uint32_t c, const ColliderPose& p) {
  float4* o = d.cposes + 5 * (size_t)c;
  o[0] = f4(p.center, p.half.x);
  o[1] = p.rotation;
  o[2] = f4(p.axes[0], p.half.y);
  o[3] = f4(p.axes[1], p.half.z);
  o[4] = f4(p.axes[2], __uint_as_float(p.body));
}

__device__ __forceinline__ ColliderPose load_pose(const Dev& d, uint32_t c) {
  const float4* o = d.cposes + 5 * (size_t)c;
  const float4 a = o[0], b = o[1], x = o[2], y = o[3], z = o[4];
  ColliderPose p;
  p.center = xyz(a);
  p.rotation = b;
  p.axes[0] = xyz(x);
  p.axes[1] = xyz(y);
  p.axes[2] = xyz(z);
  p.half = mk(a.w, x.w, y.w);
  p.half_w = 0.0f;
  p.center_w = 0.0f;
  p.body = __float_as_uint(z.w);
  p.role = 0u;
  return p;
}

__device__ __forceinline__ float comp(v3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// collider_support_point (collision.wgsl:175-191).
__device__ __forceinline__ v3 support_point(const ColliderPose& p, v3 direction) {
  v3 point = p.center;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float projection = dot(p.axes[a], direction);
    if (fabsf(projection) > 1.0e-5f) point = point + (p.axes[a] * comp(p.half, a)) * signf(projection);
  }
  return point;
}

__device__ __forceinline__ v3 normalize3(v3 v) { return v / length(v); }

// full_cylinder_contact_count (collision.wgsl:193-205).
__device__ __forceinline__ uint32_t full_cylinder_contact_count(const ColliderPose& p, v3 direction) {
  v3 axis = p.axes[1];
  float axial = dot(axis, direction);
  float radial_length = length(direction - axis * axial);
  return radial_length < 0.05f ? 4u : 1u;
}

// full_cylinder_support_point (collision.wgsl:207-238).
__device__ __forceinline__ v3 full_cylinder_support_point(const ColliderPose& p, v3 direction, uint32_t role) {
  v3 radial_axis = p.axes[0];
  v3 axis = p.axes[1];
  v3 point = p.center - radial_axis * p.center_w;
  float axial = dot(axis, direction);
  v3 radial_direction = direction - axis * axial;
  float radial_length = length(radial_direction);
  if (radial_length < 0.05f) {
    point = point + (axis * signf(axial)) * p.half.y;
    v3 tangent_a = normalize3(mk(1.0f, 0.0f, 0.0f) - axis * axis.x);
    v3 tangent_b = normalize3(cross(axis, tangent_a));
    if (role == 1u) point = point + tangent_a * p.half_w;
    else if (role == 2u) point = point - tangent_a * p.half_w;
    else if (role == 3u) point = point + tangent_b * p.half_w;
    else point = point - tangent_b * p.half_w;
  } else {
    point = point + radial_direction * (p.half_w / radial_length);
    if (fabsf(axial) >= 0.05f) point = point + (axis * signf(axial)) * p.half.y;
  }
  return point;
}

// Appends one row per lane with `pred` set, one atomic per wavefront.
// Returns the lane's row index (may be >= capacity), or kInvalid without `pred`.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
  uint64_t mask = __ballot(pred);
  if (mask == 0) return kInvalid;
  uint32_t lane = __lane_id();
  uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = __shfl(base, (int)leader);
  uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
  return pred ? base + rank : kInvalid;
}

// Appends one row per thread with `pred` set, one atomic per workgroup. Every
// thread of the block must call it (block-uniform control flow); `lds` holds
// kBlock / 64 + 1 words. Returns the row index (may be >= capacity), or
// kInvalid without `pred`.
// Look-back status words of the single-pass scans, read and written at agent scope.
__device__ __forceinline__ uint32_t load_status(const uint32_t* w) {
  return __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_status(uint32_t* w, uint32_t v) {
  __hip_atomic_store(w, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t block_append(uint32_t* counter, bool pred, uint32_t* lds) {
  const uint32_t lane = __lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  const uint32_t waves = blockDim.x >> 6;
  const uint64_t mask = __ballot(pred);
  const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
  if (lane == 0) lds[wave] = (uint32_t)__popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t total = 0;
    for (uint32_t w = 0; w < waves; ++w) {
      uint32_t n = lds[w];
      lds[w] = total;
      total += n;
    }
    lds[waves] = total ? atomicAdd(counter, total) : 0u;
  }
  __syncthreads();
  const uint32_t row = lds[waves] + lds[wave] + rank;
  __syncthreads();
  return pred ? row : kInvalid;
}

__device__ __forceinline__ void store_contact(const Dev& d, uint32_t row, uint4 meta, float4 n_pen,
                                              float4 arm_a, float4 arm_b) {
  float4* c = d.contacts + 4 * (size_t)row;
  c[0] = make_float4(__uint_as_float(meta.x), __uint_as_float(meta.y), __uint_as_float(meta.z),
                     __uint_as_float(meta.w));
  c[1] = n_pen;
  c[2] = arm_a;
  c[3] = arm_b;
}

}  // namespace mh
