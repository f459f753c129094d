// Narrowphase: 15-axis OBB SAT with one candidate pair per wavefront lane and a
// single averaged contact point, collision.wgsl:393-486, 240-287, 570-620.
#include "internal.h"
#include "mh_geometry.h"
#include "mh_math.h"

namespace mh {

struct Sat {
  v3 normal;
  float penetration;
  uint32_t near_face_axes;
};

__device__ __forceinline__ uint64_t tick_np(const Dev& d) {
  return *reinterpret_cast<const volatile uint64_t*>(&d.tick->tick_index);
}

// projection_radius (collision.wgsl:393-397).
__device__ __forceinline__ float projection_radius(const v3 (&axes)[3], v3 e, v3 axis) {
  return (fabsf(dot(axes[0], axis)) * e.x + fabsf(dot(axes[1], axis)) * e.y) + fabsf(dot(axes[2], axis)) * e.z;
}

// test_sat_axis (collision.wgsl:399-420).
__device__ __forceinline__ Sat test_axis(v3 delta, const ColliderPose& a, const ColliderPose& b, v3 raw) {
  float len2 = dot(raw, raw);
  if (len2 < 1.0e-10f) return Sat{mk(1.0f, 0.0f, 0.0f), kFltBig, 0u};
  v3 axis = raw * rsqrt_cr(len2);
  float signed_distance = dot(delta, axis);
  float penetration = (projection_radius(a.axes, a.half, axis) + projection_radius(b.axes, b.half, axis)) -
                      fabsf(signed_distance);
  if (signed_distance < 0.0f) axis = -axis;
  return Sat{axis, penetration, 0u};
}

// obb_sat (collision.wgsl:422-486): early exit on the first separating axis.
__device__ __forceinline__ Sat obb_sat(const ColliderPose& a, const ColliderPose& b) {
  const v3 delta = b.center - a.center;
  Sat minimum{mk(1.0f, 0.0f, 0.0f), kFltBig, 0u};
  uint32_t near_face = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Sat ra = test_axis(delta, a, b, a.axes[i]);
    if (ra.penetration < 0.0f) return ra;
    if (ra.penetration <= 1.0e-4f) near_face += 1;
    if (ra.penetration < minimum.penetration) minimum = ra;
    Sat rb = test_axis(delta, a, b, b.axes[i]);
    if (rb.penetration < 0.0f) return rb;
    if (rb.penetration <= 1.0e-4f) near_face += 1;
    if (rb.penetration < minimum.penetration) minimum = rb;
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      Sat r = test_axis(delta, a, b, cross(a.axes[i], b.axes[j]));
      if (r.penetration < 0.0f) return r;
      if (r.penetration < minimum.penetration) minimum = r;
    }
  }
  minimum.near_face_axes = near_face;
  return minimum;
}

// collider_vertex (collision.wgsl:240-254).
__device__ __forceinline__ v3 vertex(const ColliderPose& p, uint32_t v) {
  v3 point = p.center;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float dir = (v & (1u << a)) != 0u ? 1.0f : -1.0f;
    point = point + (p.axes[a] * comp(p.half, a)) * dir;
  }
  return point;
}

// collider_contains_point (collision.wgsl:256-264).
__device__ __forceinline__ bool contains(const ColliderPose& p, v3 point) {
  v3 local = quat_rotate(quat_conj(p.rotation), point - p.center);
  return fabsf(local.x) <= p.half.x + 1.0e-5f && fabsf(local.y) <= p.half.y + 1.0e-5f &&
         fabsf(local.z) <= p.half.z + 1.0e-5f;
}

// calculate_contact_point (collision.wgsl:266-287).
__device__ __forceinline__ v3 contact_point(const ColliderPose& a, const ColliderPose& b, v3 normal) {
  v3 sum = mk(0.0f, 0.0f, 0.0f);
  float count = 0.0f;
#pragma unroll
  for (uint32_t v = 0; v < 8; ++v) {
    v3 pa = vertex(a, v);
    if (contains(b, pa)) {
      sum = sum + pa;
      count += 1.0f;
    }
    v3 pb = vertex(b, v);
    if (contains(a, pb)) {
      sum = sum + pb;
      count += 1.0f;
    }
  }
  if (count > 0.0f) return sum / count;
  v3 sa = support_point(a, normal);
  v3 sb = support_point(b, -normal);
  return (sa + sb) * 0.5f;
}

__global__ __launch_bounds__(256) void k_narrowphase(Dev d) {
  __shared__ uint32_t lds_append[kBlock / 64 + 1];
  const uint64_t tick = tick_np(d);
  uint32_t* diag = d.diag[tick & 1];
  uint32_t requested = *reinterpret_cast<volatile uint32_t*>(&diag[kDiagPairs]);
  const uint32_t count = requested < d.pair_capacity ? requested : d.pair_capacity;
  const XcdSplit xs = grid_split(count);
  for (uint32_t base = xs.first; base < xs.end; base += xs.step) {
    const uint32_t index = base + threadIdx.x;
    bool emit = false;
    uint4 meta;
    float4 n_pen, arm_a, arm_b;
    if (index < xs.end) {
      const uint2 pair = d.pairs[index];
      const ColliderPose a = load_pose(d, pair.x);
      const ColliderPose b = load_pose(d, pair.y);
      bool skip = !d.self_collisions && d.components[a.body] == d.components[b.body];
      if (!skip) {
        Sat sat = obb_sat(a, b);
        if (!(sat.penetration < -1.0e-5f)) {
          v3 point = contact_point(a, b, sat.normal);
          emit = true;
          meta = make_uint4(a.body, b.body, pair.x, pair.y);
          n_pen = f4(sat.normal, fmx(sat.penetration, 0.0f));
          arm_a = f4(point - xyz(d.pos[a.body]), 0.0f);
          arm_b = f4(point - xyz(d.pos[b.body]), (float)sat.near_face_axes);
        }
      }
    }
    uint32_t row = block_append(&diag[kDiagContacts], emit, lds_append);
    if (emit) {
      if (row < d.pair_capacity) store_contact(d, row, meta, n_pen, arm_a, arm_b);
      else atomicOr(&diag[kDiagFlags], 1u);
    }
  }
}

void launch_narrowphase(hipStream_t s, const Dev& d) {
  uint32_t blocks = (d.pair_capacity + kBlock - 1) / kBlock;
  const uint32_t cap = d.t_narrow_blocks ? d.t_narrow_blocks : 1024u;
  if (blocks > cap) blocks = cap;
  k_narrowphase<<<blocks, kBlock, 0, s>>>(d);
}

}  // namespace mh
