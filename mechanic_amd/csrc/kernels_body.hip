// Per-body kernels: tick bookkeeping, integration, world inverse inertia,
// snapshot publication and the host-requested external impulse.
#include "internal.h"
#include "mh_math.h"

namespace mh {

__device__ __forceinline__ uint64_t tick_u64(const Dev& d) {
  return *reinterpret_cast<const volatile uint64_t*>(&d.tick->tick_index);
}

// Sets the tick that the following kernels run and clears its diagnostics.
__global__ void k_set_tick(Dev d, uint64_t tick) {
  uint32_t t = threadIdx.x;
  if (t == 0) *reinterpret_cast<uint64_t*>(&d.tick->tick_index) = tick;
  if (t < kDiagWords) d.diag[tick & 1][t] = 0u;
}

// Advances the device tick by one and clears the diagnostics the next tick uses.
__global__ void k_end_tick(Dev d) {
  uint64_t tick = tick_u64(d) + 1;
  uint32_t t = threadIdx.x;
  if (t < kDiagWords) d.diag[tick & 1][t] = 0u;
  __syncthreads();
  if (t == 0) *reinterpret_cast<uint64_t*>(&d.tick->tick_index) = tick;
}

// World-frame inverse-inertia column R * I^-1 * R^T * axis
// (collision.wgsl:134-142, the same expression as physics.wgsl:58-66).
__device__ __forceinline__ v3 world_inverse_inertia(float4 rotation, const float4* mass, v3 vector) {
  v3 local = quat_rotate(quat_conj(rotation), vector);
  v3 local_result = (xyz(mass[1]) * local.x + xyz(mass[2]) * local.y) + xyz(mass[3]) * local.z;
  return quat_rotate(rotation, local_result);
}

// integrate (physics.wgsl:92-132) fused with update_world_masses
// (collision.wgsl:144-162) and, when no later kernel moves bodies, with
// publish_snapshot (snapshot.wgsl:6-14).
__global__ __launch_bounds__(256) void k_integrate(Dev d, int collisions, int snapshot_now) {
  const uint64_t tick = tick_u64(d);
  uint32_t* diag = d.diag[tick & 1];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  // The radix sort's digit histogram and tile tickets start the tick at zero
  // (k_morton adds its workgroups' digit counts into it).
  if (d.t_sort_kind == 0u)
    for (uint32_t j = i; j < kRadixHistWords; j += gridDim.x * blockDim.x) d.rs_hist[j] = 0u;
  if (i >= d.body_count) return;

  float4 position = d.pos[i];
  float4 rotation = d.rot[i];
  float4 linear = d.lin[i];
  float4 angular = d.ang[i];
  bool stored = true;
  if (d.inv_mass[i] > 0.0f) {
    linear.y = linear.y + d.gravity_y * d.delta_seconds;
    linear.x = linear.x * d.linear_damping;
    linear.y = linear.y * d.linear_damping;
    linear.z = linear.z * d.linear_damping;
    angular.x = angular.x * d.angular_damping;
    angular.y = angular.y * d.angular_damping;
    angular.z = angular.z * d.angular_damping;
    if (d.root_flags[i] != 0u) {
      position.x = position.x + linear.x * d.delta_seconds;
      position.y = position.y + linear.y * d.delta_seconds;
      position.z = position.z + linear.z * d.delta_seconds;
      float4 spin = make_float4(angular.x, angular.y, angular.z, 0.0f);
      float4 q = quat_multiply(spin, rotation);
      float h = 0.5f * d.delta_seconds;
      rotation.x = rotation.x + q.x * h;
      rotation.y = rotation.y + q.y * h;
      rotation.z = rotation.z + q.z * h;
      rotation.w = rotation.w + q.w * h;
      float norm_squared = dot4(rotation, rotation);
      if (norm_squared > 1.0e-20f) {
        float r = rsqrt_cr(norm_squared);
        rotation.x *= r;
        rotation.y *= r;
        rotation.z *= r;
        rotation.w *= r;
      } else {
        rotation = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        atomicOr(&diag[kDiagFlags], 2u);
      }
    }
  }
  if (!(finite4(position) && finite4(rotation) && finite4(linear) && finite4(angular))) {
    atomicOr(&diag[kDiagFlags], 2u);
    stored = false;
    position = d.pos[i];
    rotation = d.rot[i];
    linear = d.lin[i];
    angular = d.ang[i];
  } else {
    d.pos[i] = position;
    d.rot[i] = rotation;
    d.lin[i] = linear;
    d.ang[i] = angular;
  }
  (void)stored;
  if (collisions) {
    d.slot_count[i] = 0u;
    const float4* m = d.masses + 4 * (size_t)i;
    v3 cx = world_inverse_inertia(rotation, m, mk(1.0f, 0.0f, 0.0f));
    v3 cy = world_inverse_inertia(rotation, m, mk(0.0f, 1.0f, 0.0f));
    v3 cz = world_inverse_inertia(rotation, m, mk(0.0f, 0.0f, 1.0f));
    float4* w = d.world_masses + 3 * (size_t)i;
    w[0] = f4(cx, m[0].x);
    w[1] = f4(cy, 0.0f);
    w[2] = f4(cz, 0.0f);
    float4* sb = d.sbody + 4 * (size_t)i;
    sb[0] = make_float4(linear.x, linear.y, linear.z, m[0].x);
    sb[1] = make_float4(angular.x, angular.y, angular.z, cx.x);
    sb[2] = make_float4(cx.y, cx.z, cy.x, cy.y);
    sb[3] = make_float4(cy.z, cz.x, cz.y, cz.z);
  }
  if (snapshot_now) {
    uint32_t slot = (uint32_t)(tick % 3u);
    d.snap_pos[slot][i] = position;
    d.snap_rot[slot][i] = rotation;
  }
}

// publish_snapshot (snapshot.wgsl:6-14) for ticks whose bodies move after integration.
__global__ __launch_bounds__(256) void k_snapshot(Dev d) {
  const uint64_t tick = tick_u64(d);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.body_count) return;
  uint32_t slot = (uint32_t)(tick % 3u);
  d.snap_pos[slot][i] = d.pos[i];
  d.snap_rot[slot][i] = d.rot[i];
}

// apply_external_impulse (physics.wgsl:72-90).
__global__ void k_external_impulse(Dev d, float4 world_point, float4 impulse4, uint32_t body) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const float4* m = d.masses + 4 * (size_t)body;
  float inverse_mass = m[0].x;
  if (inverse_mass <= 0.0f) return;
  v3 impulse = xyz(impulse4);
  v3 arm = xyz(world_point) - xyz(d.pos[body]);
  v3 lin = xyz(d.lin[body]) + impulse * inverse_mass;
  v3 ang = xyz(d.ang[body]) + world_inverse_inertia(d.rot[body], m, cross(arm, impulse));
  d.lin[body] = f4(lin, 0.0f);
  d.ang[body] = f4(ang, 0.0f);
}

static inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

void launch_set_tick(hipStream_t s, const Dev& d, uint64_t tick) {
  k_set_tick<<<1, 64, 0, s>>>(d, tick);
}
void launch_end_tick(hipStream_t s, const Dev& d) { k_end_tick<<<1, 64, 0, s>>>(d); }

void launch_integrate(hipStream_t s, const Dev& d, bool collisions, bool snapshot_now) {
  if (d.body_count == 0) return;
  k_integrate<<<blocks_for(d.body_count), kBlock, 0, s>>>(d, collisions ? 1 : 0, snapshot_now ? 1 : 0);
}
void launch_snapshot(hipStream_t s, const Dev& d) {
  if (d.body_count == 0) return;
  k_snapshot<<<blocks_for(d.body_count), kBlock, 0, s>>>(d);
}
void launch_external_impulse(hipStream_t s, const Dev& d, float4 point, float4 impulse, uint32_t body) {
  k_external_impulse<<<1, 64, 0, s>>>(d, point, impulse, body);
}

}  // namespace mh
