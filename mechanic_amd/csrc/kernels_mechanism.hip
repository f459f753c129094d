// Articulated path: bearing velocity projection, reduced coordinates, forward
// kinematics by pointer jumping, loop-closure Newton/PCG, velocity
// reconstruction and validation (articulated.wgsl, mechanism.wgsl,
// closure.wgsl, bearings.wgsl; pass order device.rs:963-1216, 1370-1416).
//
// Design on MI355X:
// * Bearing projection is a parallel Jacobi pass whose per-body deltas are Q20
//   integers summed with atomics -- the same order-independent arithmetic as the
//   reference -- applied to the bodies that some bearing references.
// * Everything that walks the forest (forward kinematics, the closure Newton
//   steps with their PCG solves, pose publication and the parent-to-child
//   velocity reconstruction) runs in ONE workgroup of 1024 threads separated by
//   workgroup barriers, instead of the reference's ~12 x 13 dispatches and its
//   single-thread loops. Reconstruction proceeds level by level (all bodies of
//   one depth in parallel), which gives the serial preorder loop's results: a
//   body reads only its parent's final velocity.
// * Closure sums are deterministic: every coordinate gathers its contributions
//   from a CSR list in (closure, side) order -- the order in which the reference's
//   serial PCG adds them -- and the PCG dot products are accumulated in FP64
//   with a fixed reduction tree.
#include <algorithm>
#include <type_traits>

#include "internal.h"
#include "mh_math.h"
#include "mh_solver.h"

namespace mh {

namespace {

constexpr float kFixedScaleM = 1048576.0f;       // articulated.wgsl:59
constexpr float kBearingRelaxation = 0.5f;       // articulated.wgsl:62
constexpr float kSleepSpeed = 0.005f;            // articulated.wgsl:63
constexpr uint32_t kDrivePassive = 0u;           // articulated.wgsl:65
constexpr uint32_t kDriveAngle = 2u;             // articulated.wgsl:66
constexpr float kAngleDeadband = 0.0005f;        // articulated.wgsl:69
constexpr uint32_t kFlagInvalid = 2u;            // INVALID_NUMERIC_FLAG
constexpr uint32_t kFlagNonConvergence = 4u;     // CONSTRAINT_NON_CONVERGENCE_FLAG
constexpr uint32_t kAnchorTolerance = 10u;       // closure.wgsl:69
constexpr uint32_t kAxisTolerance = 1000u;       // closure.wgsl:70
constexpr float kStepScale = 0.35f;              // closure.wgsl:72
constexpr float kMaxStep = 0.2f;                 // closure.wgsl:73
constexpr uint32_t kPoseThreads = 1024;

__device__ __forceinline__ uint64_t tick_m(const Dev& d) {
  return *reinterpret_cast<const volatile uint64_t*>(&d.tick->tick_index);
}
__device__ __forceinline__ uint32_t* diag_now(const Dev& d) { return d.diag[tick_m(d) & 1]; }

struct BearingRow {
  v3 anchor_a, anchor_b, axis_a, axis_b;
  uint32_t a, b, coordinate, closure;
};

__device__ __forceinline__ BearingRow load_bearing(const Dev& d, uint32_t i) {
  const float4* r = d.bearings + 5 * (size_t)i;
  BearingRow b;
  b.anchor_a = xyz(r[0]);
  b.anchor_b = xyz(r[1]);
  b.axis_a = xyz(r[2]);
  b.axis_b = xyz(r[3]);
  const float4 m = r[4];
  b.a = __float_as_uint(m.x);
  b.b = __float_as_uint(m.y);
  b.coordinate = __float_as_uint(m.z);
  b.closure = __float_as_uint(m.w);
  return b;
}

struct MechBody {
  uint32_t parent, tree_bearing, direction, root;
};
__device__ __forceinline__ MechBody load_mbody(const Dev& d, uint32_t body) {
  const float4 m = d.mbodies[4 * (size_t)body];
  return MechBody{__float_as_uint(m.x), __float_as_uint(m.y), __float_as_uint(m.z), __float_as_uint(m.w)};
}

__device__ __forceinline__ v3 normalize_v(v3 v) { return v / length(v); }
__device__ __forceinline__ float4 normalize_q(float4 q) {
  const float l = sqrtf(dot4(q, q));
  return make_float4(q.x / l, q.y / l, q.z / l, q.w / l);
}
__device__ __forceinline__ float clampf(float e, float lo, float hi) { return fmn(fmx(e, lo), hi); }
__device__ __forceinline__ bool finite3(v3 v) {
  const float lim = 3.402823e+38f;
  return v.x == v.x && v.y == v.y && v.z == v.z && fabsf(v.x) < lim && fabsf(v.y) < lim && fabsf(v.z) < lim;
}
// u32(round(x)) with WGSL's saturating conversion.
__device__ __forceinline__ uint32_t round_to_u32(float x) {
  const float r = __builtin_rintf(x);
  if (!(r > 0.0f)) return 0u;
  if (r >= 4294967296.0f) return 0xffffffffu;
  return (uint32_t)r;
}

// world_inverse_inertia (articulated.wgsl:97-105) from the body's rotation and
// body-frame inverse inertia.
__device__ __forceinline__ v3 wii(const Dev& d, uint32_t body, v3 v) {
  const float4 rotation = d.rot[body];
  const v3 local = quat_rotate(quat_conj(rotation), v);
  const float4* m = d.masses + 4 * (size_t)body;
  const v3 r = (xyz(m[1]) * local.x + xyz(m[2]) * local.y) + xyz(m[3]) * local.z;
  return quat_rotate(rotation, r);
}
__device__ __forceinline__ float inv_mass_of(const Dev& d, uint32_t body) { return d.masses[4 * (size_t)body].x; }

// Velocities: the state arrays, or the solver records while the contact solve runs.
__device__ __forceinline__ v3 lin_of(const Dev& d, uint32_t b, int in_contact) {
  return in_contact ? xyz(d.sbody[4 * (size_t)b]) : xyz(d.lin[b]);
}
__device__ __forceinline__ v3 ang_of(const Dev& d, uint32_t b, int in_contact) {
  return in_contact ? xyz(d.sbody[4 * (size_t)b + 1]) : xyz(d.ang[b]);
}


// One body's Q20 deltas from one bearing row: the rounded integers of each
// add_delta call (articulated.wgsl:107-116), summed with wrapping adds -- the
// arithmetic of the atomics, in any order -- before one atomic per component.
struct QSum {
  int32_t v[6];
};
__device__ __forceinline__ void qsum_add(QSum& q, v3 linear, v3 angular) {
  const int32_t r[6] = {round_to_i32(linear.x * kFixedScaleM),  round_to_i32(linear.y * kFixedScaleM),
                        round_to_i32(linear.z * kFixedScaleM),  round_to_i32(angular.x * kFixedScaleM),
                        round_to_i32(angular.y * kFixedScaleM), round_to_i32(angular.z * kFixedScaleM)};
#pragma unroll
  for (int k = 0; k < 6; ++k) q.v[k] = (int32_t)((uint32_t)q.v[k] + (uint32_t)r[k]);
}
__device__ __forceinline__ void qsum_flush(const Dev& d, uint32_t body, const QSum& q) {
  int32_t* p = d.mdeltas + 6 * (size_t)body;
#pragma unroll
  for (int k = 0; k < 6; ++k) atomicAdd(&p[k], q.v[k]);
}

// solve_linear_axis (articulated.wgsl:118-143).
__device__ __forceinline__ void linear_axis(const Dev& d, uint32_t a, uint32_t b, v3 arm_a, v3 arm_b, v3 relative,
                                            v3 direction, QSum& qa, QSum& qb) {
  const v3 angular_a = cross(arm_a, direction);
  const v3 angular_b = cross(arm_b, direction);
  const float denominator =
      ((inv_mass_of(d, a) + inv_mass_of(d, b)) + dot(angular_a, wii(d, a, angular_a))) + dot(angular_b, wii(d, b, angular_b));
  if (denominator <= 1.0e-12f) return;
  const v3 impulse = direction * ((kBearingRelaxation * dot(relative, direction)) / denominator);
  qsum_add(qa, impulse * inv_mass_of(d, a), wii(d, a, cross(arm_a, impulse)));
  qsum_add(qb, (-impulse) * inv_mass_of(d, b), wii(d, b, cross(arm_b, -impulse)));
}

// solve_angular_axis (articulated.wgsl:145-155).
__device__ __forceinline__ void angular_axis(const Dev& d, uint32_t a, uint32_t b, v3 relative, v3 axis, QSum& qa,
                                             QSum& qb) {
  const v3 inverse_a = wii(d, a, axis);
  const v3 inverse_b = wii(d, b, axis);
  const float denominator = dot(axis, inverse_a + inverse_b);
  if (denominator <= 1.0e-12f) return;
  const float impulse = (kBearingRelaxation * dot(relative, axis)) / denominator;
  qsum_add(qa, mk(0.0f, 0.0f, 0.0f), inverse_a * impulse);
  qsum_add(qb, mk(0.0f, 0.0f, 0.0f), (-inverse_b) * impulse);
}

// project_bearing_velocities (articulated.wgsl:281-305). The five axes of a
// bearing row all start from the velocities before it (none sees another's
// impulse), so lanes 0..4 of a group of kBearingLanes take one axis each; their
// rounded Q20 integers are summed across the group with wrapping adds (the
// arithmetic of the atomics, in any order) before lane 0's atomics.
constexpr uint32_t kBearingLanes = 8;
__global__ __launch_bounds__(256) void k_bearing_project(Dev d, int in_contact) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t i = gid / kBearingLanes, part = gid % kBearingLanes;
  const bool valid = i < d.bearing_count;
  QSum qa = {{0, 0, 0, 0, 0, 0}}, qb = {{0, 0, 0, 0, 0, 0}};
  uint32_t body_a = 0, body_b = 0;
  if (valid && part < 5u) {
    const BearingRow r = load_bearing(d, i);
    body_a = r.a;
    body_b = r.b;
    const float4 rot_a = d.rot[r.a], rot_b = d.rot[r.b];
    const v3 ang_a = ang_of(d, r.a, in_contact), ang_b = ang_of(d, r.b, in_contact);
    if (part < 3u) {
      const v3 arm_a = quat_rotate(rot_a, r.anchor_a);
      const v3 arm_b = quat_rotate(rot_b, r.anchor_b);
      const v3 velocity_a = lin_of(d, r.a, in_contact) + cross(ang_a, arm_a);
      const v3 velocity_b = lin_of(d, r.b, in_contact) + cross(ang_b, arm_b);
      const v3 relative = velocity_b - velocity_a;
      const v3 direction = part == 0u ? mk(1.0f, 0.0f, 0.0f) : part == 1u ? mk(0.0f, 1.0f, 0.0f) : mk(0.0f, 0.0f, 1.0f);
      linear_axis(d, r.a, r.b, arm_a, arm_b, relative, direction, qa, qb);
    } else {
      const v3 axis_a = normalize_v(quat_rotate(rot_a, r.axis_a));
      const v3 axis_b = normalize_v(quat_rotate(rot_b, r.axis_b));
      const v3 hinge = normalize_v(axis_a + axis_b);
      const v3 helper = fabsf(hinge.x) > 0.8f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
      const v3 tangent_a = normalize_v(cross(hinge, helper));
      const v3 tangent = part == 3u ? tangent_a : cross(hinge, tangent_a);
      angular_axis(d, r.a, r.b, ang_b - ang_a, tangent, qa, qb);
    }
  }
#pragma unroll
  for (int m = 1; m < (int)kBearingLanes; m <<= 1)
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      qa.v[k] = (int32_t)((uint32_t)qa.v[k] + (uint32_t)__shfl_xor(qa.v[k], m));
      qb.v[k] = (int32_t)((uint32_t)qb.v[k] + (uint32_t)__shfl_xor(qb.v[k], m));
    }
  if (valid && part == 0u) {
    qsum_flush(d, body_a, qa);
    qsum_flush(d, body_b, qb);
  }
}

// apply_velocity_deltas (articulated.wgsl:307-326) over the bodies some bearing
// references; every other body's delta is zero.
__global__ __launch_bounds__(256) void k_bearing_apply(Dev d, int in_contact, int write_back) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.bearing_body_count) return;
  const uint32_t body = d.blist[i];
  int32_t* p = d.mdeltas + 6 * (size_t)body;
  int32_t q[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) q[k] = atomicExch(&p[k], 0);
  const v3 linear_delta = mk((float)q[0], (float)q[1], (float)q[2]) / kFixedScaleM;
  const v3 angular_delta = mk((float)q[3], (float)q[4], (float)q[5]) / kFixedScaleM;
  const v3 lin = lin_of(d, body, in_contact) + linear_delta;
  const v3 ang = ang_of(d, body, in_contact) + angular_delta;
  if (in_contact) {
    float4* r = d.sbody + 4 * (size_t)body;
    r[0] = f4(lin, r[0].w);
    r[1] = f4(ang, r[1].w);
  }
  if (!in_contact || write_back) {
    d.lin[body] = f4(lin, 0.0f);
    d.ang[body] = f4(ang, 0.0f);
  }
}

// world_inverse_inertia (articulated.wgsl:97-105) of a body whose rotation and
// mass rows are held in registers (the arithmetic of wii).
__device__ __forceinline__ v3 wii_of(float4 rotation, const float4 (&m)[4], v3 v) {
  const v3 local = quat_rotate(quat_conj(rotation), v);
  const v3 r = (xyz(m[1]) * local.x + xyz(m[2]) * local.y) + xyz(m[3]) * local.z;
  return quat_rotate(rotation, r);
}

// `passes` passes of project_bearing_velocities and apply_velocity_deltas
// (articulated.wgsl:281-326) on the state arrays, one workgroup per group of
// bearing islands with one (row, axis) item per thread: the group's bodies
// keep their velocities and Q20 sums in LDS from pass to pass. Rotations do not
// change during the passes, so each item forms its arms, axis and denominator
// once; each pass then computes what one lane of k_bearing_project computes
// from the same values, the sums are the same wrapping integer sums and the
// update is k_bearing_apply's: the result is the one of launching the two
// kernels per pass.
//
// With `apply_first` (inside the contact solve, in the launch of the pass's
// velocity update) the group first adds each of its bodies' contact deltas of
// the pass, as solve_apply does: the same wrapping sums of its slots and
// overflow words, one thread per body; only the group that writes a body back
// clears its overflow words (a still body's are zero).
__device__ __forceinline__ void island_group(const Dev& d, uint32_t g, uint32_t passes, int in_contact,
                                             int write_back, bool apply_first) {
  __shared__ float4 s_lin[kIslandBodies], s_ang[kIslandBodies];
  __shared__ int32_t s_q[kIslandBodies][6];
  const uint32_t t = threadIdx.x;
  const uint32_t r0 = d.bi_offset[g], rows = d.bi_offset[g + 1] - r0;
  const uint32_t b0 = d.bi_boff[g], bodies = d.bi_boff[g + 1] - b0;
  uint32_t body = 0, owner = 0;
  if (t < bodies) {
    const uint32_t w = d.bi_body[b0 + t];
    body = w & 0x7fffffffu;
    owner = w >> 31;
    // Velocities: the state arrays, or the solver records (whose fourth words
    // the update keeps) while the contact solve runs.
    float4 lin = in_contact ? d.sbody[4 * (size_t)body] : d.lin[body];
    float4 ang = in_contact ? d.sbody[4 * (size_t)body + 1] : d.ang[body];
    if (apply_first && *reinterpret_cast<volatile uint32_t*>(&d.diag[tick_sv(d) & 1][kDiagActive]) != 0u) {
      const uint32_t count = d.slot_count[body];
      const uint32_t n = count < kSlotOverflow ? count : kSlotOverflow;
      uint32_t acc[6] = {0u, 0u, 0u, 0u, 0u, 0u};
      for (uint32_t k = 0; k < n; ++k) {
        const int4* e = d.cdelta + 2 * ((size_t)k * d.body_count + body);
        const int4 x = e[0];
        const int2 y = *reinterpret_cast<const int2*>(e + 1);
        acc[0] += (uint32_t)x.x;
        acc[1] += (uint32_t)x.y;
        acc[2] += (uint32_t)x.z;
        acc[3] += (uint32_t)x.w;
        acc[4] += (uint32_t)y.x;
        acc[5] += (uint32_t)y.y;
      }
      if (count > kSlotOverflow) {
        int32_t* p = d.deltas + 6 * (size_t)body;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          acc[c] += (uint32_t)p[c];
          if (owner) p[c] = 0;
        }
      }
      const v3 lin_delta = mk((float)(int32_t)acc[0], (float)(int32_t)acc[1], (float)(int32_t)acc[2]) / kFixedScale;
      const v3 ang_delta = mk((float)(int32_t)acc[3], (float)(int32_t)acc[4], (float)(int32_t)acc[5]) / kFixedScale;
      lin = f4(xyz(lin) + lin_delta, lin.w);
      ang = f4(xyz(ang) + ang_delta, ang.w);
    }
    s_lin[t] = lin;
    s_ang[t] = ang;
#pragma unroll
    for (int k = 0; k < 6; ++k) s_q[t][k] = 0;
  }
  // This thread's item: axis `part` of row j (parts 0..2 linear x, y, z; 3, 4
  // the two angular tangents).
  const bool item = t < 5u * rows;
  const uint32_t j = t / 5u, part = t % 5u;
  uint32_t la = 0, lb = 0;
  bool live = false;
  float4 rot_a = make_float4(0.f, 0.f, 0.f, 1.f), rot_b = rot_a;
  float4 ma[4], mb[4];
  float inv_a = 0.0f, inv_b = 0.0f, denominator = 1.0f;
  v3 arm_a = mk(0.0f, 0.0f, 0.0f), arm_b = arm_a, direction = arm_a, inverse_a = arm_a, inverse_b = arm_a;
  if (item) {
    const BearingRow r = load_bearing(d, d.bi_bearing[r0 + j]);
    const uint32_t side = d.bi_side[r0 + j];
    la = side & 0xffffu;
    lb = side >> 16;
    rot_a = d.rot[r.a];
    rot_b = d.rot[r.b];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      ma[k] = d.masses[4 * (size_t)r.a + k];
      mb[k] = d.masses[4 * (size_t)r.b + k];
    }
    inv_a = ma[0].x;
    inv_b = mb[0].x;
    if (part < 3u) {
      // solve_linear_axis (articulated.wgsl:118-143), the part before the impulse.
      arm_a = quat_rotate(rot_a, r.anchor_a);
      arm_b = quat_rotate(rot_b, r.anchor_b);
      direction = part == 0u ? mk(1.0f, 0.0f, 0.0f) : part == 1u ? mk(0.0f, 1.0f, 0.0f) : mk(0.0f, 0.0f, 1.0f);
      const v3 angular_a = cross(arm_a, direction);
      const v3 angular_b = cross(arm_b, direction);
      denominator = ((inv_a + inv_b) + dot(angular_a, wii_of(rot_a, ma, angular_a))) +
                    dot(angular_b, wii_of(rot_b, mb, angular_b));
    } else {
      // solve_angular_axis (articulated.wgsl:145-155), the part before the impulse.
      const v3 axis_a = normalize_v(quat_rotate(rot_a, r.axis_a));
      const v3 axis_b = normalize_v(quat_rotate(rot_b, r.axis_b));
      const v3 hinge = normalize_v(axis_a + axis_b);
      const v3 helper = fabsf(hinge.x) > 0.8f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
      const v3 tangent_a = normalize_v(cross(hinge, helper));
      direction = part == 3u ? tangent_a : cross(hinge, tangent_a);
      inverse_a = wii_of(rot_a, ma, direction);
      inverse_b = wii_of(rot_b, mb, direction);
      denominator = dot(direction, inverse_a + inverse_b);
    }
    live = !(denominator <= 1.0e-12f);
  }
  __syncthreads();
  for (uint32_t pass = 0; pass < passes; ++pass) {
    if (item && live) {
      QSum qa = {{0, 0, 0, 0, 0, 0}}, qb = {{0, 0, 0, 0, 0, 0}};
      const v3 ang_a = xyz(s_ang[la]), ang_b = xyz(s_ang[lb]);
      if (part < 3u) {
        const v3 velocity_a = xyz(s_lin[la]) + cross(ang_a, arm_a);
        const v3 velocity_b = xyz(s_lin[lb]) + cross(ang_b, arm_b);
        const v3 relative = velocity_b - velocity_a;
        const v3 impulse = direction * ((kBearingRelaxation * dot(relative, direction)) / denominator);
        qsum_add(qa, impulse * inv_a, wii_of(rot_a, ma, cross(arm_a, impulse)));
        qsum_add(qb, (-impulse) * inv_b, wii_of(rot_b, mb, cross(arm_b, -impulse)));
      } else {
        const float impulse = (kBearingRelaxation * dot(ang_b - ang_a, direction)) / denominator;
        qsum_add(qa, mk(0.0f, 0.0f, 0.0f), inverse_a * impulse);
        qsum_add(qb, mk(0.0f, 0.0f, 0.0f), (-inverse_b) * impulse);
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        atomicAdd(&s_q[la][k], qa.v[k]);
        atomicAdd(&s_q[lb][k], qb.v[k]);
      }
    }
    __syncthreads();
    if (t < bodies) {
      int32_t q[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        q[k] = s_q[t][k];
        s_q[t][k] = 0;
      }
      const v3 linear_delta = mk((float)q[0], (float)q[1], (float)q[2]) / kFixedScaleM;
      const v3 angular_delta = mk((float)q[3], (float)q[4], (float)q[5]) / kFixedScaleM;
      s_lin[t] = f4(xyz(s_lin[t]) + linear_delta, in_contact ? s_lin[t].w : 0.0f);
      s_ang[t] = f4(xyz(s_ang[t]) + angular_delta, in_contact ? s_ang[t].w : 0.0f);
    }
    __syncthreads();
  }
  if (t < bodies && owner) {
    if (in_contact) {
      d.sbody[4 * (size_t)body] = s_lin[t];
      d.sbody[4 * (size_t)body + 1] = s_ang[t];
    }
    if (!in_contact || write_back) {
      d.lin[body] = f4(xyz(s_lin[t]), 0.0f);
      d.ang[body] = f4(xyz(s_ang[t]), 0.0f);
    }
  }
}

__global__ __launch_bounds__(256) void k_bearing_islands(Dev d, uint32_t passes, int in_contact, int write_back) {
  island_group(d, blockIdx.x, passes, in_contact, write_back, false);
}

// A contact pass's velocity update and bearing projection in one launch: the
// first blocks are the island groups (dispatched first, so they overlap the
// rest), each updating its own bodies and then projecting its bearings (one
// pass on the solver records), padded to whole rounds of the XCDs; the other
// blocks update the bodies no island group holds, as k_apply does, with the
// same block-to-XCD mapping as k_apply.
__global__ __launch_bounds__(256) void k_apply_islands(Dev d, uint32_t apply_blocks, int last, uint64_t* clock) {
  const uint32_t lead = (d.bi_groups + kXcds - 1) / kXcds * kXcds;
  if (blockIdx.x < lead) {
    if (blockIdx.x < d.bi_groups) island_group(d, blockIdx.x, 1u, 1, last, true);
    return;
  }
  apply_contact_deltas(d, last, clock, d.bi_mask, blockIdx.x - lead, apply_blocks);
}

// permitted_axis / permitted_speed / stabilized_speed (articulated.wgsl:328-355).
__device__ __forceinline__ v3 permitted_axis(const Dev& d, const MechBody& m, const BearingRow& r) {
  const float4 rp = d.rot[m.parent];
  return m.direction != 0u ? -normalize_v(quat_rotate(rp, r.axis_b)) : normalize_v(quat_rotate(rp, r.axis_a));
}
__device__ __forceinline__ float stabilized(v3 axis, float speed) {
  const bool gravity_aligned = fabsf(axis.y) > 0.999f;
  return gravity_aligned && fabsf(speed) < kSleepSpeed ? 0.0f : speed;
}

// advance_coordinates (articulated.wgsl:357-407).
__global__ __launch_bounds__(256) void k_advance(Dev d) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.mech_count) return;
  const uint32_t body = d.mlist[i];
  const MechBody m = load_mbody(d, body);
  const BearingRow r = load_bearing(d, m.tree_bearing);
  const uint32_t c = r.coordinate;
  if (c == kInvalid) return;
  const v3 axis = permitted_axis(d, m, r);
  const float measured = dot(xyz(d.ang[body]) - xyz(d.ang[m.parent]), axis) * d.angular_damping;
  const float4 d0 = d.drives[2 * (size_t)c], d1 = d.drives[2 * (size_t)c + 1];
  const uint32_t mode = __float_as_uint(d0.x);
  const float max_acceleration = d0.y, max_speed = d0.z, target_speed = d0.w;
  const float target_angle = d1.x, min_angle = d1.y, max_angle = d1.z;
  const float2 coordinate = d.coords[c];
  float speed;
  if (mode == kDrivePassive) {
    speed = stabilized(axis, measured);
  } else {
    float desired = clampf(target_speed, -max_speed, max_speed);
    if (mode == kDriveAngle) {
      const float error = target_angle - coordinate.x;
      const float brake = sqrtf((2.0f * max_acceleration) * fabsf(error));
      desired = signf(error) * fmn(brake, max_speed);
      if (fabsf(error) < kAngleDeadband) desired = 0.0f;
    }
    const float budget = max_acceleration * d.delta_seconds;
    speed = measured + clampf(desired - measured, -budget, budget);
  }
  float angle = coordinate.x + speed * d.delta_seconds;
  if (angle < min_angle) {
    angle = min_angle;
    speed = 0.0f;
  } else if (angle > max_angle) {
    angle = max_angle;
    speed = 0.0f;
  }
  d.coords[c] = make_float2(angle, speed);
}

// capture_coordinates (articulated.wgsl:409-421).
__global__ __launch_bounds__(256) void k_capture(Dev d) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.mech_count) return;
  const uint32_t body = d.mlist[i];
  const MechBody m = load_mbody(d, body);
  const BearingRow r = load_bearing(d, m.tree_bearing);
  if (r.coordinate == kInvalid) return;
  const v3 axis = permitted_axis(d, m, r);
  const float speed = dot(xyz(d.ang[body]) - xyz(d.ang[m.parent]), axis);
  d.coords[r.coordinate].y = stabilized(axis, speed);
}

// ---- forest walks in one workgroup -----------------------------------------

struct Link {
  float4 position, rotation;
  uint32_t ancestor, resolved;
};
__device__ __forceinline__ Link load_link(const float4* links, uint32_t b) {
  const float4* p = links + 3 * (size_t)b;
  const float4 m = p[2];
  return Link{p[0], p[1], __float_as_uint(m.x), __float_as_uint(m.y)};
}
__device__ __forceinline__ void store_link(float4* links, uint32_t b, const Link& l) {
  float4* p = links + 3 * (size_t)b;
  p[0] = l.position;
  p[1] = l.rotation;
  p[2] = make_float4(__uint_as_float(l.ancestor), __uint_as_float(l.resolved), 0.0f, 0.0f);
}

// axis_rotation (mechanism.wgsl:70-73).
__device__ __forceinline__ float4 axis_rotation(v3 axis, float angle) {
  const float half_angle = angle * 0.5f;
  const v3 a = normalize_v(axis) * sinf(half_angle);
  return f4(a, cosf(half_angle));
}

// prepare_links (mechanism.wgsl:75-107) for one non-root body at a given angle:
// its link relative to its parent.
__device__ __forceinline__ Link make_link(const Dev& d, uint32_t body, const MechBody& m, const BearingRow& r,
                                          float angle) {
  const float4 bind_rotation = d.mbodies[4 * (size_t)body + 3];
  float4 rr;
  v3 rp;
  if (m.direction == 0u) {
    rr = quat_multiply(axis_rotation(r.axis_a, angle), bind_rotation);
    rp = r.anchor_a - quat_rotate(rr, r.anchor_b);
  } else {
    rr = quat_multiply(axis_rotation(r.axis_b, -angle), bind_rotation);
    rp = r.anchor_b - quat_rotate(rr, r.anchor_a);
  }
  return Link{f4(rp, 0.0f), rr, m.parent, 0u};
}

__device__ __forceinline__ void prepare_link(const Dev& d, uint32_t body) {
  const MechBody m = load_mbody(d, body);
  const BearingRow r = load_bearing(d, m.tree_bearing);
  const float angle = d.coords[r.coordinate].x;
  const float4 bind_rotation = d.mbodies[4 * (size_t)body + 3];
  float4 rr;
  v3 rp;
  if (m.direction == 0u) {
    rr = quat_multiply(axis_rotation(r.axis_a, angle), bind_rotation);
    rp = r.anchor_a - quat_rotate(rr, r.anchor_b);
  } else {
    rr = quat_multiply(axis_rotation(r.axis_b, -angle), bind_rotation);
    rp = r.anchor_b - quat_rotate(rr, r.anchor_a);
  }
  store_link(d.links[0], body, Link{f4(rp, 0.0f), rr, m.parent, 0u});
}

// compose (mechanism.wgsl:109-119): the child's link carried through its
// ancestor's, as jump_a_to_b / jump_b_to_a (mechanism.wgsl:121-147) apply it.
__device__ __forceinline__ Link compose_link(const Link& parent, const Link& child) {
  Link out;
  out.position = f4(xyz(parent.position) + quat_rotate(parent.rotation, xyz(child.position)), 0.0f);
  out.rotation = normalize_q(quat_multiply(parent.rotation, child.rotation));
  out.ancestor = parent.ancestor;
  out.resolved = parent.resolved;
  return out;
}

// One pointer-jumping round over kJumpBatch bodies per thread at a time: every
// child link, then every parent link, is loaded before any is composed, so a
// thread waits two memory latencies per batch instead of two per body.
constexpr int kJumpBatch = 4;

__device__ __forceinline__ void jump_round(const uint32_t* list, uint32_t begin, uint32_t end, const float4* src,
                                           float4* dst) {
  for (uint32_t i0 = begin + threadIdx.x; i0 < end; i0 += kJumpBatch * blockDim.x) {
    uint32_t body[kJumpBatch];
    Link child[kJumpBatch], parent[kJumpBatch];
#pragma unroll
    for (int k = 0; k < kJumpBatch; ++k) {
      const uint32_t i = i0 + k * blockDim.x;
      body[k] = i < end ? list[i] : kInvalid;
    }
#pragma unroll
    for (int k = 0; k < kJumpBatch; ++k)
      if (body[k] != kInvalid) child[k] = load_link(src, body[k]);
#pragma unroll
    for (int k = 0; k < kJumpBatch; ++k)
      if (body[k] != kInvalid && child[k].resolved == 0u) parent[k] = load_link(src, child[k].ancestor);
#pragma unroll
    for (int k = 0; k < kJumpBatch; ++k) {
      if (body[k] == kInvalid) continue;
      if (child[k].resolved != 0u) {
        store_link(dst, body[k], child[k]);
        continue;
      }
      store_link(dst, body[k], compose_link(parent[k], child[k]));
    }
  }
}

// Forward kinematics (prepare_links + pointer jumping, mechanism.wgsl:75-147)
// over whole trees, one workgroup per group of trees (fk_offset): a link only
// ever points at its own tree, so the jumping rounds need only the workgroup's
// barriers, and the groups run on as many CUs as there are groups. `step` >= 0
// is the pass after Newton step `step`, skipped once the loop has converged;
// the tick's first pass (step < 0) clears that flag.
__global__ __launch_bounds__(256) void k_forward_kinematics(Dev d, int step) {
  if (step >= 0 && *reinterpret_cast<volatile uint32_t*>(&d.mech_state[0]) != 0u) return;
  if (step < 0 && blockIdx.x == 0 && threadIdx.x == 0) d.mech_state[0] = 0u;
  const uint32_t begin = d.fk_offset[blockIdx.x], end = d.fk_offset[blockIdx.x + 1];
  for (uint32_t i = begin + threadIdx.x; i < end; i += blockDim.x) prepare_link(d, d.fk_list[i]);
  __syncthreads();
  for (uint32_t round = 0; round < d.jump_rounds; ++round) {
    jump_round(d.fk_list, begin, end, d.links[round & 1u], d.links[(round & 1u) ^ 1u]);
    __syncthreads();
  }
}

// reconstruct_body_velocities (articulated.wgsl:423-456), one depth level at a time.
__device__ __forceinline__ void reconstruct_body(const Dev& d, uint32_t body) {
  const MechBody m = load_mbody(d, body);
  const BearingRow r = load_bearing(d, m.tree_bearing);
  const float4 rp = d.rot[m.parent], rb = d.rot[body];
  v3 axis, parent_anchor, child_anchor;
  if (m.direction != 0u) {
    axis = -normalize_v(quat_rotate(rp, r.axis_b));
    parent_anchor = quat_rotate(rp, r.anchor_b);
    child_anchor = quat_rotate(rb, r.anchor_a);
  } else {
    axis = normalize_v(quat_rotate(rp, r.axis_a));
    parent_anchor = quat_rotate(rp, r.anchor_a);
    child_anchor = quat_rotate(rb, r.anchor_b);
  }
  const float speed = d.coords[r.coordinate].y;
  const v3 parent_ang = xyz(d.ang[m.parent]);
  const v3 angular = parent_ang + axis * speed;
  const v3 anchor_velocity = xyz(d.lin[m.parent]) + cross(parent_ang, parent_anchor);
  d.lin[body] = f4(anchor_velocity - cross(angular, child_anchor), 0.0f);
  d.ang[body] = f4(angular, 0.0f);
}

__device__ void reconstruct_levels(const Dev& d) {
  for (uint32_t level = 0; level < d.level_count; ++level) {
    const uint32_t first = d.level_offset[level], end = d.level_offset[level + 1];
    for (uint32_t i = first + threadIdx.x; i < end; i += blockDim.x) reconstruct_body(d, d.mlist[i]);
    __syncthreads();
  }
}

// publish_a / publish_b (mechanism.wgsl:149-185).
__device__ void publish(const Dev& d, const float4* links) {
  for (uint32_t i = threadIdx.x; i < d.mech_count; i += blockDim.x) {
    const uint32_t body = d.mlist[i];
    const Link l = load_link(links, body);
    const uint32_t root = l.ancestor;
    if (body == root) continue;
    const float4 root_rotation = d.rot[root];
    d.pos[body] = f4(xyz(d.pos[root]) + quat_rotate(root_rotation, xyz(l.position)), 0.0f);
    d.rot[body] = normalize_q(quat_multiply(root_rotation, l.rotation));
  }
  __syncthreads();
}

constexpr uint32_t kPathIndices = 16;  // path indices of a closure side loaded per step
constexpr uint32_t kLdsCoordinates = 4608;  // coordinates whose PCG data fits in LDS

// Deterministic FP64 sum over the workgroup: a fixed butterfly within each
// wavefront, then every thread adds the wavefronts' sums in wavefront order, so
// all threads hold the same bits. Two barriers in all.
__device__ double block_sum(double v, double* scratch) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
  if ((threadIdx.x & 63u) == 0) scratch[threadIdx.x >> 6] = v;
  __syncthreads();
  double total = 0.0;
  for (uint32_t w = 0; w < blockDim.x / 64; ++w) total += scratch[w];
  __syncthreads();
  return total;
}

// block_sum for a workgroup of kT threads: a single wavefront's butterfly
// leaves the same bits in every lane (each stage adds the same two values in
// either order), so one wavefront needs no LDS and no barrier; 0.0 + v keeps
// block_sum's result for -0.0.
template <uint32_t kT>
__device__ __forceinline__ double group_sum(double v, double* scratch) {
  if constexpr (kT == 64) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
    return 0.0 + v;
  } else {
    return block_sum(v, scratch);
  }
}

// One FP64 word of the lane `kCtrl` names, by DPP (both halves).
template <int kCtrl>
__device__ __forceinline__ double dpp_f64(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), kCtrl, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), kCtrl, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_f64(double x, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
  return __hiloint2double(hi, lo);
}

// A value every lane holds the same bits of, as a wavefront-uniform one (so
// branches on it stay scalar).
__device__ __forceinline__ double uniform_f64(double x) {
  const int lo = __builtin_amdgcn_readfirstlane(__double2loint(x));
  const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(x));
  return __hiloint2double(hi, lo);
}

// Sum over a whole wavefront (every lane active), the same bits in every lane:
// within each row of 16 lanes, four DPP rotations by 8, 4, 2 and 1 (each adds a
// pair that both of its lanes form, so the row's lanes agree at every step),
// then the four row sums as (row 0 + row 1) + (row 2 + row 3).
__device__ __forceinline__ double wave_sum_rows(double v) {
  v = v + dpp_f64<0x128>(v);  // row_ror:8
  v = v + dpp_f64<0x124>(v);  // row_ror:4
  v = v + dpp_f64<0x122>(v);  // row_ror:2
  v = v + dpp_f64<0x121>(v);  // row_ror:1
  return uniform_f64((readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48)));
}

struct Endpoint {
  v3 anchor, axis;
};

// Joint axis and anchor of a coordinate in its component root's frame
// (closure.wgsl:172-185).
__device__ __forceinline__ void joint_of(const Dev& d, const float4* links, uint32_t body, v3& axis, v3& anchor) {
  const MechBody m = load_mbody(d, body);
  const BearingRow r = load_bearing(d, m.tree_bearing);
  const Link parent = load_link(links, m.parent);
  if (m.direction == 0u) {
    axis = normalize_v(quat_rotate(parent.rotation, r.axis_a));
    anchor = xyz(parent.position) + quat_rotate(parent.rotation, r.anchor_a);
  } else {
    axis = -normalize_v(quat_rotate(parent.rotation, r.axis_b));
    anchor = xyz(parent.position) + quat_rotate(parent.rotation, r.anchor_b);
  }
}

// One step of the closure Newton loop (device.rs:976-1020, closure.wgsl:233-427)
// in one workgroup: evaluate the closures, and unless all are inside tolerance
// (then the loop is over: mech_state[0] is set and every later step and
// forward-kinematics pass of this tick returns at once), the gradient, the PCG
// solve and the clamped step of the coordinates. k_forward_kinematics follows
// each step.
__global__ __launch_bounds__(kPoseThreads) void k_newton_step(Dev d) {
  __shared__ uint32_t state[2];
  __shared__ double scratch[kPoseThreads];
  // With at most kLdsCoordinates coordinates, the step's joint frames and the
  // PCG's z and p are also held in LDS (147 KiB), where every iteration's
  // gathers read them; d.joints and d.pcg stay the full record.
  __shared__ float lds_joint[kLdsCoordinates][6];
  __shared__ float lds_zp[kLdsCoordinates][2];
  if (*reinterpret_cast<volatile uint32_t*>(&d.mech_state[0]) != 0u) return;
  uint32_t* diag = diag_now(d);
  const float4* links = d.links[d.jump_rounds & 1u];
  const uint32_t nc = d.closure_count, nq = d.coordinate_count;
  const bool in_lds = nq <= kLdsCoordinates;
  auto joint = [&](uint32_t c, v3& axis, v3& anchor) {
    if (in_lds) {
      const float* j = lds_joint[c];
      axis = mk(j[0], j[1], j[2]);
      anchor = mk(j[3], j[4], j[5]);
    } else {
      axis = xyz(d.joints[2 * (size_t)c]);
      anchor = xyz(d.joints[2 * (size_t)c + 1]);
    }
  };
  auto set_zp = [&](uint32_t c, float* q, float z, float p) {
    q[2] = z;
    q[3] = p;
    if (in_lds) {
      lds_zp[c][0] = z;
      lds_zp[c][1] = p;
    }
  };
  {
    if (threadIdx.x == 0) {
      state[0] = 0u;
      state[1] = 0u;
    }
    __syncthreads();
    // evaluate_closures: endpoints, residuals, tolerance maxima.
    for (uint32_t k = threadIdx.x; k < nc; k += blockDim.x) {
      const BearingRow r = load_bearing(d, d.closure_rows[k]);
      const Link pa = load_link(links, r.a), pb = load_link(links, r.b);
      const v3 anchor_a = xyz(pa.position) + quat_rotate(pa.rotation, r.anchor_a);
      const v3 anchor_b = xyz(pb.position) + quat_rotate(pb.rotation, r.anchor_b);
      const v3 axis_a = normalize_v(quat_rotate(pa.rotation, r.axis_a));
      const v3 axis_b = normalize_v(quat_rotate(pb.rotation, r.axis_b));
      const v3 res_pos = anchor_a - anchor_b, res_axis = axis_a - axis_b;
      float4* s = d.cl_state + 6 * (size_t)k;
      s[0] = f4(anchor_a, 0.0f);
      s[1] = f4(axis_a, 0.0f);
      s[2] = f4(anchor_b, 0.0f);
      s[3] = f4(axis_b, 0.0f);
      float active = 0.0f;
      if (!finite3(res_pos) || !finite3(res_axis)) {
        atomicOr(&diag[kDiagFlags], kFlagInvalid);
        atomicMax(&state[0], 0xffffffffu);
        atomicMax(&state[1], 0xffffffffu);
      } else {
        const uint32_t um = round_to_u32(length(res_pos) * 1000000.0f);
        const float degrees = acosf(clampf(dot(axis_a, axis_b), -1.0f, 1.0f)) * 57.295779513f;
        const uint32_t ud = round_to_u32(degrees * 1000000.0f);
        atomicMax(&state[0], um);
        atomicMax(&state[1], ud);
        active = (um <= kAnchorTolerance && ud <= kAxisTolerance) ? 0.0f : 1.0f;
      }
      s[4] = f4(res_pos, active);
      s[5] = f4(res_axis, 0.0f);
    }
    __syncthreads();
    // finalize_closures: the rest of this step, and every later step, only
    // while some closure is outside tolerance.
    if (!(state[0] > kAnchorTolerance || state[1] > kAxisTolerance)) {
      if (threadIdx.x == 0) d.mech_state[0] = 1u;
      return;
    }
    // Joint axes and anchors of every coordinate.
    for (uint32_t c = threadIdx.x; c < nq; c += blockDim.x) {
      v3 axis, anchor;
      joint_of(d, links, d.coord_body[c], axis, anchor);
      d.joints[2 * (size_t)c] = f4(axis, 0.0f);
      d.joints[2 * (size_t)c + 1] = f4(anchor, 0.0f);
      if (in_lds) {
        float* j = lds_joint[c];
        j[0] = axis.x;
        j[1] = axis.y;
        j[2] = axis.z;
        j[3] = anchor.x;
        j[4] = anchor.y;
        j[5] = anchor.z;
      }
    }
    __syncthreads();
    // Gradient and Jacobi diagonal (accumulate_branch), then the PCG start.
    double rz_part = 0.0;
    for (uint32_t c = threadIdx.x; c < nq; c += blockDim.x) {
      v3 jaxis, janchor;
      joint(c, jaxis, janchor);
      double g = 0.0, dg = 0.0;
      for (uint32_t e = d.coord_offset[c]; e < d.coord_offset[c + 1]; ++e) {
        const uint32_t entry = d.coord_entries[e], k = entry >> 1, side = entry & 1u;
        const float4* s = d.cl_state + 6 * (size_t)k;
        if (s[4].w == 0.0f) continue;
        const v3 end_anchor = xyz(s[side ? 2 : 0]), end_axis = xyz(s[side ? 3 : 1]);
        const float sign = side ? -1.0f : 1.0f;
        const v3 pd = cross(jaxis, end_anchor - janchor) * sign;
        const v3 ad = cross(jaxis, end_axis) * sign;
        const float gradient = dot(xyz(s[4]), pd) + 1.0f * dot(xyz(s[5]), ad);
        const float diagonal = dot(pd, pd) + 1.0f * dot(ad, ad);
        if (gradient == gradient && diagonal == diagonal) {
          g += (double)gradient;
          dg += (double)diagonal;
        } else {
          atomicOr(&diag[kDiagFlags], kFlagInvalid);
        }
      }
      const float gf = (float)g, df = (float)dg;
      const float residual = df > 1.0e-12f ? -gf : 0.0f;
      const float pre = df > 1.0e-12f ? residual / df : 0.0f;
      float* q = d.pcg + 8 * (size_t)c;
      q[0] = 0.0f;
      q[1] = residual;
      set_zp(c, q, pre, pre);
      q[4] = 0.0f;
      q[5] = gf;
      q[6] = df;
      rz_part += (double)residual * (double)pre;
    }
    double rz = block_sum(rz_part, scratch);
    // With a tolerance, the PCG stops once sqrt(r.z / r0.z) is at most it. rz is
    // the same in every thread, so the whole workgroup leaves together.
    const double rz_stop = (double)d.closure_tol * (double)d.closure_tol * rz;
    // The search direction of iteration `it` is p = z + beta p_prev (the
    // preconditioned start at it = 0); each phase forms it from the previous
    // iteration's z and p, which the update phase then replaces, so no pass
    // is spent storing it.
    float beta = 0.0f;
    // Closure sides of the product phase, two per closure on adjacent lanes; a
    // multiple of 64 so that lane pairs stay whole.
    const uint32_t slots = 2u * ((nc + 31u) / 32u) * 32u;
    v3 my_anchor = mk(0.0f, 0.0f, 0.0f), my_axis = my_anchor;
    uint32_t my_e0 = 0u, my_e1 = 0u;
    if ((threadIdx.x >> 1) < nc) {
      const uint32_t k = threadIdx.x >> 1, side = threadIdx.x & 1u;
      const float4* s = d.cl_state + 6 * (size_t)k;
      my_anchor = xyz(s[side ? 2 : 0]);
      my_axis = xyz(s[side ? 3 : 1]);
      my_e0 = d.cpath_offset[2 * k + side];
      my_e1 = d.cpath_offset[2 * k + side + 1];
    }
    for (uint32_t it = 0; it < d.closure_iters; ++it) {
      if (d.closure_tol > 0.0f && rz <= rz_stop) break;
      auto direction_of = [&](uint32_t c) {
        float z, p;
        if (in_lds) {
          z = lds_zp[c][0];
          p = lds_zp[c][1];
        } else {
          const float* q = d.pcg + 8 * (size_t)c;
          z = q[2];
          p = q[3];
        }
        return it == 0 ? p : z + beta * p;
      };
      // branch_product for every closure (closure.wgsl:188-220): adjacent lanes
      // take a closure's two sides. A side's end point and path range do not
      // change during the solve (the first slot of each thread keeps them in
      // registers); its path indices are loaded kPathIndices at a time and the
      // joints and directions read from LDS, summing in path order as before.
      for (uint32_t slot = threadIdx.x; slot < slots; slot += blockDim.x) {
        const uint32_t k = slot >> 1, side = slot & 1u;
        v3 pp = mk(0.0f, 0.0f, 0.0f), pa = pp;
        if (k < nc) {
          const bool own = slot == threadIdx.x;
          const float4* s = d.cl_state + 6 * (size_t)k;
          const v3 end_anchor = own ? my_anchor : xyz(s[side ? 2 : 0]);
          const v3 end_axis = own ? my_axis : xyz(s[side ? 3 : 1]);
          const float sign = side ? -1.0f : 1.0f;
          const uint32_t e0 = own ? my_e0 : d.cpath_offset[2 * k + side];
          const uint32_t e1 = own ? my_e1 : d.cpath_offset[2 * k + side + 1];
          for (uint32_t e = e0; e < e1; e += kPathIndices) {
            uint32_t cs[kPathIndices];
#pragma unroll
            for (uint32_t j = 0; j < kPathIndices; ++j) cs[j] = d.cpath[e + j < e1 ? e + j : e];
#pragma unroll
            for (uint32_t j = 0; j < kPathIndices; ++j) {
              if (e + j >= e1) break;
              v3 jaxis, janchor;
              joint(cs[j], jaxis, janchor);
              const float dir = direction_of(cs[j]);
              pp = pp + (cross(jaxis, end_anchor - janchor) * sign) * dir;
              pa = pa + (cross(jaxis, end_axis) * sign) * dir;
            }
          }
        }
        // Side 0 (even lane) adds side 1's sums from its neighbour.
        const v3 pp1 = mk(__shfl_xor(pp.x, 1), __shfl_xor(pp.y, 1), __shfl_xor(pp.z, 1));
        const v3 pa1 = mk(__shfl_xor(pa.x, 1), __shfl_xor(pa.y, 1), __shfl_xor(pa.z, 1));
        if (k < nc && side == 0u) {
          d.cl_product[2 * (size_t)k] = f4(pp + pp1, 0.0f);
          d.cl_product[2 * (size_t)k + 1] = f4(pa + pa1, 0.0f);
        }
      }
      __syncthreads();
      // branch_transpose gathered per coordinate (closure.wgsl:222-252, 365-381).
      double dad_part = 0.0;
      for (uint32_t c = threadIdx.x; c < nq; c += blockDim.x) {
        v3 jaxis, janchor;
        joint(c, jaxis, janchor);
        float op = 0.0f;
        for (uint32_t e = d.coord_offset[c]; e < d.coord_offset[c + 1]; ++e) {
          const uint32_t entry = d.coord_entries[e], k = entry >> 1, side = entry & 1u;
          const float4* s = d.cl_state + 6 * (size_t)k;
          const v3 end_anchor = xyz(s[side ? 2 : 0]), end_axis = xyz(s[side ? 3 : 1]);
          const float sign = side ? -1.0f : 1.0f;
          const v3 pd = cross(jaxis, end_anchor - janchor) * sign;
          const v3 ad = cross(jaxis, end_axis) * sign;
          op = op + (dot(pd, xyz(d.cl_product[2 * (size_t)k])) + 1.0f * dot(ad, xyz(d.cl_product[2 * (size_t)k + 1])));
        }
        float* q = d.pcg + 8 * (size_t)c;
        q[4] = op;
        dad_part += (double)direction_of(c) * (double)op;
      }
      const double dad = block_sum(dad_part, scratch);
      const float alpha = fabs(dad) > 1.0e-20 ? (float)(rz / dad) : 0.0f;
      double rz_next_part = 0.0;
      for (uint32_t c = threadIdx.x; c < nq; c += blockDim.x) {
        const float p = direction_of(c);
        float* q = d.pcg + 8 * (size_t)c;
        q[0] = q[0] + alpha * p;
        q[1] = q[1] - alpha * q[4];
        const float df = q[6];
        const float z = df > 1.0e-12f ? q[1] / df : 0.0f;
        set_zp(c, q, z, p);
        rz_next_part += (double)q[1] * (double)z;
      }
      const double rz_next = block_sum(rz_next_part, scratch);  // its barriers publish the update
      beta = fabs(rz) > 1.0e-20 ? (float)(rz_next / rz) : 0.0f;
      rz = rz_next;
    }
    // Apply the clamped Newton step (closure.wgsl:416-426).
    for (uint32_t c = threadIdx.x; c < nq; c += blockDim.x) {
      const float step_c = clampf(kStepScale * d.pcg[8 * (size_t)c], -kMaxStep, kMaxStep);
      if (step_c == step_c) d.coords[c].x = d.coords[c].x + step_c;
      else atomicOr(&diag[kDiagFlags], kFlagInvalid);
    }
  }
}

// The closure Newton loop per closure component (closure_components): one
// workgroup per component runs every Newton step of the tick -- forward
// kinematics of the component's path bodies by pointer jumping (the global
// passes' composition order, so the same link bits), evaluation, the gradient
// and Jacobi diagonal, a PCG with the component's own step lengths and
// tolerance stop, and the clamped step -- and stops once its own closures are
// inside tolerance. Everything the loop reads (bearing rows, bind rotations,
// parents, paths, per-coordinate entry lists) is staged in LDS once, so the
// dependent steps of the loop wait on LDS, not on memory. Components share no
// coordinate, so they run independently; one forward-kinematics pass over
// every tree follows the launch.
struct CompLds {
  float4 *body;   // 5 per path body: (anchor a, direction) (anchor b, parent) (axis a, coordinate) (axis b) (bind rotation)
  float4 *lpos, *lrot;  // 2 * NB each: link position / rotation, two pointer-jumping buffers
  uint2* lmeta;   // 2 * NB: local ancestor, resolved
  float4* clo;    // 4 per closure: (anchor a, end a) (anchor b, end b) (axis a) (axis b) of its bearing
  // PCG vectors in FP64: solution, residual, preconditioned residual, direction, product, diagonal;
  // 6 per closure: J d (position, axis); 6 per path entry: its column times d
  double *px, *pr, *pz, *pp, *pop, *pdg, *kprod, *ptmp;
  float *angle, *jax, *jan, *kend, *kres, *kact;
  float* col;  // 6 per coordinate entry: the Jacobian column (position, axis rows) of that entry
  uint32_t *entry_off, *entries, *path_off, *path, *coord_body;
  uint32_t* pcol;  // per path entry: its coordinate entry (the column it uses)
  // Dense form (cc_dense): J per (closure, coordinate), 6 rows each, and J^T J,
  // entry (l, m) at m * NQ + l, both FP64.
  double *dJ, *dA;
};

__device__ __forceinline__ CompLds carve_comp(float4* base, const Dev& d) {
  const uint32_t NB = d.cc_max_bodies, NQ = d.cc_max_coords, NK = d.cc_max_closures, NP = d.cc_max_path;
  CompLds c;
  c.body = base;
  c.lpos = c.body + 5 * NB;
  c.lrot = c.lpos + 2 * NB;
  c.lmeta = reinterpret_cast<uint2*>(c.lrot + 2 * NB);  // 2 * NB uint2 = NB float4
  c.clo = c.lrot + 3 * NB;
  double* g = reinterpret_cast<double*>(c.clo + 4 * NK);
  c.px = g;
  c.pr = c.px + NQ;
  c.pz = c.pr + NQ;
  c.pp = c.pz + NQ;
  c.pop = c.pp + NQ;
  c.pdg = c.pop + NQ;
  c.kprod = c.pdg + NQ;
  c.ptmp = c.kprod + 6 * NK;
  float* f = reinterpret_cast<float*>(c.ptmp + 6 * NP);
  c.angle = f;
  c.jax = c.angle + NQ;
  c.jan = c.jax + 3 * NQ;
  c.kend = c.jan + 3 * NQ;
  c.kres = c.kend + 12 * NK;
  c.kact = c.kres + 6 * NK;
  c.col = c.kact + NK;
  uint32_t* u = reinterpret_cast<uint32_t*>(c.col + 6 * NP);
  c.entry_off = u;
  c.entries = c.entry_off + NQ + 1;
  c.path_off = c.entries + NP;
  c.path = c.path_off + 2 * NK + 1;
  c.coord_body = c.path + NP;
  c.pcol = c.coord_body + NQ;
  // 16-byte aligned past the lists (base is), as an offset from base so the
  // compiler still sees an LDS pointer and emits LDS instructions for it.
  const size_t used = reinterpret_cast<const char*>(c.pcol + NP) - reinterpret_cast<const char*>(base);
  c.dJ = reinterpret_cast<double*>(reinterpret_cast<char*>(base) + ((used + 15u) & ~size_t(15u)));
  c.dA = c.dJ + 6 * NK * NQ;
  return c;
}

// (x*x' + y*y') + z*z' in FP64 of an FP32 column and an FP64 vector.
__device__ __forceinline__ double dot_d(v3 a, double bx, double by, double bz) {
  return ((double)a.x * bx + (double)a.y * by) + (double)a.z * bz;
}

// kT threads per workgroup: 64 (one wavefront, whose barriers the compiler
// reduces to wavefront barriers) or 256.
template <uint32_t kT>
__global__ __launch_bounds__(kT) void k_closure_components(Dev d) {
  extern __shared__ float4 lds4[];
  __shared__ uint32_t state[2];
  __shared__ double scratch[4];
  const uint4 o0 = d.cc_off[2 * (size_t)blockIdx.x], o1 = d.cc_off[2 * (size_t)blockIdx.x + 1];
  const uint32_t b0 = o0.x, nb = o0.y, q0 = o0.z, nq = o0.w, k0 = o1.x, nk = o1.y;
  const uint32_t NB = d.cc_max_bodies;
  const CompLds L = carve_comp(lds4, d);
  uint32_t* diag = diag_now(d);
  const uint32_t t = threadIdx.x, T = kT;
  // Stage the component's rows and lists.
  for (uint32_t i = t; i < nb; i += T) {
    const MechBody m = load_mbody(d, d.cc_body[b0 + i]);
    const BearingRow r = load_bearing(d, m.tree_bearing);
    float4* b = L.body + 5 * i;
    b[0] = f4(r.anchor_a, __uint_as_float(m.direction));
    b[1] = f4(r.anchor_b, __uint_as_float(d.cc_body_parent[b0 + i]));
    b[2] = f4(r.axis_a, __uint_as_float(d.cc_body_coord[b0 + i]));
    b[3] = f4(r.axis_b, 0.0f);
    b[4] = d.mbodies[4 * (size_t)d.cc_body[b0 + i] + 3];
  }
  for (uint32_t kk = t; kk < nk; kk += T) {
    const BearingRow r = load_bearing(d, d.closure_rows[d.cc_closure[k0 + kk]]);
    float4* c = L.clo + 4 * kk;
    c[0] = f4(r.anchor_a, __uint_as_float(d.cc_end[2 * (size_t)(k0 + kk)]));
    c[1] = f4(r.anchor_b, __uint_as_float(d.cc_end[2 * (size_t)(k0 + kk) + 1]));
    c[2] = f4(r.axis_a, 0.0f);
    c[3] = f4(r.axis_b, 0.0f);
  }
  const uint32_t e_base = d.cc_entry_off[q0], p_base = d.cc_path_off[2 * k0];
  for (uint32_t l = t; l <= nq; l += T) L.entry_off[l] = d.cc_entry_off[q0 + l] - e_base;
  for (uint32_t e = t; e < d.cc_entry_off[q0 + nq] - e_base; e += T) L.entries[e] = d.cc_entries[e_base + e];
  for (uint32_t j = t; j <= 2 * nk; j += T) L.path_off[j] = d.cc_path_off[2 * k0 + j] - p_base;
  for (uint32_t e = t; e < d.cc_path_off[2 * (k0 + nk)] - p_base; e += T) L.path[e] = d.cc_path[p_base + e];
  for (uint32_t l = t; l < nq; l += T) {
    L.angle[l] = d.coords[d.cc_coord[q0 + l]].x;
    L.coord_body[l] = d.cc_coord_body[q0 + l];
  }
  __syncthreads();
  // Each path entry's coordinate entry: the (closure, side) entry of its coordinate.
  const uint32_t np = L.path_off[2 * nk];
  for (uint32_t slot = t; slot < 2 * nk; slot += T)
    for (uint32_t pe = L.path_off[slot]; pe < L.path_off[slot + 1]; ++pe) {
      const uint32_t l = L.path[pe];
      uint32_t e = L.entry_off[l];
      while (e + 1 < L.entry_off[l + 1] && L.entries[e] != slot) ++e;
      L.pcol[pe] = e;
    }
  __syncthreads();
  uint32_t fb = 0;  // buffer holding the final links
  auto link = [&](uint32_t local, float4& position, float4& rotation) {
    if (local == kInvalid) {  // a root: the identity link of prepare_links' root rows
      position = make_float4(0.f, 0.f, 0.f, 0.f);
      rotation = make_float4(0.f, 0.f, 0.f, 1.f);
    } else {
      position = L.lpos[fb * NB + local];
      rotation = L.lrot[fb * NB + local];
    }
  };
  auto joint = [&](uint32_t l, v3& axis, v3& anchor) {
    axis = mk(L.jax[3 * l], L.jax[3 * l + 1], L.jax[3 * l + 2]);
    anchor = mk(L.jan[3 * l], L.jan[3 * l + 1], L.jan[3 * l + 2]);
  };
  auto end_of = [&](uint32_t kk, uint32_t side, v3& anchor, v3& axis) {
    const float* e = L.kend + 12 * kk + (side ? 6 : 0);
    anchor = mk(e[0], e[1], e[2]);
    axis = mk(e[3], e[4], e[5]);
  };
  // Forward kinematics of the path bodies (prepare_links + pointer jumping).
  auto forward_kinematics = [&]() {
    for (uint32_t i = t; i < nb; i += T) {
      const float4* b = L.body + 5 * i;
      const MechBody m{0u, 0u, __float_as_uint(b[0].w), 0u};
      BearingRow r;
      r.anchor_a = xyz(b[0]);
      r.anchor_b = xyz(b[1]);
      r.axis_a = xyz(b[2]);
      r.axis_b = xyz(b[3]);
      const float angle = L.angle[__float_as_uint(b[2].w)];
      // make_link's arithmetic with the staged rows (bind rotation in b[4]).
      float4 rr;
      v3 rp;
      if (m.direction == 0u) {
        rr = quat_multiply(axis_rotation(r.axis_a, angle), b[4]);
        rp = r.anchor_a - quat_rotate(rr, r.anchor_b);
      } else {
        rr = quat_multiply(axis_rotation(r.axis_b, -angle), b[4]);
        rp = r.anchor_b - quat_rotate(rr, r.anchor_a);
      }
      L.lpos[i] = f4(rp, 0.0f);
      L.lrot[i] = rr;
      L.lmeta[i] = make_uint2(__float_as_uint(b[1].w), 0u);
    }
    __syncthreads();
    uint32_t src = 0;
    for (uint32_t round = 0; round < d.jump_rounds; ++round) {
      for (uint32_t i = t; i < nb; i += T) {
        const uint2 meta = L.lmeta[src * NB + i];
        const Link child{L.lpos[src * NB + i], L.lrot[src * NB + i], meta.x, meta.y};
        Link out = child;
        if (child.resolved == 0u) {
          Link parent;
          if (child.ancestor == kInvalid) {
            parent = Link{make_float4(0.f, 0.f, 0.f, 0.f), make_float4(0.f, 0.f, 0.f, 1.f), kInvalid, 1u};
          } else {
            const uint2 pm = L.lmeta[src * NB + child.ancestor];
            parent = Link{L.lpos[src * NB + child.ancestor], L.lrot[src * NB + child.ancestor], pm.x, pm.y};
          }
          out = compose_link(parent, child);
        }
        L.lpos[(src ^ 1u) * NB + i] = out.position;
        L.lrot[(src ^ 1u) * NB + i] = out.rotation;
        L.lmeta[(src ^ 1u) * NB + i] = make_uint2(out.ancestor, out.resolved);
      }
      __syncthreads();
      src ^= 1u;
    }
    fb = src;
  };
  // evaluate_closures for the component's closures; true when all are inside tolerance.
  auto evaluate = [&]() {
    if (t == 0) {
      state[0] = 0u;
      state[1] = 0u;
    }
    __syncthreads();
    for (uint32_t kk = t; kk < nk; kk += T) {
      const float4* c = L.clo + 4 * kk;
      float4 pa_p, pa_r, pb_p, pb_r;
      link(__float_as_uint(c[0].w), pa_p, pa_r);
      link(__float_as_uint(c[1].w), pb_p, pb_r);
      const v3 anchor_a = xyz(pa_p) + quat_rotate(pa_r, xyz(c[0]));
      const v3 anchor_b = xyz(pb_p) + quat_rotate(pb_r, xyz(c[1]));
      const v3 axis_a = normalize_v(quat_rotate(pa_r, xyz(c[2])));
      const v3 axis_b = normalize_v(quat_rotate(pb_r, xyz(c[3])));
      const v3 res_pos = anchor_a - anchor_b, res_axis = axis_a - axis_b;
      float* e = L.kend + 12 * kk;
      e[0] = anchor_a.x; e[1] = anchor_a.y; e[2] = anchor_a.z;
      e[3] = axis_a.x; e[4] = axis_a.y; e[5] = axis_a.z;
      e[6] = anchor_b.x; e[7] = anchor_b.y; e[8] = anchor_b.z;
      e[9] = axis_b.x; e[10] = axis_b.y; e[11] = axis_b.z;
      float* rs = L.kres + 6 * kk;
      rs[0] = res_pos.x; rs[1] = res_pos.y; rs[2] = res_pos.z;
      rs[3] = res_axis.x; rs[4] = res_axis.y; rs[5] = res_axis.z;
      float active = 0.0f;
      if (!finite3(res_pos) || !finite3(res_axis)) {
        atomicOr(&diag[kDiagFlags], kFlagInvalid);
        atomicMax(&state[0], 0xffffffffu);
        atomicMax(&state[1], 0xffffffffu);
      } else {
        const uint32_t um = round_to_u32(length(res_pos) * 1000000.0f);
        const float degrees = acosf(clampf(dot(axis_a, axis_b), -1.0f, 1.0f)) * 57.295779513f;
        const uint32_t ud = round_to_u32(degrees * 1000000.0f);
        atomicMax(&state[0], um);
        atomicMax(&state[1], ud);
        active = (um <= kAnchorTolerance && ud <= kAxisTolerance) ? 0.0f : 1.0f;
      }
      L.kact[kk] = active;
    }
    __syncthreads();
    return !(state[0] > kAnchorTolerance || state[1] > kAxisTolerance);  // the same for every thread
  };
  // Joint axes and anchors of the component's coordinates (closure.wgsl:172-185).
  auto joint_frames = [&]() {
    for (uint32_t l = t; l < nq; l += T) {
      const float4* b = L.body + 5 * L.coord_body[l];
      float4 pp4, pr4;
      link(__float_as_uint(b[1].w), pp4, pr4);
      v3 axis, anchor;
      if (__float_as_uint(b[0].w) == 0u) {
        axis = normalize_v(quat_rotate(pr4, xyz(b[2])));
        anchor = xyz(pp4) + quat_rotate(pr4, xyz(b[0]));
      } else {
        axis = -normalize_v(quat_rotate(pr4, xyz(b[3])));
        anchor = xyz(pp4) + quat_rotate(pr4, xyz(b[1]));
      }
      L.jax[3 * l] = axis.x; L.jax[3 * l + 1] = axis.y; L.jax[3 * l + 2] = axis.z;
      L.jan[3 * l] = anchor.x; L.jan[3 * l + 1] = anchor.y; L.jan[3 * l + 2] = anchor.z;
    }
    __syncthreads();
  };
  const uint32_t slots = 2u * ((nk + 31u) / 32u) * 32u;  // whole lane pairs for the side shuffle
  // The Jacobian columns of every coordinate entry at the current joint frames
  // and closure ends (FP32, as the reference forms them).
  auto columns = [&]() {
    for (uint32_t l = t; l < nq; l += T) {
      v3 jaxis, janchor;
      joint(l, jaxis, janchor);
      for (uint32_t e = L.entry_off[l]; e < L.entry_off[l + 1]; ++e) {
        const uint32_t entry = L.entries[e], kk = entry >> 1, side = entry & 1u;
        v3 end_anchor, end_axis;
        end_of(kk, side, end_anchor, end_axis);
        const float sign = side ? -1.0f : 1.0f;
        const v3 pd = cross(jaxis, end_anchor - janchor) * sign;
        const v3 ad = cross(jaxis, end_axis) * sign;
        float* c = L.col + 6 * e;
        c[0] = pd.x; c[1] = pd.y; c[2] = pd.z;
        c[3] = ad.x; c[4] = ad.y; c[5] = ad.z;
      }
    }
    __syncthreads();
  };
  // branch_product per closure side, summed in path order (FP32 columns times the
  // FP64 vector, FP64 sums); side 0 adds side 1's. Result: J v per closure in kprod.
  // The products of all path entries are formed in parallel, the sums in path order.
  auto product = [&](auto vec_of) {
    for (uint32_t pe = t; pe < np; pe += T) {
      const float* c = L.col + 6 * L.pcol[pe];
      const double dir = vec_of(L.path[pe]);
      double* o = L.ptmp + 6 * pe;
      o[0] = (double)c[0] * dir; o[1] = (double)c[1] * dir; o[2] = (double)c[2] * dir;
      o[3] = (double)c[3] * dir; o[4] = (double)c[4] * dir; o[5] = (double)c[5] * dir;
    }
    __syncthreads();
    for (uint32_t slot = t; slot < slots; slot += T) {
      const uint32_t kk = slot >> 1, side = slot & 1u;
      double spx = 0.0, spy = 0.0, spz = 0.0, sax = 0.0, say = 0.0, saz = 0.0;
      if (kk < nk) {
        for (uint32_t pe = L.path_off[2 * kk + side]; pe < L.path_off[2 * kk + side + 1]; ++pe) {
          const double* o = L.ptmp + 6 * pe;
          spx = spx + o[0];
          spy = spy + o[1];
          spz = spz + o[2];
          sax = sax + o[3];
          say = say + o[4];
          saz = saz + o[5];
        }
      }
      const double s1[6] = {__shfl_xor(spx, 1), __shfl_xor(spy, 1), __shfl_xor(spz, 1),
                            __shfl_xor(sax, 1), __shfl_xor(say, 1), __shfl_xor(saz, 1)};
      if (kk < nk && side == 0u) {
        double* q = L.kprod + 6 * kk;
        q[0] = spx + s1[0]; q[1] = spy + s1[1]; q[2] = spz + s1[2];
        q[3] = sax + s1[3]; q[4] = say + s1[4]; q[5] = saz + s1[5];
      }
    }
    __syncthreads();
  };
  // branch_transpose of kprod gathered for coordinate l (its entries in (closure, side) order).
  auto transpose = [&](uint32_t l) {
    double op = 0.0;
    for (uint32_t e = L.entry_off[l]; e < L.entry_off[l + 1]; ++e) {
      const uint32_t kk = L.entries[e] >> 1;
      const float* c = L.col + 6 * e;
      const double* q = L.kprod + 6 * kk;
      op = op + (dot_d(mk(c[0], c[1], c[2]), q[0], q[1], q[2]) + 1.0 * dot_d(mk(c[3], c[4], c[5]), q[3], q[4], q[5]));
    }
    return op;
  };
  // J^T J of the current columns, formed once per Newton step for the PCG of a
  // one-wavefront component: J per (closure, coordinate) is the sum of the
  // coordinate's side columns of that closure (in (closure, side) order), and
  // entry (l, m) the 6-row dot of J's columns l and m summed over the closures.
  // The matrix-free product and transpose apply the same operator; this one
  // costs one row of nq products per coordinate an iteration.
  const bool dense = kT == 64 && d.cc_dense != 0u;
  const uint32_t NQ = d.cc_max_coords;
  auto form_dense = [&]() {
    for (uint32_t i = t; i < 6 * nk * NQ; i += T) L.dJ[i] = 0.0;
    __syncthreads();
    for (uint32_t l = t; l < nq; l += T)
      for (uint32_t e = L.entry_off[l]; e < L.entry_off[l + 1]; ++e) {
        const float* c = L.col + 6 * e;
        double* j = L.dJ + 6 * ((L.entries[e] >> 1) * NQ + l);
#pragma unroll
        for (int r = 0; r < 6; ++r) j[r] = j[r] + (double)c[r];
      }
    __syncthreads();
    // One entry of the lower triangle per thread and step, stored at (l, m) and
    // (m, l): the products commute, so the two entries are the same bits.
    const uint32_t pairs = nq * (nq + 1u) / 2u;
    for (uint32_t pr = t; pr < pairs; pr += T) {
      uint32_t l = (uint32_t)((sqrtf(8.0f * (float)pr + 1.0f) - 1.0f) * 0.5f);  // pr = l (l + 1) / 2 + m, m <= l
      while (l * (l + 1u) / 2u > pr) --l;
      while ((l + 1u) * (l + 2u) / 2u <= pr) ++l;
      const uint32_t m = pr - l * (l + 1u) / 2u;
      double a = 0.0;
      for (uint32_t kk = 0; kk < nk; ++kk) {
        const double* jl = L.dJ + 6 * (kk * NQ + l);
        const double* jm = L.dJ + 6 * (kk * NQ + m);
        a = a + ((((((jl[0] * jm[0]) + jl[1] * jm[1]) + jl[2] * jm[2]) + jl[3] * jm[3]) + jl[4] * jm[4]) +
                 jl[5] * jm[5]);
      }
      L.dA[m * NQ + l] = a;
      L.dA[l * NQ + m] = a;
    }
    __syncthreads();
  };
  // Jacobi-preconditioned CG on J^T J x = r from x = 0, with px, pr, pz, pp, pdg
  // set up by the caller and rz = r.z; stops at closure_iters iterations or when
  // r.z falls below closure_tol^2 of its start.
  // With the dense J^T J (one wavefront, nq <= 64) lane l keeps coordinate l's
  // vectors in registers; only the direction is shared, through LDS, for the
  // product, whose row sums run in coordinate order as in the general form.
  auto pcg_wave = [&](auto max_row, double rz) {
    // kRow > 0: lane l holds row l of J^T J in registers (nq <= kRow; rows
    // past nq are zeros, whose products add nothing: the sum never is -0.0)
    // and takes each direction entry from its lane. kRow == 0: rows and the
    // direction are read from LDS.
    constexpr uint32_t kRow = decltype(max_row)::value;
    const bool own = t < nq;
    double x = 0.0, r = 0.0, z = 0.0, p = 0.0, dg = 0.0;
    if (own) {
      r = L.pr[t];
      z = L.pz[t];
      p = L.pp[t];
      dg = L.pdg[t];
    }
    rz = uniform_f64(rz);
    const double rz_stop = (double)d.closure_tol * (double)d.closure_tol * rz;
    double beta = 0.0;
    double* shared_dir = L.pop;
    const double* row = L.dA + t;
    double arow[kRow > 0 ? kRow : 1];
    if constexpr (kRow > 0) {
#pragma unroll
      for (uint32_t m = 0; m < kRow; ++m) arow[m] = own && m < nq ? row[m * NQ] : 0.0;
    }
    for (uint32_t it = 0; it < d.closure_iters; ++it) {
      if (d.closure_tol > 0.0f && rz <= rz_stop) break;
      const double dir = it == 0 ? p : z + beta * p;
      double op = 0.0;
      if constexpr (kRow > 0) {
#pragma unroll
        for (uint32_t m = 0; m < kRow; ++m) op = op + arow[m] * readlane_f64(dir, (int)m);
      } else {
        if (own) shared_dir[t] = dir;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (own) {
          uint32_t m = 0;
          for (; m + 4u <= nq; m += 4u) {
            const double a0 = row[m * NQ], a1 = row[(m + 1u) * NQ], a2 = row[(m + 2u) * NQ], a3 = row[(m + 3u) * NQ];
            const double d0 = shared_dir[m], d1 = shared_dir[m + 1u], d2 = shared_dir[m + 2u], d3 = shared_dir[m + 3u];
            op = op + a0 * d0;
            op = op + a1 * d1;
            op = op + a2 * d2;
            op = op + a3 * d3;
          }
          for (; m < nq; ++m) op = op + row[m * NQ] * shared_dir[m];
        }
      }
      double dad_part = 0.0;
      if (own) dad_part += dir * op;
      const double dad = wave_sum_rows(dad_part);
      const double alpha = fabs(dad) > 1.0e-20 ? rz / dad : 0.0;
      double rz_next_part = 0.0;
      if (own) {
        x = x + alpha * dir;
        r = r - alpha * op;
        z = dg > (double)1.0e-12f ? r / dg : 0.0;
        p = dir;
        rz_next_part += r * z;
      }
      const double rz_next = wave_sum_rows(rz_next_part);
      beta = fabs(rz) > 1.0e-20 ? rz_next / rz : 0.0;
      rz = rz_next;
      if constexpr (kRow == 0) {
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();  // every lane has read the direction before the next is stored
      }
    }
    if (own) L.px[t] = x;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
  };
  auto pcg = [&](double rz) {
    if constexpr (kT == 64) {
      if (dense) {
        if (NQ <= 8u) pcg_wave(std::integral_constant<uint32_t, 8>{}, rz);
        else if (NQ <= 16u) pcg_wave(std::integral_constant<uint32_t, 16>{}, rz);
        else if (NQ <= 24u) pcg_wave(std::integral_constant<uint32_t, 24>{}, rz);
        else if (NQ <= 32u) pcg_wave(std::integral_constant<uint32_t, 32>{}, rz);
        else pcg_wave(std::integral_constant<uint32_t, 0>{}, rz);
        return;
      }
    }
    const double rz_stop = (double)d.closure_tol * (double)d.closure_tol * rz;
    double beta = 0.0;
    for (uint32_t it = 0; it < d.closure_iters; ++it) {
      if (d.closure_tol > 0.0f && rz <= rz_stop) break;
      auto direction_of = [&](uint32_t l) { return it == 0 ? L.pp[l] : L.pz[l] + beta * L.pp[l]; };
      double dad_part = 0.0;
      product(direction_of);
      for (uint32_t l = t; l < nq; l += T) {
        const double op = transpose(l);
        L.pop[l] = op;
        dad_part += direction_of(l) * op;
      }
      const double dad = group_sum<kT>(dad_part, scratch);
      const double alpha = fabs(dad) > 1.0e-20 ? rz / dad : 0.0;
      double rz_next_part = 0.0;
      for (uint32_t l = t; l < nq; l += T) {
        const double p = direction_of(l);
        L.px[l] = L.px[l] + alpha * p;
        L.pr[l] = L.pr[l] - alpha * L.pop[l];
        const double df = L.pdg[l];
        const double z = df > (double)1.0e-12f ? L.pr[l] / df : 0.0;
        L.pz[l] = z;
        L.pp[l] = p;
        rz_next_part += L.pr[l] * z;
      }
      const double rz_next = group_sum<kT>(rz_next_part, scratch);  // its barriers publish the update
      beta = fabs(rz) > 1.0e-20 ? rz_next / rz : 0.0;
      rz = rz_next;
    }
  };
  // The PCG's start from the right-hand side -g and diagonal dg of one coordinate.
  auto pcg_start = [&](uint32_t l, double g, double dg) {
    const double residual = dg > (double)1.0e-12f ? -g : 0.0;
    const double pre = dg > (double)1.0e-12f ? residual / dg : 0.0;
    L.px[l] = 0.0;
    L.pr[l] = residual;
    L.pz[l] = pre;
    L.pp[l] = pre;
    L.pop[l] = 0.0;
    L.pdg[l] = dg;
    return residual * pre;
  };
  bool fresh = false;  // the links hold forward kinematics of L.angle
  for (uint32_t step = 0; step < d.closure_steps; ++step) {
    forward_kinematics();
    fresh = true;
    if (evaluate()) break;
    joint_frames();
    columns();
    if (dense) form_dense();
    // Gradient and Jacobi diagonal, gathered per coordinate in (closure, side) order.
    double rz_part = 0.0;
    for (uint32_t l = t; l < nq; l += T) {
      double g = 0.0, dg = 0.0;
      for (uint32_t e = L.entry_off[l]; e < L.entry_off[l + 1]; ++e) {
        const uint32_t kk = L.entries[e] >> 1;
        if (L.kact[kk] == 0.0f) continue;
        const float* c = L.col + 6 * e;
        const v3 pd = mk(c[0], c[1], c[2]), ad = mk(c[3], c[4], c[5]);
        const float* rs = L.kres + 6 * kk;
        const float gradient = dot(mk(rs[0], rs[1], rs[2]), pd) + 1.0f * dot(mk(rs[3], rs[4], rs[5]), ad);
        const float diagonal = dot(pd, pd) + 1.0f * dot(ad, ad);
        if (gradient == gradient && diagonal == diagonal) {
          g += (double)gradient;
          dg += (double)diagonal;
        } else {
          atomicOr(&diag[kDiagFlags], kFlagInvalid);
        }
      }
      rz_part += pcg_start(l, g, dg);
    }
    pcg(group_sum<kT>(rz_part, scratch));
    // The clamped Newton step (closure.wgsl:416-426), in FP32.
    for (uint32_t l = t; l < nq; l += T) {
      const float step_c = clampf(kStepScale * (float)L.px[l], -kMaxStep, kMaxStep);
      if (step_c == step_c) L.angle[l] = L.angle[l] + step_c;
      else atomicOr(&diag[kDiagFlags], kFlagInvalid);
    }
    __syncthreads();
    fresh = false;
  }
  if (d.closure_vproj) {
    // Velocity-level closure: the coordinate speeds the tree reconstruction will
    // use, projected onto J v = 0 at the final configuration (every closure of
    // the component, inside tolerance or not): v += x with J^T J x = -J^T (J v),
    // the minimum-norm correction.
    if (!fresh) {
      forward_kinematics();
      evaluate();
    }
    joint_frames();
    columns();
    if (dense) form_dense();
    for (uint32_t l = t; l < nq; l += T) L.pz[l] = (double)d.coords[d.cc_coord[q0 + l]].y;
    __syncthreads();
    product([&](uint32_t l) { return L.pz[l]; });
    double rz_part = 0.0;
    for (uint32_t l = t; l < nq; l += T) {
      double dg = 0.0;
      for (uint32_t e = L.entry_off[l]; e < L.entry_off[l + 1]; ++e) {
        const float* c = L.col + 6 * e;
        const v3 pd = mk(c[0], c[1], c[2]), ad = mk(c[3], c[4], c[5]);
        dg += (double)(dot(pd, pd) + 1.0f * dot(ad, ad));
      }
      rz_part += pcg_start(l, transpose(l), dg);  // pz[l] is read by no other thread any more
    }
    pcg(group_sum<kT>(rz_part, scratch));
    for (uint32_t l = t; l < nq; l += T) {
      float2* c = &d.coords[d.cc_coord[q0 + l]];
      const float speed = (float)((double)c->y + L.px[l]);
      if (speed == speed) c->y = speed;
      else atomicOr(&diag[kDiagFlags], kFlagInvalid);
    }
  }
  for (uint32_t l = t; l < nq; l += T) d.coords[d.cc_coord[q0 + l]].x = L.angle[l];
}

// Pose publication and the parent-to-child velocity reconstruction of
// encode_mechanism_passes (device.rs:963-1048), in one workgroup.
__global__ __launch_bounds__(kPoseThreads) void k_publish(Dev d) {
  publish(d, d.links[d.jump_rounds & 1u]);
  reconstruct_levels(d);
}

__global__ __launch_bounds__(kPoseThreads) void k_reconstruct(Dev d) { reconstruct_levels(d); }

// k_publish per forward-kinematics group (whole trees, at most 256 bodies, one
// per thread): the same per-body arithmetic as publish and reconstruct_body,
// with the group's published rotations and reconstructed velocities kept in
// LDS for the children that read them, so a depth level costs no round trip
// to memory. A parent outside the group is a root, read from the state arrays.
__global__ __launch_bounds__(256) void k_publish_groups(Dev d) {
  __shared__ float4 s_rot[256], s_lin[256], s_ang[256];
  const uint32_t begin = d.fk_offset[blockIdx.x], n = d.fk_offset[blockIdx.x + 1] - begin;
  const uint32_t t = threadIdx.x;
  const bool own = t < n;
  uint32_t body = 0, meta = 0;
  MechBody m{};
  BearingRow r{};
  if (own) {
    body = d.fk_list[begin + t];
    meta = d.fk_meta[begin + t];
    m = load_mbody(d, body);
    r = load_bearing(d, m.tree_bearing);
    // publish_a / publish_b (mechanism.wgsl:149-185).
    const Link l = load_link(d.links[d.jump_rounds & 1u], body);
    const uint32_t root = l.ancestor;
    float4 rot = d.rot[body];
    if (body != root) {
      const float4 root_rotation = d.rot[root];
      d.pos[body] = f4(xyz(d.pos[root]) + quat_rotate(root_rotation, xyz(l.position)), 0.0f);
      rot = normalize_q(quat_multiply(root_rotation, l.rotation));
      d.rot[body] = rot;
    }
    s_rot[t] = rot;
  }
  const uint32_t depth = meta >> 16, pl = meta & 0xffffu;
  const bool local = pl != 0xffffu;
  // What this kernel does not write, loaded before the level loop: the
  // coordinate speed, and a root parent's rotation and velocities.
  float speed = 0.0f;
  float4 root_rot = make_float4(0.f, 0.f, 0.f, 1.f);
  v3 root_ang = mk(0.0f, 0.0f, 0.0f), root_lin = root_ang;
  if (own) {
    speed = d.coords[r.coordinate].y;
    if (!local) {
      root_rot = d.rot[m.parent];
      root_ang = xyz(d.ang[m.parent]);
      root_lin = xyz(d.lin[m.parent]);
    }
  }
  __syncthreads();
  // reconstruct_body_velocities (articulated.wgsl:423-456): the joint axis and
  // both anchors need only the published rotations, so every body forms them
  // at once; the velocities then follow the tree one depth level at a time.
  v3 axis = mk(0.0f, 0.0f, 0.0f), parent_anchor = axis, child_anchor = axis;
  if (own) {
    const float4 rp = local ? s_rot[pl] : root_rot, rb = s_rot[t];
    if (m.direction != 0u) {
      axis = -normalize_v(quat_rotate(rp, r.axis_b));
      parent_anchor = quat_rotate(rp, r.anchor_b);
      child_anchor = quat_rotate(rb, r.anchor_a);
    } else {
      axis = normalize_v(quat_rotate(rp, r.axis_a));
      parent_anchor = quat_rotate(rp, r.anchor_a);
      child_anchor = quat_rotate(rb, r.anchor_b);
    }
  }
  for (uint32_t level = 1; level <= d.level_count; ++level) {
    if (own && depth == level) {
      const v3 parent_ang = local ? xyz(s_ang[pl]) : root_ang;
      const v3 parent_lin = local ? xyz(s_lin[pl]) : root_lin;
      const v3 angular = parent_ang + axis * speed;
      const v3 anchor_velocity = parent_lin + cross(parent_ang, parent_anchor);
      const v3 linear = anchor_velocity - cross(angular, child_anchor);
      d.lin[body] = f4(linear, 0.0f);
      d.ang[body] = f4(angular, 0.0f);
      s_lin[t] = f4(linear, 0.0f);
      s_ang[t] = f4(angular, 0.0f);
    }
    __syncthreads();
  }
}

void launch_publish(hipStream_t s, const Dev& d) {
  if (d.fk_small) k_publish_groups<<<d.fk_groups, 256, 0, s>>>(d);
  else k_publish<<<1, kPoseThreads, 0, s>>>(d);
}

// update_world_masses (collision.wgsl:144-162) and the solver record, for ticks
// whose poses and velocities change after integration.
__device__ __forceinline__ v3 world_inverse_column(const float4* m, float4 rotation, v3 axis) {
  const v3 local = quat_rotate(quat_conj(rotation), axis);
  const v3 r = (xyz(m[1]) * local.x + xyz(m[2]) * local.y) + xyz(m[3]) * local.z;
  return quat_rotate(rotation, r);
}

__global__ __launch_bounds__(256) void k_world_masses(Dev d) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.body_count) return;
  const float4 rotation = d.rot[i];
  const float4* m = d.masses + 4 * (size_t)i;
  const v3 cx = world_inverse_column(m, rotation, mk(1.0f, 0.0f, 0.0f));
  const v3 cy = world_inverse_column(m, rotation, mk(0.0f, 1.0f, 0.0f));
  const v3 cz = world_inverse_column(m, rotation, mk(0.0f, 0.0f, 1.0f));
  float4* w = d.world_masses + 3 * (size_t)i;
  w[0] = f4(cx, m[0].x);
  w[1] = f4(cy, 0.0f);
  w[2] = f4(cz, 0.0f);
  const float4 linear = d.lin[i], angular = d.ang[i];
  float4* sb = d.sbody + 4 * (size_t)i;
  sb[0] = make_float4(linear.x, linear.y, linear.z, m[0].x);
  sb[1] = make_float4(angular.x, angular.y, angular.z, cx.x);
  sb[2] = make_float4(cx.y, cx.z, cy.x, cy.y);
  sb[3] = make_float4(cy.z, cz.x, cz.y, cz.z);
  d.slot_count[i] = 0u;
}

// validate_articulated_state (articulated.wgsl:458-484) fused with
// publish_snapshot (snapshot.wgsl:6-14).
__global__ __launch_bounds__(256) void k_validate_snapshot(Dev d) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.body_count) return;
  const uint64_t tick = tick_m(d);
  const float4 position = d.pos[i], rotation = d.rot[i];
  const float rotation_length = dot4(rotation, rotation);
  bool valid = finite4(position) && finite4(rotation) && finite4(d.lin[i]) && finite4(d.ang[i]) &&
               rotation_length > 0.5f && rotation_length < 1.5f;
  const MechBody m = load_mbody(d, i);
  if (m.root == 0u) {
    const float2 s = d.coords[load_bearing(d, m.tree_bearing).coordinate];
    valid = valid && s.x == s.x && fabsf(s.x) < 3.402823e+38f && s.y == s.y && fabsf(s.y) < 3.402823e+38f;
  }
  if (!valid) atomicOr(&d.diag[tick & 1][kDiagFlags], kFlagInvalid);
  const uint32_t slot = (uint32_t)(tick % 3u);
  d.snap_pos[slot][i] = position;
  d.snap_rot[slot][i] = rotation;
}

// record_residual (bearings.wgsl:53-72).
// Anchor residual in micrometres and axis residual in micro-degrees
// (bearings.wgsl:48-72).
__device__ __forceinline__ void residual_words(v3 anchor_a, v3 anchor_b, v3 axis_a, v3 axis_b, uint32_t& um,
                                               uint32_t& ud) {
  um = round_to_u32(length(anchor_a - anchor_b) * 1000000.0f);
  const float degrees = atan2f(length(cross(axis_a, axis_b)), clampf(dot(axis_a, axis_b), -1.0f, 1.0f)) * 57.295779513f;
  ud = round_to_u32(degrees * 1000000.0f);
}


// validate_mechanism_bearings / validate_bearings (bearings.wgsl:74-102).
// Each wavefront takes the maxima of its rows before one atomic per word (the
// maxima and the flag are order-independent): atomics on one address serialise
// device-wide.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const uint32_t o = (uint32_t)__shfl_xor((int)v, m);
    v = o > v ? o : v;
  }
  return v;
}

__global__ __launch_bounds__(256) void k_validate_bearings(Dev d, int mechanism) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if ((i & ~63u) >= d.bearing_count) return;  // whole wavefront past the last row
  uint32_t* diag = diag_now(d);
  uint32_t um = 0u, ud = 0u;
  bool open = false;
  if (i < d.bearing_count) {
  const BearingRow r = load_bearing(d, i);
  float4 pos_a, rot_a, pos_b, rot_b;
  if (mechanism) {
    const float4* links = d.links[d.jump_rounds & 1u];
    const Link la = load_link(links, r.a), lb = load_link(links, r.b);
    pos_a = la.position;
    rot_a = la.rotation;
    pos_b = lb.position;
    rot_b = lb.rotation;
  } else {
    pos_a = d.pos[r.a];
    rot_a = d.rot[r.a];
    pos_b = d.pos[r.b];
    rot_b = d.rot[r.b];
  }
  const v3 anchor_a = xyz(pos_a) + quat_rotate(rot_a, r.anchor_a);
  const v3 anchor_b = xyz(pos_b) + quat_rotate(rot_b, r.anchor_b);
  const v3 axis_a = normalize_v(quat_rotate(rot_a, r.axis_a));
  const v3 axis_b = normalize_v(quat_rotate(rot_b, r.axis_b));
  residual_words(anchor_a, anchor_b, axis_a, axis_b, um, ud);
  open = r.closure != 0u && (um > kAnchorTolerance || ud > kAxisTolerance);
  }
  um = wave_max_u32(um);
  ud = wave_max_u32(ud);
  const bool any_open = __ballot(open) != 0ull;
  if ((threadIdx.x & 63u) == 0u) {
    atomicMax(&diag[kDiagAnchor], um);
    atomicMax(&diag[kDiagAxis], ud);
    if (any_open) atomicOr(&diag[kDiagFlags], kFlagNonConvergence);
  }
}

inline uint32_t blocks_for(uint32_t n) { return (n + 255u) / 256u; }

// ---------------------------------------------------------------------------
// Serial small-mechanism solvers (device.rs:19-20, 1370-1416). With a dynamic
// mechanism root and at most 64 bearings in the scene, the passes after the
// parallel warm start are 12x as many immediate Gauss-Seidel sweeps: the active
// contacts forward then backward in their sort order (solve_accumulate_serial,
// collision.wgsl:979-1048), then the bearing rows forward then backward
// (project_bearing_velocities_serial, articulated.wgsl:277-285). The sweeps are
// one thread's work by definition; here they run as ONE launch for all passes
// (the reference dispatches two per pass) on the solver records, which hold
// the velocities while the contact solve runs.
// ---------------------------------------------------------------------------

// Sort key of an active contact: (body A, body B, manifold slot) as the
// reference selects them (collision.wgsl:991-1017), then the collider pair. The
// slot is where acquire_manifold's probe starts, pair_hash(low, high) &
// (capacity - 1) (collision.wgsl:499-504): the reference's slot whenever its
// table holds no colliding entry. Above 64 active contacts the reference sweeps
// in append order, which its atomics leave unspecified; the sorted order is
// used there too.
__global__ __launch_bounds__(256) void k_serial_keys(Dev d) {
  const uint64_t tick = tick_m(d);
  const uint32_t n = read_count(&d.diag[tick & 1][kDiagActive], d.pair_capacity);
  const float4* mrows = d.mrows[tick & 1];
  for (uint32_t ai = blockIdx.x * blockDim.x + threadIdx.x; ai < n; ai += gridDim.x * blockDim.x) {
    const uint2 bodies = row_body_pair(d, ai);
    const float4 m0 = mrows[3 * (size_t)d.active[ai]];
    const uint32_t low = __float_as_uint(m0.x), high = __float_as_uint(m0.y);
    d.serial_keys[2 * (size_t)ai] = make_uint4(bodies.x, bodies.y, pair_hash(low, high) & d.hash_mask, low);
    d.serial_keys[2 * (size_t)ai + 1] = make_uint4(high, ai, 0u, 0u);
  }
}

__device__ __forceinline__ bool key_less(uint4 x, uint32_t xh, uint4 y, uint32_t yh) {
  if (x.x != y.x) return x.x < y.x;
  if (x.y != y.y) return x.y < y.y;
  if (x.z != y.z) return x.z < y.z;
  if (x.w != y.w) return x.w < y.w;
  return xh < yh;
}

// Rank of every key among all keys (the keys are distinct: one contact per
// collider pair), so serial_order is the sorted permutation. Keys stream
// through LDS in tiles of 256.
__global__ __launch_bounds__(256) void k_serial_rank(Dev d) {
  __shared__ uint4 tile_key[256];
  __shared__ uint32_t tile_high[256];
  const uint64_t tick = tick_m(d);
  const uint32_t n = read_count(&d.diag[tick & 1][kDiagActive], d.pair_capacity);
  for (uint32_t base = blockIdx.x * 256u; base < n; base += gridDim.x * 256u) {
    const uint32_t i = base + threadIdx.x;
    const bool valid = i < n;
    uint4 mine = make_uint4(0u, 0u, 0u, 0u);
    uint32_t mine_high = 0u;
    if (valid) {
      mine = d.serial_keys[2 * (size_t)i];
      mine_high = d.serial_keys[2 * (size_t)i + 1].x;
    }
    uint32_t rank = 0;
    for (uint32_t j0 = 0; j0 < n; j0 += 256u) {
      __syncthreads();
      if (j0 + threadIdx.x < n) {
        tile_key[threadIdx.x] = d.serial_keys[2 * (size_t)(j0 + threadIdx.x)];
        tile_high[threadIdx.x] = d.serial_keys[2 * (size_t)(j0 + threadIdx.x) + 1].x;
      }
      __syncthreads();
      const uint32_t m = n - j0 < 256u ? n - j0 : 256u;
      for (uint32_t t = 0; t < m; ++t) rank += key_less(tile_key[t], tile_high[t], mine, mine_high) ? 1u : 0u;
    }
    if (valid) d.serial_order[rank] = i;
  }
}

__device__ __forceinline__ void store_velocity(const Dev& d, uint32_t b, v3 lin, v3 ang) {
  float4* r = d.sbody + 4 * (size_t)b;
  r[0] = f4(lin, r[0].w);
  r[1] = f4(ang, r[1].w);
}

// solve_contact_immediate (collision.wgsl:908-977) on active contact `ai`: the
// full projected impulse (no relaxation) against the current velocities,
// applied at once.
__device__ void contact_immediate(const Dev& d, uint32_t ai) {
  const SolverRow sr = load_row(d, ai);
  const uint32_t a = sr.body_a, b = sr.body_b;
  const bool has_b = b != kInvalid;
  const v3 normal = sr.normal, arm_a = sr.arm_a, arm_b = sr.arm_b;
  const float penetration = sr.penetration, denominator = sr.denominator;
  const SolverBody sa = load_body(d.sbody, a);
  const SolverBody sb = has_b ? load_body(d.sbody, b) : sa;
  const v3 velocity_b = has_b ? sb.lin + cross(sb.ang, arm_b) : mk(0.0f, 0.0f, 0.0f);
  if (denominator <= 0.0f) return;
  const v3 relative = velocity_b - (sa.lin + cross(sa.ang, arm_a));
  const float normal_speed = dot(relative, normal);
  const float lambda = d.lambda[ai];
  if (penetration <= 1.0e-5f && normal_speed >= -1.0e-5f && lambda <= 1.0e-6f) {
    d.lambda[ai] = 0.0f;
    return;
  }
  const float previous = fmx(lambda, 0.0f);
  const float recovery = sr.analytic ? 1.0f : 0.2f;
  const float bias = fmn((fmx(penetration - kSlop, 0.0f) * recovery) / d.delta_seconds, kMaxCorrection);
  const float accumulated = fmx(previous + (-normal_speed + bias) / denominator, 0.0f);
  const float delta = accumulated - previous;
  d.lambda[ai] = accumulated;
  v3 impulse = normal * delta;
  const v3 tangent_velocity = relative - normal * normal_speed;
  const float tangent_speed = length(tangent_velocity);
  if (tangent_speed > 1.0e-6f) {
    const v3 tangent = tangent_velocity / tangent_speed;
    const float tangent_denominator = denom_of(sa, sb, has_b, arm_a, arm_b, tangent);
    const float friction_impulse = fmn(tangent_speed / tangent_denominator, fabsf(delta) * d.friction);
    impulse = impulse - tangent * friction_impulse;
  }
  store_velocity(d, a, sa.lin - impulse * sa.inv_mass, sa.ang + inv_inertia(sa, cross(arm_a, -impulse)));
  if (has_b) store_velocity(d, b, sb.lin + impulse * sb.inv_mass, sb.ang + inv_inertia(sb, cross(arm_b, impulse)));
}

// solve_linear_axis_immediate (articulated.wgsl:157-191).
__device__ void linear_immediate(const Dev& d, uint32_t a, uint32_t b, v3 arm_a, v3 arm_b, v3 direction) {
  const v3 velocity_a = lin_of(d, a, 1) + cross(ang_of(d, a, 1), arm_a);
  const v3 velocity_b = lin_of(d, b, 1) + cross(ang_of(d, b, 1), arm_b);
  const v3 relative = velocity_b - velocity_a;
  const v3 angular_a = cross(arm_a, direction);
  const v3 angular_b = cross(arm_b, direction);
  const float denominator =
      ((inv_mass_of(d, a) + inv_mass_of(d, b)) + dot(angular_a, wii(d, a, angular_a))) + dot(angular_b, wii(d, b, angular_b));
  if (denominator <= 1.0e-12f) return;
  const v3 impulse = direction * (dot(relative, direction) / denominator);
  store_velocity(d, a, lin_of(d, a, 1) + impulse * inv_mass_of(d, a), ang_of(d, a, 1) + wii(d, a, cross(arm_a, impulse)));
  store_velocity(d, b, lin_of(d, b, 1) - impulse * inv_mass_of(d, b),
                 ang_of(d, b, 1) + wii(d, b, cross(arm_b, -impulse)));
}

// solve_angular_axis_immediate (articulated.wgsl:193-214).
__device__ void angular_immediate(const Dev& d, uint32_t a, uint32_t b, v3 axis) {
  const v3 relative = ang_of(d, b, 1) - ang_of(d, a, 1);
  const v3 inverse_a = wii(d, a, axis);
  const v3 inverse_b = wii(d, b, axis);
  const float denominator = dot(axis, inverse_a + inverse_b);
  if (denominator <= 1.0e-12f) return;
  const float impulse = dot(relative, axis) / denominator;
  store_velocity(d, a, lin_of(d, a, 1), ang_of(d, a, 1) + inverse_a * impulse);
  store_velocity(d, b, lin_of(d, b, 1), ang_of(d, b, 1) - inverse_b * impulse);
}

// project_bearing_velocity_row (articulated.wgsl:216-276).
__device__ void bearing_row_immediate(const Dev& d, uint32_t i) {
  const BearingRow r = load_bearing(d, i);
  const float4 rot_a = d.rot[r.a], rot_b = d.rot[r.b];
  const v3 arm_a = quat_rotate(rot_a, r.anchor_a);
  const v3 arm_b = quat_rotate(rot_b, r.anchor_b);
  linear_immediate(d, r.a, r.b, arm_a, arm_b, mk(1.0f, 0.0f, 0.0f));
  linear_immediate(d, r.a, r.b, arm_a, arm_b, mk(0.0f, 1.0f, 0.0f));
  linear_immediate(d, r.a, r.b, arm_a, arm_b, mk(0.0f, 0.0f, 1.0f));
  const v3 axis_a = normalize_v(quat_rotate(rot_a, r.axis_a));
  const v3 axis_b = normalize_v(quat_rotate(rot_b, r.axis_b));
  const v3 hinge = normalize_v(axis_a + axis_b);
  const v3 helper = fabsf(hinge.x) > 0.8f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
  const v3 tangent_a = normalize_v(cross(hinge, helper));
  const v3 tangent_b = cross(hinge, tangent_a);
  angular_immediate(d, r.a, r.b, tangent_a);
  angular_immediate(d, r.a, r.b, tangent_b);
}

__device__ void bearing_sweep(const Dev& d) {
  for (uint32_t i = 0; i < d.bearing_count; ++i) bearing_row_immediate(d, i);
  for (uint32_t i = d.bearing_count; i > 0; --i) bearing_row_immediate(d, i - 1);
}

// All serial passes of one tick in one thread. Without an active contact the
// reference skips its bearing sweeps (indirect arguments 12..14,
// collision.wgsl:747) and its contact sweeps have nothing to do.
__global__ __launch_bounds__(64) void k_serial_passes(Dev d, uint32_t passes) {
  if (threadIdx.x != 0) return;
  const uint64_t tick = tick_m(d);
  const uint32_t n = read_count(&d.diag[tick & 1][kDiagActive], d.pair_capacity);
  if (n == 0) return;
  bearing_sweep(d);
  for (uint32_t p = 1; p < passes; ++p) {
    for (uint32_t k = 0; k < n; ++k) contact_immediate(d, d.serial_order[k]);
    for (uint32_t k = n; k > 0; --k) contact_immediate(d, d.serial_order[k - 1]);
    bearing_sweep(d);
  }
}

// The end of the last pass: accumulated impulses into this tick's manifold rows
// (persist_contacts, collision.wgsl:1071-1088) and the records' velocities into
// the state.
__global__ __launch_bounds__(256) void k_serial_finish(Dev d) {
  const uint64_t tick = tick_m(d);
  const uint32_t n = read_count(&d.diag[tick & 1][kDiagActive], d.pair_capacity);
  if (n == 0) return;
  float4* rows_now = d.mrows[tick & 1];
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t ai = blockIdx.x * blockDim.x + threadIdx.x; ai < n; ai += stride) {
    rows_now[3 * (size_t)d.active[ai] + 2].w = d.lambda[ai];
  }
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < d.body_count; b += stride) {
    d.lin[b] = f4(xyz(d.sbody[4 * (size_t)b]), 0.0f);
    d.ang[b] = f4(xyz(d.sbody[4 * (size_t)b + 1]), 0.0f);
  }
}

}  // namespace

void launch_bearing_projection(hipStream_t s, const Dev& d, int in_contact, int write_back) {
  if (d.bearing_count == 0) return;
  if (d.bi_groups) {
    k_bearing_islands<<<d.bi_groups, 256, 0, s>>>(d, 1u, in_contact, write_back);
    return;
  }
  k_bearing_project<<<blocks_for(d.bearing_count * kBearingLanes), 256, 0, s>>>(d, in_contact);
  k_bearing_apply<<<blocks_for(d.bearing_body_count), 256, 0, s>>>(d, in_contact, write_back);
}

bool launch_apply_islands(hipStream_t s, const Dev& d, uint32_t apply_blocks, int last, uint64_t* clock) {
  if (!d.bi_groups || !d.bi_mask || apply_blocks == 0) return false;
  const uint32_t lead = (d.bi_groups + kXcds - 1) / kXcds * kXcds;
  k_apply_islands<<<lead + apply_blocks, 256, 0, s>>>(d, apply_blocks, last, clock);
  return true;
}

void launch_bearing_passes(hipStream_t s, const Dev& d, uint32_t passes) {
  if (d.bearing_count == 0) return;
  if (d.bi_groups) {
    k_bearing_islands<<<d.bi_groups, 256, 0, s>>>(d, passes, 0, 1);
    return;
  }
  for (uint32_t i = 0; i < passes; ++i) launch_bearing_projection(s, d, 0, 1);
}

void launch_serial_solve(hipStream_t s, const Dev& d, uint32_t passes) {
  const uint32_t grid = std::min(std::max(d.pair_capacity, d.body_count) / 256u + 1u, 1024u);
  k_serial_keys<<<grid, 256, 0, s>>>(d);
  k_serial_rank<<<grid, 256, 0, s>>>(d);
  k_serial_passes<<<1, 64, 0, s>>>(d, passes);
  k_serial_finish<<<grid, 256, 0, s>>>(d);
}

void launch_advance_coordinates(hipStream_t s, const Dev& d) {
  if (d.mech_count) k_advance<<<blocks_for(d.mech_count), 256, 0, s>>>(d);
}

// Forward kinematics, the closure Newton loop (each step followed by forward
// kinematics), publication and reconstruction (device.rs:963-1048).
void launch_mechanism_pose(hipStream_t s, const Dev& d, int closures) {
  if (d.fk_groups == 0) return;
  if (closures && d.cc_count) {
    // One workgroup per closure component, then forward kinematics of every tree.
    const uint32_t items = std::max(std::max(d.cc_max_bodies, d.cc_max_coords), 2u * d.cc_max_closures);
    const uint32_t threads = items <= 64u ? 64u : 256u;
    if (threads == 64u) k_closure_components<64><<<d.cc_count, 64, closure_component_lds(d), s>>>(d);
    else k_closure_components<256><<<d.cc_count, 256, closure_component_lds(d), s>>>(d);
    k_forward_kinematics<<<d.fk_groups, 256, 0, s>>>(d, -1);
    launch_publish(s, d);
    return;
  }
  k_forward_kinematics<<<d.fk_groups, 256, 0, s>>>(d, -1);
  if (closures) {
    for (uint32_t step = 0; step < d.closure_steps; ++step) {
      k_newton_step<<<1, kPoseThreads, 0, s>>>(d);
      k_forward_kinematics<<<d.fk_groups, 256, 0, s>>>(d, (int)step);
    }
  }
  launch_publish(s, d);
}

size_t closure_component_lds(const Dev& d) {
  const size_t NB = d.cc_max_bodies, NQ = d.cc_max_coords, NK = d.cc_max_closures, NP = d.cc_max_path;
  // float4 region: bodies (5), links (2 + 2) and link meta (1) per path body, 4 per closure;
  // doubles: 6 per coordinate, per closure and per path entry; floats: 7 per coordinate,
  // 19 per closure, 6 per path entry; u32: entry offsets, entries, path offsets, paths,
  // coordinate bodies, path-entry columns.
  const size_t dense = d.cc_dense ? 16 + 8 * (6 * NK * NQ + NQ * NQ) : 0;
  return 16 * (10 * NB + 4 * NK) + 8 * (6 * NQ + 6 * NK + 6 * NP) + 4 * (7 * NQ + 19 * NK + 6 * NP) +
         4 * ((NQ + 1) + NP + (2 * NK + 1) + NP + NQ + NP) + dense;
}

hipError_t prepare_closure_components(const Dev& d) {
  const size_t lds = closure_component_lds(d);
  if (lds <= 65536) return hipSuccess;
  const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_closure_components<64>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_closure_components<256>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

void launch_capture_coordinates(hipStream_t s, const Dev& d) {
  if (d.mech_count) k_capture<<<blocks_for(d.mech_count), 256, 0, s>>>(d);
}

void launch_reconstruct(hipStream_t s, const Dev& d) { k_reconstruct<<<1, kPoseThreads, 0, s>>>(d); }

void launch_world_masses(hipStream_t s, const Dev& d) {
  if (d.body_count) k_world_masses<<<blocks_for(d.body_count), 256, 0, s>>>(d);
}

void launch_validate_snapshot(hipStream_t s, const Dev& d) {
  if (d.body_count) k_validate_snapshot<<<blocks_for(d.body_count), 256, 0, s>>>(d);
}

void launch_validate_bearings(hipStream_t s, const Dev& d) {
  if (d.bearing_count) k_validate_bearings<<<blocks_for(d.bearing_count), 256, 0, s>>>(d, d.mech_count > 0 ? 1 : 0);
}

}  // namespace mh
