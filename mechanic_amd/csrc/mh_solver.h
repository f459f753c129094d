// Contact-solver state shared by the parallel passes and the serial
// small-mechanism sweeps: constants, the per-body solver record, the solver
// row planes and the reference's pair hash.
#pragma once

#include "internal.h"
#include "mh_math.h"

namespace mh {

constexpr float kFixedScale = 1048576.0f;  // collision.wgsl:63
constexpr float kRelaxation = 0.125f;      // collision.wgsl:64
constexpr float kWarmStartScale = 0.5f;    // collision.wgsl:65
constexpr float kSlop = 0.001f;            // collision.wgsl:67
constexpr float kMaxCorrection = 1.0f;     // collision.wgsl:68
constexpr float kNormalAlignment = 0.98f;  // collision.wgsl:69
constexpr float kPointMovement = 0.02f;    // collision.wgsl:70
constexpr uint32_t kMaxProbes = 256u;      // collision.wgsl:74

__device__ __forceinline__ uint64_t tick_sv(const Dev& d) {
  return *reinterpret_cast<const volatile uint64_t*>(&d.tick->tick_index);
}

__device__ __forceinline__ uint32_t read_count(uint32_t* word, uint32_t cap) {
  uint32_t v = *reinterpret_cast<volatile uint32_t*>(word);
  return v < cap ? v : cap;
}

// pair_hash (collision.wgsl:488-497).
__device__ __forceinline__ uint32_t pair_hash(uint32_t a, uint32_t b) {
  uint32_t v = a * 0x9e3779b9u;
  v ^= b * 0x85ebca6bu;
  v ^= v >> 16;
  v *= 0x7feb352du;
  v ^= v >> 15;
  v *= 0x846ca68bu;
  v ^= v >> 16;
  return v;
}

// 42-bit key of an ordered collider pair (ground: high = MH_INVALID_INDEX).
__device__ __forceinline__ uint64_t pair_key(uint32_t low, uint32_t high) {
  uint64_t hi = high == kInvalid ? 0ull : (uint64_t)high + 1ull;
  return ((uint64_t)(low + 1u) << 21) | hi;
}

// Solver state of one body as one 64-byte record (sbody): (lin, inverse mass),
// (ang, cx.x), (cx.y, cx.z, cy.x, cy.y), (cy.z, cz) where cx, cy, cz are the
// world inverse-inertia columns of update_world_masses (collision.wgsl:144-162).
struct SolverBody {
  v3 lin, ang, cx, cy, cz;
  float inv_mass;
};

__device__ __forceinline__ SolverBody unpack_body(const float4 (&w)[4]) {
  const float4 a = w[0], c = w[1], e = w[2], f = w[3];
  SolverBody s;
  s.lin = xyz(a);
  s.inv_mass = a.w;
  s.ang = xyz(c);
  s.cx = mk(c.w, e.x, e.y);
  s.cy = mk(e.z, e.w, f.x);
  s.cz = mk(f.y, f.z, f.w);
  return s;
}

__device__ __forceinline__ SolverBody load_body(const float4* records, uint32_t b) {
  const float4* r = records + 4 * (size_t)b;
  const float4 w[4] = {r[0], r[1], r[2], r[3]};
  return unpack_body(w);
}

__device__ __forceinline__ SolverBody load_body(const Dev& d, uint32_t b) { return load_body(d.sbody, b); }

// inverse_inertia (collision.wgsl:109-114).
__device__ __forceinline__ v3 inv_inertia(const SolverBody& s, v3 v) {
  return (s.cx * v.x + s.cy * v.y) + s.cz * v.z;
}

__device__ __forceinline__ float denom_of(const SolverBody& a, const SolverBody& b, bool has_b, v3 arm_a, v3 arm_b,
                                          v3 direction) {
  float result = a.inv_mass;
  v3 angular_a = cross(inv_inertia(a, cross(arm_a, direction)), arm_a);
  result += dot(angular_a, direction);
  if (has_b) {
    result += b.inv_mass;
    v3 angular_b = cross(inv_inertia(b, cross(arm_b, direction)), arm_b);
    result += dot(angular_b, direction);
  }
  return result;
}

// An active contact's solver row, stored as three 16-byte planes and one 4-byte
// plane (structure of arrays: a wavefront's load of one plane is consecutive
// rows, 1 KiB for a 16-byte plane):
//   plane 0: normal, penetration
//   plane 1: arm from body A, normal denominator
//   plane 2: arm from body B, body B
//   plane 3: body A | analytic ground contact << 30 | cache hit << 31
// 52 bytes a contact. The contact's index in the contact list is d.active.
struct SolverRow {
  v3 normal, arm_a, arm_b;
  float penetration, denominator;
  uint32_t body_a, body_b;
  bool analytic, hit;
};
constexpr uint32_t kRowAnalytic = 1u << 30, kRowHit = 1u << 31, kRowBody = kRowAnalytic - 1u;

__device__ __forceinline__ const uint32_t* row_bodies(const Dev& d) {
  return reinterpret_cast<const uint32_t*>(d.rows + 3 * (size_t)d.pair_capacity);
}
// Body A and body B of a row (the grouping and sort keys need no more).
__device__ __forceinline__ uint2 row_body_pair(const Dev& d, uint32_t ai) {
  return make_uint2(row_bodies(d)[ai] & kRowBody, __float_as_uint(d.rows[2 * (size_t)d.pair_capacity + ai].w));
}
__device__ __forceinline__ SolverRow load_row(const Dev& d, uint32_t ai) {
  const size_t cap = d.pair_capacity;
  const float4 p0 = d.rows[ai], p1 = d.rows[cap + ai], p2 = d.rows[2 * cap + ai];
  const uint32_t w = row_bodies(d)[ai];
  SolverRow r;
  r.normal = xyz(p0);
  r.penetration = p0.w;
  r.arm_a = xyz(p1);
  r.denominator = p1.w;
  r.arm_b = xyz(p2);
  r.body_b = __float_as_uint(p2.w);
  r.body_a = w & kRowBody;
  r.analytic = (w & kRowAnalytic) != 0u;
  r.hit = (w & kRowHit) != 0u;
  return r;
}
__device__ __forceinline__ void store_row(const Dev& d, uint32_t ai, const SolverRow& r) {
  const size_t cap = d.pair_capacity;
  d.rows[ai] = f4(r.normal, r.penetration);
  d.rows[cap + ai] = f4(r.arm_a, r.denominator);
  d.rows[2 * cap + ai] = f4(r.arm_b, __uint_as_float(r.body_b));
  reinterpret_cast<uint32_t*>(d.rows + 3 * cap)[ai] =
      r.body_a | (r.analytic ? kRowAnalytic : 0u) | (r.hit ? kRowHit : 0u);
}

}  // namespace mh
