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

// kApplyLanes adjacent lanes share a body: each loads every kApplyLanes-th of
// its slots, all loads in flight at once, and shuffles combine the partial
// sums. The slot loads of a body cost one memory latency instead of one per slot.
constexpr uint32_t kApplyLanes = 4;
constexpr uint32_t kApplyPerLane = (kSlotOverflow + kApplyLanes - 1) / kApplyLanes;

// solve_apply of the contact solve (collision.wgsl:1050-1069) for the
// bodies of `block` out of `nblocks` (the launch's first nblocks blocks; a
// body whose `skip` byte is set is left to another part of the launch).
__device__ __forceinline__ void apply_contact_deltas(const Dev& d, int last, uint64_t* clock, const uint8_t* skip,
                                                     uint32_t block, uint32_t nblocks) {
  if (clock && threadIdx.x == 0) clock[2 * block] = wall_clock64();
  const uint64_t tick64 = tick_sv(d);
  uint32_t* diag = d.diag[tick64 & 1];
  if (*reinterpret_cast<volatile uint32_t*>(&diag[kDiagActive]) == 0u) return;
  const uint32_t xcd = block % kXcds, local = block / kXcds;
  const uint32_t blocks = (nblocks - xcd + kXcds - 1) / kXcds;
  XcdSplit xs;
  xs.begin = (uint32_t)(((uint64_t)d.body_count * xcd) / kXcds);
  xs.end = (uint32_t)(((uint64_t)d.body_count * (xcd + 1)) / kXcds);
  const uint32_t per_block = blockDim.x / kApplyLanes;
  const uint32_t part = threadIdx.x % kApplyLanes;
  for (uint32_t b0 = xs.begin + local * per_block; b0 < xs.end; b0 += blocks * per_block) {
    const uint32_t b = b0 + threadIdx.x / kApplyLanes;
    const bool valid = b < xs.end && !(skip && skip[b]);
    uint32_t acc[6] = {0u, 0u, 0u, 0u, 0u, 0u};
    const uint32_t count = valid ? d.slot_count[b] : 0u;
    const uint32_t n = count < kSlotOverflow ? count : kSlotOverflow;
    // The body's velocity words, loaded alongside its slots.
    float4 v0 = make_float4(0.f, 0.f, 0.f, 0.f), v1 = v0;
    if (valid && part == 0u) {
      v0 = d.sbody[4 * (size_t)b];
      v1 = d.sbody[4 * (size_t)b + 1];
    }
    int4 x[kApplyPerLane];
    int2 y[kApplyPerLane];
#pragma unroll
    for (uint32_t j = 0; j < kApplyPerLane; ++j) {
      const uint32_t k = part + j * kApplyLanes;
      x[j] = make_int4(0, 0, 0, 0);
      y[j] = make_int2(0, 0);
      if (k < n) {
        const int4* e = d.cdelta + 2 * ((size_t)k * d.body_count + b);
        x[j] = e[0];
        y[j] = *reinterpret_cast<const int2*>(e + 1);
      }
    }
#pragma unroll
    for (uint32_t j = 0; j < kApplyPerLane; ++j) {
      acc[0] += (uint32_t)x[j].x;
      acc[1] += (uint32_t)x[j].y;
      acc[2] += (uint32_t)x[j].z;
      acc[3] += (uint32_t)x[j].w;
      acc[4] += (uint32_t)y[j].x;
      acc[5] += (uint32_t)y[j].y;
    }
#pragma unroll
    for (uint32_t m = 1; m < kApplyLanes; m <<= 1)
#pragma unroll
      for (int c = 0; c < 6; ++c) acc[c] += (uint32_t)__shfl_xor((int)acc[c], (int)m);
    if (!valid || part != 0u) continue;
    if (count > kSlotOverflow) {
      int32_t* p = d.deltas + 6 * (size_t)b;
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        acc[c] += (uint32_t)p[c];
        p[c] = 0;
      }
    }
    const v3 lin_delta = mk((float)(int32_t)acc[0], (float)(int32_t)acc[1], (float)(int32_t)acc[2]) / kFixedScale;
    const v3 ang_delta = mk((float)(int32_t)acc[3], (float)(int32_t)acc[4], (float)(int32_t)acc[5]) / kFixedScale;
    const v3 lin = xyz(v0) + lin_delta;
    const v3 ang = xyz(v1) + ang_delta;
    float4* r = d.sbody + 4 * (size_t)b;
    r[0] = f4(lin, v0.w);
    r[1] = f4(ang, v1.w);
    if (last) {
      d.lin[b] = f4(lin, 0.0f);
      d.ang[b] = f4(ang, 0.0f);
    }
  }
  if (clock) {
    __syncthreads();
    if (threadIdx.x == 0) clock[2 * block + 1] = wall_clock64();
  }
}

}  // namespace mh
