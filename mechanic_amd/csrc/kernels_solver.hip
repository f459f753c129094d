// Contact preparation (persistent-manifold warm start) and the relaxed projected
// Jacobi impulse solver with Q20 fixed-point velocity deltas,
// collision.wgsl:488-538, 674-906, 1050-1088.
//
// The manifold cache is keyed by the ordered collider pair. Each tick inserts
// the pairs it prepares into a fresh open-addressing index (by tick parity) and
// looks pairs up in the index of the previous tick, so whether a contact finds
// its cached impulse depends only on the previous tick's contact set, never on
// the order in which threads ran.
#include <hip/hip_ext.h>

#include "internal.h"
#include "mh_geometry.h"
#include "mh_math.h"
#include "mh_solver.h"
#include "mlp_net.h"

namespace mh {

// The six Q20 integers add_velocity_delta (collision.wgsl:764-772) adds for one
// contact side: side A gets -impulse, side B +impulse.
struct Q20 {
  int32_t v[6];
};
__device__ __forceinline__ Q20 side_delta(const SolverBody& s, v3 arm, v3 impulse) {
  const v3 linear = impulse * s.inv_mass;
  const v3 angular = inv_inertia(s, cross(arm, impulse));
  Q20 q;
  q.v[0] = round_to_i32(linear.x * kFixedScale);
  q.v[1] = round_to_i32(linear.y * kFixedScale);
  q.v[2] = round_to_i32(linear.z * kFixedScale);
  q.v[3] = round_to_i32(angular.x * kFixedScale);
  q.v[4] = round_to_i32(angular.y * kFixedScale);
  q.v[5] = round_to_i32(angular.z * kFixedScale);
  return q;
}

// One delta slot of a body: the Q20 deltas of the body's sides in one group of
// 64 contacts, summed. Slot-major: for one slot, consecutive bodies are
// adjacent, so the bodies' gather reads whole cache lines per wavefront; a
// slot's 32 bytes stay in one line.
__device__ __forceinline__ void put_side(const Dev& d, uint32_t body, uint32_t slot, const Q20& q) {
  int4* p = d.cdelta + 2 * ((size_t)slot * d.body_count + body);
  p[0] = make_int4(q.v[0], q.v[1], q.v[2], q.v[3]);
  p[1] = make_int4(q.v[4], q.v[5], 0, 0);
}

__device__ __forceinline__ void add_overflow(const Dev& d, uint32_t body, const Q20& q) {
  int32_t* p = d.deltas + 6 * (size_t)body;
#pragma unroll
  for (int k = 0; k < 6; ++k) atomicAdd(&p[k], q.v[k]);
}

// prepare_contacts (collision.wgsl:674-736) with the deterministic manifold index.
__global__ __launch_bounds__(256) void k_prepare(Dev d) {
  __shared__ uint32_t lds_append[kBlock / 64 + 1];
  const uint64_t tick64 = tick_sv(d);
  const uint32_t tick = (uint32_t)tick64;
  uint32_t* diag = d.diag[tick64 & 1];
  uint64_t* table_now = d.htab[tick64 & 1];
  const uint64_t* table_prev = d.htab[(tick64 + 1) & 1];
  float4* rows_now = d.mrows[tick64 & 1];
  const float4* rows_prev = d.mrows[(tick64 + 1) & 1];
  const uint32_t count = read_count(&diag[kDiagContacts], d.pair_capacity);
  const bool per_body = d.t_solve_kind == 1u;
  if (d.learn_mode == kLearnInfer && blockIdx.x == 0 && threadIdx.x == 0) d.learned_state[2] = 0u;
  if (per_body)
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < d.side_tiles; i += gridDim.x * blockDim.x)
      d.side_status[i] = 0u;
  const XcdSplit xs = grid_split(count);
  for (uint32_t base = xs.first; base < xs.end; base += xs.step) {
    const uint32_t ci = base + threadIdx.x;
    bool active = false;
    SolverRow row;
    float cached_lambda = 0.0f;
    if (ci < xs.end) {
      const float4* c = d.contacts + 4 * (size_t)ci;
      const float4 m = c[0];
      const uint32_t body_a = __float_as_uint(m.x), body_b = __float_as_uint(m.y);
      const uint32_t col_a = __float_as_uint(m.z), col_b = __float_as_uint(m.w);
      const float4 n_pen = c[1];
      const float4 arm_a4 = c[2];
      const float4 arm_b4 = c[3];
      const v3 arm_a = xyz(arm_a4), arm_b = xyz(arm_b4), normal = xyz(n_pen);
      const bool has_b = body_b != kInvalid;
      const SolverBody sa = load_body(d, body_a);
      SolverBody sb = sa;
      float inverse_b = 0.0f;
      v3 velocity_b = mk(0.0f, 0.0f, 0.0f);
      if (has_b) {
        sb = load_body(d, body_b);
        inverse_b = sb.inv_mass;
        velocity_b = sb.lin + cross(sb.ang, arm_b);
      }
      if (sa.inv_mass + inverse_b <= 0.0f) goto done;
      {
        const v3 relative = velocity_b - (sa.lin + cross(sa.ang, arm_a));
        const float normal_speed = dot(relative, normal);
        if (n_pen.w <= 1.0e-6f && normal_speed >= 0.0f) goto done;
        const uint32_t low = col_a < col_b ? col_a : col_b;
        const uint32_t high = col_a < col_b ? col_b : col_a;
        const uint64_t key = pair_key(low, high);
        const uint32_t mask = d.hash_mask;
        const uint32_t start = pair_hash(low, high) & mask;
        // Look the pair up among the previous tick's manifolds.
        float cached_impulse = 0.0f;
        bool hit = false;
        for (uint32_t probe = 0; probe < kMaxProbes; ++probe) {
          uint64_t e = table_prev[(start + probe) & mask];
          if (e == 0ull) break;
          if ((e >> 22) == key) {
            const float4* row = rows_prev + 3 * (size_t)(e & 0x3fffffull);
            float4 pt = row[0], cn = row[1], cp = row[2];
            bool matches = __float_as_uint(pt.z) == tick - 1u && dot(xyz(cn), normal) >= kNormalAlignment &&
                           length(xyz(cp) - arm_a) <= kPointMovement;
            cached_impulse = matches ? cp.w : 0.0f;
            hit = matches;
            break;
          }
        }
        // Insert the pair into this tick's index.
        const uint64_t entry = (key << 22) | (uint64_t)ci;
        bool inserted = false;
        for (uint32_t probe = 0; probe < kMaxProbes; ++probe) {
          unsigned long long* slot = reinterpret_cast<unsigned long long*>(&table_now[(start + probe) & mask]);
          if (atomicCAS(slot, 0ull, (unsigned long long)entry) == 0ull) {
            inserted = true;
            break;
          }
        }
        if (!inserted) {
          atomicOr(&diag[kDiagFlags], 8u);
          goto done;
        }
        float4* mr = rows_now + 3 * (size_t)ci;
        mr[0] = make_float4(__uint_as_float(low), __uint_as_float(high), __uint_as_float(tick), __uint_as_float(1u));
        mr[1] = n_pen;
        mr[2] = make_float4(arm_a4.x, arm_a4.y, arm_a4.z, cached_impulse);
        if (n_pen.w > 1.0e-5f || normal_speed < -1.0e-5f || cached_impulse > 1.0e-6f) {
          active = true;
          row.normal = normal;
          row.penetration = n_pen.w;
          row.arm_a = arm_a;
          row.denominator = denom_of(sa, sb, has_b, arm_a, arm_b, normal);
          row.arm_b = arm_b;
          row.body_a = body_a;
          row.body_b = body_b;
          row.analytic = arm_b4.w < 0.0f;  // generate_ground_contacts' analytic cylinder marker
          row.hit = hit;
          cached_lambda = cached_impulse;
        }
      }
    }
  done:
    uint32_t a = block_append(&diag[kDiagActive], active, lds_append);
    if (active) {
      const uint32_t body_a = row.body_a, body_b = row.body_b;
      // Per-body passes: both contact sides take a slot of their body, their row.
      // (The per-contact passes take slots per group of sides in their first pass.)
      if (per_body) {
        const uint32_t slot_a = atomicAdd(&d.slot_count[body_a], 1u);
        const uint32_t slot_b = body_b != kInvalid ? atomicAdd(&d.slot_count[body_b], 1u) : 0u;
        d.side_slot[a] = make_uint2(slot_a, slot_b);
      }
      d.active[a] = ci;
      d.lambda[a] = cached_lambda;
      store_row(d, a, row);
    }
  }
}

// One contact's impulse for one pass: warm_start (collision.wgsl:774-839) when
// `first`, solve_accumulate (collision.wgsl:841-906) otherwise, from the bodies'
// velocities before the pass. Updates the accumulated normal impulse `lambda`,
// sets the impulse on side B (side A gets its negation) and returns whether one
// applies. The `owner` of a contact does its learned-warm-start logging.
struct ContactRow {
  v3 normal, arm_a, arm_b;
  float penetration, denominator;
  bool analytic;
};

__device__ __forceinline__ bool contact_pass(const Dev& d, const SolverBody& sa, const SolverBody& sb, bool has_b,
                                             const ContactRow& r, int first, uint32_t ai, bool owner, float& lambda,
                                             v3& impulse) {
  const v3 velocity_b = has_b ? sb.lin + cross(sb.ang, r.arm_b) : mk(0.0f, 0.0f, 0.0f);
  impulse = mk(0.0f, 0.0f, 0.0f);
  bool applies = false;
  // The learned warm start's first pass, unless the contact's prediction was
  // not finite (then finit.w is 1 and the contact takes the cache warm start).
  float4 predicted = make_float4(0.f, 0.f, 0.f, 1.f);
  if (first && d.learn_mode == kLearnInfer) predicted = d.finit[ai];
  const bool learned_first = first && d.learn_mode == kLearnInfer && predicted.w == 0.0f;
  if (r.denominator > 0.0f) {
    const v3 relative = velocity_b - (sa.lin + cross(sa.ang, r.arm_a));
    const float normal_speed = dot(relative, r.normal);
    if (r.penetration <= 1.0e-5f && normal_speed >= -1.0e-5f && lambda <= 1.0e-6f) {
      lambda = 0.0f;
    } else {
      const float recovery = r.analytic ? 1.0f : 0.2f;
      const float bias = fmn((fmx(r.penetration - kSlop, 0.0f) * recovery) / d.delta_seconds, kMaxCorrection);
      const float step = (kRelaxation * (-normal_speed + bias)) / r.denominator;
      float bound;
      if (first) {
        // A learned prediction replaces the halved cached impulse in full.
        const float scale = (r.analytic || learned_first) ? 1.0f : kWarmStartScale;
        const float warmed = lambda * scale;
        const float accumulated = fmx(warmed + step, 0.0f);
        lambda = accumulated;
        impulse = r.normal * accumulated;
        bound = accumulated * d.friction;
      } else {
        const float previous = fmx(lambda, 0.0f);
        const float accumulated = fmx(previous + step, 0.0f);
        const float delta = accumulated - previous;
        lambda = accumulated;
        impulse = r.normal * delta;
        bound = fabsf(delta) * d.friction;
      }
      const v3 tangent_velocity = relative - r.normal * normal_speed;
      const float tangent_speed = length(tangent_velocity);
      v3 friction_vec = mk(0.0f, 0.0f, 0.0f);
      if (learned_first) {
        // Learned warm start: the predicted friction impulse, inside the cone.
        const float4 f = predicted;
        const float mag = length(xyz(f));
        const float cap = bound;
        friction_vec = mag > cap && mag > 0.0f ? xyz(f) * (cap / mag) : xyz(f);
        impulse = impulse + friction_vec;
      } else if (tangent_speed > 1.0e-6f) {
        const v3 tangent = tangent_velocity / tangent_speed;
        const float tangent_denominator = denom_of(sa, sb, has_b, r.arm_a, r.arm_b, tangent);
        const float friction_impulse = fmn(tangent_speed / tangent_denominator, bound);
        friction_vec = -(tangent * friction_impulse);
        impulse = impulse - tangent * friction_impulse;
      }
      if (owner && d.learn_mode == kLearnLog) {
        const float4 acc = first ? make_float4(0.f, 0.f, 0.f, 0.f) : d.fsum[ai];
        d.fsum[ai] = make_float4(acc.x + friction_vec.x, acc.y + friction_vec.y, acc.z + friction_vec.z, 0.0f);
      }
      applies = true;
    }
  }
  if (owner && first && !applies && d.learn_mode == kLearnLog) d.fsum[ai] = make_float4(0.f, 0.f, 0.f, 0.f);
  return applies;
}

// Contact sides grouped per wavefront (tuning solve_kind 0, the default). The
// per-contact passes run one wavefront per aligned group of 64 active contacts;
// the group's 128 sides touch far fewer bodies (about two in five on a settled
// pile), so each body of a group gets one delta slot, and its sides' Q20
// integers are summed in LDS before one store. The tick's first pass finds, for
// every side, the group's first side of the same body (its representative) and
// gives each representative a slot of its body. Packed per active contact:
// representative of side A | of side B << 7 | slot A << 14 | slot B << 19
// (kSlotOverflow: add with atomics instead).
constexpr uint32_t kGroupSlotShift = 14;

// Per-wavefront LDS table of the group's bodies: open addressing over kSideTable
// entries (at most 128 bodies, so the table never fills), each holding a body
// and the first side position seen for it.
constexpr uint32_t kSideTable = 256;

__device__ __forceinline__ uint32_t side_slot_of(uint32_t* bodies, uint32_t body) {
  uint32_t h = (body * 0x9e3779b1u) >> 24;  // top 8 bits: kSideTable entries
  for (uint32_t probe = 0; probe < kSideTable; ++probe) {
    const uint32_t prev = atomicCAS(&bodies[h], kInvalid, body);
    if (prev == kInvalid || prev == body) return h;
    h = (h + 1u) & (kSideTable - 1u);
  }
  return kInvalid;  // unreachable: at most 128 bodies in 256 entries
}

// The group word of one lane's contact (ba, bb: its bodies; kInvalid for no
// contact or no body B) in a group of 64 on one wavefront: the first position
// of each side's body over the group's sides in order (side A of lane l at 2l,
// side B at 2l + 1), from the wavefront's LDS table, and a slot of its body for
// each side that is that first position (one counter atomic per body of the
// group). All 64 lanes of the wavefront call it together.
__device__ __forceinline__ uint32_t group_word(const Dev& d, uint32_t* bodies, uint32_t* firsts, uint32_t lane,
                                               bool valid, uint32_t ba, uint32_t bb) {
  for (uint32_t e = lane; e < kSideTable; e += 64u) {
    bodies[e] = kInvalid;
    firsts[e] = kInvalid;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const uint32_t ha = ba != kInvalid ? side_slot_of(bodies, ba) : kInvalid;
  const uint32_t hb = bb != kInvalid ? side_slot_of(bodies, bb) : kInvalid;
  if (ha != kInvalid) atomicMin(&firsts[ha], 2u * lane);
  if (hb != kInvalid) atomicMin(&firsts[hb], 2u * lane + 1u);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const uint32_t rep_a = ha != kInvalid ? firsts[ha] : 2u * lane;
  const uint32_t rep_b = hb != kInvalid ? firsts[hb] : 2u * lane + 1u;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();  // the next group clears the same table
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  uint32_t own_a = 0u, own_b = 0u;
  if (valid && rep_a == 2u * lane) own_a = atomicAdd(&d.slot_count[ba], 1u);
  if (valid && bb != kInvalid && rep_b == 2u * lane + 1u) own_b = atomicAdd(&d.slot_count[bb], 1u);
  const uint32_t a_from_a = (uint32_t)__shfl((int)own_a, (int)(rep_a >> 1));
  const uint32_t a_from_b = (uint32_t)__shfl((int)own_b, (int)(rep_a >> 1));
  const uint32_t b_from_a = (uint32_t)__shfl((int)own_a, (int)(rep_b >> 1));
  const uint32_t b_from_b = (uint32_t)__shfl((int)own_b, (int)(rep_b >> 1));
  uint32_t slot_a = (rep_a & 1u) ? a_from_b : a_from_a;
  uint32_t slot_b = (rep_b & 1u) ? b_from_b : b_from_a;
  slot_a = slot_a < kSlotOverflow ? slot_a : kSlotOverflow;
  slot_b = bb != kInvalid && slot_b < kSlotOverflow ? slot_b : kSlotOverflow;
  return rep_a | (rep_b << 7) | (slot_a << kGroupSlotShift) | (slot_b << (kGroupSlotShift + 5u));
}

// warm_start (collision.wgsl:774-839) when first, solve_accumulate
// (collision.wgsl:841-906) otherwise, per contact, with the group's sides
// summed per body in LDS; the last pass also runs persist_contacts
// (collision.wgsl:1071-1088). Each XCD takes a contiguous run of whole groups.
template <bool kGroup>
__global__ __launch_bounds__(256) void k_solve(Dev d, int first, int last, uint64_t* clock, const uint32_t* gate) {
  __shared__ int32_t lds_q[kBlock / 64][128][6];
  __shared__ uint32_t lds_body[kGroup ? kBlock / 64 : 1][kGroup ? kSideTable : 1];
  __shared__ uint32_t lds_first[kGroup ? kBlock / 64 : 1][kGroup ? kSideTable : 1];
  if (gate && *reinterpret_cast<const volatile uint32_t*>(gate) == 0u) return;
  if (clock && threadIdx.x == 0) clock[2 * blockIdx.x] = wall_clock64();
  const uint64_t tick64 = tick_sv(d);
  uint32_t* diag = d.diag[tick64 & 1];
  float4* rows_now = d.mrows[tick64 & 1];
  const uint32_t count = read_count(&diag[kDiagActive], d.pair_capacity);
  const uint32_t groups = (count + 63u) / 64u;
  const uint32_t xcd = blockIdx.x % kXcds, local = blockIdx.x / kXcds;
  const uint32_t blocks = (gridDim.x - xcd + kXcds - 1) / kXcds;
  const uint32_t g_end = (uint32_t)(((uint64_t)groups * (xcd + 1)) / kXcds);
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  int32_t(*q)[6] = lds_q[wave];
  for (uint32_t g = (uint32_t)(((uint64_t)groups * xcd) / kXcds) + local * (kBlock / 64) + wave; g < g_end;
       g += blocks * (kBlock / 64)) {
    const uint32_t ai = g * 64u + lane;
    const bool valid = ai < count;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      q[2 * lane][c] = 0;
      q[2 * lane + 1][c] = 0;
    }
    uint32_t body_a = kInvalid, body_b = kInvalid, grp = 0u;
    if constexpr (kGroup) {
      // The tick's first pass groups the sides (the later passes read the words).
      if (valid) {
        const uint2 bodies = row_body_pair(d, ai);
        body_a = bodies.x;
        body_b = bodies.y;
      }
      grp = group_word(d, lds_body[wave], lds_first[wave], lane, valid, body_a, body_b);
      if (valid) d.side_group[ai] = grp;
    }
    if (valid) {
      const SolverRow sr = load_row(d, ai);
      body_a = sr.body_a;
      body_b = sr.body_b;
      if constexpr (!kGroup) grp = d.side_group[ai];
      const bool has_b = body_b != kInvalid;
      const ContactRow row{sr.normal, sr.arm_a, sr.arm_b, sr.penetration, sr.denominator, sr.analytic};
      float lambda = d.lambda[ai];
      const SolverBody sa = load_body(d.sbody, body_a);
      const SolverBody sb = has_b ? load_body(d.sbody, body_b) : sa;
      v3 impulse;
      const bool applies = contact_pass(d, sa, sb, has_b, row, first, ai, true, lambda, impulse);
      d.lambda[ai] = lambda;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      // A contact that applies no impulse adds zero deltas (collision.wgsl:793-804).
      if (applies) {
        const Q20 qa = side_delta(sa, row.arm_a, -impulse);
        const uint32_t ra = grp & 127u;
#pragma unroll
        for (int c = 0; c < 6; ++c) atomicAdd(&q[ra][c], qa.v[c]);
        if (has_b) {
          const Q20 qb = side_delta(sb, row.arm_b, impulse);
          const uint32_t rb = (grp >> 7) & 127u;
#pragma unroll
          for (int c = 0; c < 6; ++c) atomicAdd(&q[rb][c], qb.v[c]);
        }
      }
      if (last) rows_now[3 * (size_t)d.active[ai] + 2].w = lambda;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // Each body's representative side stores the group's sum for that body.
    if (valid) {
      const uint32_t slot_a = (grp >> kGroupSlotShift) & 31u, slot_b = (grp >> (kGroupSlotShift + 5u)) & 31u;
      if ((grp & 127u) == 2u * lane) {
        Q20 sum;
#pragma unroll
        for (int c = 0; c < 6; ++c) sum.v[c] = q[2 * lane][c];
        if (slot_a == kSlotOverflow) add_overflow(d, body_a, sum);
        else put_side(d, body_a, slot_a, sum);
      }
      if (body_b != kInvalid && ((grp >> 7) & 127u) == 2u * lane + 1u) {
        Q20 sum;
#pragma unroll
        for (int c = 0; c < 6; ++c) sum.v[c] = q[2 * lane + 1][c];
        if (slot_b == kSlotOverflow) add_overflow(d, body_b, sum);
        else put_side(d, body_b, slot_b, sum);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();  // the next group clears the same LDS
  }
  if (clock) {
    __syncthreads();
    if (threadIdx.x == 0) clock[2 * blockIdx.x + 1] = wall_clock64();
  }
}

// ---------------------------------------------------------------------------
// Per-body passes (tuning solve_kind 1). Once per tick each active contact's two
// sides are laid out as rows grouped by body (k_side_offsets scans the bodies'
// side counts, k_scatter_sides writes the rows); each pass then runs per body:
// its lanes take its sides, compute each contact's impulse exactly as the
// per-contact pass does (both sides of a contact see the same inputs, so they
// agree on it and on the accumulated impulse bit for bit), sum the body's own
// Q20 deltas (wrapping 32-bit sums: the arithmetic of atomicAdd on i32, in any
// order) and write the body's new velocities into the other record buffer.
// Nothing is stored per side but the accumulated impulse.
// ---------------------------------------------------------------------------
constexpr uint32_t kSidePer = kSideTile / kBlock;  // bodies per thread of the offset scan
constexpr uint32_t kSideAggregate = 1u << 30;      // status word: the tile's own side count
constexpr uint32_t kSidePrefix = 2u << 30;         // status word: side count over tiles 0..this one
constexpr uint32_t kSideCount = (1u << 30) - 1u;
constexpr uint32_t kSideSpinLimit = 1u << 22;      // reads of an unpublished word before giving up

// Exclusive scan of slot_count into side_off (decoupled look-back over tiles of
// kSideTile bodies, taken in ticket order so a tile only waits for tiles that
// started before it).
__global__ __launch_bounds__(256) void k_side_offsets(Dev d) {
  __shared__ uint32_t lds_tile, lds_before, lds_wave[kBlock / 64];
  if (threadIdx.x == 0) lds_tile = atomicAdd(&d.counters[kCountSideTicket], 1u);
  __syncthreads();
  const uint32_t tile = lds_tile;
  const uint32_t n = d.body_count;
  const uint32_t base = tile * kSideTile + threadIdx.x * kSidePer;
  uint32_t c[kSidePer];
  uint32_t run = 0;
#pragma unroll
  for (uint32_t j = 0; j < kSidePer; ++j) {
    c[j] = base + j < n ? d.slot_count[base + j] : 0u;
    run += c[j];
  }
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t incl = run;
#pragma unroll
  for (uint32_t m = 1; m < 64; m <<= 1) {
    const uint32_t up = (uint32_t)__shfl_up((int)incl, m);
    if (lane >= m) incl += up;
  }
  if (lane == 63) lds_wave[wave] = incl;
  __syncthreads();
  uint32_t wave_before = 0, total = 0;
#pragma unroll
  for (uint32_t w = 0; w < kBlock / 64; ++w) {
    wave_before += w < wave ? lds_wave[w] : 0u;
    total += lds_wave[w];
  }
  if (threadIdx.x == 0) {
    uint32_t* status = d.side_status;
    uint32_t before = 0;
    if (tile == 0) {
      store_status(&status[0], kSidePrefix | total);
    } else {
      store_status(&status[tile], kSideAggregate | total);
      uint32_t j = tile - 1u, spins = 0;
      for (;;) {
        const uint32_t v = load_status(&status[j]);
        if ((v & ~kSideCount) == 0u) {
          if (++spins >= kSideSpinLimit) {
            atomicOr(&d.diag[tick_sv(d) & 1][kDiagFlags], 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        before += v & kSideCount;
        if (v & kSidePrefix) break;
        --j;  // tile 0 always publishes a prefix, so j stays >= 0
      }
      store_status(&status[tile], kSidePrefix | (before + total));
    }
    lds_before = before;
  }
  __syncthreads();
  uint32_t off = lds_before + wave_before + incl - run;
#pragma unroll
  for (uint32_t j = 0; j < kSidePer; ++j) {
    if (base + j < n) d.side_off[base + j] = off;
    off += c[j];
  }
}


// Both sides of each active contact as rows of its bodies: (normal, penetration),
// (arm A, denominator), (arm B, analytic flag), (other body, active index,
// contact index, side), and the accumulated impulse the first pass starts from.
__global__ __launch_bounds__(256) void k_scatter_sides(Dev d) {
  const uint64_t tick64 = tick_sv(d);
  uint32_t* diag = d.diag[tick64 & 1];
  const uint32_t count = read_count(&diag[kDiagActive], d.pair_capacity);
  const XcdSplit xs = xcd_split(count);
  for (uint32_t ai = xs.first + threadIdx.x; ai < xs.end; ai += xs.step) {
    const SolverRow sr = load_row(d, ai);
    const uint32_t body_a = sr.body_a, body_b = sr.body_b;
    const uint32_t ci = d.active[ai];
    const uint2 slots = d.side_slot[ai];
    const float lambda = d.lambda[ai];
    const float4 w0 = f4(sr.normal, sr.penetration), w1 = f4(sr.arm_a, sr.denominator);
    const float4 w2 = f4(sr.arm_b, sr.analytic ? -1.0f : 0.0f);
    float4* ra = d.srow + 4 * (size_t)(d.side_off[body_a] + slots.x);
    ra[0] = w0;
    ra[1] = w1;
    ra[2] = w2;
    ra[3] = make_float4(__uint_as_float(body_b), __uint_as_float(ai), __uint_as_float(ci), __uint_as_float(0u));
    d.slam[d.side_off[body_a] + slots.x] = lambda;
    if (body_b != kInvalid) {
      float4* rb = d.srow + 4 * (size_t)(d.side_off[body_b] + slots.y);
      rb[0] = w0;
      rb[1] = w1;
      rb[2] = w2;
      rb[3] = make_float4(__uint_as_float(body_a), __uint_as_float(ai), __uint_as_float(ci), __uint_as_float(1u));
      d.slam[d.side_off[body_b] + slots.y] = lambda;
    }
  }
}

// One solver pass over groups of kGroupBodies consecutive bodies, one workgroup
// per group: solve_accumulate / warm_start for every contact side of the group
// and solve_apply (collision.wgsl:1050-1069) for its bodies, reading the records
// `src` and writing `dst`. The group's sides are contiguous, so the threads take
// them in turn, one each, with every load of a side independent of the others;
// a side finds its body by bisecting the group's offsets and adds its Q20
// integers to the body's sums in LDS. Side A owns the contact's logging and, in
// the last pass, persist_contacts (collision.wgsl:1071-1088) and the accumulated
// impulse per active contact. With no active contact the pass copies the records
// (solve_apply is skipped, finalize_active_contacts, collision.wgsl:744).
constexpr uint32_t kGroupBodies = 16;

__global__ __launch_bounds__(256) void k_solve_bodies(Dev d, const float4* __restrict__ src, float4* __restrict__ dst,
                                                      int first, int last, uint64_t* clock, const uint32_t* gate) {
  __shared__ float4 lds_own[kGroupBodies][4];
  __shared__ uint32_t lds_off[kGroupBodies + 1];
  __shared__ int32_t lds_acc[6][kGroupBodies];
  if (gate && *reinterpret_cast<const volatile uint32_t*>(gate) == 0u) return;
  if (clock && threadIdx.x == 0) clock[2 * blockIdx.x] = wall_clock64();
  const uint64_t tick64 = tick_sv(d);
  uint32_t* diag = d.diag[tick64 & 1];
  float4* rows_now = d.mrows[tick64 & 1];
  const bool any = *reinterpret_cast<volatile uint32_t*>(&diag[kDiagActive]) != 0u;
  const uint32_t n = d.body_count;
  const uint32_t groups = (n + kGroupBodies - 1) / kGroupBodies;
  // Each XCD takes a contiguous run of groups, so the bodies its sides gather
  // stay in its L2.
  const uint32_t xcd = blockIdx.x % kXcds, local = blockIdx.x / kXcds;
  const uint32_t blocks = (gridDim.x - xcd + kXcds - 1) / kXcds;
  const uint32_t g_end = (uint32_t)(((uint64_t)groups * (xcd + 1)) / kXcds);
  const uint32_t t = threadIdx.x;
  for (uint32_t g = (uint32_t)(((uint64_t)groups * xcd) / kXcds) + local; g < g_end; g += blocks) {
    const uint32_t b0 = g * kGroupBodies;
    const uint32_t nb = n - b0 < kGroupBodies ? n - b0 : kGroupBodies;
    if (t < kGroupBodies * 4) {
      const uint32_t j = t >> 2, w = t & 3u;
      if (j < nb) lds_own[j][w] = src[4 * (size_t)(b0 + j) + w];
    }
    if (t < kGroupBodies) {
#pragma unroll
      for (int c = 0; c < 6; ++c) lds_acc[c][t] = 0;
    }
    if (t < nb) lds_off[t] = d.side_off[b0 + t];
    if (t == 0) lds_off[nb] = d.side_off[b0 + nb - 1] + d.slot_count[b0 + nb - 1];  // nb >= 1
    __syncthreads();
    if (any) {
      const uint32_t side_first = lds_off[0], side_end = lds_off[nb];
      for (uint32_t s = side_first + t; s < side_end; s += blockDim.x) {
        // The body whose sides hold s: the last j with lds_off[j] <= s.
        uint32_t lo = 0, hi = nb;
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (lds_off[mid] <= s) lo = mid;
          else hi = mid;
        }
        const float4* row_s = d.srow + 4 * (size_t)s;
        const float4 w0 = row_s[0], w1 = row_s[1], w2 = row_s[2], w3 = row_s[3];
        float lambda = d.slam[s];
        const uint32_t other = __float_as_uint(w3.x), ai = __float_as_uint(w3.y);
        const bool side_b = __float_as_uint(w3.w) != 0u;
        const bool has_b = side_b || other != kInvalid;
        const float4 ow[4] = {lds_own[lo][0], lds_own[lo][1], lds_own[lo][2], lds_own[lo][3]};
        const SolverBody so = unpack_body(ow);
        const SolverBody ob = other != kInvalid ? load_body(src, other) : so;
        const SolverBody& sa = side_b ? ob : so;
        const SolverBody& sb = side_b ? so : ob;
        const ContactRow row{xyz(w0), xyz(w1), xyz(w2), w0.w, w1.w, w2.w < 0.0f};
        v3 impulse;
        const bool applies = contact_pass(d, sa, sb, has_b, row, first, ai, !side_b, lambda, impulse);
        d.slam[s] = lambda;
        if (!side_b && last) {
          d.lambda[ai] = lambda;
          rows_now[3 * (size_t)__float_as_uint(w3.z) + 2].w = lambda;
        }
        if (applies) {
          const Q20 q = side_b ? side_delta(sb, row.arm_b, impulse) : side_delta(sa, row.arm_a, -impulse);
#pragma unroll
          for (int c = 0; c < 6; ++c) atomicAdd(&lds_acc[c][lo], q.v[c]);
        }
      }
    }
    __syncthreads();
    if (t < nb) {
      float4* r = dst + 4 * (size_t)(b0 + t);
      const float4 w0 = lds_own[t][0], w1 = lds_own[t][1];
      if (!any) {
        r[0] = w0;
        r[1] = w1;
      } else {
        const v3 lin_delta = mk((float)lds_acc[0][t], (float)lds_acc[1][t], (float)lds_acc[2][t]) / kFixedScale;
        const v3 ang_delta = mk((float)lds_acc[3][t], (float)lds_acc[4][t], (float)lds_acc[5][t]) / kFixedScale;
        const v3 lin = xyz(w0) + lin_delta;
        const v3 ang = xyz(w1) + ang_delta;
        r[0] = f4(lin, w0.w);
        r[1] = f4(ang, w1.w);
        if (last) {
          d.lin[b0 + t] = f4(lin, 0.0f);
          d.ang[b0 + t] = f4(ang, 0.0f);
        }
      }
      r[2] = lds_own[t][2];
      r[3] = lds_own[t][3];
    }
    __syncthreads();  // the next group reuses the LDS
  }
  if (clock) {
    __syncthreads();
    if (threadIdx.x == 0) clock[2 * blockIdx.x + 1] = wall_clock64();
  }
}

// The records after an odd number of per-body passes, back into sbody.
__global__ __launch_bounds__(256) void k_copy_records(const float4* __restrict__ src, float4* __restrict__ dst,
                                                      uint32_t words, const uint32_t* gate) {
  if (gate && *reinterpret_cast<const volatile uint32_t*>(gate) == 0u) return;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < words) dst[i] = src[i];
}

// solve_apply (collision.wgsl:1050-1069); skipped when no contact is active
// (finalize_active_contacts, collision.wgsl:744). Each body sums the Q20 deltas
// in its slots (wrapping 32-bit sums: the arithmetic of atomicAdd on i32, in any
// order) and adds them to its velocities. The last pass also writes
// the body velocities back to the state arrays.
//
__global__ __launch_bounds__(256) void k_apply(Dev d, int last, uint64_t* clock, const uint32_t* gate) {
  if (gate && *reinterpret_cast<const volatile uint32_t*>(gate) == 0u) return;
  apply_contact_deltas(d, last, clock, nullptr, blockIdx.x, gridDim.x);
}

// Inputs of the learned warm start for one active contact, from the state the
// first solver pass sees: relative normal and tangential speed, penetration,
// bias, both inverse masses, both inverse inertias along the contact frame
// (n, t1, t2), the normal denominator, the cached impulse, whether the cache
// matched, and whether the contact is against the ground. Also returns the
// tangent basis.
__device__ __forceinline__ void contact_features(const Dev& d, uint32_t ai, float f[kFeatures], v3& e1, v3& e2) {
  const SolverRow sr = load_row(d, ai);
  const uint32_t body_a = sr.body_a, body_b = sr.body_b;
  const bool has_b = body_b != kInvalid;
  const v3 normal = sr.normal, arm_a = sr.arm_a, arm_b = sr.arm_b;
  const SolverBody sa = load_body(d, body_a);
  const SolverBody sb = has_b ? load_body(d, body_b) : sa;
  const v3 vb = has_b ? sb.lin + cross(sb.ang, arm_b) : mk(0.0f, 0.0f, 0.0f);
  const v3 relative = vb - (sa.lin + cross(sa.ang, arm_a));
  const float normal_speed = dot(relative, normal);
  const float tangent_speed = length(relative - normal * normal_speed);
  const bool analytic = sr.analytic;
  const float bias =
      fmn((fmx(sr.penetration - kSlop, 0.0f) * (analytic ? 1.0f : 0.2f)) / d.delta_seconds, kMaxCorrection);
  float t1[3], t2[3];
  tangent_basis(normal.x, normal.y, normal.z, t1, t2);
  e1 = mk(t1[0], t1[1], t1[2]);
  e2 = mk(t2[0], t2[1], t2[2]);
  f[0] = normal_speed;
  f[1] = tangent_speed;
  f[2] = sr.penetration;
  f[3] = bias;
  f[4] = sa.inv_mass;
  f[5] = has_b ? sb.inv_mass : 0.0f;
  f[6] = dot(inv_inertia(sa, normal), normal);
  f[7] = dot(inv_inertia(sa, e1), e1);
  f[8] = dot(inv_inertia(sa, e2), e2);
  f[9] = has_b ? dot(inv_inertia(sb, normal), normal) : 0.0f;
  f[10] = has_b ? dot(inv_inertia(sb, e1), e1) : 0.0f;
  f[11] = has_b ? dot(inv_inertia(sb, e2), e2) : 0.0f;
  f[12] = sr.denominator;
  f[13] = d.lambda[ai];
  f[14] = sr.hit ? 1.0f : 0.0f;
  f[15] = has_b ? 0.0f : 1.0f;
}

// Logging: the features of every active contact, row-major.
__global__ __launch_bounds__(256) void k_features(Dev d) {
  const uint64_t tick64 = tick_sv(d);
  uint32_t* diag = d.diag[tick64 & 1];
  const uint32_t count = read_count(&diag[kDiagActive], d.pair_capacity);
  const XcdSplit xs = xcd_split(count);
  for (uint32_t ai = xs.first + threadIdx.x; ai < xs.end; ai += xs.step) {
    float f[kFeatures];
    v3 e1, e2;
    contact_features(d, ai, f, e1, e2);
    float4* o = reinterpret_cast<float4*>(d.feat + (size_t)ai * kFeatures);
#pragma unroll
    for (int k = 0; k < kFeatures / 4; ++k) o[k] = make_float4(f[4 * k], f[4 * k + 1], f[4 * k + 2], f[4 * k + 3]);
  }
}

// Logging: the training targets of every active contact after the last pass --
// the accumulated normal impulse and the friction impulse applied over the
// passes in the contact's tangent basis.
__global__ __launch_bounds__(256) void k_targets(Dev d) {
  const uint64_t tick64 = tick_sv(d);
  uint32_t* diag = d.diag[tick64 & 1];
  const uint32_t count = read_count(&diag[kDiagActive], d.pair_capacity);
  const XcdSplit xs = xcd_split(count);
  for (uint32_t ai = xs.first + threadIdx.x; ai < xs.end; ai += xs.step) {
    const float4 n = d.rows[ai];  // plane 0: the normal
    float t1[3], t2[3];
    tangent_basis(n.x, n.y, n.z, t1, t2);
    const v3 f = xyz(d.fsum[ai]);
    d.finit[ai] = make_float4(d.lambda[ai], dot(f, mk(t1[0], t1[1], t1[2])), dot(f, mk(t2[0], t2[1], t2[2])), 0.0f);
  }
}

// Inference: features in registers, staged through LDS into the MFMA operand
// layout, the network on MFMA, and the predicted initial normal impulse and
// friction impulse written for the first solver pass. Each wave takes 64
// contacts at a time: lane l computes contact l's features, then four groups
// of 16 contacts run through the network.
__global__ __launch_bounds__(256) void k_learned(Dev d) {
  constexpr int kWaves = kBlock / 64;
  __shared__ float sfeat[kWaves][64][kFeatures + 1];
  __shared__ float sout[kWaves][64][4];
  const uint64_t tick64 = tick_sv(d);
  uint32_t* diag = d.diag[tick64 & 1];
  const uint32_t count = read_count(&diag[kDiagActive], d.pair_capacity);
  MlpRegs m;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  mlp_load(d.mlp, lane, m);
  const uint32_t chunks = (count + 63) / 64;
  const uint32_t waves = gridDim.x * kWaves;
  for (uint32_t chunk = blockIdx.x * kWaves + wave; chunk < chunks; chunk += waves) {
    const uint32_t ai = chunk * 64 + lane;
    v3 e1 = mk(0.f, 0.f, 0.f), e2 = e1;
    float f[kFeatures];
    if (ai < count) {
      contact_features(d, ai, f, e1, e2);
    } else {
#pragma unroll
      for (int k = 0; k < kFeatures; ++k) f[k] = 0.0f;
    }
#pragma unroll
    for (int k = 0; k < kFeatures; ++k) sfeat[wave][lane][k] = f[k];
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const uint32_t c = g * 16 + lane % 16, k0 = 4 * (lane / 16);
      const float x[4] = {sfeat[wave][c][k0], sfeat[wave][c][k0 + 1], sfeat[wave][c][k0 + 2], sfeat[wave][c][k0 + 3]};
      float y[3];
      mlp_group(m, x, y);
      if (lane < 16) {
        sout[wave][c][0] = y[0];
        sout[wave][c][1] = y[1];
        sout[wave][c][2] = y[2];
      }
    }
    __builtin_amdgcn_wave_barrier();
    if (ai < count) {
      const float y0 = sout[wave][lane][0], y1 = sout[wave][lane][1], y2 = sout[wave][lane][2];
      if (isfinite(y0) && isfinite(y1) && isfinite(y2)) {
        d.lambda[ai] = fmx(y0, 0.0f);
        d.finit[ai] = f4(e1 * y1 + e2 * y2, 0.0f);
      } else {
        // Fallback for this contact: keep its cached impulse (the cache warm
        // start), flagged for the first pass, and count it.
        d.finit[ai] = make_float4(0.f, 0.f, 0.f, 1.f);
        atomicAdd(&d.learned_state[2], 1u);
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// Normal complementarity residual after the last pass, per active contact:
// r = |max(0, lambda + (bias - v_n) / d) - lambda|, the impulse change an
// unrelaxed projected step would still make (zero when the contact is either
// separating with no impulse or holding exactly its bias speed). Per-block
// partial sums in FP64; the host adds the blocks in order.
__global__ __launch_bounds__(256) void k_residual(Dev d, const uint32_t* gate) {
  __shared__ double part[4][kBlock / 64];
  if (gate && *reinterpret_cast<const volatile uint32_t*>(gate) == 0u) return;
  const uint64_t tick64 = tick_sv(d);
  uint32_t* diag = d.diag[tick64 & 1];
  const uint32_t count = read_count(&diag[kDiagActive], d.pair_capacity);
  double sum_r2 = 0.0, n = 0.0, sum_pen = 0.0, max_pen = 0.0;
  for (uint32_t ai = blockIdx.x * blockDim.x + threadIdx.x; ai < count; ai += gridDim.x * blockDim.x) {
    const SolverRow sr = load_row(d, ai);
    const uint32_t body_a = sr.body_a, body_b = sr.body_b;
    const bool has_b = body_b != kInvalid;
    const v3 normal = sr.normal, arm_a = sr.arm_a, arm_b = sr.arm_b;
    const SolverBody sa = load_body(d, body_a);
    const v3 vb = has_b ? [&] { const SolverBody sb = load_body(d, body_b); return sb.lin + cross(sb.ang, arm_b); }()
                        : mk(0.0f, 0.0f, 0.0f);
    const float normal_speed = dot(vb - (sa.lin + cross(sa.ang, arm_a)), normal);
    const bool analytic = sr.analytic;
    const double bias = (double)fmn((fmx(sr.penetration - kSlop, 0.0f) * (analytic ? 1.0f : 0.2f)) / d.delta_seconds,
                                    kMaxCorrection);
    const double lambda = (double)d.lambda[ai];
    double r = 0.0;
    if (sr.denominator > 0.0f) {
      double next = lambda + (bias - (double)normal_speed) / (double)sr.denominator;
      if (next < 0.0) next = 0.0;
      r = next - lambda;
    }
    sum_r2 += r * r;
    n += 1.0;
    sum_pen += (double)sr.penetration;
    max_pen = (double)sr.penetration > max_pen ? (double)sr.penetration : max_pen;
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    sum_r2 += __shfl_xor(sum_r2, m);
    n += __shfl_xor(n, m);
    sum_pen += __shfl_xor(sum_pen, m);
    const double o = __shfl_xor(max_pen, m);
    max_pen = o > max_pen ? o : max_pen;
  }
  const uint32_t wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63u) == 0) {
    part[0][wave] = sum_r2;
    part[1][wave] = n;
    part[2][wave] = sum_pen;
    part[3][wave] = max_pen;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0, c = 0.0, m = 0.0;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
      a += part[0][w];
      b += part[1][w];
      c += part[2][w];
      m = part[3][w] > m ? part[3][w] : m;
    }
    double* o = d.resid + 4 * (size_t)blockIdx.x;
    o[0] = a;
    o[1] = b;
    o[2] = c;
    o[3] = m;
  }
}

// The learned warm start's judge: the RMS residual after the learned passes,
// from the residual kernel's block partials summed in block order. Above the
// limit, or with any non-finite prediction, the tick's extra passes run.
__global__ void k_learned_judge(Dev d) {
  if (threadIdx.x != 0) return;
  double r2 = 0.0, n = 0.0;
  for (uint32_t b = 0; b < kResidBlocks; ++b) {
    r2 += d.resid[4 * b];
    n += d.resid[4 * b + 1];
  }
  const double rms = n > 0.0 ? sqrt(r2 / n) : 0.0;
  uint32_t* st = d.learned_state;
  const uint32_t nonfinite = *reinterpret_cast<volatile uint32_t*>(&st[2]);
  const uint32_t run = (rms > (double)d.learned_limit || nonfinite > 0u) ? 1u : 0u;
  st[0] = run;
  st[1] += run;
  st[3] += nonfinite;
}

static inline uint32_t grid_for(uint32_t n, uint32_t cap) {
  uint32_t b = (n + kBlock - 1) / kBlock;
  return b < cap ? b : cap;
}

__global__ __launch_bounds__(256) void k_round_i32(const float* values, int32_t* out, uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = round_to_i32(values[i]);
}

void launch_round_i32(const float* values, int32_t* out, uint32_t count) {
  k_round_i32<<<(count + kBlock - 1) / kBlock, kBlock>>>(values, out, count);
}

void launch_prepare(hipStream_t s, const Dev& d) {
  k_prepare<<<grid_for(d.pair_capacity, d.t_prepare_blocks ? d.t_prepare_blocks : 1024u), kBlock, 0, s>>>(d);
}

// The contact passes of encode_collision_passes (device.rs:1351-1424); with
// `bearings`, one bearing projection follows every pass (a mechanism with a
// dynamic root, device.rs:1373-1416).
uint32_t solve_grid(const Dev& d) {
  const uint32_t cap = d.t_solve_blocks ? d.t_solve_blocks : 4096u;
  if (d.t_solve_kind != 1u) return grid_for(d.pair_capacity, cap);
  if (d.body_count == 0) return 0;
  const uint32_t b = (d.body_count + kGroupBodies - 1) / kGroupBodies + kXcds;
  return b < cap ? b : cap;
}
// The velocity update's grid: one block per 64 bodies, at most four blocks per
// CU (one resident round at the kernel's occupancy; the blocks loop over the
// rest), a multiple of the XCD count. Measured on the 100k pile: 1571 blocks
// 8.9 us a pass, 1024 blocks 8.1 us.
constexpr uint32_t kApplyMaxBlocks = 1024;
uint32_t apply_grid(const Dev& d) {
  if (d.t_solve_kind == 1u || d.body_count == 0) return 0;
  const uint32_t g = (d.body_count * kApplyLanes + kBlock - 1) / kBlock + kXcds;
  return g < kApplyMaxBlocks ? g : kApplyMaxBlocks;
}

void launch_solve(hipStream_t s, const Dev& d, int bearings, uint64_t* clock, hipEvent_t* events) {
  const uint32_t iterations = d.solver_iterations < 1 ? 1 : d.solver_iterations;
  const uint32_t sgrid = solve_grid(d), bgrid = apply_grid(d);
  // The learned warm start's fallback: after the passes, the residual is judged
  // on the device and `extra` more passes run only when it is over the limit
  // (every kernel of theirs returns at once otherwise, so the captured graph
  // is the same every tick). Not with bearing projections between passes.
  const bool fallback = learned_fallback_on(d) && !bearings;
  const uint32_t extra = fallback ? (d.learned_extra ? d.learned_extra : iterations) : 0u;
  const uint32_t* gate = d.learned_state;
  if (bearings && d.serial_mech) {
    // device.rs:1352-1416 with serial_mechanism: the parallel warm-start pass,
    // then 12x the passes as serial sweeps (per-contact passes only).
    k_solve<true><<<sgrid, kBlock, 0, s>>>(d, 1, 0, nullptr, nullptr);
    if (bgrid) k_apply<<<bgrid, kBlock, 0, s>>>(d, 0, nullptr, nullptr);
    launch_serial_solve(s, d, iterations * kSerialPassMultiplier);
    return;
  }
  if (d.t_solve_kind != 1u) {
    for (uint32_t it = 0; it < iterations; ++it) {
      const int last = it + 1 == iterations ? 1 : 0;
      uint64_t* pass_clock = clock ? clock + 2 * (size_t)it * sgrid : nullptr;
      // The first pass also groups the tick's contact sides.
      if (events)  // start and stop events carry this dispatch's own timestamps
        hipExtLaunchKernelGGL(it == 0 ? k_solve<true> : k_solve<false>, dim3(sgrid), dim3(kBlock), 0, s,
                              events[2 * it], events[2 * it + 1], 0, d, it == 0 ? 1 : 0, last, pass_clock,
                              static_cast<const uint32_t*>(nullptr));
      else if (it == 0)
        k_solve<true><<<sgrid, kBlock, 0, s>>>(d, 1, last, pass_clock, nullptr);
      else
        k_solve<false><<<sgrid, kBlock, 0, s>>>(d, 0, last, pass_clock, nullptr);
      uint64_t* apply_clock = clock ? clock + 2 * ((size_t)iterations * sgrid + (size_t)it * bgrid) : nullptr;
      // With bearing islands the pass's velocity update and bearing projection are one launch.
      if (!(bearings && launch_apply_islands(s, d, bgrid, last, apply_clock))) {
        if (bgrid) k_apply<<<bgrid, kBlock, 0, s>>>(d, last, apply_clock, nullptr);
        if (bearings) launch_bearing_projection(s, d, 1, last);
      }
    }
    if (fallback) {
      launch_residual(s, d);
      k_learned_judge<<<1, 64, 0, s>>>(d);
      for (uint32_t it = 0; it < extra; ++it) {
        const int last = it + 1 == extra ? 1 : 0;
        k_solve<false><<<sgrid, kBlock, 0, s>>>(d, 0, last, nullptr, gate);
        if (bgrid) k_apply<<<bgrid, kBlock, 0, s>>>(d, last, nullptr, gate);
      }
    }
    return;
  }
  if (d.body_count == 0) return;
  k_side_offsets<<<d.side_tiles, kBlock, 0, s>>>(d);
  k_scatter_sides<<<grid_for(d.pair_capacity, d.t_prepare_blocks ? d.t_prepare_blocks : 1024u), kBlock, 0, s>>>(d);
  float4* records[2] = {d.sbody, d.sbody_alt};
  const uint32_t words = 4 * d.body_count;
  for (uint32_t it = 0; it < iterations; ++it) {
    const int last = it + 1 == iterations ? 1 : 0;
    float4* src = records[it & 1u];
    float4* dst = records[(it + 1u) & 1u];
    uint64_t* pass_clock = clock ? clock + 2 * (size_t)it * sgrid : nullptr;
    if (events)
      hipExtLaunchKernelGGL(k_solve_bodies, dim3(sgrid), dim3(kBlock), 0, s, events[2 * it], events[2 * it + 1], 0, d,
                            static_cast<const float4*>(src), dst, it == 0 ? 1 : 0, last, pass_clock,
                            static_cast<const uint32_t*>(nullptr));
    else
      k_solve_bodies<<<sgrid, kBlock, 0, s>>>(d, src, dst, it == 0 ? 1 : 0, last, pass_clock, nullptr);
    if (bearings) {
      // The bearing projection works on the records the pass just wrote.
      Dev dk = d;
      dk.sbody = dst;
      launch_bearing_projection(s, dk, 1, last);
    }
  }
  if (iterations & 1u)
    k_copy_records<<<(words + kBlock - 1) / kBlock, kBlock, 0, s>>>(d.sbody_alt, d.sbody, words, nullptr);
  if (fallback) {
    launch_residual(s, d);
    k_learned_judge<<<1, 64, 0, s>>>(d);
    for (uint32_t it = 0; it < extra; ++it) {
      const int last = it + 1 == extra ? 1 : 0;
      k_solve_bodies<<<sgrid, kBlock, 0, s>>>(d, records[it & 1u], records[(it + 1u) & 1u], 0, last, nullptr, gate);
    }
    if (extra & 1u)
      k_copy_records<<<(words + kBlock - 1) / kBlock, kBlock, 0, s>>>(d.sbody_alt, d.sbody, words, gate);
  }
}

}  // namespace mh

namespace mh {

void launch_features(hipStream_t s, const Dev& d) {
  k_features<<<grid_for(d.pair_capacity, 2048), kBlock, 0, s>>>(d);
}

void launch_targets(hipStream_t s, const Dev& d) { k_targets<<<grid_for(d.pair_capacity, 2048), kBlock, 0, s>>>(d); }

void launch_mlp_infer(hipStream_t s, const Dev& d) { k_learned<<<1024, kBlock, 0, s>>>(d); }

void launch_residual(hipStream_t s, const Dev& d, const uint32_t* gate) {
  k_residual<<<kResidBlocks, kBlock, 0, s>>>(d, gate);
}

}  // namespace mh
