// Device-side state shared by the step kernels and the host runtime.
//
// All per-body and per-collider state is structure-of-arrays: one array per
// field, each element a 16-byte vector (or a scalar), so a wavefront touches
// one contiguous 1 KiB run per field and every load is a dwordx4.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mh {

constexpr uint32_t kInvalid = 0xffffffffu;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kPosMask = 0x7fffffffu;
constexpr uint32_t kBucketBit = 0x40000000u;  // internal node whose leaves share one Morton code
constexpr uint32_t kBlock = 256;
constexpr uint32_t kSideTile = 2048;  // bodies per tile of the contact-side offset scan
// Contact sides a body can gather without atomics per tick; further sides
// accumulate through the atomic overflow path.
constexpr uint32_t kMaxSlots = 32;      // contact-side delta slots per body (body-major)
constexpr uint32_t kSlotOverflow = 31;  // slot code of a side past the body's slots (atomics instead)
constexpr uint32_t kIslandBodies = 256u;    // bodies of one fused bearing-projection group (one per thread)
constexpr uint32_t kIslandBearings = 51u;   // bearing rows of one such group (5 axis items each, one per thread)
constexpr uint32_t kSerialBearingLimit = 64u;    // device.rs:19: serial sweeps up to this many bearings
constexpr uint32_t kSerialPassMultiplier = 12u;  // device.rs:20: their passes per solver iteration
constexpr uint32_t kRadixTile = 1024;  // keys per radix-sort workgroup
constexpr uint32_t kRadixMaxPasses = 8;  // 8-bit digit passes at most (64-bit keys)
constexpr uint32_t kRadixTicket = kRadixMaxPasses * 256;  // rs_hist word of pass 0's tile ticket
constexpr uint32_t kRadixHistWords = kRadixTicket + kRadixMaxPasses;
constexpr uint32_t kFeatures = 16;      // learned warm-start inputs per contact
constexpr uint32_t kResidBlocks = 512;  // residual kernel grid
// Largest closure component the per-component solve holds in LDS.
constexpr uint32_t kCompMaxBodies = 256, kCompMaxCoords = 256, kCompMaxClosures = 128, kCompMaxPath = 2048;
enum LearnMode : uint32_t { kLearnOff = 0, kLearnLog = 1, kLearnInfer = 2 };

// Diagnostics words (abi.rs:56-73) plus private counters after them.
enum DiagWord : uint32_t {
  kDiagFlags = 0,
  kDiagPairs = 1,
  kDiagContacts = 2,
  kDiagAnchor = 3,
  kDiagAxis = 4,
  kDiagActive = 5,
  kDiagWords = 8,
};

// Private per-tick counters (cleared by the Morton kernel).
enum CounterWord : uint32_t {
  kCountFrontier = 0,
  kCountReserved = 1,
  kCountTop = 2,
  kCountSideTicket = 3,  // tile tickets of the contact-side offset scan
  kCountWalked = 4,      // packets whose walk has finished (the last one forms the item prefix)
  kCounterWords = 16,
};

// Tick state kept on the device so a captured tick can be replayed unchanged.
struct TickState {
  uint64_t tick_index;  // tick being run; its low 32 bits are tick_index of abi.rs:21
  uint64_t pad;
};

struct Dev {
  // ---- sizes and constants ----
  uint32_t body_count;
  uint32_t collider_count;
  uint32_t sort_count;
  uint32_t suppression_count;
  uint32_t pair_capacity;
  uint32_t hash_mask;
  uint32_t solver_iterations;
  uint32_t self_collisions;
  float delta_seconds;
  float gravity_y;
  float linear_damping;
  float angular_damping;
  float friction;
  uint32_t pad0;

  // ---- bodies ----
  float4* pos;
  float4* rot;
  float4* lin;
  float4* ang;
  const float* inv_mass;
  const float4* masses;   // 4 per body: inverse mass, inverse inertia columns (body frame)
  float4* world_masses;   // 3 per body: world inverse inertia columns, inverse mass in [0].w
  const uint32_t* root_flags;
  const uint32_t* components;

  // ---- colliders ----
  const float4* colliders;  // 4 per collider: centre, rotation, half extents, metadata bits
  const uint2* suppressed;
  const uint32_t* supp_off;   // body_count + 1: each body's suppressed partners in supp_adj
  const uint32_t* supp_adj;   // 2 * suppression_count: partners, sorted per body

  // ---- broadphase ----
  float4* aabbs;          // 2 per collider: min, max
  float4* cposes;         // 5 per collider: (centre, hx) (rotation) (axis x, hy) (axis y, hz) (axis z, body)
  uint64_t* keys;         // sort_count Morton entries (code << 32 | collider)
  float4* inodes;         // 2 per internal node: (min, left child) (max, escape)
  float4* leaves;         // 2 per leaf: (min, collider) (max, escape)
  uint2* children;        // reference encoding of node_children (lbvh.wgsl:275)
  uint2* ranges;          // sorted leaf range (first, last) per internal node
  uint32_t* leaf_body;    // body of the collider at each sorted leaf
  uint32_t* top;          // internal nodes whose leaves span several refit workgroups
  float4* block_box;      // 2 per refit workgroup: box of its 256 leaves
  float4* leaf_scan;      // 4 per leaf: box of its workgroup's leaves up to it (min, max), from it on (min, max)
  uint4* work;            // traversal work items, kItemStage per packet: (packet first leaf, first, last, 0)
  uint32_t work_capacity;
  uint32_t* work_count;   // items recorded per packet
  uint32_t* work_first;   // packets + 1: exclusive prefix of work_count (the last walk workgroup forms it)
  uint32_t* counters;     // per-tick private counters, see CounterWord
  uint32_t* parents;      // reference encoding of node_parents (parent + 1)
  uint2* pairs;

  // ---- contacts and solver ----
  float4* contacts;       // 4 per contact, mh_contact layout
  uint32_t* active;       // contact index per active contact
  float4* rows;           // solver rows: 3 planes of pair_capacity float4, then a u32 plane (SolverRow)
  float* lambda;          // accumulated normal impulse per active contact
  int4* cdelta;           // slot-major: 2 int4 per (slot, body) -- one body's Q20 deltas of a group of sides, last pass
  float4* sbody;          // 4 per body: solver record (velocities, inverse mass and inertia)
  int32_t* deltas;        // 6 per body, Q20 fixed-point velocity deltas (overflow path)
  uint32_t* slot_count;   // delta slots per body this tick (per group of sides; per side for the per-body passes)
  float4* sbody_alt;      // second solver record buffer: per-body passes read one and write the other
  uint2* side_slot;       // per active contact: its sides' slots in their bodies (uncapped)
  uint32_t* side_group;   // per active contact: group representatives and delta slots of its sides
  uint32_t* side_off;     // per body: first of its contact-side rows (exclusive scan of slot_count)
  uint32_t* side_status;  // per kSideTile bodies: look-back status of the offset scan
  uint32_t side_tiles;
  float4* srow;           // 4 per contact side (2 * pair_capacity sides): side rows, each body's sides contiguous
  float* slam;            // accumulated normal impulse per contact side (both sides of a contact agree)

  // ---- persistent manifolds (ping-pong by tick parity) ----
  uint64_t* htab[2];      // open-addressing index: key -> manifold row
  float4* mrows[2];       // 3 per contact: pair_tick bits, normal/penetration, point/impulse

  // ---- learned warm start and contact residual ----
  uint32_t learn_mode;    // kLearnOff, kLearnLog or kLearnInfer
  uint32_t residual_on;
  float* feat;            // kFeatures per active contact
  float4* fsum;           // friction impulse applied over the passes, per active contact
  float4* finit;          // predicted initial friction impulse, per active contact
  double* resid;          // 4 per residual block: sum r^2, active contacts, sum penetration, max penetration
  const float* mlp;       // network parameters (mlp.hip layout), or null
  float learned_limit;    // RMS residual above which a learned tick runs learned_extra more passes (0: never)
  uint32_t learned_extra; // extra passes of such a tick
  uint32_t* learned_state;  // [0] extra passes run this tick, [1] ticks that ran them, [2] non-finite
                            // predictions this tick, [3] non-finite predictions since creation

  // ---- scene batches: several independent scenes in one handle ----
  uint32_t scene_count;       // 1 for a single scene
  uint32_t key_index_bits;    // collider-index bits of a sort key (32 for a single scene)
  const uint32_t* cscene;     // scene of each collider (batches only)
  const uint32_t* scene_first;// scene_count + 1: first sorted leaf of each scene (batches only)

  // ---- radix sort of the keys ----
  uint64_t* keys_alt;         // collider_count: ping-pong buffer
  uint32_t* rs_hist;          // kRadixHistWords: 256 digit counts per pass, then one tile ticket per pass
  uint32_t* rs_status;        // rs_passes * rs_blocks * 256: per-tile digit counts published for look-back
  uint32_t rs_blocks;         // key tiles of kRadixTile
  uint32_t rs_passes;         // 8-bit digit passes over the (scene, code) bits

  // ---- launch shape (tunable; 0 = default) ----
  uint32_t t_sort_kind;       // 0: radix sort, 1: bitonic sort
  uint32_t t_sort_tile;       // bitonic tile width: 512, 1024 or 2048 keys
  uint32_t t_sort_levels;     // butterfly levels per global sort kernel: 1..4
  uint32_t t_narrow_blocks;   // workgroup cap of the narrowphase
  uint32_t t_prepare_blocks;  // workgroup cap of contact preparation
  uint32_t t_solve_blocks;    // workgroup cap of the solver passes
  uint32_t t_scan_blocks;     // workgroups of k_scan
  uint32_t t_traverse_kind;   // 0: packet walk + k_scan, 1: one walk per leaf, 2: packet walk scanning inline
  uint32_t t_solve_kind;      // 0: per-contact passes + per-body delta gather, 1: per-body passes over side rows

  // ---- articulated path (bearings, coordinates, forward kinematics, closures) ----
  uint32_t bearing_count;
  uint32_t coordinate_count;
  uint32_t closure_count;
  uint32_t mech_count;        // non-root mechanism bodies (mlist)
  uint32_t bearing_body_count;// bodies some bearing references (blist)
  uint32_t level_count;       // depth levels 1..max in mlist
  uint32_t jump_rounds;       // pointer-jump rounds of one forward-kinematics pass
  uint32_t has_dynamic_root;  // a root with inverse mass > 0 exists
  uint32_t serial_mech;       // dynamic root and <= 64 bearings: serial contact and bearing sweeps
  uint4* serial_keys;         // 2 per active contact: (body a, body b, manifold slot, low), (high, active index)
  uint32_t* serial_order;     // active contacts in the serial sweep order
  const float4* bearings;     // 5 per bearing: anchor a, anchor b, axis a, axis b, metadata bits
  const float4* mbodies;      // 4 per body (mh_mechanism_body): metadata, traversal, bind position, bind rotation
  float2* coords;             // angle, angular velocity per coordinate
  const float4* drives;       // 2 per coordinate (mh_mechanism_drive)
  float4* links[2];           // 3 per body (mh_link_state): position, rotation, metadata bits
  int32_t* mdeltas;           // 6 Q20 integers per body
  const uint32_t* mlist;      // non-root mechanism bodies sorted by (depth, preorder)
  const uint32_t* level_offset;  // level_count + 1 offsets into mlist
  const uint32_t* fk_list;    // non-root mechanism bodies grouped by tree
  const uint32_t* fk_offset;  // fk_groups + 1 offsets into fk_list: whole trees per forward-kinematics workgroup
  uint32_t fk_groups;
  const uint32_t* fk_meta;    // per fk_list entry: depth << 16 | its parent's index in the group (0xffff: a root)
  uint32_t fk_small;          // every group holds at most 256 bodies (publication per group)
  uint32_t* mech_state;       // [0]: this tick's closure Newton loop has converged
  const uint32_t* blist;      // bodies referenced by bearings
  // Bearing islands in groups for the fused state-array projection passes
  // (bi_groups 0: one launch pair per pass instead).
  uint32_t bi_groups;
  const uint32_t* bi_offset;  // bi_groups + 1 offsets into bi_bearing / bi_side
  const uint32_t* bi_bearing; // bearing rows, grouped
  const uint32_t* bi_side;    // group-local body of side a | of side b << 16
  const uint32_t* bi_boff;    // bi_groups + 1 offsets into bi_body
  const uint32_t* bi_body;    // group bodies; bit 31: this group writes the body back
  const uint8_t* bi_mask;     // per body: 1 when an island group holds it
  const uint32_t* closure_rows;  // bearing row of each closure
  const uint32_t* cpath_offset;  // 2 * closure_count + 1 offsets: branch a, branch b of each closure
  const uint32_t* cpath;         // coordinates along each branch, walking to the root
  const uint32_t* coord_offset;  // coordinate_count + 1 offsets into coord_entries
  const uint32_t* coord_entries; // (closure << 1 | side) in (closure, side) order
  const uint32_t* coord_body;    // child body of each coordinate's tree bearing
  float4* cl_state;           // 5 per closure: anchor a, axis a, anchor b, axis b, (residual active flag)
  float4* cl_product;         // 2 per closure: J d (position, axis)
  float4* joints;             // 2 per coordinate: joint axis, joint anchor (component-root frame)
  float* pcg;                 // 8 per coordinate: solution, residual, preconditioned, direction,
                              // operator product, gradient, diagonal, unused
  // Per-component closure solve (closure_components): per component two uint4,
  // (first path body, path bodies, first coordinate, coordinates) and (first
  // closure, closures, 0, 0), into the flattened lists below. Bodies: global
  // id, local parent (kInvalid: a root), local coordinate; coordinates: global
  // id, local child body, (closure << 1 | side) entries in (closure, side)
  // order; closures: global closure index, local end body per side (kInvalid: a
  // root), local coordinates along each side's path to the root.
  uint32_t cc_count, cc_max_bodies, cc_max_coords, cc_max_closures, cc_max_path;
  uint32_t cc_dense;  // one-wavefront components hold J^T J densely in LDS for the PCG products
  const uint4* cc_off;
  const uint32_t *cc_body, *cc_body_parent, *cc_body_coord, *cc_coord, *cc_coord_body, *cc_entry_off, *cc_entries,
      *cc_closure, *cc_end, *cc_path_off, *cc_path;
  uint32_t closure_steps;     // Newton steps per tick (the reference's 12 by default)
  uint32_t closure_iters;     // PCG iterations per Newton step (the reference's 8 by default)
  float closure_tol;          // PCG stop when sqrt(r.z / r0.z) falls below it (0: run every iteration)
  uint32_t closure_vproj;     // project coordinate speeds onto J v = 0 before reconstruction (component solve)
  uint32_t pad_closure;

  // ---- publication ----
  float4* snap_pos[3];
  float4* snap_rot[3];

  // ---- diagnostics ----
  uint32_t* diag[2];      // kDiagWords each, by tick parity
  TickState* tick;
};

// Work split for grid-stride kernels on gfx950: workgroups are dispatched to the
// 8 XCDs round-robin by block index, so block b runs on XCD b % 8. Giving each
// XCD a contiguous eighth of the (spatially ordered) items keeps the bodies it
// touches in its own L2.
constexpr uint32_t kXcds = 8;
struct XcdSplit {
  uint32_t begin, end;   // this XCD's item range
  uint32_t first, step;  // first item of this block, block stride within the XCD
};
__device__ __forceinline__ XcdSplit xcd_split(uint32_t n) {
  const uint32_t xcd = blockIdx.x % kXcds;
  const uint32_t local = blockIdx.x / kXcds;
  const uint32_t blocks = (gridDim.x - xcd + kXcds - 1) / kXcds;
  XcdSplit s;
  s.begin = (uint32_t)(((uint64_t)n * xcd) / kXcds);
  s.end = (uint32_t)(((uint64_t)n * (xcd + 1)) / kXcds);
  s.first = s.begin + local * blockDim.x;
  s.step = blocks * blockDim.x;
  return s;
}

// Plain grid order for kernels that append what they produce (narrowphase,
// contact preparation): all XCDs work on neighbouring items at any moment, so
// the appended rows come out in nearly the items' own (spatial) order, and the
// solver's contiguous XCD eighths of them stay spatially compact. Split by XCD
// instead, the eight streams of appends interleave rows from eight regions.
__device__ __forceinline__ XcdSplit grid_split(uint32_t n) {
  XcdSplit s;
  s.begin = 0;
  s.end = n;
  s.first = blockIdx.x * blockDim.x;
  s.step = gridDim.x * blockDim.x;
  return s;
}

// Launchers (one per kernel group), all on `stream`.
void launch_set_tick(hipStream_t s, const Dev& d, uint64_t tick);
void launch_end_tick(hipStream_t s, const Dev& d);
void launch_integrate(hipStream_t s, const Dev& d, bool collisions, bool snapshot_now);
void launch_snapshot(hipStream_t s, const Dev& d);
void launch_external_impulse(hipStream_t s, const Dev& d, float4 point, float4 impulse, uint32_t body);
void launch_morton(hipStream_t s, const Dev& d);
void launch_sort(hipStream_t s, const Dev& d);
void launch_lbvh(hipStream_t s, const Dev& d);
void launch_traverse(hipStream_t s, const Dev& d);
void launch_narrowphase(hipStream_t s, const Dev& d);
void launch_prepare(hipStream_t s, const Dev& d);
// round_to_i32 over `count` device values (the conversion's self-test).
void launch_round_i32(const float* values, int32_t* out, uint32_t count);
// With `clock`, the pass kernels stamp the device wall clock per workgroup:
// pass i of k_solve at clock + 2 * i * solve_grid(d) (start, end per block),
// pass i of k_apply at clock + 2 * (passes * solve_grid(d) + i * apply_grid(d)).
// With `events` (2 per pass), pass i's solve kernel is launched with
// events[2 i] and events[2 i + 1] as its start and stop events: they carry the
// dispatch's own start and end timestamps.
void launch_solve(hipStream_t s, const Dev& d, int bearings, uint64_t* clock = nullptr, hipEvent_t* events = nullptr);
uint32_t solve_grid(const Dev& d);
uint32_t apply_grid(const Dev& d);
void launch_features(hipStream_t s, const Dev& d);
void launch_targets(hipStream_t s, const Dev& d);
// Articulated path (kernels_mechanism.hip). `in_contact`: velocities live in the
// solver records during the contact solve; `write_back` also stores them to the
// state arrays.
void launch_bearing_projection(hipStream_t s, const Dev& d, int in_contact, int write_back);
// `passes` projection passes on the state arrays (in_contact 0, write_back 1).
void launch_bearing_passes(hipStream_t s, const Dev& d, uint32_t passes);
// A contact pass's velocity update fused with its bearing projection when the
// bearings form island groups; false (nothing launched) otherwise.
bool launch_apply_islands(hipStream_t s, const Dev& d, uint32_t apply_blocks, int last, uint64_t* clock);
// The serial small-mechanism sweeps after the parallel warm-start pass:
// passes - 1 contact sweeps, each bearing sweep after the warm start and after
// every contact sweep, then the final velocities and accumulated impulses.
void launch_serial_solve(hipStream_t s, const Dev& d, uint32_t passes);
void launch_advance_coordinates(hipStream_t s, const Dev& d);
void launch_mechanism_pose(hipStream_t s, const Dev& d, int closures);
// LDS bytes of the per-component closure kernel for `d`'s largest component;
// raises the kernel's dynamic LDS limit when it needs more than 64 KiB.
size_t closure_component_lds(const Dev& d);
hipError_t prepare_closure_components(const Dev& d);
void launch_capture_coordinates(hipStream_t s, const Dev& d);
void launch_reconstruct(hipStream_t s, const Dev& d);
void launch_world_masses(hipStream_t s, const Dev& d);
void launch_validate_snapshot(hipStream_t s, const Dev& d);
void launch_validate_bearings(hipStream_t s, const Dev& d);
void launch_mlp_infer(hipStream_t s, const Dev& d);
// With `gate`, the kernel returns at once unless *gate is non-zero.
void launch_residual(hipStream_t s, const Dev& d, const uint32_t* gate = nullptr);
// Whether a tick's solve ends with the learned warm start's residual check.
inline bool learned_fallback_on(const Dev& d) { return d.learn_mode == kLearnInfer && d.learned_limit > 0.0f; }

}  // namespace mh

struct mh_mlp;

namespace mh {

// Device parameters of a learned warm-start network (mlp.hip).
const float* mlp_device_net(const mh_mlp* m);

// Tangent basis of a contact normal: t1 = normalize(n x e), t2 = n x t1, with e
// the x axis unless n is within ~25 degrees of it (then the y axis).
__device__ __forceinline__ void tangent_basis(float nx, float ny, float nz, float* t1, float* t2) {
  float ex = 1.0f, ey = 0.0f;
  if (fabsf(nx) > 0.9f) {
    ex = 0.0f;
    ey = 1.0f;
  }
  // n x e with e = (ex, ey, 0)
  float ax = ny * 0.0f - nz * ey, ay = nz * ex - nx * 0.0f, az = nx * ey - ny * ex;
  float inv = 1.0f / sqrtf((ax * ax + ay * ay) + az * az);
  t1[0] = ax * inv;
  t1[1] = ay * inv;
  t1[2] = az * inv;
  t2[0] = ny * t1[2] - nz * t1[1];
  t2[1] = nz * t1[0] - nx * t1[2];
  t2[2] = nx * t1[1] - ny * t1[0];
}

}  // namespace mh
