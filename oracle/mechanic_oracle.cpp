// ORACLE — test infrastructure only.
//
// A CPU restatement of the reference's per-tick rigid-body step
// (GpuPhysics::dispatch_tick, crates/mechanic-gpu/src/device.rs:859-961, and the
// WGSL kernels it encodes), articulated scenes included. Only tests/, the smoke check and the bench's
// CPU-baseline leg load it, always as the checker; the product never does.
//
// Scalar type: `Real` is float for the FP32 parity oracle and double for the FP64
// accuracy oracle. Build with -ffp-contract=off and without -ffast-math so each
// float expression rounds once per operation, in source order; every expression
// below keeps the evaluation order of the WGSL line it cites.
//
// Semantics that the reference leaves to scheduling are fixed here:
//  * broadphase pairs and contacts are produced in leaf / pair order (the GPU
//    appends them with atomics, so only their set is defined);
//  * the persistent-manifold cache is an exact map keyed by the ordered collider
//    pair. The reference's open-addressing table (collision.wgsl:499-538) can
//    reclaim a stale slot that precedes a pair's own slot on its probe path, and
//    then that pair loses its warm start in an order-dependent way; the map never
//    does. With no stale slot on the path both agree;
//  * the closure solve's gradient / diagonal sums and its PCG dot products
//    accumulate in FP64 (the reference adds FP32 values with compare-and-swap
//    atomics in scheduling order, closure.wgsl:94-126, and runs its dots in FP32).
//
// Pinning: see tests/test_oracle.py (closed forms, brute force, the reference's
// own SAT known-answer vectors, collision.rs:199-232).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <type_traits>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

constexpr uint32_t INVALID = 0xffffffffu;
constexpr uint32_t PAIR_OVERFLOW = 1u, INVALID_NUMERIC = 2u, MANIFOLD_OVERFLOW = 8u;

template <typename R>
struct V3 {
  R x, y, z;
};
template <typename R>
struct V4 {
  R x, y, z, w;
};

template <typename R> V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <typename R> V3<R> xyz(const V4<R>& v) { return V3<R>{v.x, v.y, v.z}; }
template <typename R> V3<R> add(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R> V3<R> sub(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R> V3<R> neg(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <typename R> V3<R> mul(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename R> V3<R> divs(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
// WGSL dot: (x*x' + y*y') + z*z'.
template <typename R> R dot(V3<R> a, V3<R> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <typename R> R dot4(V4<R> a, V4<R> b) { return ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w; }
template <typename R> V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename R> R mx(R a, R b) { return a > b ? a : b; }
template <typename R> R mn(R a, R b) { return a < b ? a : b; }
template <typename R> R length(V3<R> a) { return std::sqrt(dot(a, a)); }
// Box-union min/max ordering -0 below +0, so a union does not depend on the
// grouping (WGSL leaves min/max of signed zeros to the implementation).
template <typename R> R umin(R a, R b) { return a < b ? a : (b < a ? b : (std::signbit(a) ? a : b)); }
template <typename R> R umax(R a, R b) { return a > b ? a : (b > a ? b : (std::signbit(a) ? b : a)); }
template <typename R> R inverse_sqrt(R x) { return R(1) / std::sqrt(x); }
template <typename R> R sgn(R x) { return x > R(0) ? R(1) : (x < R(0) ? R(-1) : R(0)); }
template <typename R> V3<R> vabs(V3<R> a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }

// quat_multiply, quat_rotate (physics.wgsl:46-56).
template <typename R> V4<R> quat_multiply(V4<R> a, V4<R> b) {
  V3<R> av = xyz(a), bv = xyz(b);
  V3<R> v = add(add(mul(bv, a.w), mul(av, b.w)), cross(av, bv));
  return {v.x, v.y, v.z, a.w * b.w - dot(av, bv)};
}
template <typename R> V3<R> quat_rotate(V4<R> q, V3<R> v) {
  V3<R> qv = xyz(q);
  V3<R> c = cross(qv, v);
  V3<R> t = {R(2) * c.x, R(2) * c.y, R(2) * c.z};
  return add(add(v, mul(t, q.w)), cross(qv, t));
}
template <typename R> V4<R> conj(V4<R> q) { return {-q.x, -q.y, -q.z, q.w}; }

template <typename R> bool finite4(V4<R> v) {
  const R lim = R(3.402823e+38f);
  return v.x == v.x && v.y == v.y && v.z == v.z && v.w == v.w && std::fabs(v.x) < lim && std::fabs(v.y) < lim &&
         std::fabs(v.z) < lim && std::fabs(v.w) < lim;
}

// i32(round(x)) with WGSL's round-half-to-even and a saturating conversion.
template <typename R> int32_t q20(R x) {
  R r = std::nearbyint(x);
  if (!(r == r)) return 0;
  if (r >= R(2147483648.0)) return 2147483647;
  if (r < R(-2147483648.0)) return (int32_t)0x80000000u;
  return (int32_t)r;
}

// Row layouts of the C ABI (abi.rs:75-187), always f32 on the boundary.
struct Transform { float position[4], rotation[4]; };
struct Mass { float inverse_mass[4], ix[4], iy[4], iz[4]; };
struct Collider { float local_center[4], local_rotation[4], half_extents[4]; uint32_t metadata[4]; };
struct ContactRow { uint32_t metadata[4]; float normal_penetration[4], arm_a_impulse[4], arm_b[4]; };

template <typename R>
struct Contact {
  uint32_t body_a, body_b, m_z, m_w;  // metadata (collision.wgsl:27-32)
  V4<R> normal_penetration;
  V4<R> arm_a_impulse;
  V4<R> arm_b;
};

template <typename R>
struct Manifold {
  uint32_t tick, count;
  V4<R> normal_penetration;
  V4<R> point_impulse;
};

template <typename R>
struct Aabb {
  V3<R> mn, mx;
};

template <typename R>
struct ColliderPose {
  V3<R> center;
  V4<R> rotation;
  V3<R> axes[3];
  V3<R> half;
  R half_w, center_w;
  uint32_t body, role;
};

template <typename R>
struct Sat {
  V3<R> normal;
  R penetration;
  uint32_t near_face_axes;
};

// Phases of a tick, for the per-phase wall time of the CPU baseline.
enum Phase {
  kPhaseIntegrate, kPhaseMechanism, kPhaseWorldMasses, kPhaseMorton, kPhaseSort, kPhaseLbvh, kPhaseTraverse,
  kPhaseNarrowphase, kPhasePrepare, kPhaseSolve, kPhaseValidate, kPhaseCount
};

template <typename R>
struct World {
  // configuration (GpuTickConfig, device.rs:866-883)
  uint32_t body_count = 0, collider_count = 0, sort_count = 0;
  uint32_t pair_capacity = 1u << 21, iterations = 8;
  bool collisions = true, self_collisions = true;
  R dt = R(1.0f / 60.0f), gravity = R(-9.81f), lin_damp = R(0.999f), ang_damp = R(0.98f);
  R friction = R(0.05f);
  int threads = 1;
  // bodies
  std::vector<V4<R>> pos, rot, lin, ang;
  std::vector<R> inv_mass;
  std::vector<V4<R>> masses;  // 4 per body
  std::vector<V4<R>> world_masses;  // 3 per body
  std::vector<uint32_t> roots, components;
  std::vector<Collider> colliders;
  std::vector<std::pair<uint32_t, uint32_t>> suppressed;
  // per-tick products
  std::vector<Aabb<R>> aabbs;
  std::vector<uint64_t> keys;  // sorted (code << 32 | collider)
  std::vector<std::pair<uint32_t, uint32_t>> children;
  std::vector<uint32_t> parents;
  std::vector<Aabb<R>> node_aabbs;
  std::vector<std::pair<uint32_t, uint32_t>> pairs;
  std::vector<Contact<R>> contacts;
  std::vector<uint32_t> active;
  std::vector<int32_t> deltas;
  uint32_t diag[8] = {};
  // persistent manifolds keyed by the ordered collider pair (low << 32 | high)
  std::unordered_map<uint64_t, Manifold<R>> manifolds;
  std::vector<std::vector<std::pair<uint32_t, uint32_t>>> pair_bufs;
  // articulated rows (device.rs:1684-1808)
  struct BearingRow { V3<R> anchor_a, anchor_b, axis_a, axis_b; uint32_t a, b, coordinate, closure; };
  struct MechBody { uint32_t parent, tree_bearing, direction, root; V3<R> bind_position; V4<R> bind_rotation; };
  struct Drive { uint32_t mode; R max_acceleration, max_speed, target_speed, target_angle, min_angle, max_angle; };
  struct Link { V4<R> position, rotation; uint32_t ancestor, resolved; };
  uint32_t bearing_count = 0, coordinate_count = 0, max_depth = 0, jump_rounds = 0;
  bool mech_active = false, has_dynamic_root = false;
  std::vector<BearingRow> bearings;
  std::vector<MechBody> mbodies;
  std::vector<uint32_t> preorder;
  std::vector<R> coord_angle, coord_speed;
  std::vector<Drive> drives;
  std::vector<Link> links_a, links_b;
  std::vector<int32_t> mdeltas;
  double phase_ns[kPhaseCount] = {};  // accumulated wall time per tick phase
  // Closure Newton solve: steps per tick (device.rs:977), PCG iterations per
  // step (closure.wgsl:358) and the optional PCG tolerance.
  uint32_t closure_steps = 12, closure_iters = 8;
  float closure_tol = 0.0f;
  bool closure_components = false;  // one PCG per closure component instead of one global system
  bool closure_vproj = false;       // project coordinate speeds onto J v = 0 before reconstruction
  uint64_t pcg_iterations_run = 0;  // PCG iterations run since creation
  // Largest bearing count that takes the serial small-mechanism sweeps (the
  // reference's 64; tests lower it to run the same scene without them).
  uint32_t serial_limit = 64;  // device.rs:19
};

template <typename R> V4<R> load4(const float* f) { return {R(f[0]), R(f[1]), R(f[2]), R(f[3])}; }

template <typename R> ColliderPose<R> collider_pose(const World<R>& w, uint32_t c) {
  const Collider& row = w.colliders[c];
  ColliderPose<R> p;
  p.body = row.metadata[0];
  p.role = row.metadata[3];
  V4<R> lc = load4<R>(row.local_center), lr = load4<R>(row.local_rotation), he = load4<R>(row.half_extents);
  // collider_center / collider_rotation (collision.wgsl:164-173)
  p.center = add(xyz(w.pos[p.body]), quat_rotate(w.rot[p.body], xyz(lc)));
  p.rotation = quat_multiply(w.rot[p.body], lr);
  p.axes[0] = quat_rotate(p.rotation, v3<R>(1, 0, 0));
  p.axes[1] = quat_rotate(p.rotation, v3<R>(0, 1, 0));
  p.axes[2] = quat_rotate(p.rotation, v3<R>(0, 0, 1));
  p.half = xyz(he);
  p.half_w = he.w;
  p.center_w = lc.w;
  return p;
}

template <typename R> R comp(V3<R> v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// ---- integrate (physics.wgsl:92-132) -------------------------------------
template <typename R> void integrate(World<R>& w) {
  uint32_t flags = 0;
#pragma omp parallel for num_threads(w.threads) reduction(| : flags) schedule(static)
  for (int64_t ii = 0; ii < (int64_t)w.body_count; ++ii) {
    size_t i = (size_t)ii;
    V4<R> position = w.pos[i], rotation = w.rot[i], linear = w.lin[i], angular = w.ang[i];
    if (w.inv_mass[i] > R(0)) {
      linear.y += w.gravity * w.dt;
      linear.x = linear.x * w.lin_damp;
      linear.y = linear.y * w.lin_damp;
      linear.z = linear.z * w.lin_damp;
      angular.x = angular.x * w.ang_damp;
      angular.y = angular.y * w.ang_damp;
      angular.z = angular.z * w.ang_damp;
      if (w.roots[i] != 0u) {
        position.x = position.x + linear.x * w.dt;
        position.y = position.y + linear.y * w.dt;
        position.z = position.z + linear.z * w.dt;
        V4<R> spin = {angular.x, angular.y, angular.z, R(0)};
        V4<R> q = quat_multiply(spin, rotation);
        R h = R(0.5) * w.dt;
        rotation = {rotation.x + q.x * h, rotation.y + q.y * h, rotation.z + q.z * h, rotation.w + q.w * h};
        R n2 = dot4(rotation, rotation);
        if (n2 > R(1.0e-20f)) {
          R r = inverse_sqrt(n2);
          rotation = {rotation.x * r, rotation.y * r, rotation.z * r, rotation.w * r};
        } else {
          rotation = {R(0), R(0), R(0), R(1)};
          flags |= INVALID_NUMERIC;
        }
      }
    }
    if (!(finite4(position) && finite4(rotation) && finite4(linear) && finite4(angular))) {
      flags |= INVALID_NUMERIC;
      continue;
    }
    w.pos[i] = position;
    w.rot[i] = rotation;
    w.lin[i] = linear;
    w.ang[i] = angular;
  }
  w.diag[0] |= flags;
}

// ---- update_world_masses (collision.wgsl:134-162) ------------------------
template <typename R> V3<R> world_inverse_inertia_column(const World<R>& w, uint32_t b, V3<R> axis) {
  V4<R> rotation = w.rot[b];
  V3<R> local = quat_rotate(conj(rotation), axis);
  const V4<R>* m = &w.masses[4 * (size_t)b];
  V3<R> lr = add(add(mul(xyz(m[1]), local.x), mul(xyz(m[2]), local.y)), mul(xyz(m[3]), local.z));
  return quat_rotate(rotation, lr);
}

template <typename R> void update_world_masses(World<R>& w) {
#pragma omp parallel for num_threads(w.threads) schedule(static)
  for (int64_t ii = 0; ii < (int64_t)w.body_count; ++ii) {
    uint32_t b = (uint32_t)ii;
    V3<R> cx = world_inverse_inertia_column(w, b, v3<R>(1, 0, 0));
    V3<R> cy = world_inverse_inertia_column(w, b, v3<R>(0, 1, 0));
    V3<R> cz = world_inverse_inertia_column(w, b, v3<R>(0, 0, 1));
    w.world_masses[3 * (size_t)b] = {cx.x, cx.y, cx.z, w.masses[4 * (size_t)b].x};
    w.world_masses[3 * (size_t)b + 1] = {cy.x, cy.y, cy.z, R(0)};
    w.world_masses[3 * (size_t)b + 2] = {cz.x, cz.y, cz.z, R(0)};
  }
}

// ---- compute_morton (lbvh.wgsl:84-130) ------------------------------------
inline uint32_t expand_bits(uint32_t v) {
  v &= 0x000003ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
template <typename R> uint32_t quantize(R c) { return (uint32_t)mn<R>(mx<R>(std::floor(c + R(512)), R(0)), R(1023)); }
template <typename R> uint32_t morton_code(V3<R> c) {
  return expand_bits(quantize(c.x)) | (expand_bits(quantize(c.y)) << 1) | (expand_bits(quantize(c.z)) << 2);
}

template <typename R> void compute_morton(World<R>& w) {
  w.keys.assign(w.sort_count, 0);
#pragma omp parallel for num_threads(w.threads) schedule(static)
  for (int64_t ii = 0; ii < (int64_t)w.sort_count; ++ii) {
    uint32_t i = (uint32_t)ii;
    if (i >= w.collider_count) {
      w.keys[i] = (0xffffffffull << 32) | i;
      continue;
    }
    ColliderPose<R> p = collider_pose(w, i);
    V3<R> we = add(add(mul(vabs(p.axes[0]), p.half.x), mul(vabs(p.axes[1]), p.half.y)), mul(vabs(p.axes[2]), p.half.z));
    w.aabbs[i] = {sub(p.center, we), add(p.center, we)};
    w.keys[i] = ((uint64_t)morton_code(p.center) << 32) | i;
  }
}

// ---- sort: the bitonic network sorts unique keys ascending (lbvh.wgsl:132-217)
template <typename R> void sort_keys(World<R>& w) { std::sort(w.keys.begin(), w.keys.end()); }

// ---- build_topology (lbvh.wgsl:219-278) -----------------------------------
template <typename R> int common_prefix(const World<R>& w, int left, int right) {
  int count = (int)w.collider_count;
  if (right < 0 || right >= count) return -1;
  uint32_t lc = (uint32_t)(w.keys[left] >> 32), rc = (uint32_t)(w.keys[right] >> 32);
  uint32_t li = (uint32_t)w.keys[left], ri = (uint32_t)w.keys[right];
  uint32_t diff = lc ^ rc;
  if (diff != 0u) return __builtin_clz(diff);
  return 32 + __builtin_clz(li ^ ri);
}

template <typename R> void build_topology(World<R>& w) {
  const uint32_t c = w.collider_count;
  w.children.assign(c > 1 ? c - 1 : 0, {0, 0});
  w.parents.assign(c > 0 ? 2 * (size_t)c - 1 : 0, 0);
  if (c <= 1) return;
  for (uint32_t node = 0; node < c - 1; ++node) {
    int index = (int)node;
    int direction = common_prefix(w, index, index + 1) > common_prefix(w, index, index - 1) ? 1 : -1;
    int minimum_prefix = common_prefix(w, index, index - direction);
    int maximum_length = 2;
    while (common_prefix(w, index, index + maximum_length * direction) > minimum_prefix) maximum_length *= 2;
    int length = 0;
    int step = maximum_length / 2;
    while (step >= 1) {
      if (common_prefix(w, index, index + (length + step) * direction) > minimum_prefix) length += step;
      step /= 2;
    }
    int other = index + length * direction;
    int node_prefix = common_prefix(w, index, other);
    int split_offset = 0;
    step = (length + 1) / 2;
    while (step >= 1) {
      if (common_prefix(w, index, index + (split_offset + step) * direction) > node_prefix) split_offset += step;
      if (step == 1) break;
      step = (step + 1) / 2;
    }
    int split = index + split_offset * direction + std::min(direction, 0);
    int first = std::min(index, other), last = std::max(index, other);
    uint32_t leaf_base = c - 1;
    uint32_t left_child = split == first ? leaf_base + (uint32_t)split : (uint32_t)split;
    uint32_t right_child = split + 1 == last ? leaf_base + (uint32_t)(split + 1) : (uint32_t)(split + 1);
    w.children[node] = {left_child, right_child};
    w.parents[left_child] = node + 1;
    w.parents[right_child] = node + 1;
  }
}

// ---- prepare_leaves + build_bounds (lbvh.wgsl:280-319): min/max are exact, so
// a serial climb with arrival counters gives the same boxes as the atomic one.
template <typename R> void build_bounds(World<R>& w) {
  const uint32_t c = w.collider_count;
  w.node_aabbs.assign(c > 0 ? 2 * (size_t)c - 1 : 0, Aabb<R>{});
  for (uint32_t leaf = 0; leaf < c; ++leaf) w.node_aabbs[c - 1 + leaf] = w.aabbs[(uint32_t)w.keys[leaf]];
  if (c <= 1) return;
  std::vector<uint32_t> visits(c - 1, 0);
  for (uint32_t leaf = 0; leaf < c; ++leaf) {
    uint32_t current = c - 1 + leaf;
    while (true) {
      uint32_t encoded = w.parents[current];
      if (encoded == 0) break;
      uint32_t parent = encoded - 1;
      uint32_t arrival = visits[parent]++;
      if (arrival == 0) break;
      if (arrival > 1) {
        w.diag[0] |= PAIR_OVERFLOW;
        break;
      }
      const Aabb<R>& l = w.node_aabbs[w.children[parent].first];
      const Aabb<R>& r = w.node_aabbs[w.children[parent].second];
      w.node_aabbs[parent] = {{umin(l.mn.x, r.mn.x), umin(l.mn.y, r.mn.y), umin(l.mn.z, r.mn.z)},
                              {umax(l.mx.x, r.mx.x), umax(l.mx.y, r.mx.y), umax(l.mx.z, r.mx.z)}};
      current = parent;
    }
  }
}

// ---- traverse (lbvh.wgsl:321-398) -----------------------------------------
template <typename R> bool overlaps(const Aabb<R>& a, const Aabb<R>& b) {
  return a.mn.x <= b.mx.x && a.mn.y <= b.mx.y && a.mn.z <= b.mx.z && b.mn.x <= a.mx.x && b.mn.y <= a.mx.y &&
         b.mn.z <= a.mx.z;
}

template <typename R> bool pair_is_suppressed(const World<R>& w, uint32_t a, uint32_t b) {
  std::pair<uint32_t, uint32_t> key{std::min(a, b), std::max(a, b)};
  return std::binary_search(w.suppressed.begin(), w.suppressed.end(), key);
}

template <typename R> void traverse(World<R>& w) {
  const uint32_t c = w.collider_count;
  w.pairs.clear();
  uint32_t requested = 0, flags = 0;
  int nt = w.threads;
  w.pair_bufs.assign((size_t)nt, {});
  std::vector<uint32_t> leaf_start(c + 1, 0);
  std::vector<std::vector<uint32_t>> leaf_counts((size_t)nt);
  // Per-leaf candidate lists in leaf order (deterministic for any thread count).
  std::vector<std::vector<std::pair<uint32_t, uint32_t>>> per_leaf(c);
#pragma omp parallel for num_threads(nt) schedule(dynamic, 256) reduction(| : flags)
  for (int64_t ll = 0; ll < (int64_t)c; ++ll) {
    uint32_t leaf = (uint32_t)ll;
    uint32_t collider_index = (uint32_t)w.keys[leaf];
    const Aabb<R> query = w.aabbs[collider_index];
    uint32_t leaf_base = c - 1;
    uint32_t stack[64];
    uint32_t stack_size = 1;
    stack[0] = c > 1 ? 0u : leaf_base;
    while (stack_size != 0) {
      stack_size -= 1;
      uint32_t node = stack[stack_size];
      if (!overlaps(query, w.node_aabbs[node])) continue;
      if (node >= leaf_base) {
        uint32_t candidate = (uint32_t)w.keys[node - leaf_base];
        if (candidate > collider_index) {
          uint32_t body_a = w.colliders[collider_index].metadata[0];
          uint32_t body_b = w.colliders[candidate].metadata[0];
          if (body_a != body_b && !pair_is_suppressed(w, body_a, body_b))
            per_leaf[leaf].push_back({collider_index, candidate});
        }
        continue;
      }
      if (stack_size > 61) {
        flags |= PAIR_OVERFLOW;
        break;
      }
      stack[stack_size] = w.children[node].first;
      stack[stack_size + 1] = w.children[node].second;
      stack_size += 2;
    }
  }
  w.diag[0] |= flags;
  for (uint32_t leaf = 0; leaf < c; ++leaf)
    for (auto& pr : per_leaf[leaf]) {
      uint32_t out = requested++;
      if (out < w.pair_capacity) w.pairs.push_back(pr);
      else w.diag[0] |= PAIR_OVERFLOW;
    }
  w.diag[1] = requested;
}

// ---- narrowphase (collision.wgsl:393-620) ---------------------------------
template <typename R> V3<R> support_point(const ColliderPose<R>& p, V3<R> direction) {
  V3<R> point = p.center;
  for (int a = 0; a < 3; ++a) {
    R projection = dot(p.axes[a], direction);
    if (std::fabs(projection) > R(1.0e-5f)) point = add(point, mul(mul(p.axes[a], comp(p.half, a)), sgn(projection)));
  }
  return point;
}

template <typename R> R projection_radius(const V3<R> (&axes)[3], V3<R> e, V3<R> axis) {
  return (std::fabs(dot(axes[0], axis)) * e.x + std::fabs(dot(axes[1], axis)) * e.y) + std::fabs(dot(axes[2], axis)) * e.z;
}

template <typename R> Sat<R> test_sat_axis(V3<R> delta, const ColliderPose<R>& a, const ColliderPose<R>& b, V3<R> raw) {
  R len2 = dot(raw, raw);
  if (len2 < R(1.0e-10f)) return {v3<R>(1, 0, 0), R(3.402823e+38f), 0u};
  V3<R> axis = mul(raw, inverse_sqrt(len2));
  R signed_distance = dot(delta, axis);
  R penetration = (projection_radius(a.axes, a.half, axis) + projection_radius(b.axes, b.half, axis)) -
                  std::fabs(signed_distance);
  if (signed_distance < R(0)) axis = neg(axis);
  return {axis, penetration, 0u};
}

template <typename R> Sat<R> obb_sat(const ColliderPose<R>& a, const ColliderPose<R>& b) {
  V3<R> delta = sub(b.center, a.center);
  Sat<R> minimum{v3<R>(1, 0, 0), R(3.402823e+38f), 0u};
  uint32_t near_face_axes = 0;
  for (int i = 0; i < 3; ++i) {
    Sat<R> ra = test_sat_axis(delta, a, b, a.axes[i]);
    if (ra.penetration < R(0)) return ra;
    if (ra.penetration <= R(1.0e-4f)) near_face_axes += 1;
    if (ra.penetration < minimum.penetration) minimum = ra;
    Sat<R> rb = test_sat_axis(delta, a, b, b.axes[i]);
    if (rb.penetration < R(0)) return rb;
    if (rb.penetration <= R(1.0e-4f)) near_face_axes += 1;
    if (rb.penetration < minimum.penetration) minimum = rb;
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      Sat<R> r = test_sat_axis(delta, a, b, cross(a.axes[i], b.axes[j]));
      if (r.penetration < R(0)) return r;
      if (r.penetration < minimum.penetration) minimum = r;
    }
  minimum.near_face_axes = near_face_axes;
  return minimum;
}

template <typename R> V3<R> collider_vertex(const ColliderPose<R>& p, uint32_t vertex) {
  V3<R> point = p.center;
  for (int a = 0; a < 3; ++a) {
    R direction = (vertex & (1u << a)) != 0u ? R(1) : R(-1);
    point = add(point, mul(mul(p.axes[a], comp(p.half, a)), direction));
  }
  return point;
}

template <typename R> bool collider_contains_point(const ColliderPose<R>& p, V3<R> point) {
  V3<R> local = quat_rotate(conj(p.rotation), sub(point, p.center));
  return std::fabs(local.x) <= p.half.x + R(1.0e-5f) && std::fabs(local.y) <= p.half.y + R(1.0e-5f) &&
         std::fabs(local.z) <= p.half.z + R(1.0e-5f);
}

template <typename R> V3<R> calculate_contact_point(const ColliderPose<R>& a, const ColliderPose<R>& b, V3<R> normal) {
  V3<R> sum = v3<R>(0, 0, 0);
  R count = R(0);
  for (uint32_t vertex = 0; vertex < 8; ++vertex) {
    V3<R> pa = collider_vertex(a, vertex);
    if (collider_contains_point(b, pa)) {
      sum = add(sum, pa);
      count += R(1);
    }
    V3<R> pb = collider_vertex(b, vertex);
    if (collider_contains_point(a, pb)) {
      sum = add(sum, pb);
      count += R(1);
    }
  }
  if (count > R(0)) return divs(sum, count);
  V3<R> sa = support_point(a, normal);
  V3<R> sb = support_point(b, neg(normal));
  return mul(add(sa, sb), R(0.5));
}

template <typename R> void push_contact(World<R>& w, const Contact<R>& c) {
  uint32_t out = w.diag[2]++;
  if (out < w.pair_capacity) w.contacts.push_back(c);
  else w.diag[0] |= PAIR_OVERFLOW;
}

template <typename R> void narrowphase(World<R>& w) {
  w.contacts.clear();
  const size_t count = std::min<size_t>(w.pairs.size(), w.pair_capacity);
  std::vector<Contact<R>> out(count);
  std::vector<uint8_t> emit(count, 0);
#pragma omp parallel for num_threads(w.threads) schedule(dynamic, 512)
  for (int64_t ii = 0; ii < (int64_t)count; ++ii) {
    auto pr = w.pairs[(size_t)ii];
    ColliderPose<R> a = collider_pose(w, pr.first), b = collider_pose(w, pr.second);
    if (!w.self_collisions && w.components[a.body] == w.components[b.body]) continue;
    Sat<R> sat = obb_sat(a, b);
    if (sat.penetration < R(-1.0e-5f)) continue;
    V3<R> point = calculate_contact_point(a, b, sat.normal);
    Contact<R>& c = out[(size_t)ii];
    c.body_a = a.body;
    c.body_b = b.body;
    c.m_z = pr.first;
    c.m_w = pr.second;
    c.normal_penetration = {sat.normal.x, sat.normal.y, sat.normal.z, mx(sat.penetration, R(0))};
    V3<R> arm_a = sub(point, xyz(w.pos[a.body])), arm_b = sub(point, xyz(w.pos[b.body]));
    c.arm_a_impulse = {arm_a.x, arm_a.y, arm_a.z, R(0)};
    c.arm_b = {arm_b.x, arm_b.y, arm_b.z, R((float)sat.near_face_axes)};
    emit[(size_t)ii] = 1;
  }
  for (size_t i = 0; i < count; ++i)
    if (emit[i]) push_contact(w, out[i]);
}

// ---- generate_ground_contacts (collision.wgsl:193-238, 622-664) -----------
template <typename R> V3<R> normalize(V3<R> v) { return divs(v, length(v)); }

template <typename R> void generate_ground_contacts(World<R>& w) {
  const V3<R> down = v3<R>(0, -1, 0);
  for (uint32_t ci = 0; ci < w.collider_count; ++ci) {
    ColliderPose<R> p = collider_pose(w, ci);
    V3<R> support = support_point(p, down);
    if (p.role != 0u) {
      V3<R> axis = p.axes[1];
      R axial = dot(axis, down);
      R radial_length = length(sub(down, mul(axis, axial)));
      uint32_t count = radial_length < R(0.05f) ? 4u : 1u;
      if (p.role > count) continue;
      V3<R> point = sub(p.center, mul(p.axes[0], p.center_w));
      V3<R> radial_direction = sub(down, mul(axis, axial));
      if (radial_length < R(0.05f)) {
        point = add(point, mul(mul(axis, sgn(axial)), p.half.y));
        V3<R> ta = normalize(sub(v3<R>(1, 0, 0), mul(axis, axis.x)));
        V3<R> tb = normalize(cross(axis, ta));
        if (p.role == 1u) point = add(point, mul(ta, p.half_w));
        else if (p.role == 2u) point = sub(point, mul(ta, p.half_w));
        else if (p.role == 3u) point = add(point, mul(tb, p.half_w));
        else point = sub(point, mul(tb, p.half_w));
      } else {
        point = add(point, mul(radial_direction, p.half_w / radial_length));
        if (std::fabs(axial) >= R(0.05f)) point = add(point, mul(mul(axis, sgn(axial)), p.half.y));
      }
      support = point;
    }
    R bottom = support.y;
    if (bottom > R(1.0e-5f)) continue;
    Contact<R> c;
    c.body_a = p.body;
    c.body_b = INVALID;
    c.m_z = ci;
    c.m_w = INVALID;
    c.normal_penetration = {R(0), R(-1), R(0), mx(-bottom, R(0))};
    V3<R> arm = sub(support, xyz(w.pos[p.body]));
    c.arm_a_impulse = {arm.x, arm.y, arm.z, R(0)};
    c.arm_b = {R(0), R(0), R(0), p.role != 0u ? R(-1) : R(0)};
    push_contact(w, c);
  }
}

// ---- contact solver (collision.wgsl:109-132, 674-906, 1050-1088) ----------
template <typename R> V3<R> inverse_inertia(const World<R>& w, uint32_t b, V3<R> v) {
  const V4<R>* m = &w.world_masses[3 * (size_t)b];
  return add(add(mul(xyz(m[0]), v.x), mul(xyz(m[1]), v.y)), mul(xyz(m[2]), v.z));
}

template <typename R> R impulse_denominator(const World<R>& w, uint32_t a, uint32_t b, V3<R> arm_a, V3<R> arm_b, V3<R> d) {
  R result = w.world_masses[3 * (size_t)a].w;
  V3<R> angular_a = cross(inverse_inertia(w, a, cross(arm_a, d)), arm_a);
  result += dot(angular_a, d);
  if (b != INVALID) {
    result += w.world_masses[3 * (size_t)b].w;
    V3<R> angular_b = cross(inverse_inertia(w, b, cross(arm_b, d)), arm_b);
    result += dot(angular_b, d);
  }
  return result;
}

template <typename R> V3<R> contact_velocity(const World<R>& w, uint32_t b, V3<R> arm) {
  return add(xyz(w.lin[b]), cross(xyz(w.ang[b]), arm));
}

template <typename R> void prepare_contacts(World<R>& w, uint32_t tick) {
  w.active.clear();
  std::unordered_map<uint64_t, Manifold<R>> next;
  next.reserve(w.contacts.size() * 2 + 16);
  for (uint32_t ci = 0; ci < (uint32_t)w.contacts.size(); ++ci) {
    Contact<R>& c = w.contacts[ci];
    V3<R> arm_a = xyz(c.arm_a_impulse), arm_b = xyz(c.arm_b), normal = xyz(c.normal_penetration);
    R inverse_b = R(0);
    V3<R> velocity_b = v3<R>(0, 0, 0);
    if (c.body_b != INVALID) {
      inverse_b = w.world_masses[3 * (size_t)c.body_b].w;
      velocity_b = contact_velocity(w, c.body_b, arm_b);
    }
    if (w.world_masses[3 * (size_t)c.body_a].w + inverse_b <= R(0)) continue;
    V3<R> relative = sub(velocity_b, contact_velocity(w, c.body_a, arm_a));
    R normal_speed = dot(relative, normal);
    if (c.normal_penetration.w <= R(1.0e-6f) && normal_speed >= R(0)) {
      c.m_z = INVALID;
      c.m_w = 0;
      continue;
    }
    const uint64_t key = ((uint64_t)std::min(c.m_z, c.m_w) << 32) | std::max(c.m_z, c.m_w);
    R cached_impulse = R(0);
    auto it = w.manifolds.find(key);
    if (it != w.manifolds.end()) {
      const Manifold<R>& m = it->second;
      bool matches = m.tick == tick - 1u && dot(xyz(m.normal_penetration), normal) >= R(0.98f) &&
                     length(sub(xyz(m.point_impulse), arm_a)) <= R(0.02f);
      cached_impulse = matches ? m.point_impulse.w : R(0);
    }
    c.m_z = ci;  // manifold row
    c.m_w = 1;
    c.arm_a_impulse.w = cached_impulse;
    Manifold<R> m;
    m.tick = tick;
    m.count = 1;
    m.normal_penetration = c.normal_penetration;
    m.point_impulse = {arm_a.x, arm_a.y, arm_a.z, cached_impulse};
    next[key] = m;
    if (c.normal_penetration.w > R(1.0e-5f) || normal_speed < R(-1.0e-5f) || cached_impulse > R(1.0e-6f)) {
      w.diag[5] += 1;
      w.active.push_back(ci);
    }
  }
  // An entry not prepared this tick can never match again (its tick would
  // differ from the next tick - 1), so only this tick's entries are kept.
  w.manifolds = std::move(next);
}

template <typename R> R penetration_bias(const World<R>& w, R penetration, bool analytic) {
  R recovery = analytic ? R(1) : R(0.2f);
  return mn((mx(penetration - R(0.001f), R(0)) * recovery) / w.dt, R(1));
}

template <typename R> void add_delta(int32_t* d, V3<R> linear, V3<R> angular) {
  const R s = R(1048576.0f);
  int32_t v[6] = {q20(linear.x * s), q20(linear.y * s), q20(linear.z * s),
                  q20(angular.x * s), q20(angular.y * s), q20(angular.z * s)};
  for (int k = 0; k < 6; ++k) d[k] = (int32_t)((uint32_t)d[k] + (uint32_t)v[k]);  // wrapping, like atomicAdd
}

// One projected Jacobi pass: warm_start when `first`, solve_accumulate otherwise.
template <typename R> void solve_pass(World<R>& w, bool first) {
  const int nt = w.threads;
  std::vector<std::vector<int32_t>> local((size_t)nt);
#pragma omp parallel num_threads(nt)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    std::vector<int32_t>& acc = local[(size_t)t];
    acc.assign(6 * (size_t)w.body_count, 0);
#pragma omp for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)w.active.size(); ++ii) {
      Contact<R>& c = w.contacts[w.active[(size_t)ii]];
      uint32_t a = c.body_a, b = c.body_b;
      V3<R> arm_a = xyz(c.arm_a_impulse), arm_b = xyz(c.arm_b), normal = xyz(c.normal_penetration);
      V3<R> velocity_b = v3<R>(0, 0, 0);
      if (b != INVALID) velocity_b = contact_velocity(w, b, arm_b);
      R denominator = impulse_denominator(w, a, b, arm_a, arm_b, normal);
      if (denominator <= R(0)) continue;
      V3<R> relative = sub(velocity_b, contact_velocity(w, a, arm_a));
      R normal_speed = dot(relative, normal);
      if (c.normal_penetration.w <= R(1.0e-5f) && normal_speed >= R(-1.0e-5f) && c.arm_a_impulse.w <= R(1.0e-6f)) {
        c.arm_a_impulse.w = R(0);
        continue;
      }
      bool analytic = c.arm_b.w < R(0);
      R bias = penetration_bias(w, c.normal_penetration.w, analytic);
      V3<R> impulse;
      R bound;
      if (first) {
        R warmed = c.arm_a_impulse.w * (analytic ? R(1) : R(0.5f));
        R accumulated = mx(warmed + (R(0.125f) * (-normal_speed + bias)) / denominator, R(0));
        c.arm_a_impulse.w = accumulated;
        impulse = mul(normal, accumulated);
        bound = accumulated * w.friction;
      } else {
        R previous = mx(c.arm_a_impulse.w, R(0));
        R accumulated = mx(previous + (R(0.125f) * (-normal_speed + bias)) / denominator, R(0));
        R delta = accumulated - previous;
        c.arm_a_impulse.w = accumulated;
        impulse = mul(normal, delta);
        bound = std::fabs(delta) * w.friction;
      }
      V3<R> tangent_velocity = sub(relative, mul(normal, normal_speed));
      R tangent_speed = length(tangent_velocity);
      if (tangent_speed > R(1.0e-6f)) {
        V3<R> tangent = divs(tangent_velocity, tangent_speed);
        R tangent_denominator = impulse_denominator(w, a, b, arm_a, arm_b, tangent);
        R friction_impulse = mn(tangent_speed / tangent_denominator, bound);
        impulse = sub(impulse, mul(tangent, friction_impulse));
      }
      V3<R> ni = neg(impulse);
      add_delta(&acc[6 * (size_t)a], mul(ni, w.world_masses[3 * (size_t)a].w), inverse_inertia(w, a, cross(arm_a, ni)));
      if (b != INVALID)
        add_delta(&acc[6 * (size_t)b], mul(impulse, w.world_masses[3 * (size_t)b].w),
                  inverse_inertia(w, b, cross(arm_b, impulse)));
    }
  }
  for (int t = 0; t < nt; ++t) {
    const std::vector<int32_t>& acc = local[(size_t)t];
    if (acc.empty()) continue;
    for (size_t k = 0; k < acc.size(); ++k) w.deltas[k] = (int32_t)((uint32_t)w.deltas[k] + (uint32_t)acc[k]);
  }
}

template <typename R> void solve_apply(World<R>& w) {
  const R s = R(1048576.0f);
#pragma omp parallel for num_threads(w.threads) schedule(static)
  for (int64_t bb = 0; bb < (int64_t)w.body_count; ++bb) {
    size_t b = (size_t)bb;
    int32_t* d = &w.deltas[6 * b];
    // f32(i32) rounds to nearest even; the division by 2^20 is exact.
    V3<R> ld = divs(v3<R>(R(d[0]), R(d[1]), R(d[2])), s);
    V3<R> ad = divs(v3<R>(R(d[3]), R(d[4]), R(d[5])), s);
    for (int k = 0; k < 6; ++k) d[k] = 0;
    V3<R> l = add(xyz(w.lin[b]), ld), a = add(xyz(w.ang[b]), ad);
    w.lin[b] = {l.x, l.y, l.z, R(0)};
    w.ang[b] = {a.x, a.y, a.z, R(0)};
  }
}

// ---- articulated path (articulated.wgsl, mechanism.wgsl, closure.wgsl, bearings.wgsl) ----
constexpr uint32_t NON_CONVERGENCE = 4u;

template <typename R> V3<R> normalize_v(V3<R> v) { return divs(v, length(v)); }
template <typename R> V4<R> normalize_q(V4<R> q) {
  R l = std::sqrt(dot4(q, q));
  return {q.x / l, q.y / l, q.z / l, q.w / l};
}
template <typename R> R clampr(R e, R lo, R hi) { return mn(mx(e, lo), hi); }
template <typename R> bool finite3v(V3<R> v) {
  const R lim = R(3.402823e+38f);
  return v.x == v.x && v.y == v.y && v.z == v.z && std::fabs(v.x) < lim && std::fabs(v.y) < lim &&
         std::fabs(v.z) < lim;
}
template <typename R> uint32_t round_u32(R x) {
  R r = std::nearbyint(x);
  if (!(r > R(0))) return 0u;
  if (r >= R(4294967296.0)) return 0xffffffffu;
  return (uint32_t)r;
}

// world_inverse_inertia (articulated.wgsl:97-105).
template <typename R> V3<R> wii(const World<R>& w, uint32_t b, V3<R> v) {
  V4<R> rotation = w.rot[b];
  V3<R> local = quat_rotate(conj(rotation), v);
  const V4<R>* m = &w.masses[4 * (size_t)b];
  V3<R> r = add(add(mul(xyz(m[1]), local.x), mul(xyz(m[2]), local.y)), mul(xyz(m[3]), local.z));
  return quat_rotate(rotation, r);
}
template <typename R> R mass_inv(const World<R>& w, uint32_t b) { return w.masses[4 * (size_t)b].x; }

template <typename R> void add_mdelta(World<R>& w, uint32_t b, V3<R> linear, V3<R> angular) {
  add_delta(&w.mdeltas[6 * (size_t)b], linear, angular);
}

// solve_linear_axis / solve_angular_axis (articulated.wgsl:118-155).
template <typename R> void bearing_linear(World<R>& w, uint32_t a, uint32_t b, V3<R> arm_a, V3<R> arm_b, V3<R> rel,
                                          V3<R> dir) {
  V3<R> angular_a = cross(arm_a, dir), angular_b = cross(arm_b, dir);
  R den = ((mass_inv(w, a) + mass_inv(w, b)) + dot(angular_a, wii(w, a, angular_a))) + dot(angular_b, wii(w, b, angular_b));
  if (den <= R(1.0e-12f)) return;
  V3<R> impulse = mul(dir, (R(0.5f) * dot(rel, dir)) / den);
  add_mdelta(w, a, mul(impulse, mass_inv(w, a)), wii(w, a, cross(arm_a, impulse)));
  add_mdelta(w, b, mul(neg(impulse), mass_inv(w, b)), wii(w, b, cross(arm_b, neg(impulse))));
}
template <typename R> void bearing_angular(World<R>& w, uint32_t a, uint32_t b, V3<R> rel, V3<R> axis) {
  V3<R> ia = wii(w, a, axis), ib = wii(w, b, axis);
  R den = dot(axis, add(ia, ib));
  if (den <= R(1.0e-12f)) return;
  R impulse = (R(0.5f) * dot(rel, axis)) / den;
  add_mdelta(w, a, v3<R>(0, 0, 0), mul(ia, impulse));
  add_mdelta(w, b, v3<R>(0, 0, 0), mul(neg(ib), impulse));
}

// project_bearing_velocities + apply_velocity_deltas (articulated.wgsl:281-326).
template <typename R> void bearing_projection(World<R>& w) {
  for (uint32_t i = 0; i < w.bearing_count; ++i) {
    const auto& r = w.bearings[i];
    V3<R> arm_a = quat_rotate(w.rot[r.a], r.anchor_a), arm_b = quat_rotate(w.rot[r.b], r.anchor_b);
    V3<R> va = add(xyz(w.lin[r.a]), cross(xyz(w.ang[r.a]), arm_a));
    V3<R> vb = add(xyz(w.lin[r.b]), cross(xyz(w.ang[r.b]), arm_b));
    V3<R> rel = sub(vb, va);
    bearing_linear(w, r.a, r.b, arm_a, arm_b, rel, v3<R>(1, 0, 0));
    bearing_linear(w, r.a, r.b, arm_a, arm_b, rel, v3<R>(0, 1, 0));
    bearing_linear(w, r.a, r.b, arm_a, arm_b, rel, v3<R>(0, 0, 1));
    V3<R> axis_a = normalize_v(quat_rotate(w.rot[r.a], r.axis_a));
    V3<R> axis_b = normalize_v(quat_rotate(w.rot[r.b], r.axis_b));
    V3<R> hinge = normalize_v(add(axis_a, axis_b));
    V3<R> helper = std::fabs(hinge.x) > R(0.8f) ? v3<R>(0, 1, 0) : v3<R>(1, 0, 0);
    V3<R> ta = normalize_v(cross(hinge, helper));
    V3<R> tb = cross(hinge, ta);
    V3<R> rel_ang = sub(xyz(w.ang[r.b]), xyz(w.ang[r.a]));
    bearing_angular(w, r.a, r.b, rel_ang, ta);
    bearing_angular(w, r.a, r.b, rel_ang, tb);
  }
  const R s = R(1048576.0f);
  for (uint32_t b = 0; b < w.body_count; ++b) {
    int32_t* d = &w.mdeltas[6 * (size_t)b];
    V3<R> ld = divs(v3<R>(R(d[0]), R(d[1]), R(d[2])), s);
    V3<R> ad = divs(v3<R>(R(d[3]), R(d[4]), R(d[5])), s);
    for (int k = 0; k < 6; ++k) d[k] = 0;
    V3<R> l = add(xyz(w.lin[b]), ld), a = add(xyz(w.ang[b]), ad);
    w.lin[b] = {l.x, l.y, l.z, R(0)};
    w.ang[b] = {a.x, a.y, a.z, R(0)};
  }
}

// ---- serial small-mechanism solvers (device.rs:19-20, 1370-1416) ----------
// A scene whose mechanisms have a dynamic root and at most 64 bearings in all
// runs, after the parallel warm-start pass, 12x the passes as immediate
// Gauss-Seidel sweeps: contacts forward then backward in their sort order
// (collision.wgsl:979-1048), each followed by a forward-then-backward sweep of
// the bearing rows (articulated.wgsl:216-285) when any contact is active.
constexpr uint32_t kSerialPassMultiplier = 12;  // device.rs:20

// pair_hash (collision.wgsl:488-497).
inline uint32_t pair_hash(uint32_t a, uint32_t b) {
  uint32_t v = a * 0x9e3779b9u;
  v ^= b * 0x85ebca6bu;
  v ^= v >> 16;
  v *= 0x7feb352du;
  v ^= v >> 15;
  v *= 0x846ca68bu;
  v ^= v >> 16;
  return v;
}

// Order of the serial contact sweep: ascending (body A, body B, manifold slot)
// as solve_accumulate_serial selects them (collision.wgsl:991-1017). The slot
// is where acquire_manifold's probe starts, pair_hash(low, high) & (capacity - 1)
// (collision.wgsl:499-504): the reference's slot whenever its table holds no
// colliding entry; the collider pair breaks the remaining ties. Above 64
// active contacts the reference sweeps them in append order, which its atomics
// leave unspecified; the same sorted order is used there.
template <typename R>
std::vector<uint32_t> serial_order(const World<R>& w, const std::vector<uint64_t>& pair_keys) {
  struct Key { uint32_t a, b, slot, low, high, ci; };
  std::vector<Key> k;
  k.reserve(w.active.size());
  for (uint32_t ci : w.active) {
    const Contact<R>& c = w.contacts[ci];
    const uint32_t low = (uint32_t)(pair_keys[ci] >> 32), high = (uint32_t)pair_keys[ci];
    k.push_back({c.body_a, c.body_b, pair_hash(low, high) & (w.pair_capacity - 1u), low, high, ci});
  }
  std::sort(k.begin(), k.end(), [](const Key& x, const Key& y) {
    if (x.a != y.a) return x.a < y.a;
    if (x.b != y.b) return x.b < y.b;
    if (x.slot != y.slot) return x.slot < y.slot;
    if (x.low != y.low) return x.low < y.low;
    return x.high < y.high;
  });
  std::vector<uint32_t> out;
  out.reserve(k.size());
  for (const Key& x : k) out.push_back(x.ci);
  return out;
}

// solve_contact_immediate (collision.wgsl:908-977): the full projected impulse
// (no relaxation) against the current velocities, applied at once.
template <typename R> void contact_immediate(World<R>& w, Contact<R>& c) {
  const uint32_t a = c.body_a, b = c.body_b;
  const V3<R> arm_a = xyz(c.arm_a_impulse), arm_b = xyz(c.arm_b), normal = xyz(c.normal_penetration);
  V3<R> velocity_b = v3<R>(0, 0, 0);
  if (b != INVALID) velocity_b = contact_velocity(w, b, arm_b);
  const R denominator = impulse_denominator(w, a, b, arm_a, arm_b, normal);
  if (denominator <= R(0)) return;
  const V3<R> relative = sub(velocity_b, contact_velocity(w, a, arm_a));
  const R normal_speed = dot(relative, normal);
  if (c.normal_penetration.w <= R(1.0e-5f) && normal_speed >= R(-1.0e-5f) && c.arm_a_impulse.w <= R(1.0e-6f)) {
    c.arm_a_impulse.w = R(0);
    return;
  }
  const R previous = mx(c.arm_a_impulse.w, R(0));
  const R bias = penetration_bias(w, c.normal_penetration.w, c.arm_b.w < R(0));
  const R accumulated = mx(previous + (-normal_speed + bias) / denominator, R(0));
  const R delta = accumulated - previous;
  c.arm_a_impulse.w = accumulated;
  V3<R> impulse = mul(normal, delta);
  const V3<R> tangent_velocity = sub(relative, mul(normal, normal_speed));
  const R tangent_speed = length(tangent_velocity);
  if (tangent_speed > R(1.0e-6f)) {
    const V3<R> tangent = divs(tangent_velocity, tangent_speed);
    const R tangent_denominator = impulse_denominator(w, a, b, arm_a, arm_b, tangent);
    const R friction_impulse = mn(tangent_speed / tangent_denominator, std::fabs(delta) * w.friction);
    impulse = sub(impulse, mul(tangent, friction_impulse));
  }
  const V3<R> la = sub(xyz(w.lin[a]), mul(impulse, w.world_masses[3 * (size_t)a].w));
  w.lin[a] = {la.x, la.y, la.z, R(0)};
  const V3<R> aa = add(xyz(w.ang[a]), inverse_inertia(w, a, cross(arm_a, neg(impulse))));
  w.ang[a] = {aa.x, aa.y, aa.z, R(0)};
  if (b != INVALID) {
    const V3<R> lb = add(xyz(w.lin[b]), mul(impulse, w.world_masses[3 * (size_t)b].w));
    w.lin[b] = {lb.x, lb.y, lb.z, R(0)};
    const V3<R> ab = add(xyz(w.ang[b]), inverse_inertia(w, b, cross(arm_b, impulse)));
    w.ang[b] = {ab.x, ab.y, ab.z, R(0)};
  }
}

// solve_accumulate_serial (collision.wgsl:979-1048) over the sorted order.
template <typename R> void serial_contacts(World<R>& w, const std::vector<uint32_t>& order) {
  for (size_t i = 0; i < order.size(); ++i) contact_immediate(w, w.contacts[order[i]]);
  for (size_t i = order.size(); i > 0; --i) contact_immediate(w, w.contacts[order[i - 1]]);
}

// solve_linear_axis_immediate / solve_angular_axis_immediate (articulated.wgsl:157-214).
template <typename R> void bearing_linear_immediate(World<R>& w, uint32_t a, uint32_t b, V3<R> arm_a, V3<R> arm_b,
                                                    V3<R> rel, V3<R> dir) {
  const V3<R> angular_a = cross(arm_a, dir), angular_b = cross(arm_b, dir);
  const R den =
      ((mass_inv(w, a) + mass_inv(w, b)) + dot(angular_a, wii(w, a, angular_a))) + dot(angular_b, wii(w, b, angular_b));
  if (den <= R(1.0e-12f)) return;
  const V3<R> impulse = mul(dir, dot(rel, dir) / den);
  const V3<R> la = add(xyz(w.lin[a]), mul(impulse, mass_inv(w, a)));
  w.lin[a] = {la.x, la.y, la.z, R(0)};
  const V3<R> aa = add(xyz(w.ang[a]), wii(w, a, cross(arm_a, impulse)));
  w.ang[a] = {aa.x, aa.y, aa.z, R(0)};
  const V3<R> lb = sub(xyz(w.lin[b]), mul(impulse, mass_inv(w, b)));
  w.lin[b] = {lb.x, lb.y, lb.z, R(0)};
  const V3<R> ab = add(xyz(w.ang[b]), wii(w, b, cross(arm_b, neg(impulse))));
  w.ang[b] = {ab.x, ab.y, ab.z, R(0)};
}
template <typename R> void bearing_angular_immediate(World<R>& w, uint32_t a, uint32_t b, V3<R> rel, V3<R> axis) {
  const V3<R> ia = wii(w, a, axis), ib = wii(w, b, axis);
  const R den = dot(axis, add(ia, ib));
  if (den <= R(1.0e-12f)) return;
  const R impulse = dot(rel, axis) / den;
  const V3<R> aa = add(xyz(w.ang[a]), mul(ia, impulse));
  w.ang[a] = {aa.x, aa.y, aa.z, R(0)};
  const V3<R> ab = sub(xyz(w.ang[b]), mul(ib, impulse));
  w.ang[b] = {ab.x, ab.y, ab.z, R(0)};
}

// project_bearing_velocity_row (articulated.wgsl:216-276).
template <typename R> void bearing_row_immediate(World<R>& w, uint32_t i) {
  const auto& r = w.bearings[i];
  const V3<R> arm_a = quat_rotate(w.rot[r.a], r.anchor_a), arm_b = quat_rotate(w.rot[r.b], r.anchor_b);
  const V3<R> dirs[3] = {v3<R>(1, 0, 0), v3<R>(0, 1, 0), v3<R>(0, 0, 1)};
  for (const V3<R>& dir : dirs) {
    const V3<R> va = add(xyz(w.lin[r.a]), cross(xyz(w.ang[r.a]), arm_a));
    const V3<R> vb = add(xyz(w.lin[r.b]), cross(xyz(w.ang[r.b]), arm_b));
    bearing_linear_immediate(w, r.a, r.b, arm_a, arm_b, sub(vb, va), dir);
  }
  const V3<R> axis_a = normalize_v(quat_rotate(w.rot[r.a], r.axis_a));
  const V3<R> axis_b = normalize_v(quat_rotate(w.rot[r.b], r.axis_b));
  const V3<R> hinge = normalize_v(add(axis_a, axis_b));
  const V3<R> helper = std::fabs(hinge.x) > R(0.8f) ? v3<R>(0, 1, 0) : v3<R>(1, 0, 0);
  const V3<R> ta = normalize_v(cross(hinge, helper));
  const V3<R> tb = cross(hinge, ta);
  bearing_angular_immediate(w, r.a, r.b, sub(xyz(w.ang[r.b]), xyz(w.ang[r.a])), ta);
  bearing_angular_immediate(w, r.a, r.b, sub(xyz(w.ang[r.b]), xyz(w.ang[r.a])), tb);
}

// project_bearing_velocities_serial (articulated.wgsl:277-285).
template <typename R> void serial_bearings(World<R>& w) {
  for (uint32_t i = 0; i < w.bearing_count; ++i) bearing_row_immediate(w, i);
  for (uint32_t i = w.bearing_count; i > 0; --i) bearing_row_immediate(w, i - 1);
}

template <typename R> V3<R> permitted_axis(const World<R>& w, uint32_t body) {
  const auto& m = w.mbodies[body];
  const auto& r = w.bearings[m.tree_bearing];
  return m.direction != 0u ? neg(normalize_v(quat_rotate(w.rot[m.parent], r.axis_b)))
                           : normalize_v(quat_rotate(w.rot[m.parent], r.axis_a));
}
template <typename R> R stabilized_speed(V3<R> axis, R speed) {
  bool aligned = std::fabs(axis.y) > R(0.999f);
  return aligned && std::fabs(speed) < R(0.005f) ? R(0) : speed;
}

// advance_coordinates (articulated.wgsl:357-407).
template <typename R> void advance_coordinates(World<R>& w) {
  for (uint32_t body = 0; body < w.body_count; ++body) {
    const auto& m = w.mbodies[body];
    if (m.root != 0u) continue;
    uint32_t c = w.bearings[m.tree_bearing].coordinate;
    if (c == INVALID) continue;
    V3<R> axis = permitted_axis(w, body);
    R measured = dot(sub(xyz(w.ang[body]), xyz(w.ang[m.parent])), axis) * w.ang_damp;
    const auto& dr = w.drives[c];
    R speed;
    if (dr.mode == 0u) {
      speed = stabilized_speed(axis, measured);
    } else {
      R desired = clampr(dr.target_speed, -dr.max_speed, dr.max_speed);
      if (dr.mode == 2u) {
        R error = dr.target_angle - w.coord_angle[c];
        R brake = std::sqrt((R(2) * dr.max_acceleration) * std::fabs(error));
        desired = sgn(error) * mn(brake, dr.max_speed);
        if (std::fabs(error) < R(0.0005f)) desired = R(0);
      }
      R budget = dr.max_acceleration * w.dt;
      speed = measured + clampr(desired - measured, -budget, budget);
    }
    R angle = w.coord_angle[c] + speed * w.dt;
    if (angle < dr.min_angle) {
      angle = dr.min_angle;
      speed = R(0);
    } else if (angle > dr.max_angle) {
      angle = dr.max_angle;
      speed = R(0);
    }
    w.coord_angle[c] = angle;
    w.coord_speed[c] = speed;
  }
}

// capture_coordinates (articulated.wgsl:409-421).
template <typename R> void capture_coordinates(World<R>& w) {
  for (uint32_t body = 0; body < w.body_count; ++body) {
    const auto& m = w.mbodies[body];
    if (m.root != 0u) continue;
    uint32_t c = w.bearings[m.tree_bearing].coordinate;
    if (c == INVALID) continue;
    V3<R> axis = permitted_axis(w, body);
    w.coord_speed[c] = stabilized_speed(axis, dot(sub(xyz(w.ang[body]), xyz(w.ang[m.parent])), axis));
  }
}

// reconstruct_body_velocities (articulated.wgsl:423-456), serial over preorder.
template <typename R> void reconstruct(World<R>& w) {
  for (uint32_t t = 0; t < w.body_count; ++t) {
    uint32_t body = w.preorder[t];
    const auto& m = w.mbodies[body];
    if (m.root != 0u) continue;
    const auto& r = w.bearings[m.tree_bearing];
    V4<R> rp = w.rot[m.parent], rb = w.rot[body];
    V3<R> axis = normalize_v(quat_rotate(rp, r.axis_a));
    V3<R> parent_anchor = quat_rotate(rp, r.anchor_a);
    V3<R> child_anchor = quat_rotate(rb, r.anchor_b);
    if (m.direction != 0u) {
      axis = neg(normalize_v(quat_rotate(rp, r.axis_b)));
      parent_anchor = quat_rotate(rp, r.anchor_b);
      child_anchor = quat_rotate(rb, r.anchor_a);
    }
    R speed = w.coord_speed[r.coordinate];
    V3<R> angular = add(xyz(w.ang[m.parent]), mul(axis, speed));
    V3<R> anchor_velocity = add(xyz(w.lin[m.parent]), cross(xyz(w.ang[m.parent]), parent_anchor));
    V3<R> l = sub(anchor_velocity, cross(angular, child_anchor));
    w.lin[body] = {l.x, l.y, l.z, R(0)};
    w.ang[body] = {angular.x, angular.y, angular.z, R(0)};
  }
}

// prepare_links, jump_a_to_b / jump_b_to_a (mechanism.wgsl:75-147).
template <typename R> V4<R> axis_rotation(V3<R> axis, R angle) {
  R half = angle * R(0.5f);
  V3<R> a = mul(normalize_v(axis), std::sin(half));
  return {a.x, a.y, a.z, std::cos(half)};
}
template <typename R> void forward_kinematics(World<R>& w) {
  using L = typename World<R>::Link;
  for (uint32_t body = 0; body < w.body_count; ++body) {
    const auto& m = w.mbodies[body];
    if (m.root != 0u) {
      w.links_a[body] = L{{R(0), R(0), R(0), R(0)}, {R(0), R(0), R(0), R(1)}, body, 1u};
      continue;
    }
    const auto& r = w.bearings[m.tree_bearing];
    R angle = w.coord_angle[r.coordinate];
    V4<R> rr;
    V3<R> rp;
    if (m.direction == 0u) {
      rr = quat_multiply(axis_rotation(r.axis_a, angle), m.bind_rotation);
      rp = sub(r.anchor_a, quat_rotate(rr, r.anchor_b));
    } else {
      rr = quat_multiply(axis_rotation(r.axis_b, -angle), m.bind_rotation);
      rp = sub(r.anchor_b, quat_rotate(rr, r.anchor_a));
    }
    w.links_a[body] = L{{rp.x, rp.y, rp.z, R(0)}, rr, m.parent, 0u};
  }
  bool in_a = true;
  for (uint32_t round = 0; round < w.jump_rounds; ++round) {
    const std::vector<L>& src = in_a ? w.links_a : w.links_b;
    std::vector<L>& dst = in_a ? w.links_b : w.links_a;
    for (uint32_t body = 0; body < w.body_count; ++body) {
      const L child = src[body];
      if (child.resolved != 0u) {
        dst[body] = child;
        continue;
      }
      const L parent = src[child.ancestor];
      V3<R> pos = add(xyz(parent.position), quat_rotate(parent.rotation, xyz(child.position)));
      dst[body] = L{{pos.x, pos.y, pos.z, R(0)}, normalize_q(quat_multiply(parent.rotation, child.rotation)),
                    parent.ancestor, parent.resolved};
    }
    in_a = !in_a;
  }
}
template <typename R> const std::vector<typename World<R>::Link>& final_links(const World<R>& w) {
  return (w.jump_rounds % 2u == 0u) ? w.links_a : w.links_b;
}

// publish_a / publish_b (mechanism.wgsl:149-185).
template <typename R> void publish(World<R>& w) {
  const auto& links = final_links(w);
  for (uint32_t body = 0; body < w.body_count; ++body) {
    const auto& l = links[body];
    uint32_t root = l.ancestor;
    if (body == root) continue;
    V3<R> p = add(xyz(w.pos[root]), quat_rotate(w.rot[root], xyz(l.position)));
    w.pos[body] = {p.x, p.y, p.z, R(0)};
    w.rot[body] = normalize_q(quat_multiply(w.rot[root], l.rotation));
  }
}

// Joint axis / anchor of a tree body in its root's frame (closure.wgsl:172-185).
template <typename R> void joint_of(const World<R>& w, uint32_t body, V3<R>& axis, V3<R>& anchor) {
  const auto& m = w.mbodies[body];
  const auto& r = w.bearings[m.tree_bearing];
  const auto& pp = final_links(w)[m.parent];
  if (m.direction == 0u) {
    axis = normalize_v(quat_rotate(pp.rotation, r.axis_a));
    anchor = add(xyz(pp.position), quat_rotate(pp.rotation, r.anchor_a));
  } else {
    axis = neg(normalize_v(quat_rotate(pp.rotation, r.axis_b)));
    anchor = add(xyz(pp.position), quat_rotate(pp.rotation, r.anchor_b));
  }
}

// The closure Newton loop (device.rs:976-1020; closure.wgsl:233-427). Gradient
// and diagonal sums and the PCG dot products accumulate in FP64; vectors stay in R.
//
// Two organisations of the PCG: one system over every coordinate (the
// reference's), or, with `closure_components`, one independent PCG per closure
// component -- closures linked by shared coordinates -- each with its own
// step lengths, tolerance stop and Newton convergence test. JᵀJ is block
// diagonal over components, so both solve the same system; the second
// converges each block in a few iterations instead of coupling every loop's
// step lengths.
template <typename R> void closure_newton(World<R>& w, int max_steps = -1) {
  if (max_steps < 0) max_steps = (int)w.closure_steps;
  const uint32_t nq = w.coordinate_count;
  std::vector<uint32_t> closures;
  for (uint32_t i = 0; i < w.bearing_count; ++i)
    if (w.bearings[i].closure != 0u) closures.push_back(i);
  if (closures.empty()) return;
  const size_t nk = closures.size();
  struct End { V3<R> anchor_a, axis_a, anchor_b, axis_b, res_pos, res_axis; bool active; };
  std::vector<End> ends(nk);
  std::vector<R> solution(nq), residual(nq), pre(nq), direction(nq), op(nq), diagonal(nq);
  // Coordinates on the walk from `body` to its root.
  auto walk = [&](uint32_t body, auto&& visit) {
    while (w.mbodies[body].root == 0u) {
      visit(body, w.bearings[w.mbodies[body].tree_bearing].coordinate);
      body = w.mbodies[body].parent;
    }
  };
  // Component of each closure and coordinate (one component for the whole
  // system unless closure_components); coordinates on no closure path: none.
  const uint32_t kNone = 0xffffffffu;
  std::vector<uint32_t> comp_closure(nk, 0), comp_coord(nq, kNone);
  uint32_t ncomp = 1;
  {
    std::vector<uint32_t> parent(nk);
    for (size_t k = 0; k < nk; ++k) parent[k] = (uint32_t)k;
    auto find = [&](uint32_t x) {
      while (parent[x] != x) x = parent[x] = parent[parent[x]];
      return x;
    };
    std::vector<uint32_t> owner(nq, kNone);
    for (size_t k = 0; k < nk; ++k) {
      const auto& r = w.bearings[closures[k]];
      for (uint32_t side = 0; side < 2; ++side)
        walk(side ? r.b : r.a, [&](uint32_t, uint32_t c) {
          if (owner[c] == kNone) owner[c] = (uint32_t)k;
          else if (w.closure_components) parent[find((uint32_t)k)] = find(owner[c]);
        });
    }
    if (w.closure_components) {
      std::vector<uint32_t> id(nk, kNone);
      ncomp = 0;
      for (size_t k = 0; k < nk; ++k) {
        const uint32_t root = find((uint32_t)k);
        if (id[root] == kNone) id[root] = ncomp++;
        comp_closure[k] = id[root];
      }
    }
    for (uint32_t c = 0; c < nq; ++c)
      if (owner[c] != kNone) comp_coord[c] = comp_closure[owner[c]];
  }
  std::vector<uint8_t> comp_active(ncomp), comp_done(ncomp);
  std::vector<uint32_t> um_max(ncomp), ud_max(ncomp);
  std::vector<double> rz(ncomp), rz_stop(ncomp), dad(ncomp), rz_next(ncomp);
  std::vector<R> alpha(ncomp), beta(ncomp);
  for (int step = 0; step < max_steps; ++step) {
    std::fill(um_max.begin(), um_max.end(), 0u);
    std::fill(ud_max.begin(), ud_max.end(), 0u);
    const auto& links = final_links(w);
    for (size_t k = 0; k < nk; ++k) {
      const auto& r = w.bearings[closures[k]];
      const auto& pa = links[r.a];
      const auto& pb = links[r.b];
      End& e = ends[k];
      const uint32_t cp = comp_closure[k];
      e.anchor_a = add(xyz(pa.position), quat_rotate(pa.rotation, r.anchor_a));
      e.anchor_b = add(xyz(pb.position), quat_rotate(pb.rotation, r.anchor_b));
      e.axis_a = normalize_v(quat_rotate(pa.rotation, r.axis_a));
      e.axis_b = normalize_v(quat_rotate(pb.rotation, r.axis_b));
      e.res_pos = sub(e.anchor_a, e.anchor_b);
      e.res_axis = sub(e.axis_a, e.axis_b);
      e.active = false;
      if (!finite3v(e.res_pos) || !finite3v(e.res_axis)) {
        w.diag[0] |= INVALID_NUMERIC;
        um_max[cp] = ud_max[cp] = 0xffffffffu;
        continue;
      }
      uint32_t um = round_u32(length(e.res_pos) * R(1000000.0f));
      R degrees = std::acos(clampr(dot(e.axis_a, e.axis_b), R(-1), R(1))) * R(57.295779513f);
      uint32_t ud = round_u32(degrees * R(1000000.0f));
      um_max[cp] = std::max(um_max[cp], um);
      ud_max[cp] = std::max(ud_max[cp], ud);
      e.active = !(um <= 10u && ud <= 1000u);
    }
    bool any = false;
    for (uint32_t cp = 0; cp < ncomp; ++cp) {
      comp_active[cp] = (um_max[cp] > 10u || ud_max[cp] > 1000u) ? 1 : 0;
      any = any || comp_active[cp];
    }
    if (!any) break;
    // Coordinates the step moves: all (one global system), or those of the
    // components still outside tolerance.
    auto live = [&](uint32_t c) {
      if (!w.closure_components) return true;
      return comp_coord[c] != kNone && comp_active[comp_coord[c]] != 0;
    };
    std::vector<double> g(nq, 0.0), dg(nq, 0.0);
    for (size_t k = 0; k < nk; ++k) {
      const End& e = ends[k];
      if (!e.active || !comp_active[comp_closure[k]]) continue;
      const auto& r = w.bearings[closures[k]];
      for (uint32_t side = 0; side < 2; ++side) {
        V3<R> end_anchor = side ? e.anchor_b : e.anchor_a, end_axis = side ? e.axis_b : e.axis_a;
        R sign = side ? R(-1) : R(1);
        walk(side ? r.b : r.a, [&](uint32_t body, uint32_t c) {
          V3<R> jaxis, janchor;
          joint_of(w, body, jaxis, janchor);
          V3<R> pd = mul(cross(jaxis, sub(end_anchor, janchor)), sign);
          V3<R> ad = mul(cross(jaxis, end_axis), sign);
          R gradient = dot(e.res_pos, pd) + R(1) * dot(e.res_axis, ad);
          R diag = dot(pd, pd) + R(1) * dot(ad, ad);
          if (gradient == gradient && diag == diag) {
            g[c] += (double)gradient;
            dg[c] += (double)diag;
          } else {
            w.diag[0] |= INVALID_NUMERIC;
          }
        });
      }
    }
    // The PCG: vectors of type P -- R for the global system (the reference's
    // FP32 vectors, FP64 dots), double for the per-component solve, whose
    // tolerance stop FP32 recursive residuals do not reach.
    auto pcg = [&](auto zero) {
      using P = decltype(zero);
      std::vector<P> sol(nq), res(nq), pz(nq), dir(nq), opp(nq), dia(nq);
      std::fill(rz.begin(), rz.end(), 0.0);
      for (uint32_t c = 0; c < nq; ++c) {
        const P gf = P(std::is_same<P, double>::value ? g[c] : (double)R(g[c]));
        const P df = P(std::is_same<P, double>::value ? dg[c] : (double)R(dg[c]));
        dia[c] = df;
        res[c] = df > P(1.0e-12f) ? -gf : P(0);
        pz[c] = df > P(1.0e-12f) ? res[c] / df : P(0);
        sol[c] = P(0);
        dir[c] = pz[c];
        if (comp_coord[c] != kNone) rz[comp_coord[c]] += (double)res[c] * (double)pz[c];
      }
      std::vector<P> al(ncomp), be(ncomp);
      for (uint32_t cp = 0; cp < ncomp; ++cp) {
        rz_stop[cp] = (double)w.closure_tol * (double)w.closure_tol * rz[cp];
        comp_done[cp] = comp_active[cp] ? 0 : 1;
      }
      for (uint32_t it = 0; it < w.closure_iters; ++it) {
        bool running = false;
        for (uint32_t cp = 0; cp < ncomp; ++cp) {
          if (!comp_done[cp] && w.closure_tol > 0.0f && rz[cp] <= rz_stop[cp]) comp_done[cp] = 1;
          running = running || !comp_done[cp];
        }
        if (!running) break;
        w.pcg_iterations_run += 1;
        std::fill(opp.begin(), opp.end(), P(0));
        for (size_t k = 0; k < nk; ++k) {
          if (comp_done[comp_closure[k]]) continue;
          const End& e = ends[k];
          const auto& r = w.bearings[closures[k]];
          V3<P> prod_pos[2], prod_axis[2];
          for (uint32_t side = 0; side < 2; ++side) {
            V3<R> end_anchor = side ? e.anchor_b : e.anchor_a, end_axis = side ? e.axis_b : e.axis_a;
            R sign = side ? R(-1) : R(1);
            V3<P> pp = v3<P>(0, 0, 0), pa = pp;
            walk(side ? r.b : r.a, [&](uint32_t body, uint32_t c) {
              V3<R> jaxis, janchor;
              joint_of(w, body, jaxis, janchor);
              const V3<R> pd = mul(cross(jaxis, sub(end_anchor, janchor)), sign), ad = mul(cross(jaxis, end_axis), sign);
              pp = add(pp, mul(v3<P>(P(pd.x), P(pd.y), P(pd.z)), dir[c]));
              pa = add(pa, mul(v3<P>(P(ad.x), P(ad.y), P(ad.z)), dir[c]));
            });
            prod_pos[side] = pp;
            prod_axis[side] = pa;
          }
          V3<P> product_pos = add(prod_pos[0], prod_pos[1]), product_axis = add(prod_axis[0], prod_axis[1]);
          for (uint32_t side = 0; side < 2; ++side) {
            V3<R> end_anchor = side ? e.anchor_b : e.anchor_a, end_axis = side ? e.axis_b : e.axis_a;
            R sign = side ? R(-1) : R(1);
            walk(side ? r.b : r.a, [&](uint32_t body, uint32_t c) {
              V3<R> jaxis, janchor;
              joint_of(w, body, jaxis, janchor);
              const V3<R> pd = mul(cross(jaxis, sub(end_anchor, janchor)), sign), ad = mul(cross(jaxis, end_axis), sign);
              opp[c] = opp[c] + (dot(v3<P>(P(pd.x), P(pd.y), P(pd.z)), product_pos) +
                                 P(1) * dot(v3<P>(P(ad.x), P(ad.y), P(ad.z)), product_axis));
            });
          }
        }
        std::fill(dad.begin(), dad.end(), 0.0);
        for (uint32_t c = 0; c < nq; ++c)
          if (comp_coord[c] != kNone) dad[comp_coord[c]] += (double)dir[c] * (double)opp[c];
        for (uint32_t cp = 0; cp < ncomp; ++cp) al[cp] = std::fabs(dad[cp]) > 1.0e-20 ? P(rz[cp] / dad[cp]) : P(0);
        std::fill(rz_next.begin(), rz_next.end(), 0.0);
        for (uint32_t c = 0; c < nq; ++c) {
          const uint32_t cp = comp_coord[c];
          if (cp == kNone || comp_done[cp]) continue;
          sol[c] = sol[c] + al[cp] * dir[c];
          res[c] = res[c] - al[cp] * opp[c];
          pz[c] = dia[c] > P(1.0e-12f) ? res[c] / dia[c] : P(0);
          rz_next[cp] += (double)res[c] * (double)pz[c];
        }
        for (uint32_t cp = 0; cp < ncomp; ++cp) {
          if (comp_done[cp]) continue;
          be[cp] = std::fabs(rz[cp]) > 1.0e-20 ? P(rz_next[cp] / rz[cp]) : P(0);
          rz[cp] = rz_next[cp];
        }
        for (uint32_t c = 0; c < nq; ++c) {
          const uint32_t cp = comp_coord[c];
          if (cp == kNone || comp_done[cp]) continue;
          dir[c] = pz[c] + be[cp] * dir[c];
        }
      }
      // The solution in the step's type: FP32 for the FP32 oracle.
      for (uint32_t c = 0; c < nq; ++c) solution[c] = R(sol[c]);
    };
    if (w.closure_components) pcg(0.0);
    else pcg(R(0));
    for (uint32_t c = 0; c < nq; ++c) {
      if (!live(c)) continue;
      R step_c = clampr(R(0.35f) * solution[c], R(-0.2f), R(0.2f));
      if (step_c == step_c) w.coord_angle[c] = w.coord_angle[c] + step_c;
      else w.diag[0] |= INVALID_NUMERIC;
    }
    forward_kinematics(w);
  }
}

// Velocity-level closure (closure_velocity_projection, component solve only):
// the coordinate speeds are projected onto J v = 0 at the final configuration,
// v += x with J^T J x = -J^T (J v) solved per closure component by the
// Jacobi-preconditioned CG of the Newton steps (FP32 Jacobian columns, FP64
// vectors and sums, the same tolerance and iteration cap), before the tree
// reconstructs body velocities from the speeds. Every closure takes part.
template <typename R> void closure_velocity_projection(World<R>& w) {
  const uint32_t nq = w.coordinate_count;
  std::vector<uint32_t> closures;
  for (uint32_t i = 0; i < w.bearing_count; ++i)
    if (w.bearings[i].closure != 0u) closures.push_back(i);
  const size_t nk = closures.size();
  if (nk == 0) return;
  auto walk = [&](uint32_t body, auto&& visit) {
    while (w.mbodies[body].root == 0u) {
      visit(body, w.bearings[w.mbodies[body].tree_bearing].coordinate);
      body = w.mbodies[body].parent;
    }
  };
  // Components: closures linked by shared coordinates.
  const uint32_t kNone = 0xffffffffu;
  std::vector<uint32_t> parent(nk), owner(nq, kNone), comp_closure(nk), comp_coord(nq, kNone);
  for (size_t k = 0; k < nk; ++k) parent[k] = (uint32_t)k;
  auto find = [&](uint32_t x) {
    while (parent[x] != x) x = parent[x] = parent[parent[x]];
    return x;
  };
  for (size_t k = 0; k < nk; ++k) {
    const auto& r = w.bearings[closures[k]];
    for (uint32_t side = 0; side < 2; ++side)
      walk(side ? r.b : r.a, [&](uint32_t, uint32_t c) {
        if (owner[c] == kNone) owner[c] = (uint32_t)k;
        else parent[find((uint32_t)k)] = find(owner[c]);
      });
  }
  std::vector<uint32_t> id(nk, kNone);
  uint32_t ncomp = 0;
  for (size_t k = 0; k < nk; ++k) {
    const uint32_t root = find((uint32_t)k);
    if (id[root] == kNone) id[root] = ncomp++;
    comp_closure[k] = id[root];
  }
  for (uint32_t c = 0; c < nq; ++c)
    if (owner[c] != kNone) comp_coord[c] = comp_closure[owner[c]];
  // Closure ends at the final configuration.
  const auto& links = final_links(w);
  struct End { V3<R> anchor[2], axis[2]; };
  std::vector<End> ends(nk);
  for (size_t k = 0; k < nk; ++k) {
    const auto& r = w.bearings[closures[k]];
    const auto& pa = links[r.a];
    const auto& pb = links[r.b];
    ends[k].anchor[0] = add(xyz(pa.position), quat_rotate(pa.rotation, r.anchor_a));
    ends[k].anchor[1] = add(xyz(pb.position), quat_rotate(pb.rotation, r.anchor_b));
    ends[k].axis[0] = normalize_v(quat_rotate(pa.rotation, r.axis_a));
    ends[k].axis[1] = normalize_v(quat_rotate(pb.rotation, r.axis_b));
  }
  // Columns of one path entry: position and axis rows, signed by side.
  auto column = [&](size_t k, uint32_t side, uint32_t body, V3<R>& pd, V3<R>& ad) {
    V3<R> jaxis, janchor;
    joint_of(w, body, jaxis, janchor);
    const R sign = side ? R(-1) : R(1);
    pd = mul(cross(jaxis, sub(ends[k].anchor[side], janchor)), sign);
    ad = mul(cross(jaxis, ends[k].axis[side]), sign);
  };
  auto dot_d = [](V3<R> a, const double* b) {
    return ((double)a.x * b[0] + (double)a.y * b[1]) + (double)a.z * b[2];
  };
  // J v per closure: each side summed in path order, side 0 plus side 1.
  std::vector<double> prod(6 * nk);
  auto product = [&](const std::vector<double>& v) {
    for (size_t k = 0; k < nk; ++k) {
      const auto& r = w.bearings[closures[k]];
      double acc[2][6] = {};
      for (uint32_t side = 0; side < 2; ++side)
        walk(side ? r.b : r.a, [&](uint32_t body, uint32_t c) {
          V3<R> pd, ad;
          column(k, side, body, pd, ad);
          const double x[6] = {(double)pd.x, (double)pd.y, (double)pd.z, (double)ad.x, (double)ad.y, (double)ad.z};
          for (int j = 0; j < 6; ++j) acc[side][j] = acc[side][j] + x[j] * v[c];
        });
      for (int j = 0; j < 6; ++j) prod[6 * k + j] = acc[0][j] + acc[1][j];
    }
  };
  // J^T prod per coordinate, gathered in (closure, side) order.
  auto transpose = [&](std::vector<double>& out) {
    std::fill(out.begin(), out.end(), 0.0);
    for (size_t k = 0; k < nk; ++k) {
      const auto& r = w.bearings[closures[k]];
      for (uint32_t side = 0; side < 2; ++side)
        walk(side ? r.b : r.a, [&](uint32_t body, uint32_t c) {
          V3<R> pd, ad;
          column(k, side, body, pd, ad);
          out[c] = out[c] + (dot_d(pd, &prod[6 * k]) + 1.0 * dot_d(ad, &prod[6 * k + 3]));
        });
    }
  };
  std::vector<double> speed(nq), dg(nq, 0.0), g(nq), x(nq, 0.0), res(nq), z(nq), dir(nq), op(nq);
  for (uint32_t c = 0; c < nq; ++c) speed[c] = (double)w.coord_speed[c];
  for (size_t k = 0; k < nk; ++k) {
    const auto& r = w.bearings[closures[k]];
    for (uint32_t side = 0; side < 2; ++side)
      walk(side ? r.b : r.a, [&](uint32_t body, uint32_t c) {
        V3<R> pd, ad;
        column(k, side, body, pd, ad);
        dg[c] += (double)(dot(pd, pd) + R(1) * dot(ad, ad));
      });
  }
  product(speed);
  transpose(g);
  std::vector<double> rz(ncomp, 0.0), rz_stop(ncomp), dad(ncomp), rz_next(ncomp), beta(ncomp, 0.0);
  std::vector<uint8_t> done(ncomp, 0);
  for (uint32_t c = 0; c < nq; ++c) {
    res[c] = dg[c] > (double)1.0e-12f ? -g[c] : 0.0;
    z[c] = dg[c] > (double)1.0e-12f ? res[c] / dg[c] : 0.0;
    dir[c] = z[c];
    if (comp_coord[c] != kNone) rz[comp_coord[c]] += res[c] * z[c];
  }
  for (uint32_t cp = 0; cp < ncomp; ++cp) rz_stop[cp] = (double)w.closure_tol * (double)w.closure_tol * rz[cp];
  for (uint32_t it = 0; it < w.closure_iters; ++it) {
    bool running = false;
    for (uint32_t cp = 0; cp < ncomp; ++cp) {
      if (!done[cp] && w.closure_tol > 0.0f && rz[cp] <= rz_stop[cp]) done[cp] = 1;
      running = running || !done[cp];
    }
    if (!running) break;
    if (it > 0)
      for (uint32_t c = 0; c < nq; ++c)
        if (comp_coord[c] != kNone && !done[comp_coord[c]]) dir[c] = z[c] + beta[comp_coord[c]] * dir[c];
    product(dir);
    transpose(op);
    std::fill(dad.begin(), dad.end(), 0.0);
    for (uint32_t c = 0; c < nq; ++c)
      if (comp_coord[c] != kNone) dad[comp_coord[c]] += dir[c] * op[c];
    std::fill(rz_next.begin(), rz_next.end(), 0.0);
    for (uint32_t c = 0; c < nq; ++c) {
      const uint32_t cp = comp_coord[c];
      if (cp == kNone || done[cp]) continue;
      const double alpha = std::fabs(dad[cp]) > 1.0e-20 ? rz[cp] / dad[cp] : 0.0;
      x[c] = x[c] + alpha * dir[c];
      res[c] = res[c] - alpha * op[c];
      z[c] = dg[c] > (double)1.0e-12f ? res[c] / dg[c] : 0.0;
      rz_next[cp] += res[c] * z[c];
    }
    for (uint32_t cp = 0; cp < ncomp; ++cp) {
      if (done[cp]) continue;
      beta[cp] = std::fabs(rz[cp]) > 1.0e-20 ? rz_next[cp] / rz[cp] : 0.0;
      rz[cp] = rz_next[cp];
    }
  }
  for (uint32_t c = 0; c < nq; ++c) {
    if (comp_coord[c] == kNone) continue;
    const R v = R(speed[c] + x[c]);
    if (v == v) w.coord_speed[c] = v;
    else w.diag[0] |= INVALID_NUMERIC;
  }
}

// validate_articulated_state (articulated.wgsl:458-484).
template <typename R> void validate_articulated(World<R>& w) {
  for (uint32_t b = 0; b < w.body_count; ++b) {
    R rl = dot4(w.rot[b], w.rot[b]);
    bool valid = finite4(w.pos[b]) && finite4(w.rot[b]) && finite4(w.lin[b]) && finite4(w.ang[b]) && rl > R(0.5f) &&
                 rl < R(1.5f);
    const auto& m = w.mbodies[b];
    if (m.root == 0u) {
      uint32_t c = w.bearings[m.tree_bearing].coordinate;
      R a = w.coord_angle[c], v = w.coord_speed[c];
      const R lim = R(3.402823e+38f);
      valid = valid && a == a && std::fabs(a) < lim && v == v && std::fabs(v) < lim;
    }
    if (!valid) w.diag[0] |= INVALID_NUMERIC;
  }
}

// validate_bearings / validate_mechanism_bearings (bearings.wgsl:53-102).
template <typename R> void validate_bearings(World<R>& w) {
  for (uint32_t i = 0; i < w.bearing_count; ++i) {
    const auto& r = w.bearings[i];
    V4<R> pa, ra, pb, rb;
    if (w.mech_active) {
      const auto& links = final_links(w);
      pa = links[r.a].position; ra = links[r.a].rotation; pb = links[r.b].position; rb = links[r.b].rotation;
    } else {
      pa = w.pos[r.a]; ra = w.rot[r.a]; pb = w.pos[r.b]; rb = w.rot[r.b];
    }
    V3<R> anchor_a = add(xyz(pa), quat_rotate(ra, r.anchor_a)), anchor_b = add(xyz(pb), quat_rotate(rb, r.anchor_b));
    V3<R> axis_a = normalize_v(quat_rotate(ra, r.axis_a)), axis_b = normalize_v(quat_rotate(rb, r.axis_b));
    uint32_t um = round_u32(length(sub(anchor_a, anchor_b)) * R(1000000.0f));
    R degrees = std::atan2(length(cross(axis_a, axis_b)), clampr(dot(axis_a, axis_b), R(-1), R(1))) *
                R(57.295779513f);
    uint32_t ud = round_u32(degrees * R(1000000.0f));
    w.diag[3] = std::max(w.diag[3], um);
    w.diag[4] = std::max(w.diag[4], ud);
    if (r.closure != 0u && (um > 10u || ud > 1000u)) w.diag[0] |= NON_CONVERGENCE;
  }
}

// encode_mechanism_passes (device.rs:963-1048).
template <typename R> void mechanism_passes(World<R>& w) {
  const uint32_t iterations = std::max(1u, w.iterations);
  for (uint32_t i = 0; i < iterations; ++i) bearing_projection(w);
  advance_coordinates(w);
  forward_kinematics(w);
  closure_newton(w);
  if (w.closure_components && w.closure_vproj) closure_velocity_projection(w);
  publish(w);
  reconstruct(w);
}

// One tick, pass order of dispatch_tick (device.rs:859-961) and
// encode_collision_passes (device.rs:1219-1431).
// `stop` ends the tick early for phase tests: 1 after integration, 2 after the
// broadphase, 3 after contact generation, anything else runs the whole tick.
template <typename R> void step(World<R>& w, uint64_t tick64, int stop) {
  const uint32_t tick = (uint32_t)tick64;
  std::memset(w.diag, 0, sizeof(w.diag));
  auto t0 = std::chrono::steady_clock::now();
  // Wall time of each phase of the tick, for the CPU baseline's per-phase report.
  auto lap = [&](int phase) {
    const auto t1 = std::chrono::steady_clock::now();
    w.phase_ns[phase] += (double)std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
    t0 = t1;
  };
  integrate(w);
  lap(kPhaseIntegrate);
  if (w.mech_active) mechanism_passes(w);
  lap(kPhaseMechanism);
  if (stop == 1) return;
  if (w.collisions && w.collider_count > 0) {
    update_world_masses(w);
    lap(kPhaseWorldMasses);
    compute_morton(w);
    lap(kPhaseMorton);
    sort_keys(w);
    lap(kPhaseSort);
    build_topology(w);
    build_bounds(w);
    lap(kPhaseLbvh);
    traverse(w);
    lap(kPhaseTraverse);
    if (stop == 2) return;
    narrowphase(w);
    generate_ground_contacts(w);
    lap(kPhaseNarrowphase);
    if (stop == 3) return;
    // Collider pair of each contact before prepare overwrites metadata z/w.
    std::vector<uint64_t> keys(w.contacts.size());
    for (size_t i = 0; i < w.contacts.size(); ++i)
      keys[i] = ((uint64_t)std::min(w.contacts[i].m_z, w.contacts[i].m_w) << 32) |
                std::max(w.contacts[i].m_z, w.contacts[i].m_w);
    prepare_contacts(w, tick);
    lap(kPhasePrepare);
    w.deltas.assign(6 * (size_t)w.body_count, 0);
    const uint32_t iterations = std::max(1u, w.iterations);
    const bool interleave = w.mech_active && w.has_dynamic_root;
    if (interleave && w.bearing_count <= w.serial_limit) {
      // device.rs:1352-1416 with serial_mechanism: the parallel warm start, then
      // the serial sweeps; the bearing sweeps run only with active contacts.
      solve_pass(w, true);
      if (!w.active.empty()) {
        solve_apply(w);
        const std::vector<uint32_t> order = serial_order(w, keys);
        serial_bearings(w);
        for (uint32_t p = 1; p < iterations * kSerialPassMultiplier; ++p) {
          serial_contacts(w, order);
          serial_bearings(w);
        }
      }
    } else {
      for (uint32_t it = 0; it < iterations; ++it) {
        solve_pass(w, it == 0);
        if (!w.active.empty()) solve_apply(w);
        if (interleave) bearing_projection(w);
      }
    }
    // persist_contacts (collision.wgsl:1071-1088)
    for (uint32_t ci : w.active) {
      const Contact<R>& c = w.contacts[ci];
      Manifold<R>& m = w.manifolds[keys[ci]];
      m.tick = tick;
      m.count = c.m_w;
      m.normal_penetration = c.normal_penetration;
      m.point_impulse = c.arm_a_impulse;
    }
    if (w.mech_active) {
      // encode_post_contact_mechanism (device.rs:1128-1155).
      if (!w.has_dynamic_root)
        for (uint32_t i = 0; i < iterations; ++i) bearing_projection(w);
      capture_coordinates(w);
      if (!w.has_dynamic_root) reconstruct(w);
    }
    lap(kPhaseSolve);
  }
  if (w.mech_active) validate_articulated(w);
  if (w.bearing_count > 0) validate_bearings(w);
  lap(kPhaseValidate);
}

}  // namespace oracle

// ---------------------------------------------------------------------------
// C interface for ctypes.
// ---------------------------------------------------------------------------
using namespace oracle;

struct OracleHandle {
  bool fp64;
  World<float> f;
  World<double> d;
};

template <typename R>
static void init_world(World<R>& w, uint32_t n, uint32_t c, uint32_t ns, const Transform* poses, const float* inv_mass,
                       const Mass* masses, const Collider* colliders, const uint32_t* suppressed,
                       const uint32_t* components, const uint32_t* roots) {
  w.body_count = n;
  w.collider_count = c;
  uint32_t s = 1;
  while (s < (c < 256 ? 256u : c)) s <<= 1;
  w.sort_count = s;
  w.pos.resize(n); w.rot.resize(n); w.lin.assign(n, {0, 0, 0, 0}); w.ang.assign(n, {0, 0, 0, 0});
  w.inv_mass.resize(n); w.masses.resize(4 * (size_t)n); w.world_masses.assign(3 * (size_t)n, {0, 0, 0, 0});
  w.roots.resize(n); w.components.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    w.pos[i] = {R(poses[i].position[0]), R(poses[i].position[1]), R(poses[i].position[2]), R(0)};
    w.rot[i] = load4<R>(poses[i].rotation);
    w.inv_mass[i] = R(inv_mass[i]);
    w.masses[4 * (size_t)i] = load4<R>(masses[i].inverse_mass);
    w.masses[4 * (size_t)i + 1] = load4<R>(masses[i].ix);
    w.masses[4 * (size_t)i + 2] = load4<R>(masses[i].iy);
    w.masses[4 * (size_t)i + 3] = load4<R>(masses[i].iz);
    w.roots[i] = roots ? roots[i] : 1u;
    w.components[i] = components ? components[i] : i;
  }
  w.colliders.assign(colliders, colliders + c);
  w.suppressed.clear();
  for (uint32_t i = 0; i < ns; ++i) w.suppressed.push_back({suppressed[2 * i], suppressed[2 * i + 1]});
  std::sort(w.suppressed.begin(), w.suppressed.end());
  w.aabbs.assign(c, Aabb<R>{});
  w.deltas.assign(6 * (size_t)n, 0);
}

template <typename R> static void put4(float* out, const V4<R>& v) {
  out[0] = (float)v.x; out[1] = (float)v.y; out[2] = (float)v.z; out[3] = (float)v.w;
}

extern "C" {

OracleHandle* oracle_create(uint32_t n, uint32_t c, uint32_t ns, const Transform* poses, const float* inv_mass,
                            const Mass* masses, const Collider* colliders, const uint32_t* suppressed,
                            const uint32_t* components, const uint32_t* roots, uint32_t iterations, float friction,
                            uint32_t pair_capacity, uint32_t collisions, uint32_t self_collisions, uint32_t fp64,
                            int threads) {
  OracleHandle* h = new OracleHandle();
  h->fp64 = fp64 != 0;
  auto setup = [&](auto& w) {
    using R = typename std::remove_reference<decltype(w.dt)>::type;
    init_world(w, n, c, ns, poses, inv_mass, masses, colliders, suppressed, components, roots);
    w.iterations = iterations;
    w.friction = R(friction);
    w.pair_capacity = pair_capacity;
    w.collisions = collisions != 0;
    w.self_collisions = self_collisions != 0;
    w.threads = threads < 1 ? 1 : threads;
  };
  if (h->fp64) setup(h->d);
  else setup(h->f);
  return h;
}

void oracle_destroy(OracleHandle* h) { delete h; }

void oracle_step(OracleHandle* h, uint64_t tick) {
  if (h->fp64) step(h->d, tick, 0);
  else step(h->f, tick, 0);
}

void oracle_step_partial(OracleHandle* h, uint64_t tick, int stop) {
  if (h->fp64) step(h->d, tick, stop);
  else step(h->f, tick, stop);
}

void oracle_set_threads(OracleHandle* h, int threads) {
  h->f.threads = h->d.threads = threads < 1 ? 1 : threads;
}

void oracle_apply_impulse(OracleHandle* h, uint32_t body, const float* point, const float* impulse) {
  auto apply = [&](auto& w) {
    using R = typename std::remove_reference<decltype(w.dt)>::type;
    const V4<R>* m = &w.masses[4 * (size_t)body];
    R inverse_mass = m[0].x;
    if (inverse_mass <= R(0)) return;
    V3<R> j = {R(impulse[0]), R(impulse[1]), R(impulse[2])};
    V3<R> arm = sub(V3<R>{R(point[0]), R(point[1]), R(point[2])}, xyz(w.pos[body]));
    V3<R> l = add(xyz(w.lin[body]), mul(j, inverse_mass));
    // world_inverse_inertia (physics.wgsl:58-66) on the body's current rotation.
    V4<R> rotation = w.rot[body];
    V3<R> v = cross(arm, j);
    V3<R> local = quat_rotate(conj(rotation), v);
    V3<R> lr = add(add(mul(xyz(m[1]), local.x), mul(xyz(m[2]), local.y)), mul(xyz(m[3]), local.z));
    V3<R> a = add(xyz(w.ang[body]), quat_rotate(rotation, lr));
    w.lin[body] = {l.x, l.y, l.z, R(0)};
    w.ang[body] = {a.x, a.y, a.z, R(0)};
  };
  if (h->fp64) apply(h->d);
  else apply(h->f);
}

void oracle_get_bodies(OracleHandle* h, float* poses, float* velocities) {
  auto get = [&](auto& w) {
    for (uint32_t i = 0; i < w.body_count; ++i) {
      if (poses) { put4(poses + 8 * (size_t)i, w.pos[i]); put4(poses + 8 * (size_t)i + 4, w.rot[i]); }
      if (velocities) { put4(velocities + 8 * (size_t)i, w.lin[i]); put4(velocities + 8 * (size_t)i + 4, w.ang[i]); }
    }
  };
  if (h->fp64) get(h->d);
  else get(h->f);
}

void oracle_get_bodies_f64(OracleHandle* h, double* poses, double* velocities) {
  World<double>& w = h->d;
  for (uint32_t i = 0; i < w.body_count; ++i) {
    double* p = poses + 8 * (size_t)i;
    double* v = velocities + 8 * (size_t)i;
    p[0] = w.pos[i].x; p[1] = w.pos[i].y; p[2] = w.pos[i].z; p[3] = w.pos[i].w;
    p[4] = w.rot[i].x; p[5] = w.rot[i].y; p[6] = w.rot[i].z; p[7] = w.rot[i].w;
    v[0] = w.lin[i].x; v[1] = w.lin[i].y; v[2] = w.lin[i].z; v[3] = w.lin[i].w;
    v[4] = w.ang[i].x; v[5] = w.ang[i].y; v[6] = w.ang[i].z; v[7] = w.ang[i].w;
  }
}

void oracle_set_bodies(OracleHandle* h, const float* poses, const float* velocities) {
  auto set = [&](auto& w) {
    using R = typename std::remove_reference<decltype(w.dt)>::type;
    for (uint32_t i = 0; i < w.body_count; ++i) {
      if (poses) { w.pos[i] = load4<R>(poses + 8 * (size_t)i); w.rot[i] = load4<R>(poses + 8 * (size_t)i + 4); }
      if (velocities) { w.lin[i] = load4<R>(velocities + 8 * (size_t)i); w.ang[i] = load4<R>(velocities + 8 * (size_t)i + 4); }
    }
  };
  if (h->fp64) set(h->d);
  else set(h->f);
}

// Accumulated wall time (ms) of each tick phase since the last reset, in Phase
// order; returns the phase count.
int oracle_phase_ms(OracleHandle* h, double* out, int reset) {
  double* ns = h->fp64 ? h->d.phase_ns : h->f.phase_ns;
  for (int k = 0; k < kPhaseCount; ++k) {
    if (out) out[k] = ns[k] * 1e-6;
    if (reset) ns[k] = 0.0;
  }
  return kPhaseCount;
}

void oracle_get_diag(OracleHandle* h, uint32_t* out) {
  std::memcpy(out, h->fp64 ? h->d.diag : h->f.diag, 8 * sizeof(uint32_t));
}

size_t oracle_get_keys(OracleHandle* h, uint64_t* out, size_t cap) {
  const std::vector<uint64_t>& k = h->fp64 ? h->d.keys : h->f.keys;
  size_t n = std::min(cap, k.size());
  std::memcpy(out, k.data(), n * 8);
  return k.size();
}

size_t oracle_get_children(OracleHandle* h, uint32_t* out, size_t cap) {
  auto& ch = h->fp64 ? h->d.children : h->f.children;
  size_t n = std::min(cap, ch.size());
  for (size_t i = 0; i < n; ++i) { out[2 * i] = ch[i].first; out[2 * i + 1] = ch[i].second; }
  return ch.size();
}

size_t oracle_get_parents(OracleHandle* h, uint32_t* out, size_t cap) {
  auto& p = h->fp64 ? h->d.parents : h->f.parents;
  size_t n = std::min(cap, p.size());
  std::memcpy(out, p.data(), n * 4);
  return p.size();
}

size_t oracle_get_aabbs(OracleHandle* h, float* out, size_t cap, int nodes) {
  auto get = [&](auto& w) {
    auto& src = nodes ? w.node_aabbs : w.aabbs;
    size_t n = std::min(cap, src.size());
    for (size_t i = 0; i < n; ++i) {
      float* o = out + 8 * i;
      o[0] = (float)src[i].mn.x; o[1] = (float)src[i].mn.y; o[2] = (float)src[i].mn.z; o[3] = 0.0f;
      o[4] = (float)src[i].mx.x; o[5] = (float)src[i].mx.y; o[6] = (float)src[i].mx.z; o[7] = 0.0f;
    }
    return src.size();
  };
  return h->fp64 ? get(h->d) : get(h->f);
}

size_t oracle_get_pairs(OracleHandle* h, uint32_t* out, size_t cap) {
  auto& p = h->fp64 ? h->d.pairs : h->f.pairs;
  size_t n = std::min(cap, p.size());
  for (size_t i = 0; i < n; ++i) { out[2 * i] = p[i].first; out[2 * i + 1] = p[i].second; }
  return p.size();
}

size_t oracle_get_contacts(OracleHandle* h, ContactRow* out, size_t cap) {
  auto get = [&](auto& w) {
    size_t n = std::min(cap, w.contacts.size());
    for (size_t i = 0; i < n; ++i) {
      const auto& c = w.contacts[i];
      ContactRow& r = out[i];
      r.metadata[0] = c.body_a; r.metadata[1] = c.body_b; r.metadata[2] = c.m_z; r.metadata[3] = c.m_w;
      put4(r.normal_penetration, c.normal_penetration);
      put4(r.arm_a_impulse, c.arm_a_impulse);
      put4(r.arm_b, c.arm_b);
    }
    return w.contacts.size();
  };
  return h->fp64 ? get(h->d) : get(h->f);
}

size_t oracle_get_world_masses(OracleHandle* h, float* out, size_t cap) {
  auto get = [&](auto& w) {
    size_t n = std::min(cap, w.world_masses.size());
    for (size_t i = 0; i < n; ++i) put4(out + 4 * i, w.world_masses[i]);
    return w.world_masses.size();
  };
  return h->fp64 ? get(h->d) : get(h->f);
}

// Articulated rows: bearings (80-byte mh_bearing rows), mechanism bodies
// (64-byte rows), preorder, drives (32-byte rows; NULL = passive).
void oracle_set_mechanisms(OracleHandle* h, uint32_t nb, const float* bearings, const float* mbodies,
                           const uint32_t* preorder, uint32_t nq, const float* drives) {
  auto set = [&](auto& w) {
    using W = typename std::remove_reference<decltype(w)>::type;
    using R = typename std::remove_reference<decltype(w.dt)>::type;
    w.bearing_count = nb;
    w.coordinate_count = nq;
    w.bearings.resize(nb);
    for (uint32_t i = 0; i < nb; ++i) {
      const float* r = bearings + 20 * (size_t)i;
      uint32_t meta[4];
      std::memcpy(meta, r + 16, 16);
      w.bearings[i] = typename W::BearingRow{{R(r[0]), R(r[1]), R(r[2])}, {R(r[4]), R(r[5]), R(r[6])},
                                             {R(r[8]), R(r[9]), R(r[10])}, {R(r[12]), R(r[13]), R(r[14])},
                                             meta[0], meta[1], meta[2], meta[3]};
    }
    const uint32_t n = w.body_count;
    w.mbodies.resize(n);
    w.max_depth = 0;
    w.has_dynamic_root = false;
    for (uint32_t b = 0; b < n; ++b) {
      const float* r = mbodies + 16 * (size_t)b;
      uint32_t meta[4], trav[4];
      std::memcpy(meta, r, 16);
      std::memcpy(trav, r + 4, 16);
      w.mbodies[b] = typename W::MechBody{meta[0], meta[1], meta[2], meta[3], {R(r[8]), R(r[9]), R(r[10])},
                                          {R(r[12]), R(r[13]), R(r[14]), R(r[15])}};
      if (meta[3] == 0u) w.max_depth = std::max(w.max_depth, trav[1]);
      if (meta[3] != 0u && w.inv_mass[b] > R(0)) w.has_dynamic_root = true;
    }
    w.preorder.assign(preorder, preorder + n);
    w.jump_rounds = 0;
    uint32_t covered = 1;
    while (covered < w.max_depth) { covered *= 2; ++w.jump_rounds; }
    if (w.max_depth > 0) w.jump_rounds = std::max(w.jump_rounds, 1u);
    w.mech_active = w.max_depth > 0;
    w.coord_angle.assign(nq, R(0));
    w.coord_speed.assign(nq, R(0));
    w.drives.resize(nq);
    for (uint32_t q = 0; q < nq; ++q) {
      if (drives) {
        const float* r = drives + 8 * (size_t)q;
        uint32_t mode;
        std::memcpy(&mode, r, 4);
        w.drives[q] = typename W::Drive{mode, R(r[1]), R(r[2]), R(r[3]), R(r[4]), R(r[5]), R(r[6])};
      } else {
        w.drives[q] = typename W::Drive{0u, R(0), R(0), R(0), R(0), -R(INFINITY), R(INFINITY)};
      }
    }
    typename W::Link identity{{R(0), R(0), R(0), R(0)}, {R(0), R(0), R(0), R(1)}, 0u, 1u};
    w.links_a.assign(n, identity);
    w.links_b.assign(n, identity);
    for (uint32_t b = 0; b < n; ++b) w.links_a[b].ancestor = w.links_b[b].ancestor = b;
    w.mdeltas.assign(6 * (size_t)n, 0);
  };
  if (h->fp64) set(h->d);
  else set(h->f);
}

void oracle_get_coordinates(OracleHandle* h, float* out) {
  auto get = [&](auto& w) {
    for (uint32_t q = 0; q < w.coordinate_count; ++q) {
      out[2 * q] = (float)w.coord_angle[q];
      out[2 * q + 1] = (float)w.coord_speed[q];
    }
  };
  if (h->fp64) get(h->d);
  else get(h->f);
}

void oracle_set_coordinates(OracleHandle* h, const float* in) {
  auto set = [&](auto& w) {
    using R = typename std::remove_reference<decltype(w.dt)>::type;
    for (uint32_t q = 0; q < w.coordinate_count; ++q) {
      w.coord_angle[q] = R(in[2 * q]);
      w.coord_speed[q] = R(in[2 * q + 1]);
    }
  };
  if (h->fp64) set(h->d);
  else set(h->f);
}

// write_mechanism_drives (device.rs:1622-1644): replaces every coordinate's
// drive row (8 words each, the GpuMechanismDrive layout). Returns 0, or -1 when
// the row count differs from the coordinate count (DriveStateCount).
int oracle_set_drives(OracleHandle* h, uint32_t nq, const float* drives) {
  int rc = 0;
  auto set = [&](auto& w) {
    using W = typename std::remove_reference<decltype(w)>::type;
    using R = typename std::remove_reference<decltype(w.dt)>::type;
    if (nq != w.coordinate_count) {
      rc = -1;
      return;
    }
    for (uint32_t q = 0; q < nq; ++q) {
      const float* r = drives + 8 * (size_t)q;
      uint32_t mode;
      std::memcpy(&mode, r, 4);
      w.drives[q] = typename W::Drive{mode, R(r[1]), R(r[2]), R(r[3]), R(r[4]), R(r[5]), R(r[6])};
    }
  };
  if (h->fp64) set(h->d);
  else set(h->f);
  return rc;
}

// Forward kinematics from the current coordinates, then `steps` closure Newton
// steps (no dynamics), for the closure-solver checks.
void oracle_fk_newton(OracleHandle* h, int steps) {
  auto run = [&](auto& w) {
    forward_kinematics(w);
    if (steps > 0) closure_newton(w, steps);
  };
  if (h->fp64) run(h->d);
  else run(h->f);
}

// Closure solve settings (0 keeps the reference's 12 steps / 8 iterations; a
// tolerance of 0 runs every iteration); returns the PCG iterations run so far.
uint64_t oracle_set_closure(OracleHandle* h, uint32_t steps, uint32_t iterations, float tolerance, int components,
                            int velocity_projection) {
  auto set = [&](auto& w) {
    if (steps) w.closure_steps = steps;
    if (iterations) w.closure_iters = iterations;
    w.closure_tol = tolerance;
    w.closure_components = components != 0;
    w.closure_vproj = velocity_projection != 0;
    return w.pcg_iterations_run;
  };
  return h->fp64 ? set(h->d) : set(h->f);
}

void oracle_set_serial_limit(OracleHandle* h, uint32_t bearings) {
  if (h->fp64) h->d.serial_limit = bearings;
  else h->f.serial_limit = bearings;
}

// Per closure: anchor residual (3) and axis residual (3) from the final links.
size_t oracle_closure_residuals(OracleHandle* h, double* out, size_t cap) {
  auto get = [&](auto& w) {
    size_t k = 0;
    const auto& links = final_links(w);
    for (uint32_t i = 0; i < w.bearing_count; ++i) {
      const auto& r = w.bearings[i];
      if (r.closure == 0u) continue;
      if (k < cap) {
        const auto& pa = links[r.a];
        const auto& pb = links[r.b];
        auto aa = add(xyz(pa.position), quat_rotate(pa.rotation, r.anchor_a));
        auto ab = add(xyz(pb.position), quat_rotate(pb.rotation, r.anchor_b));
        auto xa = normalize_v(quat_rotate(pa.rotation, r.axis_a));
        auto xb = normalize_v(quat_rotate(pb.rotation, r.axis_b));
        double* o = out + 6 * k;
        o[0] = aa.x - ab.x; o[1] = aa.y - ab.y; o[2] = aa.z - ab.z;
        o[3] = xa.x - xb.x; o[4] = xa.y - xb.y; o[5] = xa.z - xb.z;
      }
      ++k;
    }
    return k;
  };
  return h->fp64 ? get(h->d) : get(h->f);
}

// ---- pure helpers for known-answer tests ----------------------------------
uint32_t oracle_morton_code(float x, float y, float z) { return morton_code(V3<float>{x, y, z}); }

// SAT between two boxes given centre, rotation (xyzw) and half extents (f64).
int oracle_obb_sat_f64(const double* ca, const double* ra, const double* ha, const double* cb, const double* rb,
                       const double* hb, double* normal_pen) {
  auto pose = [](const double* c, const double* r, const double* h) {
    ColliderPose<double> p;
    p.center = {c[0], c[1], c[2]};
    p.rotation = {r[0], r[1], r[2], r[3]};
    p.axes[0] = quat_rotate(p.rotation, v3<double>(1, 0, 0));
    p.axes[1] = quat_rotate(p.rotation, v3<double>(0, 1, 0));
    p.axes[2] = quat_rotate(p.rotation, v3<double>(0, 0, 1));
    p.half = {h[0], h[1], h[2]};
    p.half_w = p.center_w = 0;
    p.body = p.role = 0;
    return p;
  };
  ColliderPose<double> a = pose(ca, ra, ha), b = pose(cb, rb, hb);
  Sat<double> s = obb_sat(a, b);
  normal_pen[0] = s.normal.x; normal_pen[1] = s.normal.y; normal_pen[2] = s.normal.z; normal_pen[3] = s.penetration;
  V3<double> p = calculate_contact_point(a, b, s.normal);
  normal_pen[4] = p.x; normal_pen[5] = p.y; normal_pen[6] = p.z;
  return (int)s.near_face_axes;
}

}  // extern "C"
