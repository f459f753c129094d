// Host runtime behind the C ABI: owns the device buffers and one HIP stream per
// handle, encodes a tick as a fixed kernel sequence and replays it as a captured
// hipGraph. Mirrors GpuPhysics (crates/mechanic-gpu/src/device.rs:199-1663).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/mechanic_hip.h"
#include "internal.h"

using namespace mh;

namespace {

struct Timing {
  mh_phase_timings t;
  bool valid = false;
};

// Event marks of a profiled tick.
enum Mark {
  kMarkStart, kMarkIntegrate, kMarkMechanism, kMarkMorton, kMarkSort, kMarkLbvh, kMarkTraverse, kMarkNarrow,
  kMarkPrepare, kMarkSolve, kMarkValidate, kMarkEnd, kEvents
};

}  // namespace

struct mh_physics {
  int device = 0;
  hipStream_t stream = nullptr;
  Dev d{};
  std::vector<void*> allocations;
  mh_config config{};
  bool collisions = false;
  bool graph_enabled = false;
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  uint64_t expected_tick = 0;
  bool tick_known = false;
  uint64_t last_tick = 0;
  bool has_last_tick = false;
  uint32_t* host_diag = nullptr;  // pinned
  hipEvent_t events[kEvents] = {};
  std::vector<hipEvent_t> solve_events;  // 2 per solver pass of a profiled tick
  uint64_t* kclock = nullptr;  // device: per-workgroup clock stamps of profiled solver passes
  size_t kclock_words = 0;
  Timing timing;
  const mh_mlp* mlp = nullptr;
  bool mech_active = false;  // some body has a parent (device.rs:2232)
  std::vector<uint32_t> body_first, scene_first;  // batch layout (empty for one scene)
};

namespace {

int fail(mh_error_detail* e, int code, const char* msg) {
  if (e) {
    e->code = code;
    std::snprintf(e->message, sizeof(e->message), "%s", msg);
  }
  return code;
}

int hip_fail(mh_error_detail* e, hipError_t err, const char* what) {
  char buf[160];
  std::snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(err));
  return fail(e, MH_ERR_DEVICE, buf);
}

#define HIP_TRY(expr)                                  \
  do {                                                 \
    hipError_t err_ = (expr);                          \
    if (err_ != hipSuccess) return hip_fail(error, err_, #expr); \
  } while (0)

#define HIP_TRY_NOERR(expr)                 \
  do {                                      \
    hipError_t err_ = (expr);               \
    if (err_ != hipSuccess) return MH_ERR_DEVICE; \
  } while (0)

template <typename T>
int alloc(mh_physics* p, T** out, size_t count, mh_error_detail* error) {
  if (count == 0) count = 1;
  void* ptr = nullptr;
  HIP_TRY(hipMalloc(&ptr, count * sizeof(T)));
  HIP_TRY(hipMemset(ptr, 0, count * sizeof(T)));
  p->allocations.push_back(ptr);
  *out = static_cast<T*>(ptr);
  return MH_OK;
}

template <typename T>
int upload(mh_physics* p, T* dst, const void* src, size_t bytes, mh_error_detail* error) {
  if (bytes == 0 || src == nullptr) return MH_OK;
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  (void)p;
  return MH_OK;
}

uint32_t next_pow2(uint32_t v) {
  uint32_t r = 1;
  while (r < v) r <<= 1;
  return r;
}

// Enqueues one tick in the pass order of dispatch_tick (device.rs:886-951).
// `stop` ends it early for phase tests (1: after integration and the mechanism
// passes, 2: after the broadphase, 3: after contact generation).
void encode_tick(mh_physics* p, hipStream_t s, hipEvent_t* ev, int stop = 0, uint64_t* clock = nullptr,
                 hipEvent_t* solve_events = nullptr) {
  const Dev& d = p->d;
  const bool mech = p->mech_active;
  auto mark = [&](int i) {
    if (ev) (void)hipEventRecord(ev[i], s);
  };
  mark(kMarkStart);
  // Unarticulated ticks fuse world masses and the snapshot into integration:
  // no later kernel moves a body.
  launch_integrate(s, d, p->collisions && !mech, !mech);
  mark(kMarkIntegrate);
  if (mech) {
    // encode_mechanism_passes (device.rs:963-1048).
    const uint32_t iterations = d.solver_iterations ? d.solver_iterations : 1u;
    launch_bearing_passes(s, d, iterations);
    launch_advance_coordinates(s, d);
    launch_mechanism_pose(s, d, d.closure_count > 0 ? 1 : 0);
  }
  mark(kMarkMechanism);
  if (p->collisions && stop != 1) {
    if (mech) launch_world_masses(s, d);
    launch_morton(s, d);
    mark(kMarkMorton);
    launch_sort(s, d);
    mark(kMarkSort);
    launch_lbvh(s, d);
    mark(kMarkLbvh);
    launch_traverse(s, d);
    mark(kMarkTraverse);
    if (stop == 2) return;
    launch_narrowphase(s, d);
    mark(kMarkNarrow);
    if (stop == 3) return;
    launch_prepare(s, d);
    if (d.learn_mode == kLearnLog) launch_features(s, d);
    if (d.learn_mode == kLearnInfer) launch_mlp_infer(s, d);
    mark(kMarkPrepare);
    launch_solve(s, d, mech && d.has_dynamic_root, clock, solve_events);
    if (d.learn_mode == kLearnLog) launch_targets(s, d);
    // With the learned warm start's fallback the solve already measured the
    // residual; it is measured again only if the extra passes ran.
    if (d.residual_on)
      launch_residual(s, d, learned_fallback_on(d) && !(mech && d.has_dynamic_root) ? d.learned_state : nullptr);
    if (mech) {
      // encode_post_contact_mechanism (device.rs:1128-1155).
      if (!d.has_dynamic_root) {
        const uint32_t iterations = d.solver_iterations ? d.solver_iterations : 1u;
        launch_bearing_passes(s, d, iterations);
      }
      launch_capture_coordinates(s, d);
      if (!d.has_dynamic_root) launch_reconstruct(s, d);
    }
    mark(kMarkSolve);
  } else {
    if (stop != 0) return;
    for (int i = kMarkMorton; i <= kMarkSolve; ++i) mark(i);
  }
  if (mech) launch_validate_snapshot(s, d);
  if (d.bearing_count > 0) launch_validate_bearings(s, d);
  mark(kMarkValidate);
  launch_end_tick(s, d);
  mark(kMarkEnd);
}

int begin_tick(mh_physics* p, uint64_t tick) {
  if (p->d.learn_mode == kLearnInfer && p->d.mlp == nullptr) return MH_ERR_INVALID_ARGUMENT;
  if (!p->tick_known || p->expected_tick != tick) {
    launch_set_tick(p->stream, p->d, tick);
    HIP_TRY_NOERR(hipGetLastError());
  }
  p->expected_tick = tick + 1;
  p->tick_known = true;
  p->last_tick = tick;
  p->has_last_tick = true;
  return MH_OK;
}

int ensure_graph(mh_physics* p) {
  if (p->graph_exec) return MH_OK;
  HIP_TRY_NOERR(hipStreamBeginCapture(p->stream, hipStreamCaptureModeRelaxed));
  encode_tick(p, p->stream, nullptr);
  hipGraph_t g = nullptr;
  HIP_TRY_NOERR(hipStreamEndCapture(p->stream, &g));
  p->graph = g;
  HIP_TRY_NOERR(hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0));
  return MH_OK;
}

bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

// Host-side derivation of the forest lists the articulated kernels walk
// (what create_mechanism_resources derives, device.rs:1684-2232).
struct Forest {
  std::vector<uint32_t> mlist, level_offset, blist, closure_rows, cpath_offset, cpath, coord_offset, coord_entries,
      coord_body, fk_list, fk_offset, fk_meta;
  uint32_t max_depth = 0, jump_rounds = 0;
  bool fk_small = false;
  // Bearing islands packed into groups for the fused state-array projection
  // passes: per group its bearing rows (bi_offset), each row's two bodies as
  // group-local indices (bi_side: a | b << 16), and its bodies (bi_boff,
  // bi_body; bit 31: this group writes the body back). Empty when some group
  // would exceed kIslandBodies bodies or kIslandBearings rows.
  std::vector<uint32_t> bi_offset, bi_bearing, bi_side, bi_boff, bi_body;
  bool has_dynamic_root = false;
  // Closure components (closures linked by shared coordinates), each with its
  // path bodies, coordinates and closures in local numbering (see Dev).
  std::vector<uint32_t> cc_off, cc_body, cc_body_parent, cc_body_coord, cc_coord, cc_coord_body, cc_entry_off,
      cc_entries, cc_closure, cc_end, cc_path_off, cc_path;
  uint32_t cc_count = 0, cc_max_bodies = 0, cc_max_coords = 0, cc_max_closures = 0, cc_max_path = 0;
};

// Closure components of a forest: closures joined when their branches share a
// coordinate, with every component's path bodies (closed under parents up to
// the roots), coordinates and closures renumbered locally.
void build_components(const mh_creation* c, Forest& f, const std::vector<uint32_t>& parent,
                      const std::vector<uint32_t>& root, const std::vector<uint32_t>& tree_row) {
  const uint32_t nk = (uint32_t)f.closure_rows.size(), nq = c->coordinate_count;
  if (nk == 0) return;
  std::vector<uint32_t> up(nk);
  for (uint32_t k = 0; k < nk; ++k) up[k] = k;
  auto find = [&](uint32_t x) {
    while (up[x] != x) x = up[x] = up[up[x]];
    return x;
  };
  std::vector<uint32_t> owner(nq, kInvalid);
  for (uint32_t k = 0; k < nk; ++k)
    for (uint32_t e = f.cpath_offset[2 * k]; e < f.cpath_offset[2 * k + 2]; ++e) {
      const uint32_t q = f.cpath[e];
      if (owner[q] == kInvalid) owner[q] = k;
      else up[find(k)] = find(owner[q]);
    }
  std::vector<std::vector<uint32_t>> members;
  std::vector<uint32_t> id(nk, kInvalid);
  for (uint32_t k = 0; k < nk; ++k) {
    const uint32_t r = find(k);
    if (id[r] == kInvalid) {
      id[r] = (uint32_t)members.size();
      members.emplace_back();
    }
    members[id[r]].push_back(k);
  }
  f.cc_count = (uint32_t)members.size();
  std::vector<uint32_t> local_body(c->body_count, kInvalid), local_coord(nq, kInvalid);
  f.cc_entry_off.push_back(0);
  f.cc_path_off.push_back(0);
  for (const auto& ks : members) {
    const uint32_t body_first = (uint32_t)f.cc_body.size(), coord_first = (uint32_t)f.cc_coord.size();
    const uint32_t closure_first = (uint32_t)f.cc_closure.size();
    // Path bodies: every non-root body from a closure end up to its root.
    std::vector<uint32_t> bodies, coords;
    for (uint32_t k : ks)
      for (uint32_t side = 0; side < 2; ++side) {
        uint32_t b = c->bearings[f.closure_rows[k]].metadata[side];
        while (!root[b] && local_body[b] == kInvalid) {
          local_body[b] = 0u;  // marked; numbered below
          bodies.push_back(b);
          b = parent[b];
        }
      }
    std::sort(bodies.begin(), bodies.end());
    for (uint32_t i = 0; i < (uint32_t)bodies.size(); ++i) local_body[bodies[i]] = i;
    for (uint32_t b : bodies) coords.push_back(c->bearings[tree_row[b]].metadata[2]);
    std::sort(coords.begin(), coords.end());
    for (uint32_t i = 0; i < (uint32_t)coords.size(); ++i) local_coord[coords[i]] = i;
    for (uint32_t b : bodies) {
      f.cc_body.push_back(b);
      f.cc_body_parent.push_back(root[parent[b]] ? kInvalid : local_body[parent[b]]);
      f.cc_body_coord.push_back(local_coord[c->bearings[tree_row[b]].metadata[2]]);
    }
    std::vector<std::vector<uint32_t>> entries(coords.size());
    for (uint32_t kk = 0; kk < (uint32_t)ks.size(); ++kk) {
      const uint32_t k = ks[kk];
      f.cc_closure.push_back(k);
      for (uint32_t side = 0; side < 2; ++side) {
        const uint32_t b = c->bearings[f.closure_rows[k]].metadata[side];
        f.cc_end.push_back(root[b] ? kInvalid : local_body[b]);
        for (uint32_t e = f.cpath_offset[2 * k + side]; e < f.cpath_offset[2 * k + side + 1]; ++e) {
          const uint32_t l = local_coord[f.cpath[e]];
          f.cc_path.push_back(l);
          entries[l].push_back((kk << 1) | side);
        }
        f.cc_path_off.push_back((uint32_t)f.cc_path.size());
      }
    }
    for (uint32_t l = 0; l < (uint32_t)coords.size(); ++l) {
      f.cc_coord.push_back(coords[l]);
      f.cc_coord_body.push_back(local_body[f.coord_body[coords[l]]]);
      f.cc_entries.insert(f.cc_entries.end(), entries[l].begin(), entries[l].end());
      f.cc_entry_off.push_back((uint32_t)f.cc_entries.size());
    }
    const uint32_t nb = (uint32_t)bodies.size(), ncq = (uint32_t)coords.size(), ncl = (uint32_t)ks.size();
    f.cc_off.insert(f.cc_off.end(), {body_first, nb, coord_first, ncq, closure_first, ncl, 0u, 0u});
    f.cc_max_bodies = std::max(f.cc_max_bodies, nb);
    f.cc_max_coords = std::max(f.cc_max_coords, ncq);
    f.cc_max_closures = std::max(f.cc_max_closures, ncl);
    f.cc_max_path = std::max(f.cc_max_path, f.cc_path_off.back() - f.cc_path_off[2 * closure_first]);
    for (uint32_t b : bodies) local_body[b] = kInvalid;
    for (uint32_t q : coords) local_coord[q] = kInvalid;
  }
}

// Bodies joined by bearings through bodies that move. A body with zero inverse
// mass and inverse inertia gets no velocity delta from any bearing (its Q20
// terms are all zero), so it joins nothing: each group that reads it holds a
// copy, which stays bit-identical to the others, and one of them writes it.
void build_islands(const mh_creation* c, Forest& f) {
  const uint32_t n = c->body_count, nb = c->bearing_count;
  auto still = [&](uint32_t b) {
    const mh_mass& m = c->masses[b];
    if (m.inverse_mass[0] != 0.0f) return false;
    for (int k = 0; k < 3; ++k)
      if (m.inverse_inertia_x[k] != 0.0f || m.inverse_inertia_y[k] != 0.0f || m.inverse_inertia_z[k] != 0.0f)
        return false;
    return true;
  };
  std::vector<uint32_t> up(n);
  for (uint32_t b = 0; b < n; ++b) up[b] = b;
  auto find = [&](uint32_t x) {
    while (up[x] != x) x = up[x] = up[up[x]];
    return x;
  };
  std::vector<uint8_t> moving(n, 0);
  for (uint32_t b = 0; b < n; ++b) moving[b] = still(b) ? 0 : 1;
  for (uint32_t i = 0; i < nb; ++i) {
    const uint32_t a = c->bearings[i].metadata[0], b = c->bearings[i].metadata[1];
    if (moving[a] && moving[b]) up[find(a)] = find(b);
  }
  // Island key of a row: its moving body's root, or the row itself when both are still.
  std::vector<std::vector<uint32_t>> rows_of;
  std::vector<uint32_t> island_of(n, kInvalid);
  std::vector<uint32_t> key(nb);
  for (uint32_t i = 0; i < nb; ++i) {
    const uint32_t a = c->bearings[i].metadata[0], b = c->bearings[i].metadata[1];
    uint32_t k;
    if (moving[a] || moving[b]) {
      const uint32_t r = find(moving[a] ? a : b);
      if (island_of[r] == kInvalid) {
        island_of[r] = (uint32_t)rows_of.size();
        rows_of.emplace_back();
      }
      k = island_of[r];
    } else {
      k = (uint32_t)rows_of.size();
      rows_of.emplace_back();
    }
    rows_of[k].push_back(i);
  }
  std::vector<uint32_t> local(n, kInvalid);
  std::vector<uint8_t> written(n, 0);
  std::vector<uint32_t> group_bodies;
  f.bi_offset.assign(1, 0u);
  f.bi_boff.assign(1, 0u);
  auto close_group = [&]() {
    for (uint32_t b : group_bodies) {
      const uint32_t owner = written[b] ? 0u : 0x80000000u;
      written[b] = 1;
      f.bi_body.push_back(b | owner);
      local[b] = kInvalid;
    }
    group_bodies.clear();
    f.bi_offset.push_back((uint32_t)f.bi_bearing.size());
    f.bi_boff.push_back((uint32_t)f.bi_body.size());
  };
  // Bodies of `rows` not yet in the open group (all of them when `all`).
  std::vector<uint8_t> mark(n, 0);
  auto count_new = [&](const std::vector<uint32_t>& rows, bool all) {
    uint32_t count = 0;
    for (uint32_t i : rows)
      for (int side = 0; side < 2; ++side) {
        const uint32_t b = c->bearings[i].metadata[side];
        if ((all || local[b] == kInvalid) && !mark[b]) {
          mark[b] = 1;
          ++count;
        }
      }
    for (uint32_t i : rows)
      for (int side = 0; side < 2; ++side) mark[c->bearings[i].metadata[side]] = 0;
    return count;
  };
  for (const std::vector<uint32_t>& rows : rows_of) {
    if (count_new(rows, true) > kIslandBodies || rows.size() > kIslandBearings) {
      f.bi_offset.clear();
      f.bi_bearing.clear();
      f.bi_side.clear();
      f.bi_boff.clear();
      f.bi_body.clear();
      return;
    }
    const uint32_t extra = count_new(rows, false);
    const uint32_t open_rows = (uint32_t)f.bi_bearing.size() - f.bi_offset.back();
    if (group_bodies.size() + extra > kIslandBodies || open_rows + rows.size() > kIslandBearings) close_group();
    for (uint32_t i : rows) {
      uint32_t idx[2];
      for (int side = 0; side < 2; ++side) {
        const uint32_t b = c->bearings[i].metadata[side];
        if (local[b] == kInvalid) {
          local[b] = (uint32_t)group_bodies.size();
          group_bodies.push_back(b);
        }
        idx[side] = local[b];
      }
      f.bi_bearing.push_back(i);
      f.bi_side.push_back(idx[0] | idx[1] << 16);
    }
  }
  if (f.bi_bearing.size() > f.bi_offset.back()) close_group();
}

int build_forest(const mh_creation* c, Forest& f, mh_error_detail* error) {
  const uint32_t n = c->body_count, nb = c->bearing_count, nq = c->coordinate_count;
  if (!c->bearings || !c->mechanism_bodies || (n > 0 && !c->preorder))
    return fail(error, MH_ERR_INVALID_ARGUMENT, "bearing rows need mechanism_bodies and preorder");
  for (uint32_t i = 0; i < nb; ++i) {
    const uint32_t* m = c->bearings[i].metadata;
    if (m[0] >= n || m[1] >= n || m[0] == m[1])
      return fail(error, MH_ERR_INVALID_ARGUMENT, "bearing joins bodies outside the scene");
    if (m[3] == 0 && m[2] >= nq) return fail(error, MH_ERR_INVALID_ARGUMENT, "tree bearing without a coordinate");
  }
  std::vector<uint32_t> parent(n), depth(n, 0), root(n, 1), preorder_index(n, 0), tree_row(n, 0);
  for (uint32_t b = 0; b < n; ++b) {
    const mh_mechanism_body& m = c->mechanism_bodies[b];
    root[b] = m.metadata[3] != 0 ? 1u : 0u;
    parent[b] = root[b] ? b : m.metadata[0];
    tree_row[b] = m.metadata[1];
    depth[b] = m.traversal[1];
    if (!root[b] && (parent[b] >= n || tree_row[b] >= nb || c->bearings[tree_row[b]].metadata[3] != 0 ||
                     depth[b] == 0 || depth[parent[b]] + 1 != depth[b]))
      return fail(error, MH_ERR_INVALID_ARGUMENT, "inconsistent mechanism body row");
    if (root[b] && c->inverse_masses[b] > 0.0f) f.has_dynamic_root = true;
    f.max_depth = std::max(f.max_depth, root[b] ? 0u : depth[b]);
  }
  for (uint32_t k = 0; k < n; ++k) {
    if (c->preorder[k] >= n) return fail(error, MH_ERR_INVALID_ARGUMENT, "preorder row outside the scene");
    preorder_index[c->preorder[k]] = k;
  }
  // Non-root bodies by (depth, preorder) and the offsets of each depth level.
  for (uint32_t b = 0; b < n; ++b)
    if (!root[b]) f.mlist.push_back(b);
  std::sort(f.mlist.begin(), f.mlist.end(), [&](uint32_t x, uint32_t y) {
    return depth[x] != depth[y] ? depth[x] < depth[y] : preorder_index[x] < preorder_index[y];
  });
  f.level_offset.push_back(0);
  for (uint32_t level = 1; level <= f.max_depth; ++level) {
    uint32_t end = f.level_offset.back();
    while (end < f.mlist.size() && depth[f.mlist[end]] == level) ++end;
    f.level_offset.push_back(end);
  }
  // Trees for the forward-kinematics workgroups: each non-root body's tree root
  // (parents come first in mlist's depth order), bodies of one tree together,
  // whole trees packed into groups of about kFkGroupBodies bodies.
  constexpr uint32_t kFkGroupBodies = 256;
  std::vector<uint32_t> tree(n, kInvalid);
  for (uint32_t b : f.mlist) tree[b] = root[parent[b]] ? parent[b] : tree[parent[b]];
  f.fk_list = f.mlist;
  std::stable_sort(f.fk_list.begin(), f.fk_list.end(), [&](uint32_t x, uint32_t y) { return tree[x] < tree[y]; });
  f.fk_offset.push_back(0);
  for (uint32_t i = 0; i < f.fk_list.size();) {
    uint32_t j = i;
    while (j < f.fk_list.size() && tree[f.fk_list[j]] == tree[f.fk_list[i]]) ++j;  // one tree: [i, j)
    if (f.fk_offset.back() < i && j - f.fk_offset.back() > kFkGroupBodies) f.fk_offset.push_back(i);
    i = j;
  }
  if (!f.fk_list.empty()) f.fk_offset.push_back((uint32_t)f.fk_list.size());
  // Per group entry: its depth and its parent's index in the group (a tree's
  // bodies share a group; a parent outside it is a root), for the publication
  // kernel that keeps a group's rotations and velocities in LDS. Only when
  // every group holds at most 256 bodies (one per thread).
  f.fk_meta.assign(f.fk_list.size(), 0u);
  f.fk_small = f.max_depth < 0xffffu;
  {
    std::vector<uint32_t> local(n, kInvalid);
    for (size_t g = 0; g + 1 < f.fk_offset.size(); ++g) {
      const uint32_t g0 = f.fk_offset[g], g1 = f.fk_offset[g + 1];
      if (g1 - g0 > 256u) f.fk_small = false;
      for (uint32_t i = g0; i < g1; ++i) local[f.fk_list[i]] = i - g0;
      for (uint32_t i = g0; i < g1; ++i) {
        const uint32_t b = f.fk_list[i], pb = parent[b];
        const uint32_t pl = root[pb] ? 0xffffu : local[pb];
        if (!root[pb] && pl == kInvalid) f.fk_small = false;  // not expected: trees are never split
        f.fk_meta[i] = (depth[b] << 16) | (pl & 0xffffu);
      }
      for (uint32_t i = g0; i < g1; ++i) local[f.fk_list[i]] = kInvalid;
    }
  }
  // pointer_jump_rounds (device.rs:2103-2111).
  uint32_t covered = 1;
  while (covered < f.max_depth) {
    covered *= 2;
    ++f.jump_rounds;
  }
  if (f.max_depth > 0) f.jump_rounds = std::max(f.jump_rounds, 1u);
  // Bodies any bearing touches.
  std::vector<uint8_t> touched(n, 0);
  for (uint32_t i = 0; i < nb; ++i) touched[c->bearings[i].metadata[0]] = touched[c->bearings[i].metadata[1]] = 1;
  for (uint32_t b = 0; b < n; ++b)
    if (touched[b]) f.blist.push_back(b);
  build_islands(c, f);
  // Coordinates and their child bodies.
  f.coord_body.assign(nq, kInvalid);
  for (uint32_t b = 0; b < n; ++b)
    if (!root[b]) f.coord_body[c->bearings[tree_row[b]].metadata[2]] = b;
  for (uint32_t q = 0; q < nq; ++q)
    if (f.coord_body[q] == kInvalid) return fail(error, MH_ERR_INVALID_ARGUMENT, "coordinate without a tree body");
  // Closure branches: coordinates from each endpoint body to its root, and
  // the inverse lists in (closure, side) order.
  std::vector<std::vector<uint32_t>> per_coord(nq);
  f.cpath_offset.push_back(0);
  for (uint32_t i = 0; i < nb; ++i) {
    if (c->bearings[i].metadata[3] == 0) continue;
    const uint32_t k = (uint32_t)f.closure_rows.size();
    f.closure_rows.push_back(i);
    for (uint32_t side = 0; side < 2; ++side) {
      uint32_t body = c->bearings[i].metadata[side];
      uint32_t guard = 0;
      while (!root[body]) {
        const uint32_t q = c->bearings[tree_row[body]].metadata[2];
        f.cpath.push_back(q);
        per_coord[q].push_back((k << 1) | side);
        body = parent[body];
        if (++guard > n) return fail(error, MH_ERR_INVALID_ARGUMENT, "mechanism parents form a cycle");
      }
      f.cpath_offset.push_back((uint32_t)f.cpath.size());
    }
  }
  f.coord_offset.push_back(0);
  for (uint32_t q = 0; q < nq; ++q) {
    f.coord_entries.insert(f.coord_entries.end(), per_coord[q].begin(), per_coord[q].end());
    f.coord_offset.push_back((uint32_t)f.coord_entries.size());
  }
  build_components(c, f, parent, root, tree_row);
  return MH_OK;
}


template <typename T>
int alloc_upload(mh_physics* p, T** out, const std::vector<T>& host, mh_error_detail* error) {
  int rc = alloc(p, out, host.size(), error);
  if (rc != MH_OK) return rc;
  return upload(p, *out, host.data(), host.size() * sizeof(T), error);
}

int upload_forest(mh_physics* p, const mh_creation* c, const Forest& f, mh_error_detail* error) {
  Dev& d = p->d;
  const uint32_t n = c->body_count, nb = c->bearing_count, nq = c->coordinate_count;
  d.bearing_count = nb;
  d.coordinate_count = nq;
  d.closure_count = (uint32_t)f.closure_rows.size();
  d.mech_count = (uint32_t)f.mlist.size();
  d.bearing_body_count = (uint32_t)f.blist.size();
  d.level_count = f.max_depth;
  d.jump_rounds = f.jump_rounds;
  d.has_dynamic_root = f.has_dynamic_root ? 1u : 0u;
  p->mech_active = f.max_depth > 0;
  // The serial small-mechanism solvers (device.rs:1370-1373).
  d.serial_mech = (p->mech_active && f.has_dynamic_root && nb <= kSerialBearingLimit) ? 1u : 0u;
  int rc;
#define TRY(x)                  \
  do {                          \
    rc = (x);                   \
    if (rc != MH_OK) return rc; \
  } while (0)
  float4 *bearings, *mbodies, *drives, *l0, *l1, *cl_state, *cl_product, *joints;
  float2* coords;
  int32_t* mdeltas;
  float* pcg;
  TRY(alloc(p, &bearings, 5 * (size_t)nb, error));
  TRY(upload(p, bearings, c->bearings, (size_t)nb * sizeof(mh_bearing), error));
  TRY(alloc(p, &mbodies, 4 * (size_t)n, error));
  TRY(upload(p, mbodies, c->mechanism_bodies, (size_t)n * sizeof(mh_mechanism_body), error));
  TRY(alloc(p, &coords, nq, error));
  // Passive rows when no drives are given (device.rs:1799-1808).
  std::vector<mh_mechanism_drive> dr(nq);
  for (uint32_t q = 0; q < nq; ++q) {
    if (c->drives) {
      dr[q] = c->drives[q];
    } else {
      dr[q] = mh_mechanism_drive{MH_DRIVE_MODE_PASSIVE, 0.0f, 0.0f, 0.0f, 0.0f, -INFINITY, INFINITY, 0.0f};
    }
  }
  TRY(alloc(p, &drives, 2 * (size_t)nq, error));
  TRY(upload(p, drives, dr.data(), (size_t)nq * sizeof(mh_mechanism_drive), error));
  // Links: every body starts as its own resolved root (prepare_links' root rows).
  std::vector<float4> links(3 * (size_t)n);
  for (uint32_t b = 0; b < n; ++b) {
    links[3 * (size_t)b] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    links[3 * (size_t)b + 1] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    uint32_t meta[4] = {b, 1u, 0u, 0u};
    std::memcpy(&links[3 * (size_t)b + 2], meta, 16);
  }
  TRY(alloc_upload(p, &l0, links, error));
  TRY(alloc_upload(p, &l1, links, error));
  TRY(alloc(p, &mdeltas, 6 * (size_t)n, error));
  if (d.serial_mech) {
    TRY(alloc(p, &d.serial_keys, 2 * (size_t)d.pair_capacity, error));
    TRY(alloc(p, &d.serial_order, d.pair_capacity, error));
  }
  uint32_t *mlist, *level_offset, *blist, *closure_rows, *cpath_offset, *cpath, *coord_offset, *coord_entries,
      *coord_body;
  TRY(alloc_upload(p, &mlist, f.mlist, error));
  uint32_t *fk_list, *fk_offset, *mech_state;
  TRY(alloc_upload(p, &fk_list, f.fk_list.empty() ? std::vector<uint32_t>{0u} : f.fk_list, error));
  TRY(alloc_upload(p, &fk_offset, f.fk_offset.empty() ? std::vector<uint32_t>{0u} : f.fk_offset, error));
  TRY(alloc_upload(p, &mech_state, std::vector<uint32_t>(4, 0u), error));
  d.fk_list = fk_list;
  d.fk_offset = fk_offset;
  d.fk_groups = f.fk_offset.empty() ? 0u : (uint32_t)f.fk_offset.size() - 1u;
  uint32_t* fk_meta;
  TRY(alloc_upload(p, &fk_meta, f.fk_meta.empty() ? std::vector<uint32_t>{0u} : f.fk_meta, error));
  d.fk_meta = fk_meta;
  d.fk_small = f.fk_small ? 1u : 0u;
  d.mech_state = mech_state;
  TRY(alloc_upload(p, &level_offset, f.level_offset, error));
  TRY(alloc_upload(p, &blist, f.blist, error));
  d.bi_groups = f.bi_offset.size() > 1 ? (uint32_t)f.bi_offset.size() - 1u : 0u;
  if (d.bi_groups) {
    uint32_t *bi_offset, *bi_bearing, *bi_side, *bi_boff, *bi_body;
    TRY(alloc_upload(p, &bi_offset, f.bi_offset, error));
    TRY(alloc_upload(p, &bi_bearing, f.bi_bearing, error));
    TRY(alloc_upload(p, &bi_side, f.bi_side, error));
    TRY(alloc_upload(p, &bi_boff, f.bi_boff, error));
    TRY(alloc_upload(p, &bi_body, f.bi_body, error));
    d.bi_offset = bi_offset;
    d.bi_bearing = bi_bearing;
    d.bi_side = bi_side;
    d.bi_boff = bi_boff;
    d.bi_body = bi_body;
    std::vector<uint8_t> mask(n, 0);
    for (uint32_t w : f.bi_body) mask[w & 0x7fffffffu] = 1;
    uint8_t* bi_mask;
    TRY(alloc_upload(p, &bi_mask, mask, error));
    d.bi_mask = bi_mask;
  }
  TRY(alloc_upload(p, &closure_rows, f.closure_rows, error));
  TRY(alloc_upload(p, &cpath_offset, f.cpath_offset, error));
  TRY(alloc_upload(p, &cpath, f.cpath, error));
  TRY(alloc_upload(p, &coord_offset, f.coord_offset, error));
  TRY(alloc_upload(p, &coord_entries, f.coord_entries, error));
  TRY(alloc_upload(p, &coord_body, f.coord_body, error));
  TRY(alloc(p, &cl_state, 6 * (size_t)d.closure_count, error));
  TRY(alloc(p, &cl_product, 2 * (size_t)d.closure_count, error));
  TRY(alloc(p, &joints, 2 * (size_t)nq, error));
  TRY(alloc(p, &pcg, 8 * (size_t)nq, error));
  if (p->config.closure_components && f.cc_count > 0) {
    if (f.cc_max_bodies > kCompMaxBodies || f.cc_max_coords > kCompMaxCoords || f.cc_max_closures > kCompMaxClosures ||
        f.cc_max_path > kCompMaxPath)
      return fail(error, MH_ERR_UNSUPPORTED,
                  "a closure component exceeds 256 path bodies, 256 coordinates, 128 closures or 2048 path entries: use "
                  "closure_components = 0");
    d.cc_count = f.cc_count;
    d.cc_max_bodies = f.cc_max_bodies;
    d.cc_max_coords = f.cc_max_coords;
    d.cc_max_closures = f.cc_max_closures;
    d.cc_max_path = f.cc_max_path;
    // One-wavefront components keep J^T J in LDS when it (with J per closure and
    // coordinate) takes at most 48 KiB.
    const uint32_t items = std::max(std::max(f.cc_max_bodies, f.cc_max_coords), 2u * f.cc_max_closures);
    const size_t dense_bytes = 8 * (6 * (size_t)f.cc_max_closures * f.cc_max_coords +
                                    (size_t)f.cc_max_coords * f.cc_max_coords);
    d.cc_dense = items <= 64u && dense_bytes <= 48u * 1024u ? 1u : 0u;
    if (d.cc_dense && closure_component_lds(d) > 160u * 1024u) d.cc_dense = 0u;  // the CU's LDS
    if (prepare_closure_components(d) != hipSuccess)
      return fail(error, MH_ERR_DEVICE, "closure component kernel: LDS attribute");
    uint32_t *off, *body, *body_parent, *body_coord, *coord, *coord_body, *entry_off, *entries, *closure, *end, *path_off,
        *path;
    TRY(alloc_upload(p, &off, f.cc_off, error));
    TRY(alloc_upload(p, &body, f.cc_body, error));
    TRY(alloc_upload(p, &body_parent, f.cc_body_parent, error));
    TRY(alloc_upload(p, &body_coord, f.cc_body_coord, error));
    TRY(alloc_upload(p, &coord, f.cc_coord, error));
    TRY(alloc_upload(p, &coord_body, f.cc_coord_body, error));
    TRY(alloc_upload(p, &entry_off, f.cc_entry_off, error));
    TRY(alloc_upload(p, &entries, f.cc_entries.empty() ? std::vector<uint32_t>{0u} : f.cc_entries, error));
    TRY(alloc_upload(p, &closure, f.cc_closure, error));
    TRY(alloc_upload(p, &end, f.cc_end, error));
    TRY(alloc_upload(p, &path_off, f.cc_path_off, error));
    TRY(alloc_upload(p, &path, f.cc_path.empty() ? std::vector<uint32_t>{0u} : f.cc_path, error));
    d.cc_off = reinterpret_cast<const uint4*>(off);
    d.cc_body = body;
    d.cc_body_parent = body_parent;
    d.cc_body_coord = body_coord;
    d.cc_coord = coord;
    d.cc_coord_body = coord_body;
    d.cc_entry_off = entry_off;
    d.cc_entries = entries;
    d.cc_closure = closure;
    d.cc_end = end;
    d.cc_path_off = path_off;
    d.cc_path = path;
  }
#undef TRY
  d.bearings = bearings;
  d.mbodies = mbodies;
  d.coords = coords;
  d.drives = drives;
  d.links[0] = l0;
  d.links[1] = l1;
  d.mdeltas = mdeltas;
  d.mlist = mlist;
  d.level_offset = level_offset;
  d.blist = blist;
  d.closure_rows = closure_rows;
  d.cpath_offset = cpath_offset;
  d.cpath = cpath;
  d.coord_offset = coord_offset;
  d.coord_entries = coord_entries;
  d.coord_body = coord_body;
  d.cl_state = cl_state;
  d.cl_product = cl_product;
  d.joints = joints;
  d.pcg = pcg;
  return MH_OK;
}

}  // namespace

extern "C" {

mh_config mh_config_default(void) {
  mh_config c{};
  c.collisions_enabled = 1;
  c.mechanism_self_collisions = 1;
  c.solver_iterations = 8;
  c.friction = 0.05f;
  c.pair_capacity = MH_MAX_CONTACT_PAIRS;
  c.warm_start_mode = 0;
  c.use_graph = 1;
  return c;
}

const char* mh_status_string(int code) {
  switch (code) {
    case MH_OK: return "ok";
    case MH_ERR_BODY_CAPACITY: return "body capacity exceeded";
    case MH_ERR_BEARING_CAPACITY: return "bearing capacity exceeded";
    case MH_ERR_COLLIDER_CAPACITY: return "collider capacity exceeded";
    case MH_ERR_DRIVE_STATE_COUNT: return "drive state row count mismatch";
    case MH_ERR_COORDINATE_STATE_COUNT: return "coordinate state row count mismatch";
    case MH_ERR_BODY_INDEX_OUT_OF_RANGE: return "body index out of range";
    case MH_ERR_NON_FINITE: return "non-finite impulse";
    case MH_ERR_INVALID_SNAPSHOT_SLOT: return "invalid snapshot slot";
    case MH_ERR_DEVICE: return "HIP device error";
    case MH_ERR_INVALID_ARGUMENT: return "invalid argument";
    case MH_ERR_NO_DEVICE: return "no HIP device";
    case MH_ERR_UNSUPPORTED: return "unsupported in this build";
    default: return "unknown status";
  }
}

int mh_device_arch(char* out, size_t capacity) {
  if (!out || capacity == 0) return MH_ERR_INVALID_ARGUMENT;
  out[0] = 0;
  int device = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess)
    return MH_ERR_NO_DEVICE;
  std::snprintf(out, capacity, "%s", prop.gcnArchName);
  return MH_OK;
}

int mh_device_round_i32(const float* values, int32_t* out, size_t count) {
  if ((!values || !out) && count) return MH_ERR_INVALID_ARGUMENT;
  if (count == 0) return MH_OK;
  if (count > (1u << 24)) return MH_ERR_INVALID_ARGUMENT;
  float* dv = nullptr;
  int32_t* dq = nullptr;
  HIP_TRY_NOERR(hipMalloc(&dv, count * sizeof(float)));
  hipError_t e = hipMalloc(&dq, count * sizeof(int32_t));
  if (e == hipSuccess) e = hipMemcpy(dv, values, count * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    launch_round_i32(dv, dq, (uint32_t)count);
    e = hipDeviceSynchronize();
  }
  if (e == hipSuccess) e = hipMemcpy(out, dq, count * sizeof(int32_t), hipMemcpyDeviceToHost);
  (void)hipFree(dv);
  if (dq) (void)hipFree(dq);
  HIP_TRY_NOERR(e);
  return MH_OK;
}

int mh_device_count(int* count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (count) *count = (e == hipSuccess) ? n : 0;
  return e == hipSuccess ? MH_OK : MH_ERR_NO_DEVICE;
}

}  // extern "C"

namespace {

// Scene layout of a batch handle.
struct BatchInfo {
  std::vector<uint32_t> cscene;       // scene of each collider
  std::vector<uint32_t> scene_first;  // first collider (= first sorted leaf) of each scene, + total
  std::vector<uint32_t> body_first;   // first body of each scene, + total
};

int create_impl(const mh_creation* creation, const mh_config* config_in, const BatchInfo* batch, mh_physics** out,
                mh_error_detail* error) {
  if (error) std::memset(error, 0, sizeof(*error));
  if (!creation || !out) return fail(error, MH_ERR_INVALID_ARGUMENT, "null creation or output");
  *out = nullptr;
  const uint32_t body_cap = batch ? MH_MAX_BATCH_BODIES : MH_MAX_BODIES;
  const uint32_t collider_cap = batch ? MH_MAX_BATCH_BODIES : MH_MAX_COLLIDERS;
  // Capacity checks before any device work (device.rs:405-422).
  if (creation->body_count > body_cap) {
    if (error) { error->required = creation->body_count; error->capacity = body_cap; }
    return fail(error, MH_ERR_BODY_CAPACITY, "scene exceeds the body capacity");
  }
  if (creation->bearing_count > MH_MAX_BEARINGS) {
    if (error) { error->required = creation->bearing_count; error->capacity = MH_MAX_BEARINGS; }
    return fail(error, MH_ERR_BEARING_CAPACITY, "scene exceeds the bearing capacity");
  }
  if (creation->collider_count > collider_cap) {
    if (error) { error->required = creation->collider_count; error->capacity = collider_cap; }
    return fail(error, MH_ERR_COLLIDER_CAPACITY, "scene exceeds the collider capacity");
  }
  Forest forest;
  if (creation->bearing_count > 0) {
    int frc = build_forest(creation, forest, error);
    if (frc != MH_OK) return frc;
  }
  const uint32_t n = creation->body_count, c = creation->collider_count;
  if ((n > 0 && (!creation->poses || !creation->inverse_masses || !creation->masses)) ||
      (c > 0 && !creation->colliders) || (creation->suppression_count > 0 && !creation->suppressed_pairs))
    return fail(error, MH_ERR_INVALID_ARGUMENT, "missing row arrays");
  for (uint32_t i = 0; i < c; ++i)
    if (creation->colliders[i].metadata[0] >= n)
      return fail(error, MH_ERR_INVALID_ARGUMENT, "collider refers to a body outside the scene");
  mh_config cfg = config_in ? *config_in : mh_config_default();
  if (cfg.solver_iterations == 0) cfg.solver_iterations = 8;
  if (cfg.pair_capacity == 0) cfg.pair_capacity = MH_MAX_CONTACT_PAIRS;
  if (!(cfg.friction >= 0.0f) || !std::isfinite(cfg.friction))
    return fail(error, MH_ERR_INVALID_ARGUMENT, "friction must be finite and non-negative");
  if (cfg.warm_start_mode > MH_WARM_START_LEARNED)
    return fail(error, MH_ERR_INVALID_ARGUMENT, "unknown warm_start_mode");
  if (!(cfg.learned_residual_limit >= 0.0f) || !std::isfinite(cfg.learned_residual_limit) ||
      cfg.learned_fallback_passes > 64)
    return fail(error, MH_ERR_INVALID_ARGUMENT, "learned_residual_limit must be finite and >= 0, passes <= 64");
  if (cfg.closure_newton_steps > 64 || cfg.closure_pcg_iterations > 256 || !(cfg.closure_pcg_tolerance >= 0.0f) ||
      !std::isfinite(cfg.closure_pcg_tolerance))
    return fail(error, MH_ERR_INVALID_ARGUMENT, "closure steps <= 64, PCG iterations <= 256, tolerance >= 0");
  if (cfg.closure_velocity_projection && !cfg.closure_components)
    return fail(error, MH_ERR_INVALID_ARGUMENT, "closure_velocity_projection needs closure_components");
  if ((cfg.pair_capacity & (cfg.pair_capacity - 1)) != 0 || cfg.pair_capacity < 256 ||
      cfg.pair_capacity > (1u << 22))
    return fail(error, MH_ERR_INVALID_ARGUMENT, "pair_capacity must be a power of two in [256, 2^22]");

  int devices = 0;
  if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0)
    return fail(error, MH_ERR_NO_DEVICE, "no HIP device visible");

  mh_physics* p = new mh_physics();
  p->config = cfg;
  p->collisions = cfg.collisions_enabled != 0 && c > 0;
  p->graph_enabled = cfg.use_graph != 0;
  HIP_TRY(hipGetDevice(&p->device));
  int rc = MH_OK;
#define ALLOC(ptr, count)                            \
  do {                                               \
    rc = alloc(p, &(ptr), (count), error);           \
    if (rc != MH_OK) { mh_physics_destroy(p); return rc; } \
  } while (0)
#define UPLOAD(dst, src, bytes)                          \
  do {                                                   \
    rc = upload(p, (dst), (src), (bytes), error);        \
    if (rc != MH_OK) { mh_physics_destroy(p); return rc; } \
  } while (0)

  hipError_t se = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
  if (se != hipSuccess) { delete p; return hip_fail(error, se, "hipStreamCreateWithFlags"); }
  for (auto& e : p->events) (void)hipEventCreate(&e);

  Dev& d = p->d;
  d.body_count = n;
  d.collider_count = c;
  d.sort_count = next_pow2(c < 256 ? 256 : c);
  d.suppression_count = creation->suppression_count;
  d.pair_capacity = cfg.pair_capacity;
  d.hash_mask = cfg.pair_capacity - 1;
  d.solver_iterations = cfg.solver_iterations;
  d.self_collisions = cfg.mechanism_self_collisions ? 1u : 0u;
  d.delta_seconds = 1.0f / 60.0f;  // FIXED_DT_SECONDS, lib.rs:33
  d.gravity_y = -9.81f;            // device.rs:872
  d.linear_damping = 0.999f;       // device.rs:873
  d.angular_damping = 0.98f;       // device.rs:874
  d.friction = cfg.friction;
  d.learn_mode = cfg.warm_start_mode;
  d.residual_on = cfg.residual ? 1u : 0u;
  d.learned_limit = cfg.learned_residual_limit;
  d.learned_extra = cfg.learned_fallback_passes;
  d.closure_steps = cfg.closure_newton_steps ? cfg.closure_newton_steps : 12u;  // device.rs:977
  d.closure_iters = cfg.closure_pcg_iterations ? cfg.closure_pcg_iterations : 8u;  // closure.wgsl:358
  d.closure_tol = cfg.closure_pcg_tolerance;
  d.closure_vproj = cfg.closure_velocity_projection ? 1u : 0u;

  float4 *pos, *rot, *lin, *ang, *masses, *wm, *cols, *aabbs, *cposes, *inodes, *leaves, *contacts, *rows;
  float* inv_mass;
  float* lambda;
  float4 *sbody, *sbody_alt, *srow;
  float* slam;
  uint2* side_slot;
  uint32_t *side_off, *side_status;
  int4* cdelta;
  uint32_t *roots, *comps, *parents, *active, *diag;
  uint2 *supp, *children, *pairs, *ranges;
  uint32_t *leaf_body, *top, *counters;
  float4 *block_box, *leaf_scan;
  uint4* work;
  uint64_t *keys, *h0, *h1;
  int32_t* deltas;
  uint32_t* slot_count;
  float4 *m0, *m1;
  TickState* tick;
  ALLOC(pos, n); ALLOC(rot, n); ALLOC(lin, n); ALLOC(ang, n);
  ALLOC(inv_mass, n); ALLOC(masses, 4 * (size_t)n); ALLOC(wm, 3 * (size_t)n);
  ALLOC(roots, n); ALLOC(comps, n);
  ALLOC(cols, 4 * (size_t)c); ALLOC(supp, creation->suppression_count);
  ALLOC(aabbs, 2 * (size_t)c); ALLOC(cposes, 5 * (size_t)c); ALLOC(keys, d.sort_count);
  ALLOC(inodes, 2 * (size_t)(c > 1 ? c - 1 : 1)); ALLOC(leaves, 2 * (size_t)c);
  ALLOC(children, c > 1 ? c - 1 : 1); ALLOC(parents, c > 0 ? 2 * (size_t)c - 1 : 1);
  ALLOC(ranges, c > 1 ? c - 1 : 1); ALLOC(leaf_body, c);
  ALLOC(top, c); ALLOC(block_box, 2 * (size_t)((c + kBlock - 1) / kBlock)); ALLOC(leaf_scan, 4 * (size_t)(c > 0 ? c : 1)); ALLOC(counters, kCounterWords);
  const uint32_t packets = (c + 63) / 64;
  d.work_capacity = 64 * packets < 4096 ? 4096 : 64 * packets;
  ALLOC(work, d.work_capacity);
  uint32_t *work_count, *work_first;
  ALLOC(work_count, packets);
  ALLOC(work_first, (size_t)packets + 1);
  d.work_count = work_count;
  d.work_first = work_first;
  ALLOC(pairs, cfg.pair_capacity);
  ALLOC(contacts, 4 * (size_t)cfg.pair_capacity);
  ALLOC(active, cfg.pair_capacity);
  ALLOC(rows, 4 * (size_t)cfg.pair_capacity);
  ALLOC(lambda, cfg.pair_capacity);
  ALLOC(cdelta, 2 * (size_t)kMaxSlots * n);
  ALLOC(sbody, 4 * (size_t)n); ALLOC(sbody_alt, 4 * (size_t)n);
  d.side_tiles = (n + kSideTile - 1) / kSideTile;
  ALLOC(side_off, n); ALLOC(side_status, d.side_tiles ? d.side_tiles : 1u);
  ALLOC(side_slot, cfg.pair_capacity);
  ALLOC(d.side_group, cfg.pair_capacity);
  ALLOC(srow, 8 * (size_t)cfg.pair_capacity); ALLOC(slam, 2 * (size_t)cfg.pair_capacity);
  ALLOC(deltas, 6 * (size_t)n);
  ALLOC(slot_count, n);
  ALLOC(h0, cfg.pair_capacity); ALLOC(h1, cfg.pair_capacity);
  ALLOC(m0, 3 * (size_t)cfg.pair_capacity); ALLOC(m1, 3 * (size_t)cfg.pair_capacity);
  for (int s = 0; s < 3; ++s) {
    ALLOC(d.snap_pos[s], n);
    ALLOC(d.snap_rot[s], n);
  }
  ALLOC(diag, 2 * kDiagWords);
  ALLOC(tick, 1);
  if (d.learn_mode != kLearnOff) {
    ALLOC(d.feat, (size_t)kFeatures * cfg.pair_capacity);
    ALLOC(d.fsum, cfg.pair_capacity);
    ALLOC(d.finit, cfg.pair_capacity);
  }
  ALLOC(d.resid, 4 * (size_t)kResidBlocks);
  ALLOC(d.learned_state, 4);
  d.pos = pos; d.rot = rot; d.lin = lin; d.ang = ang;
  d.inv_mass = inv_mass; d.masses = masses; d.world_masses = wm;
  d.root_flags = roots; d.components = comps;
  d.colliders = cols; d.suppressed = supp;
  d.aabbs = aabbs; d.cposes = cposes; d.keys = keys; d.inodes = inodes; d.leaves = leaves;
  d.children = children; d.parents = parents; d.pairs = pairs;
  d.ranges = ranges; d.leaf_body = leaf_body; d.top = top; d.block_box = block_box; d.leaf_scan = leaf_scan; d.work = work; d.counters = counters;
  d.contacts = contacts; d.active = active; d.rows = rows; d.deltas = deltas;
 d.slot_count = slot_count; d.lambda = lambda; d.cdelta = cdelta; d.sbody = sbody;
  d.sbody_alt = sbody_alt; d.side_off = side_off; d.side_status = side_status; d.side_slot = side_slot;
  d.srow = srow; d.slam = slam;
  d.htab[0] = h0; d.htab[1] = h1; d.mrows[0] = m0; d.mrows[1] = m1;
  d.diag[0] = diag; d.diag[1] = diag + kDiagWords;
  d.tick = tick;

  // Lowering (device.rs:424-523): poses split into position/rotation planes,
  // zero velocities, inverse masses, mass rows, colliders, suppression pairs.
  if (n > 0) {
    std::vector<float4> hp(n), hr(n);
    for (uint32_t i = 0; i < n; ++i) {
      const float* q = creation->poses[i].position;
      const float* r = creation->poses[i].rotation;
      hp[i] = make_float4(q[0], q[1], q[2], 0.0f);
      hr[i] = make_float4(r[0], r[1], r[2], r[3]);
    }
    UPLOAD(pos, hp.data(), n * sizeof(float4));
    UPLOAD(rot, hr.data(), n * sizeof(float4));
    UPLOAD(inv_mass, creation->inverse_masses, n * sizeof(float));
    UPLOAD(masses, creation->masses, n * sizeof(mh_mass));
    std::vector<uint32_t> hroots(n, 1u), hcomps(n);
    if (creation->root_flags)
      for (uint32_t i = 0; i < n; ++i) hroots[i] = creation->root_flags[i];
    for (uint32_t i = 0; i < n; ++i) hcomps[i] = creation->body_components ? creation->body_components[i] : i;
    UPLOAD(roots, hroots.data(), n * sizeof(uint32_t));
    UPLOAD(comps, hcomps.data(), n * sizeof(uint32_t));
  }
  UPLOAD(cols, creation->colliders, (size_t)c * sizeof(mh_collider));
  UPLOAD(supp, creation->suppressed_pairs, (size_t)creation->suppression_count * sizeof(mh_pair));
  {
    // Each body's suppressed partners, sorted: the traversal's membership test.
    std::vector<uint32_t> off(n + 1, 0), adj(2 * (size_t)creation->suppression_count);
    for (uint32_t i = 0; i < creation->suppression_count; ++i) {
      const mh_pair& q = creation->suppressed_pairs[i];
      if (q.collider_a >= n || q.collider_b >= n) {
        mh_physics_destroy(p);
        return fail(error, MH_ERR_INVALID_ARGUMENT, "a suppressed pair names no body");
      }
      ++off[q.collider_a + 1];
      ++off[q.collider_b + 1];
    }
    for (uint32_t b = 0; b < n; ++b) off[b + 1] += off[b];
    std::vector<uint32_t> fill(off.begin(), off.end() - 1);
    for (uint32_t i = 0; i < creation->suppression_count; ++i) {
      const mh_pair& q = creation->suppressed_pairs[i];
      adj[fill[q.collider_a]++] = q.collider_b;
      adj[fill[q.collider_b]++] = q.collider_a;
    }
    for (uint32_t b = 0; b < n; ++b) std::sort(adj.begin() + off[b], adj.begin() + off[b + 1]);
    uint32_t *soff, *sadj;
    ALLOC(soff, (size_t)n + 1);
    ALLOC(sadj, adj.size());
    UPLOAD(soff, off.data(), off.size() * sizeof(uint32_t));
    if (!adj.empty()) UPLOAD(sadj, adj.data(), adj.size() * sizeof(uint32_t));
    d.supp_off = soff;
    d.supp_adj = sadj;
  }
  if (creation->bearing_count > 0) {
    rc = upload_forest(p, creation, forest, error);
    if (rc != MH_OK) { mh_physics_destroy(p); return rc; }
  }
  {
    uint64_t* keys_alt;
    ALLOC(keys_alt, c);
    d.keys_alt = keys_alt;
    d.rs_blocks = (c + kRadixTile - 1) / kRadixTile;
    uint32_t* rs_hist;
    ALLOC(rs_hist, kRadixHistWords);
    d.rs_hist = rs_hist;
  }
  d.scene_count = 1;
  d.key_index_bits = 32;
  if (batch) {
    uint32_t *cscene, *scene_first;
    ALLOC(cscene, c);
    ALLOC(scene_first, batch->scene_first.size());
    UPLOAD(cscene, batch->cscene.data(), (size_t)c * sizeof(uint32_t));
    UPLOAD(scene_first, batch->scene_first.data(), batch->scene_first.size() * sizeof(uint32_t));
    d.cscene = cscene;
    d.scene_first = scene_first;
    d.scene_count = (uint32_t)batch->scene_first.size() - 1u;
    uint32_t bits = 1;
    while ((1u << bits) < c) ++bits;
    d.key_index_bits = bits;
    p->body_first = batch->body_first;
    p->scene_first = batch->scene_first;
  }
  {
    // Digit passes over the (scene, code) bits above the collider field.
    uint32_t scene_bits = 0;
    while ((1u << scene_bits) < d.scene_count) ++scene_bits;
    d.rs_passes = (30u + scene_bits + 7u) / 8u;
    uint32_t* rs_status;
    ALLOC(rs_status, (size_t)d.rs_passes * 256u * (d.rs_blocks ? d.rs_blocks : 1u));
    d.rs_status = rs_status;
  }
  hipError_t he = hipHostMalloc(reinterpret_cast<void**>(&p->host_diag), kDiagWords * sizeof(uint32_t),
                                hipHostMallocDefault);
  if (he != hipSuccess) { mh_physics_destroy(p); return hip_fail(error, he, "hipHostMalloc"); }
  he = hipDeviceSynchronize();
  if (he != hipSuccess) { mh_physics_destroy(p); return hip_fail(error, he, "upload"); }
  *out = p;
  return MH_OK;
#undef ALLOC
#undef UPLOAD
}

}  // namespace

extern "C" {

int mh_physics_create(const mh_creation* creation, const mh_config* config, mh_physics** out,
                      mh_error_detail* error) {
  return create_impl(creation, config, nullptr, out, error);
}

// Several independent scenes in one handle: bodies, colliders and suppression
// pairs are concatenated scene after scene; scene ids in the sort keys keep
// every scene's broadphase to itself, so each scene steps exactly as it would
// alone.
int mh_batch_create(const mh_creation* scenes, uint32_t scene_count, const mh_config* config, mh_physics** out,
                    mh_error_detail* error) {
  if (error) std::memset(error, 0, sizeof(*error));
  if (!scenes || !out || scene_count == 0 || scene_count > MH_MAX_BATCH_SCENES)
    return fail(error, MH_ERR_INVALID_ARGUMENT, "batch needs 1..MH_MAX_BATCH_SCENES scenes");
  *out = nullptr;
  BatchInfo info;
  uint64_t bodies = 0, colliders = 0, supp = 0;
  for (uint32_t s = 0; s < scene_count; ++s) {
    const mh_creation& c = scenes[s];
    if (c.bearing_count > 0) return fail(error, MH_ERR_UNSUPPORTED, "batched scenes cannot hold bearings");
    if (c.body_count > MH_MAX_BODIES) {
      if (error) { error->required = c.body_count; error->capacity = MH_MAX_BODIES; error->body_index = s; }
      return fail(error, MH_ERR_BODY_CAPACITY, "a batched scene exceeds the body capacity");
    }
    if (c.collider_count > MH_MAX_COLLIDERS) {
      if (error) { error->required = c.collider_count; error->capacity = MH_MAX_COLLIDERS; error->body_index = s; }
      return fail(error, MH_ERR_COLLIDER_CAPACITY, "a batched scene exceeds the collider capacity");
    }
    if ((c.body_count > 0 && (!c.poses || !c.inverse_masses || !c.masses)) ||
        (c.collider_count > 0 && !c.colliders) || (c.suppression_count > 0 && !c.suppressed_pairs))
      return fail(error, MH_ERR_INVALID_ARGUMENT, "missing row arrays");
    bodies += c.body_count;
    colliders += c.collider_count;
    supp += c.suppression_count;
  }
  if (bodies > MH_MAX_BATCH_BODIES || colliders > MH_MAX_BATCH_BODIES) {
    if (error) { error->required = bodies > colliders ? bodies : colliders; error->capacity = MH_MAX_BATCH_BODIES; }
    return fail(error, MH_ERR_BODY_CAPACITY, "batch exceeds the batch body capacity");
  }
  std::vector<mh_transform> poses;
  std::vector<float> inv;
  std::vector<mh_mass> masses;
  std::vector<mh_collider> cols;
  std::vector<mh_pair> pairs;
  std::vector<uint32_t> comps, roots;
  poses.reserve(bodies); inv.reserve(bodies); masses.reserve(bodies); cols.reserve(colliders); pairs.reserve(supp);
  comps.reserve(bodies); roots.reserve(bodies);
  info.cscene.reserve(colliders);
  for (uint32_t s = 0; s < scene_count; ++s) {
    const mh_creation& c = scenes[s];
    const uint32_t b0 = (uint32_t)poses.size();
    info.body_first.push_back(b0);
    info.scene_first.push_back((uint32_t)cols.size());
    for (uint32_t i = 0; i < c.body_count; ++i) {
      poses.push_back(c.poses[i]);
      inv.push_back(c.inverse_masses[i]);
      masses.push_back(c.masses[i]);
      comps.push_back(c.body_components ? b0 + c.body_components[i] : b0 + i);
      roots.push_back(c.root_flags ? c.root_flags[i] : 1u);
    }
    for (uint32_t i = 0; i < c.collider_count; ++i) {
      mh_collider row = c.colliders[i];
      if (row.metadata[0] >= c.body_count)
        return fail(error, MH_ERR_INVALID_ARGUMENT, "collider refers to a body outside its scene");
      row.metadata[0] += b0;
      cols.push_back(row);
      info.cscene.push_back(s);
    }
    for (uint32_t i = 0; i < c.suppression_count; ++i)
      pairs.push_back(mh_pair{c.suppressed_pairs[i].collider_a + b0, c.suppressed_pairs[i].collider_b + b0});
  }
  info.body_first.push_back((uint32_t)poses.size());
  info.scene_first.push_back((uint32_t)cols.size());
  mh_creation merged{};
  merged.body_count = (uint32_t)poses.size();
  merged.collider_count = (uint32_t)cols.size();
  merged.suppression_count = (uint32_t)pairs.size();
  merged.poses = poses.data();
  merged.inverse_masses = inv.data();
  merged.masses = masses.data();
  merged.colliders = cols.data();
  merged.suppressed_pairs = pairs.data();
  merged.body_components = comps.data();
  merged.root_flags = roots.data();
  return create_impl(&merged, config, &info, out, error);
}

int mh_batch_scene_range(const mh_physics* p, uint32_t scene, uint32_t* body_first, uint32_t* body_count,
                         uint32_t* collider_first, uint32_t* collider_count) {
  if (!p) return MH_ERR_INVALID_ARGUMENT;
  if (p->body_first.empty()) {
    if (scene != 0) return MH_ERR_INVALID_ARGUMENT;
    if (body_first) *body_first = 0;
    if (body_count) *body_count = p->d.body_count;
    if (collider_first) *collider_first = 0;
    if (collider_count) *collider_count = p->d.collider_count;
    return MH_OK;
  }
  if (scene + 1 >= p->body_first.size()) return MH_ERR_INVALID_ARGUMENT;
  if (body_first) *body_first = p->body_first[scene];
  if (body_count) *body_count = p->body_first[scene + 1] - p->body_first[scene];
  if (collider_first) *collider_first = p->scene_first[scene];
  if (collider_count) *collider_count = p->scene_first[scene + 1] - p->scene_first[scene];
  return MH_OK;
}

uint32_t mh_batch_scene_count(const mh_physics* p) {
  return p ? (p->body_first.empty() ? 1u : (uint32_t)p->body_first.size() - 1u) : 0u;
}

int mh_physics_set_tuning(mh_physics* p, const mh_tuning* t) {
  if (!p || !t) return MH_ERR_INVALID_ARGUMENT;
  if (t->sort_tile && (t->sort_tile < 256 || t->sort_tile > 2048 || (t->sort_tile & (t->sort_tile - 1))))
    return MH_ERR_INVALID_ARGUMENT;
  if (t->sort_levels > 4 || t->sort_kind > 1 || t->traverse_kind > 2 || t->solve_kind > 1)
    return MH_ERR_INVALID_ARGUMENT;
  if (t->narrow_blocks > (1u << 16) || t->prepare_blocks > (1u << 16) || t->solve_blocks > (1u << 16) ||
      t->scan_blocks > (1u << 16))
    return MH_ERR_INVALID_ARGUMENT;
  HIP_TRY_NOERR(hipStreamSynchronize(p->stream));
  Dev& d = p->d;
  d.t_sort_kind = t->sort_kind;
  d.t_sort_tile = t->sort_tile;
  d.t_sort_levels = t->sort_levels;
  d.t_narrow_blocks = t->narrow_blocks;
  d.t_prepare_blocks = t->prepare_blocks;
  d.t_solve_blocks = t->solve_blocks;
  d.t_scan_blocks = t->scan_blocks;
  d.t_traverse_kind = t->traverse_kind;
  // The serial small-mechanism path warms up with the per-contact pass.
  d.t_solve_kind = d.serial_mech ? 0u : t->solve_kind;
  if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->graph_exec = nullptr;
  p->graph = nullptr;
  return MH_OK;
}

int mh_physics_get_tuning(const mh_physics* p, mh_tuning* t) {
  if (!p || !t) return MH_ERR_INVALID_ARGUMENT;
  std::memset(t, 0, sizeof(*t));
  const Dev& d = p->d;
  t->sort_kind = d.t_sort_kind;
  t->sort_tile = d.t_sort_tile ? d.t_sort_tile : 2048u;
  t->sort_levels = d.t_sort_levels ? d.t_sort_levels : 4u;
  t->narrow_blocks = d.t_narrow_blocks ? d.t_narrow_blocks : 1024u;
  t->prepare_blocks = d.t_prepare_blocks ? d.t_prepare_blocks : 1024u;
  t->solve_blocks = d.t_solve_blocks ? d.t_solve_blocks : 4096u;
  t->scan_blocks = d.t_scan_blocks ? d.t_scan_blocks : 4096u;
  t->traverse_kind = d.t_traverse_kind;
  t->solve_kind = d.t_solve_kind;
  return MH_OK;
}

void mh_physics_destroy(mh_physics* p) {
  if (!p) return;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  for (auto& e : p->events)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : p->solve_events)
    if (e) (void)hipEventDestroy(e);
  for (void* a : p->allocations) (void)hipFree(a);
  if (p->kclock) (void)hipFree(p->kclock);
  if (p->host_diag) (void)hipHostFree(p->host_diag);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

uint32_t mh_physics_body_count(const mh_physics* p) { return p ? p->d.body_count : 0; }

int mh_physics_has_gpu_timestamps(const mh_physics* p) { return p ? 1 : 0; }

int mh_physics_apply_impulse(mh_physics* p, uint32_t body, const float world_point[3], const float impulse[3],
                             mh_error_detail* error) {
  if (error) std::memset(error, 0, sizeof(*error));
  if (!p || !world_point || !impulse) return fail(error, MH_ERR_INVALID_ARGUMENT, "null argument");
  if (body >= p->d.body_count) {
    if (error) { error->body_index = body; error->body_count = p->d.body_count; }
    return fail(error, MH_ERR_BODY_INDEX_OUT_OF_RANGE, "body index outside the uploaded body count");
  }
  if (!finite3(world_point) || !finite3(impulse))
    return fail(error, MH_ERR_NON_FINITE, "external impulse point and vector must be finite");
  launch_external_impulse(p->stream, p->d, make_float4(world_point[0], world_point[1], world_point[2], 0.0f),
                          make_float4(impulse[0], impulse[1], impulse[2], 0.0f), body);
  HIP_TRY(hipGetLastError());
  return MH_OK;
}

int mh_physics_dispatch_tick(mh_physics* p, uint64_t tick, uint8_t* slot) {
  if (!p) return MH_ERR_INVALID_ARGUMENT;
  int rc = begin_tick(p, tick);
  if (rc != MH_OK) return rc;
  if (p->graph_enabled) {
    rc = ensure_graph(p);
    if (rc != MH_OK) return rc;
    HIP_TRY_NOERR(hipGraphLaunch(p->graph_exec, p->stream));
  } else {
    encode_tick(p, p->stream, nullptr);
    HIP_TRY_NOERR(hipGetLastError());
  }
  if (slot) *slot = (uint8_t)(tick % 3u);
  return MH_OK;
}

int mh_physics_dispatch_ticks(mh_physics* p, uint64_t first_tick, uint32_t count) {
  for (uint32_t i = 0; i < count; ++i) {
    int rc = mh_physics_dispatch_tick(p, first_tick + i, nullptr);
    if (rc != MH_OK) return rc;
  }
  return MH_OK;
}

int mh_physics_profile_tick(mh_physics* p, uint64_t tick, mh_phase_timings* out) {
  if (!p) return MH_ERR_INVALID_ARGUMENT;
  const uint32_t passes = p->d.solver_iterations ? p->d.solver_iterations : 1u;
  const size_t sgrid = solve_grid(p->d), bgrid = apply_grid(p->d);
  const size_t words = 2 * (size_t)passes * (sgrid + bgrid);
  if (p->collisions && p->kclock_words < words) {
    if (p->kclock) (void)hipFree(p->kclock);
    p->kclock = nullptr;
    p->kclock_words = 0;
    HIP_TRY_NOERR(hipMalloc(&p->kclock, words * sizeof(uint64_t)));
    p->kclock_words = words;
  }
  while (p->solve_events.size() < 2 * (size_t)passes) {
    hipEvent_t e = nullptr;
    HIP_TRY_NOERR(hipEventCreate(&e));
    p->solve_events.push_back(e);
  }
  int rc = begin_tick(p, tick);
  if (rc != MH_OK) return rc;
  uint64_t* clock = p->collisions ? p->kclock : nullptr;
  if (clock) HIP_TRY_NOERR(hipMemsetAsync(clock, 0, words * sizeof(uint64_t), p->stream));
  encode_tick(p, p->stream, p->events, 0, clock, clock ? p->solve_events.data() : nullptr);
  HIP_TRY_NOERR(hipGetLastError());
  HIP_TRY_NOERR(hipEventSynchronize(p->events[kMarkEnd]));
  float ms[kEvents - 1];
  for (int i = 0; i + 1 < kEvents; ++i) HIP_TRY_NOERR(hipEventElapsedTime(&ms[i], p->events[i], p->events[i + 1]));
  mh_phase_timings t{};
  t.integrate_ms = ms[kMarkStart];
  t.mechanism_ms = ms[kMarkIntegrate];
  t.morton_ms = ms[kMarkMechanism];
  t.sort_ms = ms[kMarkMorton];
  t.lbvh_ms = ms[kMarkSort];
  t.traverse_ms = ms[kMarkLbvh];
  t.narrowphase_ms = ms[kMarkTraverse];
  t.prepare_ms = ms[kMarkNarrow];
  t.solve_ms = ms[kMarkPrepare];
  t.persist_ms = 0.0;  // persist is fused into the last solver pass
  t.validation_ms = ms[kMarkSolve];
  t.snapshot_ms = 0.0;  // fused into integration, or into validation for articulated ticks
  float total = 0.0f;
  HIP_TRY_NOERR(hipEventElapsedTime(&total, p->events[kMarkStart], p->events[kMarkEnd]));
  t.total_ms = total;
  if (clock) {
    // Span of each pass kernel on the device wall clock: earliest workgroup start
    // to latest workgroup end (workgroups that returned early stamp no end).
    std::vector<uint64_t> h(words);
    HIP_TRY_NOERR(hipMemcpy(h.data(), clock, words * sizeof(uint64_t), hipMemcpyDeviceToHost));
    int khz = 0;
    HIP_TRY_NOERR(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, p->device));
    auto mean_span = [&](size_t base, size_t grid) {
      double sum = 0.0;
      uint32_t n = 0;
      for (uint32_t i = 0; i < passes; ++i) {
        uint64_t lo = ~0ull, hi = 0;
        for (size_t b = 0; b < grid; ++b) {
          const uint64_t s0 = h[base + 2 * (i * grid + b)], s1 = h[base + 2 * (i * grid + b) + 1];
          if (s1 == 0) continue;
          lo = s0 < lo ? s0 : lo;
          hi = s1 > hi ? s1 : hi;
        }
        if (hi > lo && lo != ~0ull) {
          sum += (double)(hi - lo);
          ++n;
        }
      }
      return n && khz > 0 ? sum / n / (double)khz : 0.0;
    };
    t.solve_kernel_ms = mean_span(0, sgrid);
    t.apply_kernel_ms = mean_span(2 * (size_t)passes * sgrid, bgrid);
    double event_sum = 0.0;
    for (uint32_t i = 0; i < passes; ++i) {
      float e = 0.0f;
      HIP_TRY_NOERR(hipEventElapsedTime(&e, p->solve_events[2 * i], p->solve_events[2 * i + 1]));
      event_sum += e;
    }
    t.solve_kernel_event_ms = event_sum / passes;
  }
  p->timing.t = t;
  p->timing.valid = true;
  if (out) *out = t;
  return MH_OK;
}

int mh_physics_debug_tick(mh_physics* p, uint64_t tick, int stop) {
  if (!p || stop < 0 || stop > 3) return MH_ERR_INVALID_ARGUMENT;
  int rc = begin_tick(p, tick);
  if (rc != MH_OK) return rc;
  encode_tick(p, p->stream, nullptr, stop);
  HIP_TRY_NOERR(hipGetLastError());
  // The device tick did not advance: the next tick must set it again.
  if (stop != 0) p->tick_known = false;
  HIP_TRY_NOERR(hipStreamSynchronize(p->stream));
  return MH_OK;
}

int mh_physics_synchronize(mh_physics* p) {
  if (!p) return MH_ERR_INVALID_ARGUMENT;
  HIP_TRY_NOERR(hipStreamSynchronize(p->stream));
  return MH_OK;
}

int mh_physics_read_last_tick(mh_physics* p, mh_tick_readback* out) {
  if (!p || !out) return MH_ERR_INVALID_ARGUMENT;
  std::memset(out, 0, sizeof(*out));
  HIP_TRY_NOERR(hipStreamSynchronize(p->stream));
  const uint64_t tick = p->has_last_tick ? p->last_tick : 0;
  HIP_TRY_NOERR(hipMemcpyAsync(p->host_diag, p->d.diag[tick & 1], kDiagWords * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, p->stream));
  HIP_TRY_NOERR(hipStreamSynchronize(p->stream));
  out->error_flags = p->host_diag[kDiagFlags];
  out->pair_count = p->host_diag[kDiagPairs];
  out->contact_count = p->host_diag[kDiagContacts];
  out->active_contact_count = p->host_diag[kDiagActive];
  out->anchor_residual_meters = (float)p->host_diag[kDiagAnchor] * 1.0e-6f;
  out->axis_residual_degrees = (float)p->host_diag[kDiagAxis] * 1.0e-6f;
  out->tick_index = tick;
  out->snapshot_slot = (uint32_t)(tick % 3u);
  if (p->timing.valid) {
    const mh_phase_timings& t = p->timing.t;
    out->has_timings = 1;
    out->integration_ms = t.integrate_ms;
    out->mechanism_ms = t.mechanism_ms;
    out->broadphase_ms = t.morton_ms + t.sort_ms + t.lbvh_ms + t.traverse_ms;
    out->narrowphase_ms = t.narrowphase_ms;
    out->contact_solver_ms = t.prepare_ms + t.solve_ms + t.persist_ms;
    out->bearings_ms = t.validation_ms;
    out->snapshot_ms = t.snapshot_ms;
    out->gpu_tick_ms = out->integration_ms + out->mechanism_ms + out->broadphase_ms + out->narrowphase_ms +
                       out->contact_solver_ms + out->bearings_ms + out->snapshot_ms;
  } else {
    out->gpu_tick_ms = -1.0;
  }
  return MH_OK;
}

int mh_physics_read_snapshot(mh_physics* p, uint8_t slot, mh_transform* out, size_t capacity) {
  if (!p) return MH_ERR_INVALID_ARGUMENT;
  if (slot >= MH_SNAPSHOT_RING_SIZE) return MH_ERR_INVALID_SNAPSHOT_SLOT;
  const uint32_t n = p->d.body_count;
  if (!out || capacity < n) return MH_ERR_INVALID_ARGUMENT;
  std::vector<float4> hp(n), hr(n);
  HIP_TRY_NOERR(hipStreamSynchronize(p->stream));
  if (n) {
    HIP_TRY_NOERR(hipMemcpy(hp.data(), p->d.snap_pos[slot], n * sizeof(float4), hipMemcpyDeviceToHost));
    HIP_TRY_NOERR(hipMemcpy(hr.data(), p->d.snap_rot[slot], n * sizeof(float4), hipMemcpyDeviceToHost));
  }
  for (uint32_t i = 0; i < n; ++i) {
    std::memcpy(out[i].position, &hp[i], 16);
    std::memcpy(out[i].rotation, &hr[i], 16);
  }
  return MH_OK;
}

int mh_physics_snapshot_device_ptrs(mh_physics* p, uint8_t slot, const void** positions, const void** rotations) {
  if (!p) return MH_ERR_INVALID_ARGUMENT;
  if (slot >= MH_SNAPSHOT_RING_SIZE) return MH_ERR_INVALID_SNAPSHOT_SLOT;
  if (positions) *positions = p->d.snap_pos[slot];
  if (rotations) *rotations = p->d.snap_rot[slot];
  return MH_OK;
}

int mh_physics_read_bodies(mh_physics* p, mh_transform* poses, mh_velocity* velocities, size_t capacity) {
  if (!p) return MH_ERR_INVALID_ARGUMENT;
  const uint32_t n = p->d.body_count;
  if (capacity < n) return MH_ERR_INVALID_ARGUMENT;
  HIP_TRY_NOERR(hipStreamSynchronize(p->stream));
  std::vector<float4> a(n), b(n);
  if (poses && n) {
    HIP_TRY_NOERR(hipMemcpy(a.data(), p->d.pos, n * sizeof(float4), hipMemcpyDeviceToHost));
    HIP_TRY_NOERR(hipMemcpy(b.data(), p->d.rot, n * sizeof(float4), hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; ++i) {
      std::memcpy(poses[i].position, &a[i], 16);
      std::memcpy(poses[i].rotation, &b[i], 16);
    }
  }
  if (velocities && n) {
    HIP_TRY_NOERR(hipMemcpy(a.data(), p->d.lin, n * sizeof(float4), hipMemcpyDeviceToHost));
    HIP_TRY_NOERR(hipMemcpy(b.data(), p->d.ang, n * sizeof(float4), hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; ++i) {
      std::memcpy(velocities[i].linear, &a[i], 16);
      std::memcpy(velocities[i].angular, &b[i], 16);
    }
  }
  return MH_OK;
}

int mh_physics_write_bodies(mh_physics* p, const mh_transform* poses, const mh_velocity* velocities, size_t count) {
  if (!p) return MH_ERR_INVALID_ARGUMENT;
  const uint32_t n = p->d.body_count;
  if (count != n) return MH_ERR_INVALID_ARGUMENT;
  HIP_TRY_NOERR(hipStreamSynchronize(p->stream));
  std::vector<float4> a(n), b(n);
  if (poses && n) {
    for (uint32_t i = 0; i < n; ++i) {
      std::memcpy(&a[i], poses[i].position, 16);
      std::memcpy(&b[i], poses[i].rotation, 16);
    }
    HIP_TRY_NOERR(hipMemcpy(p->d.pos, a.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    HIP_TRY_NOERR(hipMemcpy(p->d.rot, b.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  }
  if (velocities && n) {
    for (uint32_t i = 0; i < n; ++i) {
      std::memcpy(&a[i], velocities[i].linear, 16);
      std::memcpy(&b[i], velocities[i].angular, 16);
    }
    HIP_TRY_NOERR(hipMemcpy(p->d.lin, a.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    HIP_TRY_NOERR(hipMemcpy(p->d.ang, b.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  }
  return MH_OK;
}

int mh_physics_read_buffer(mh_physics* p, int buffer, void* dst, size_t bytes, size_t* written) {
  if (!p || !dst) return MH_ERR_INVALID_ARGUMENT;
  HIP_TRY_NOERR(hipStreamSynchronize(p->stream));
  const Dev& d = p->d;
  const uint64_t tick = p->has_last_tick ? p->last_tick : 0;
  const void* src = nullptr;
  size_t size = 0;
  uint32_t diag_host[kDiagWords] = {};
  HIP_TRY_NOERR(hipMemcpy(diag_host, d.diag[tick & 1], sizeof(diag_host), hipMemcpyDeviceToHost));
  auto capped = [&](uint32_t v) { return (size_t)(v < d.pair_capacity ? v : d.pair_capacity); };
  const uint32_t c = d.collider_count;
  std::vector<float4> node_aabbs;
  switch (buffer) {
    case MH_BUFFER_MORTON_KEYS: src = d.keys; size = (size_t)d.sort_count * 8; break;
    case MH_BUFFER_COLLIDER_AABBS: src = d.aabbs; size = (size_t)c * 32; break;
    case MH_BUFFER_NODE_CHILDREN: src = d.children; size = c > 1 ? (size_t)(c - 1) * 8 : 0; break;
    case MH_BUFFER_NODE_PARENTS: src = d.parents; size = c > 0 ? (size_t)(2 * c - 1) * 4 : 0; break;
    case MH_BUFFER_PAIRS: src = d.pairs; size = capped(diag_host[kDiagPairs]) * 8; break;
    case MH_BUFFER_CONTACTS: src = d.contacts; size = capped(diag_host[kDiagContacts]) * 64; break;
    case MH_BUFFER_DIAGNOSTICS: src = d.diag[tick & 1]; size = 32; break;
    case MH_BUFFER_WORLD_MASSES: src = d.world_masses; size = (size_t)d.body_count * 48; break;
    case MH_BUFFER_BODY_SLOTS: src = d.slot_count; size = (size_t)d.body_count * 4; break;
    case MH_BUFFER_MANIFOLDS: src = d.mrows[tick & 1]; size = capped(diag_host[kDiagContacts]) * 48; break;
    case MH_BUFFER_NODE_AABBS: {
      // Reference layout (lbvh.wgsl:19, 287): internal nodes then leaves, min/max with w = 0.
      size_t nodes = c > 0 ? 2 * (size_t)c - 1 : 0;
      size = nodes * 32;
      if (bytes < size) return MH_ERR_INVALID_ARGUMENT;
      std::vector<float4> in(2 * (size_t)(c > 1 ? c - 1 : 0)), lv(2 * (size_t)c);
      if (!in.empty())
        HIP_TRY_NOERR(hipMemcpy(in.data(), d.inodes, in.size() * sizeof(float4), hipMemcpyDeviceToHost));
      if (!lv.empty())
        HIP_TRY_NOERR(hipMemcpy(lv.data(), d.leaves, lv.size() * sizeof(float4), hipMemcpyDeviceToHost));
      float4* o = static_cast<float4*>(dst);
      size_t k = 0;
      for (size_t i = 0; i < in.size(); ++i) { o[k] = in[i]; o[k].w = 0.0f; ++k; }
      for (size_t i = 0; i < lv.size(); ++i) { o[k] = lv[i]; o[k].w = 0.0f; ++k; }
      if (written) *written = size;
      return MH_OK;
    }
    default: return MH_ERR_INVALID_ARGUMENT;
  }
  if (bytes < size) return MH_ERR_INVALID_ARGUMENT;
  if (size) HIP_TRY_NOERR(hipMemcpy(dst, src, size, hipMemcpyDeviceToHost));
  if (written) *written = size;
  return MH_OK;
}

int mh_physics_write_drives(mh_physics* p, const mh_mechanism_drive* drives, size_t count, mh_error_detail* error) {
  if (error) std::memset(error, 0, sizeof(*error));
  if (!p) return fail(error, MH_ERR_INVALID_ARGUMENT, "null handle");
  // write_mechanism_drives (device.rs:1622-1644): the row count must match the coordinates.
  const size_t required = p->d.coordinate_count;
  if (count != required) {
    if (error) { error->provided = count; error->required = required; }
    return fail(error, MH_ERR_DRIVE_STATE_COUNT, "drive state row count does not match the scene");
  }
  if (count == 0) return MH_OK;
  if (!drives) return fail(error, MH_ERR_INVALID_ARGUMENT, "null drive rows");
  HIP_TRY(hipMemcpyAsync(const_cast<float4*>(p->d.drives), drives, count * sizeof(mh_mechanism_drive),
                         hipMemcpyHostToDevice, p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  return MH_OK;
}

int mh_physics_init_coordinates(mh_physics* p, const mh_mechanism_coordinate* coordinates, size_t count,
                                mh_error_detail* error) {
  if (error) std::memset(error, 0, sizeof(*error));
  if (!p) return fail(error, MH_ERR_INVALID_ARGUMENT, "null handle");
  // initialize_mechanism_coordinates (device.rs:1646-1662).
  const size_t required = p->d.coordinate_count;
  if (count != required) {
    if (error) { error->provided = count; error->required = required; }
    return fail(error, MH_ERR_COORDINATE_STATE_COUNT, "coordinate state row count does not match the scene");
  }
  if (count == 0) return MH_OK;
  if (!coordinates) return fail(error, MH_ERR_INVALID_ARGUMENT, "null coordinate rows");
  HIP_TRY(hipMemcpyAsync(p->d.coords, coordinates, count * sizeof(mh_mechanism_coordinate), hipMemcpyHostToDevice,
                         p->stream));
  HIP_TRY(hipStreamSynchronize(p->stream));
  return MH_OK;
}

int mh_physics_read_coordinates(mh_physics* p, mh_mechanism_coordinate* out, size_t capacity) {
  if (!p || (!out && p->d.coordinate_count > 0)) return MH_ERR_INVALID_ARGUMENT;
  if (capacity < p->d.coordinate_count) return MH_ERR_INVALID_ARGUMENT;
  HIP_TRY_NOERR(hipStreamSynchronize(p->stream));
  if (p->d.coordinate_count)
    HIP_TRY_NOERR(hipMemcpy(out, p->d.coords, p->d.coordinate_count * sizeof(mh_mechanism_coordinate),
                            hipMemcpyDeviceToHost));
  return MH_OK;
}

int mh_physics_set_mlp(mh_physics* p, const mh_mlp* mlp) {
  if (!p) return MH_ERR_INVALID_ARGUMENT;
  HIP_TRY_NOERR(hipStreamSynchronize(p->stream));
  p->mlp = mlp;
  p->d.mlp = mlp_device_net(mlp);
  // Kernel arguments are captured by value: capture again with the new pointer.
  if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->graph_exec = nullptr;
  p->graph = nullptr;
  return MH_OK;
}

int mh_physics_read_training_rows(mh_physics* p, float* features, float* targets, size_t capacity, size_t* rows) {
  if (!p || !features || !targets) return MH_ERR_INVALID_ARGUMENT;
  if (p->d.learn_mode != kLearnLog) return MH_ERR_INVALID_ARGUMENT;
  HIP_TRY_NOERR(hipStreamSynchronize(p->stream));
  const uint64_t tick = p->has_last_tick ? p->last_tick : 0;
  uint32_t diag_host[kDiagWords] = {};
  HIP_TRY_NOERR(hipMemcpy(diag_host, p->d.diag[tick & 1], sizeof(diag_host), hipMemcpyDeviceToHost));
  const size_t n = diag_host[kDiagActive] < p->d.pair_capacity ? diag_host[kDiagActive] : p->d.pair_capacity;
  if (rows) *rows = n;
  if (capacity < n) return MH_ERR_INVALID_ARGUMENT;
  if (n == 0) return MH_OK;
  HIP_TRY_NOERR(hipMemcpy(features, p->d.feat, n * kFeatures * sizeof(float), hipMemcpyDeviceToHost));
  std::vector<float4> t(n);
  HIP_TRY_NOERR(hipMemcpy(t.data(), p->d.finit, n * sizeof(float4), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; ++i) {
    targets[3 * i] = t[i].x;
    targets[3 * i + 1] = t[i].y;
    targets[3 * i + 2] = t[i].z;
  }
  return MH_OK;
}

int mh_physics_read_learned_stats(mh_physics* p, mh_learned_stats* out) {
  if (!p || !out) return MH_ERR_INVALID_ARGUMENT;
  std::memset(out, 0, sizeof(*out));
  if (p->d.learn_mode != kLearnInfer) return MH_ERR_INVALID_ARGUMENT;
  HIP_TRY_NOERR(hipStreamSynchronize(p->stream));
  uint32_t st[4] = {};
  HIP_TRY_NOERR(hipMemcpy(st, p->d.learned_state, sizeof(st), hipMemcpyDeviceToHost));
  out->fallback_this_tick = st[0];
  out->fallback_ticks = st[1];
  out->nonfinite_this_tick = st[2];
  out->nonfinite_total = st[3];
  return MH_OK;
}

int mh_physics_read_residual(mh_physics* p, mh_contact_residual* out) {
  if (!p || !out) return MH_ERR_INVALID_ARGUMENT;
  std::memset(out, 0, sizeof(*out));
  if (!p->d.residual_on) return MH_ERR_INVALID_ARGUMENT;
  HIP_TRY_NOERR(hipStreamSynchronize(p->stream));
  if (learned_fallback_on(p->d)) {
    uint32_t gate = 0;
    HIP_TRY_NOERR(hipMemcpy(&gate, p->d.learned_state, sizeof(gate), hipMemcpyDeviceToHost));
    const uint32_t passes = p->d.solver_iterations ? p->d.solver_iterations : 1u;
    out->fallback_passes = gate ? (p->d.learned_extra ? p->d.learned_extra : passes) : 0u;
  }
  std::vector<double> part(4 * (size_t)kResidBlocks);
  HIP_TRY_NOERR(hipMemcpy(part.data(), p->d.resid, part.size() * sizeof(double), hipMemcpyDeviceToHost));
  double r2 = 0.0, n = 0.0, pen = 0.0, mx = 0.0;
  for (uint32_t b = 0; b < kResidBlocks; ++b) {
    r2 += part[4 * b];
    n += part[4 * b + 1];
    pen += part[4 * b + 2];
    mx = part[4 * b + 3] > mx ? part[4 * b + 3] : mx;
  }
  out->active_contacts = (uint32_t)n;
  out->rms = n > 0.0 ? std::sqrt(r2 / n) : 0.0;
  out->mean_penetration = n > 0.0 ? pen / n : 0.0;
  out->max_penetration = mx;
  return MH_OK;
}

}  // extern "C"
