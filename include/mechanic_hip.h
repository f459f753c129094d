/*
 * mechanic_hip.h — C ABI of the MI355X (gfx950) rigid-body step backend.
 *
 * This library stands in for `mechanic_gpu::GpuPhysics`
 * (crates/mechanic-gpu/src/device.rs:199-1663) behind a flat C interface:
 * plain pointers and sizes, no HIP, torch or wgpu types in any signature.
 *
 * Row layouts mirror the reference's fixed-size Pod rows byte for byte
 * (crates/mechanic-gpu/src/abi.rs:16-296, size asserts abi.rs:298-315), and the
 * error-flag bits and capacities are the reference's (abi.rs:3-13,
 * crates/mechanic-gpu/src/lib.rs:29-51).
 *
 * Threading: a handle is not re-entrant; separate handles are independent.
 * All work of one handle is ordered on one HIP stream owned by the handle.
 */
#ifndef MECHANIC_HIP_H
#define MECHANIC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MH_API __attribute__((visibility("default")))

/* ---- flags and capacities (abi.rs:3-13, lib.rs:29-51) ------------------- */

#define MH_PAIR_OVERFLOW_FLAG (1u << 0)              /* abi.rs:4 */
#define MH_INVALID_NUMERIC_FLAG (1u << 1)            /* abi.rs:7 */
#define MH_CONSTRAINT_NON_CONVERGENCE_FLAG (1u << 2) /* abi.rs:10 */
#define MH_MANIFOLD_OVERFLOW_FLAG (1u << 3)          /* abi.rs:13 */

#define MH_MAX_BODIES 131072u         /* lib.rs:36 */
#define MH_MAX_BEARINGS 262144u       /* lib.rs:39 */
#define MH_MAX_COLLIDERS 131072u      /* lib.rs:42 */
#define MH_MAX_CONTACT_PAIRS 2097152u /* lib.rs:45 */
#define MH_SNAPSHOT_RING_SIZE 3u      /* lib.rs:51 */
#define MH_MAX_BATCH_BODIES 1048576u  /* bodies (and colliders) of a scene batch */
#define MH_MAX_BATCH_SCENES 4096u
#define MH_INVALID_INDEX 0xffffffffu

#define MH_DRIVE_MODE_PASSIVE 0u /* abi.rs:256 */
#define MH_DRIVE_MODE_SPEED 1u   /* abi.rs:259 */
#define MH_DRIVE_MODE_ANGLE 2u   /* abi.rs:262 */

/* ---- status codes ---------------------------------------------------------
 * One code per variant of GpuPhysicsError (device.rs:130-173),
 * GpuImpulseError (device.rs:175-189) and GpuReadbackError (device.rs:113-128),
 * plus codes for problems a Rust caller would meet as panics or wgpu errors. */
enum {
  MH_OK = 0,
  MH_ERR_BODY_CAPACITY = 1,             /* GpuPhysicsError::BodyCapacity */
  MH_ERR_BEARING_CAPACITY = 2,          /* GpuPhysicsError::BearingCapacity */
  MH_ERR_COLLIDER_CAPACITY = 3,         /* GpuPhysicsError::ColliderCapacity */
  MH_ERR_DRIVE_STATE_COUNT = 4,         /* GpuPhysicsError::DriveStateCount */
  MH_ERR_COORDINATE_STATE_COUNT = 5,    /* GpuPhysicsError::CoordinateStateCount */
  MH_ERR_BODY_INDEX_OUT_OF_RANGE = 6,   /* GpuImpulseError::BodyIndexOutOfRange */
  MH_ERR_NON_FINITE = 7,                /* GpuImpulseError::NonFinite */
  MH_ERR_INVALID_SNAPSHOT_SLOT = 8,     /* GpuReadbackError::InvalidSnapshotSlot */
  MH_ERR_DEVICE = 9,                    /* GpuReadbackError::DevicePoll (a HIP error) */
  MH_ERR_INVALID_ARGUMENT = 10,         /* null pointer, short buffer, bad config */
  MH_ERR_NO_DEVICE = 11,                /* no HIP device visible */
  MH_ERR_UNSUPPORTED = 12               /* a feature this build does not run */
};

/* ---- rows (abi.rs) ------------------------------------------------------- */

typedef struct { /* GpuTickConfig, abi.rs:18-51 (64 B) */
  uint32_t body_count, tick_index, snapshot_slot, collider_count;
  float delta_seconds, gravity_y, linear_damping, angular_damping;
  uint32_t bearing_count, suppression_count, pair_capacity, flags;
  uint32_t hash_capacity, solver_iterations, reserved_a, reserved_b;
} mh_tick_config;

typedef struct { /* GpuDiagnostics, abi.rs:56-73 (32 B) */
  uint32_t error_flags, pair_count, contact_count;
  uint32_t max_anchor_micrometers, max_axis_microdegrees;
  uint32_t active_contact_count, reserved_b, reserved_c;
} mh_diagnostics;

typedef struct { float position[4]; float rotation[4]; } mh_transform; /* abi.rs:78-83 */
typedef struct { float linear[4]; float angular[4]; } mh_velocity;    /* abi.rs:88-93 */

typedef struct { /* GpuMass, abi.rs:98-107 */
  float inverse_mass[4], inverse_inertia_x[4], inverse_inertia_y[4], inverse_inertia_z[4];
} mh_mass;

typedef struct { /* GpuSpatialInertia, abi.rs:112-121 */
  float mass[4], inertia_x[4], inertia_y[4], inertia_z[4];
} mh_spatial_inertia;

typedef struct { /* GpuCollider, abi.rs:126-135 */
  float local_center[4];   /* w: full-cylinder segment-centre radius */
  float local_rotation[4]; /* xyzw */
  float half_extents[4];   /* w: full-cylinder outer radius */
  uint32_t metadata[4];    /* compound, source slot, source generation, ground role */
} mh_collider;

typedef struct { /* GpuBearing, abi.rs:140-151 */
  float local_anchor_a[4], local_anchor_b[4], local_axis_a[4], local_axis_b[4];
  uint32_t metadata[4]; /* compound a, compound b, coordinate or MH_INVALID_INDEX, flags */
} mh_bearing;

typedef struct { uint32_t collider_a, collider_b; } mh_pair; /* abi.rs:156-161 */

typedef struct { /* GpuContact, abi.rs:166-175 */
  uint32_t metadata[4];         /* body a, body b, manifold slot, point count */
  float normal_penetration[4];  /* normal a->b, depth */
  float arm_a_impulse[4];       /* lever arm from body a, accumulated normal impulse */
  float arm_b[4];               /* lever arm from body b, w: near-face axes / -1 analytic */
} mh_contact;

typedef struct { /* GpuPersistentManifold, abi.rs:180-187 */
  uint32_t pair_tick[4];       /* collider low, collider high, last tick, point count */
  float normal_penetration[4];
  float point_impulse[4];      /* arm from body a, accumulated normal impulse */
} mh_persistent_manifold;

typedef struct { /* GpuMechanismBody, abi.rs:191-201 */
  uint32_t metadata[4];  /* parent body, tree bearing row, direction, root flag */
  uint32_t traversal[4]; /* component, depth, preorder, postorder */
  float bind_relative_position[4];
  float bind_relative_rotation[4];
} mh_mechanism_body;

typedef struct { float angle, angular_velocity; } mh_mechanism_coordinate; /* abi.rs:206-211 */

typedef struct { /* GpuMechanismDrive, abi.rs:220-238 */
  uint32_t mode;
  float max_acceleration, max_speed, target_speed, target_angle, min_angle, max_angle, padding;
} mh_mechanism_drive;

typedef struct { float position[4]; float rotation[4]; uint32_t metadata[4]; } mh_link_state; /* abi.rs:280-288 */
typedef struct { uint32_t metadata[4]; } mh_contraction_node; /* abi.rs:292-296 */

/* ---- scene upload (what GpuPhysics::new_with_config lowers, device.rs:424-523) */

typedef struct {
  uint32_t body_count;
  uint32_t collider_count;
  uint32_t bearing_count;     /* bearing rows (tree and loop-closure) */
  uint32_t suppression_count;
  const mh_transform* poses;          /* root pose per body (device.rs:429-441) */
  const float* inverse_masses;        /* per body (device.rs:443-447) */
  const mh_mass* masses;              /* body-frame inverse mass/inertia (device.rs:454-466) */
  const mh_spatial_inertia* inertias; /* optional, may be NULL (device.rs:467-479) */
  const mh_collider* colliders;       /* device.rs:480-499 */
  const mh_bearing* bearings;         /* device.rs:500-515 */
  const mh_pair* suppressed_pairs;    /* sorted (low, high) body pairs, device.rs:516-523 */
  const uint32_t* body_components;    /* NULL: every body its own component (device.rs:448-453) */
  const uint32_t* root_flags;         /* NULL: every body a root (device.rs:1699-1704) */
  /* Articulated rows (device.rs:1684-1808); all NULL / 0 when bearing_count == 0. */
  const mh_mechanism_body* mechanism_bodies; /* per body: parent, tree bearing, direction, root flag */
  const uint32_t* preorder;                  /* bodies with every parent before its children */
  const mh_mechanism_drive* drives;          /* per coordinate; NULL: every coordinate passive */
  uint32_t coordinate_count;                 /* tree bearings (one coordinate each) */
  uint32_t reserved;
} mh_creation;

/* GpuPhysicsConfig (device.rs:22-41) plus the constants the reference hard-codes,
 * exposed so the BASELINE scenes (friction 0.6) can run. Zero-initialised fields
 * take the reference value; call mh_config_default() for an explicit default. */
typedef struct {
  uint32_t collisions_enabled;        /* device.rs:26, default 1 */
  uint32_t mechanism_self_collisions; /* device.rs:28, default 1 */
  uint32_t solver_iterations;         /* device.rs:30, default 8 */
  float friction;                     /* CONTACT_FRICTION collision.wgsl:66, default 0.05 */
  uint32_t pair_capacity;             /* MAX_CONTACT_PAIRS lib.rs:45 (power of two) */
  uint32_t warm_start_mode;           /* MH_WARM_START_*; default MH_WARM_START_CACHE */
  uint32_t use_graph;                 /* 1 = replay each tick as a captured hipGraph */
  uint32_t residual;                  /* 1 = measure the contact residual after the solve */
  /* Learned warm start fallback: after the learned passes the RMS contact
   * residual is measured on the device; a tick above this limit (N s) runs
   * learned_fallback_passes more passes. 0 disables the check. */
  float learned_residual_limit;
  uint32_t learned_fallback_passes;   /* 0: as many as the learned run's passes */
  /* Loop-closure Newton solve (device.rs:976-1020, closure.wgsl:333-427): Newton
   * steps per tick (0: the reference's 12), PCG iterations per step (0: the
   * reference's 8), and a PCG tolerance on the preconditioned residual
   * sqrt(r.z / r0.z) (0: run every iteration, as the reference does). */
  uint32_t closure_newton_steps;
  uint32_t closure_pcg_iterations;
  float closure_pcg_tolerance;
  /* 1: one PCG per closure component (closures linked by shared coordinates),
   * each with its own step lengths, tolerance stop and Newton convergence
   * test, one small workgroup per component running every Newton step and its
   * forward kinematics in one launch. The system is block diagonal over
   * components, so both organisations solve the same equations. Components may
   * hold at most 256 path bodies, 256 coordinates and 128 closures
   * (MH_ERR_UNSUPPORTED otherwise). 0: one global PCG, as the reference. */
  uint32_t closure_components;
  /* 1 (needs closure_components): after the Newton steps, the coordinate speeds
   * are projected onto the loops' velocity constraint J v = 0 (minimum-norm
   * correction, the same per-component PCG and tolerance) before the tree
   * reconstructs body velocities from them. The reference reconstructs from the
   * unprojected speeds (articulated.wgsl:423-456), which leaves loop-violating
   * velocities for the bearing projection to absorb. 0: as the reference. */
  uint32_t closure_velocity_projection;
} mh_config;

/* Warm start of the contact solver's first pass. */
enum {
  MH_WARM_START_CACHE = 0,   /* persistent-manifold cache, halved impulse (collision.wgsl:699-733, 783) */
  MH_WARM_START_LOG = 1,     /* as CACHE, and records training rows (features, solved impulses) */
  MH_WARM_START_LEARNED = 2  /* initial impulses predicted by the network set with mh_physics_set_mlp */
};

/* GpuTickReadback (device.rs:74-92) with GpuKernelTimings (device.rs:95-111). */
typedef struct {
  double gpu_tick_ms;   /* sum of the phase timings below; -1 when not profiled */
  double integration_ms, mechanism_ms, broadphase_ms, narrowphase_ms;
  double contact_solver_ms, bearings_ms, snapshot_ms;
  uint32_t error_flags, pair_count, contact_count, active_contact_count;
  float anchor_residual_meters, axis_residual_degrees;
  uint64_t tick_index;
  uint32_t snapshot_slot, has_timings;
} mh_tick_readback;

/* Per-kernel-group timings of an event-bracketed (uncaptured) tick. */
typedef struct {
  double integrate_ms, morton_ms, sort_ms, lbvh_ms, traverse_ms;
  double narrowphase_ms, prepare_ms, solve_ms, persist_ms, snapshot_ms, total_ms;
  double mechanism_ms;   /* bearing projection, coordinates, forward kinematics, closures */
  double validation_ms;  /* articulated-state and bearing validation (with the snapshot) */
  double solve_kernel_ms;  /* mean span of one contact-solve pass kernel (k_solve), device clock */
  double apply_kernel_ms;  /* mean span of one velocity-gather kernel (k_apply), device clock */
  double solve_kernel_event_ms;  /* mean time between the start and stop HIP events each k_solve
                                    dispatch carries (hipExtLaunchKernelGGL): the dispatch's own
                                    start and end, what rocprofv3 records */
} mh_phase_timings;

typedef struct {
  int code;
  uint64_t required, capacity, provided;
  uint32_t body_index, body_count;
  char message[160];
} mh_error_detail;

typedef struct mh_physics mh_physics;

MH_API mh_config mh_config_default(void);
MH_API const char* mh_status_string(int code);
MH_API int mh_device_count(int* count);
/* Architecture name of the current device (e.g. "gfx950:sramecc+:xnack-"). */
MH_API int mh_device_arch(char* out, size_t capacity);
/* Converts `count` values on the current device with the rounding of the
   solvers' Q20 velocity deltas (add_velocity_delta, collision.wgsl:764-772;
   articulated.wgsl:104-112): round half to even, then a saturating cast (NaN ->
   0). A self-test of the device conversion; no counterpart in the reference. */
MH_API int mh_device_round_i32(const float* values, int32_t* out, size_t count);

/* GpuPhysics::new_with_config (device.rs:399-811). */
MH_API int mh_physics_create(const mh_creation* creation, const mh_config* config,
                             mh_physics** out, mh_error_detail* error);
MH_API void mh_physics_destroy(mh_physics* physics);
/* GpuPhysics::body_count (device.rs:1582). */
MH_API uint32_t mh_physics_body_count(const mh_physics* physics);
/* GpuPhysics::has_gpu_timestamps (device.rs:1573): always 1 (HIP events). */
MH_API int mh_physics_has_gpu_timestamps(const mh_physics* physics);
/* GpuPhysics::apply_impulse (device.rs:821-856). */
MH_API int mh_physics_apply_impulse(mh_physics* physics, uint32_t body, const float world_point[3],
                                    const float impulse[3], mh_error_detail* error);
/* GpuPhysics::dispatch_tick (device.rs:859-961): enqueues one tick, returns at once. */
MH_API int mh_physics_dispatch_tick(mh_physics* physics, uint64_t tick_index, uint8_t* snapshot_slot);
/* `count` consecutive dispatch_tick calls starting at `first_tick`. */
MH_API int mh_physics_dispatch_ticks(mh_physics* physics, uint64_t first_tick, uint32_t count);
/* Uncaptured tick with HIP events between kernel groups (these include host
 * launch gaps); the solver-pass kernels also stamp the device wall clock at the
 * start and end of every workgroup, giving their own spans. Fills `timings`. */
MH_API int mh_physics_profile_tick(mh_physics* physics, uint64_t tick_index, mh_phase_timings* timings);
/* GpuPhysics::read_last_tick (device.rs:1439-1524): waits for the stream. */
MH_API int mh_physics_read_last_tick(mh_physics* physics, mh_tick_readback* out);
/* GpuPhysics::read_snapshot_transforms (device.rs:1534-1570). */
MH_API int mh_physics_read_snapshot(mh_physics* physics, uint8_t slot, mh_transform* out, size_t capacity);
/* GpuPhysics::snapshot (device.rs:1578-1580): device pointers of one ring slot. */
MH_API int mh_physics_snapshot_device_ptrs(mh_physics* physics, uint8_t slot, const void** positions,
                                           const void** rotations);
/* Body state read/write (the reference's raw buffer getters, device.rs:1582-1620). */
MH_API int mh_physics_read_bodies(mh_physics* physics, mh_transform* poses, mh_velocity* velocities,
                                  size_t capacity);
MH_API int mh_physics_write_bodies(mh_physics* physics, const mh_transform* poses,
                                   const mh_velocity* velocities, size_t count);
/* Raw read of an internal buffer for parity tests (see MH_BUFFER_*). */
MH_API int mh_physics_read_buffer(mh_physics* physics, int buffer, void* dst, size_t bytes, size_t* written);
MH_API int mh_physics_synchronize(mh_physics* physics);
/* Runs one uncaptured tick up to a phase boundary, for phase parity tests:
 * stop 1 = after integration, 2 = after the broadphase, 3 = after contact
 * generation, 0 = whole tick. A stopped tick leaves the handle mid-tick. */
MH_API int mh_physics_debug_tick(mh_physics* physics, uint64_t tick_index, int stop);
/* GpuPhysics::write_mechanism_drives / initialize_mechanism_coordinates
 * (device.rs:1622-1662): replace every coordinate's drive row or state; the row
 * count must equal the scene's coordinate count (MH_ERR_DRIVE_STATE_COUNT /
 * MH_ERR_COORDINATE_STATE_COUNT otherwise). Ordered after the handle's queued ticks. */
MH_API int mh_physics_write_drives(mh_physics* physics, const mh_mechanism_drive* drives, size_t count,
                                   mh_error_detail* error);
MH_API int mh_physics_init_coordinates(mh_physics* physics, const mh_mechanism_coordinate* coordinates,
                                       size_t count, mh_error_detail* error);
/* Current coordinate rows (angle, angular velocity), for tests and tools. */
MH_API int mh_physics_read_coordinates(mh_physics* physics, mh_mechanism_coordinate* out, size_t capacity);

/* ---- scene batches and launch tuning (no counterpart in the reference) ----
 * A batch handle steps many independent scenes -- each with its own ground
 * plane and coordinates -- in one tick (one graph launch). Each scene's results
 * are those of the same scene in a handle of its own. Bodies and colliders are
 * concatenated scene after scene; dispatch, readback and impulses address the
 * concatenated rows. Scenes may not hold bearings. */
MH_API int mh_batch_create(const mh_creation* scenes, uint32_t scene_count, const mh_config* config,
                           mh_physics** out, mh_error_detail* error);
MH_API uint32_t mh_batch_scene_count(const mh_physics* physics);
MH_API int mh_batch_scene_range(const mh_physics* physics, uint32_t scene, uint32_t* body_first,
                                uint32_t* body_count, uint32_t* collider_first, uint32_t* collider_count);

/* Launch shape of a handle's kernels; 0 keeps a field's default. Changing it
 * re-captures the tick graph; results do not depend on it. */
typedef struct {
  uint32_t sort_tile;       /* bitonic: tile width 256..2048, power of two (default 2048) */
  uint32_t sort_levels;     /* bitonic: butterfly levels fused per global launch, 1..4 (default 4) */
  uint32_t narrow_blocks;   /* workgroups of the narrowphase (default 1024) */
  uint32_t prepare_blocks;  /* workgroups of contact preparation (default 1024) */
  uint32_t solve_blocks;    /* workgroups of each solver pass (default 4096) */
  uint32_t scan_blocks;     /* workgroups of the broadphase candidate scan (default 4096) */
  uint32_t sort_kind;       /* 0: radix sort (default), 1: bitonic sort */
  uint32_t traverse_kind;   /* 0: packet walk + separate scan of its leaf ranges (default), 1: one
                               walk per leaf, 2: packet walk scanning its ranges itself */
  uint32_t solve_kind;      /* 0: each solver pass runs per contact, then a per-body gather of
                               the stored deltas (default); 1: per body over its contact sides */
} mh_tuning;

MH_API int mh_physics_set_tuning(mh_physics* physics, const mh_tuning* tuning);
MH_API int mh_physics_get_tuning(const mh_physics* physics, mh_tuning* tuning);

/* ---- learned warm start (no counterpart in the reference) --------------
 * A 16 -> 64 -> 64 -> 3 GELU network (tanh form) predicts, per active contact,
 * the accumulated normal impulse and the friction impulse (two components in
 * the contact's tangent basis) from 16 contact features. Inference inside a
 * tick runs on MFMA in BF16 with FP32 accumulation; training runs on the GPU in
 * FP32 (forward, backward, Adam). Parameter layout, floats: W1[16][64] b1[64]
 * W2[64][64] b2[64] W3[64][3] b3[3] (5443), then input mean[16], 1/std[16],
 * output mean[3], std[3] (5481 in all). */
typedef struct mh_mlp mh_mlp;

#define MH_MLP_FEATURES 16
#define MH_MLP_OUTPUTS 3

MH_API int mh_mlp_create(uint64_t seed, mh_mlp** out);
MH_API void mh_mlp_destroy(mh_mlp* mlp);
MH_API uint32_t mh_mlp_param_count(void);
MH_API uint32_t mh_mlp_net_floats(void);
MH_API int mh_mlp_get_net(mh_mlp* mlp, float* out, size_t count);
MH_API int mh_mlp_set_net(mh_mlp* mlp, const float* net, size_t count);
/* Uploads `rows` training rows (16 features, 3 targets each) and sets the
 * normalisation to the dataset's per-column mean and standard deviation. */
MH_API int mh_mlp_set_dataset(mh_mlp* mlp, const float* features, const float* targets, size_t rows);
/* Mini-batch Adam (beta 0.9 / 0.999) over the dataset, shuffled per epoch by
 * the handle's seeded generator; losses[e] = mean batch loss of epoch e. */
MH_API int mh_mlp_train(mh_mlp* mlp, uint32_t epochs, uint32_t batch, float learning_rate, float* losses);
/* Loss and gradient over the whole dataset (no update), FP32. */
MH_API int mh_mlp_gradient(mh_mlp* mlp, float* gradient, float* loss);
/* MFMA inference of host feature rows; out receives 3 floats per row. */
MH_API int mh_mlp_predict(mh_mlp* mlp, const float* features, size_t rows, float* out);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (online fine-tuning inside a serving or stepping loop): the loss and
 * gradient of mh_mlp_gradient, with the same bits, over `rows` caller-owned
 * DEVICE rows (features: 16 floats per row, targets: 3 floats per row), written
 * to caller-owned DEVICE outputs (gradient: mh_mlp_param_count() floats, loss:
 * 1 float). Allocates nothing: the per-block partial gradients pass through a
 * fixed 2.8 MB window in the library's static device memory, held by every
 * GPU the library runs on until the process ends. rows must be 1..2^28
 * (MH_ERR_INVALID_ARGUMENT otherwise). Runs on the network's stream and
 * returns when the outputs are written. Calls that overlap on one GPU (several
 * threads or networks at once) share that window and race. */
MH_API int mh_mlp_gradient_noalloc(mh_mlp* mlp, const float* features, const float* targets, size_t rows,
                                   float* gradient, float* loss);

/* Attaches a network to a handle (NULL detaches). The network must outlive its
 * use; weights changed by training are seen by the next tick. */
MH_API int mh_physics_set_mlp(mh_physics* physics, const mh_mlp* mlp);
/* MH_WARM_START_LOG handles: the last tick's training rows (features, then the
 * solved normal impulse and friction impulse in the tangent basis). */
MH_API int mh_physics_read_training_rows(mh_physics* physics, float* features, float* targets, size_t capacity,
                                         size_t* rows);

/* Normal complementarity residual of the last tick, per active contact
 * r = |max(0, lambda + (bias - v_n) / d) - lambda| after the last pass. */
typedef struct {
  double rms;                 /* sqrt(mean r^2) (N s) */
  double mean_penetration;    /* (m) */
  double max_penetration;     /* (m) */
  uint32_t active_contacts;
  uint32_t fallback_passes;   /* learned warm start: extra passes this tick ran (0 when within the limit) */
} mh_contact_residual;

/* Handles created with config.residual = 1. */
MH_API int mh_physics_read_residual(mh_physics* physics, mh_contact_residual* out);

/* Learned warm start fallback counts of a MH_WARM_START_LEARNED handle. */
typedef struct {
  uint32_t fallback_this_tick;   /* 1 if the last tick's residual check ran the extra passes */
  uint32_t fallback_ticks;       /* ticks that ran them since creation */
  uint32_t nonfinite_this_tick;  /* contacts whose prediction was not finite (cache warm start instead) */
  uint32_t nonfinite_total;      /* such contacts since creation */
} mh_learned_stats;
MH_API int mh_physics_read_learned_stats(mh_physics* physics, mh_learned_stats* out);

enum {
  MH_BUFFER_MORTON_KEYS = 1,     /* u64 (code << 32 | collider) per sort slot, sorted after a tick */
  MH_BUFFER_COLLIDER_AABBS = 2,  /* float4 min, float4 max per collider */
  MH_BUFFER_NODE_CHILDREN = 3,   /* uint2 per internal node, reference encoding (leaf = C-1+i) */
  MH_BUFFER_NODE_AABBS = 4,      /* float4 min, float4 max per node (internal then leaves) */
  MH_BUFFER_PAIRS = 5,           /* mh_pair per broadphase pair (unordered) */
  MH_BUFFER_CONTACTS = 6,        /* mh_contact per contact as generated (unordered); prepared state: MANIFOLDS */
  MH_BUFFER_DIAGNOSTICS = 7,     /* mh_diagnostics of the last tick */
  MH_BUFFER_WORLD_MASSES = 8,    /* 3 x float4 per body */
  MH_BUFFER_MANIFOLDS = 9,       /* mh_persistent_manifold per contact of the last tick */
  MH_BUFFER_NODE_PARENTS = 10,   /* u32 parent+1 per node, reference encoding */
  MH_BUFFER_BODY_SLOTS = 11      /* u32 per body: delta slots its contact sides took in the last tick (one
                                    per 64-contact group with a side of the body; per side for the
                                    per-body passes). No counterpart in the reference. */
};

/* ---- synthetic scenes (seeded; SplitMix64, u = (x >> 40) * 2^-24) ------- */

typedef struct {
  uint32_t nx, ny, nz;          /* lattice extent along x, y, z */
  float edge;                   /* cuboid edge (m) */
  float spacing;                /* lattice spacing (m) */
  float drop_height;            /* lowest faces above y = 0 (m) */
  float jitter;                 /* uniform +-jitter per axis (m) */
  float density;                /* kg / m^3 */
  uint64_t seed;
  uint32_t yaw;                 /* 1: uniform yaw in [0, 2pi) */
  uint32_t reserved;
} mh_pile_params;

/* Fills body_count = nx*ny*nz rows of a jittered cuboid lattice. Every body is
 * one compound with one cuboid collider (root rotation identity, yaw carried by
 * the collider's local rotation, as the reference's compile step lowers a
 * rotated part: crates/mechanic-core/src/compile.rs:333-334, 889-895). */
MH_API int mh_scene_cuboid_pile(const mh_pile_params* params, mh_transform* poses, float* inverse_masses,
                                mh_mass* masses, mh_spatial_inertia* inertias, mh_collider* colliders,
                                size_t capacity);

#ifdef __cplusplus
} /* extern "C" */

static_assert(sizeof(mh_tick_config) == 64, "abi.rs:299");
static_assert(sizeof(mh_diagnostics) == 32, "abi.rs:300");
static_assert(sizeof(mh_transform) == 32, "abi.rs:301");
static_assert(sizeof(mh_velocity) == 32, "abi.rs:302");
static_assert(sizeof(mh_mass) == 64, "abi.rs:303");
static_assert(sizeof(mh_spatial_inertia) == 64, "abi.rs:304");
static_assert(sizeof(mh_collider) == 64, "abi.rs:305");
static_assert(sizeof(mh_bearing) == 80, "abi.rs:306");
static_assert(sizeof(mh_pair) == 8, "abi.rs:307");
static_assert(sizeof(mh_contact) == 64, "abi.rs:308");
static_assert(sizeof(mh_persistent_manifold) == 48, "abi.rs:309");
static_assert(sizeof(mh_mechanism_body) == 64, "abi.rs:310");
static_assert(sizeof(mh_mechanism_coordinate) == 8, "abi.rs:311");
static_assert(sizeof(mh_mechanism_drive) == 32, "abi.rs:312");
static_assert(sizeof(mh_link_state) == 48, "abi.rs:313");
static_assert(sizeof(mh_contraction_node) == 16, "abi.rs:314");
#endif

#endif /* MECHANIC_HIP_H */
