"""Parity of the HIP step with the CPU oracle, through the C ABI.

Integer and index results (Morton keys, sorted order, LBVH topology, pair set)
must be bit-exact. Floating-point results are compared bit-exactly as well: the
kernels and the oracle evaluate the same IEEE-754 binary32 operations in the
same order (no contraction), so any difference is a defect. The north-star
tolerances (1e-5 m manifolds, 1e-5 relative velocities) are asserted too, as
the documented bar.
"""
import numpy as np
import pytest

import mechanic_amd as mh
from mechanic_amd import _lib
from tests._oracle import Oracle

pytestmark = pytest.mark.gpu

INVALID = 0xFFFFFFFF

# Broadphase variants selected by the launch tuning: the key sort (radix or
# bitonic) and the pair walk (packet walk handing its merged leaf ranges to a
# separate scan, one walk per leaf, or the packet walk scanning them itself).
# Every variant must give the same keys, tree and pair set.
BROADPHASE = {"radix-scan": {}, "bitonic": {"sort_kind": 1}, "lane": {"traverse_kind": 1},
              "bitonic-lane": {"sort_kind": 1, "traverse_kind": 1}, "inline": {"traverse_kind": 2}}
# Solver pass organisations: per contact with a per-body gather of the stored
# deltas (default), or per body over its contact-side rows. Both must give the
# oracle's bits.
SOLVER = {"per-contact": {}, "per-body": {"solve_kind": 1}}


def make(scene, **cfg):
    iterations = cfg.pop("solver_iterations", 8)
    friction = cfg.pop("friction", 0.05)
    collisions = cfg.pop("collisions_enabled", True)
    pair_capacity = cfg.pop("pair_capacity", _lib.MAX_CONTACT_PAIRS)
    use_graph = cfg.pop("use_graph", True)
    tuning = cfg.pop("tuning", None)
    assert not cfg
    gpu = mh.GpuPhysics.new_with_config(scene, mh.GpuPhysicsConfig(
        collisions_enabled=collisions, solver_iterations=iterations, friction=friction,
        pair_capacity=pair_capacity, use_graph=use_graph))
    if tuning:
        gpu.set_tuning(**tuning)
    ora = Oracle(scene, iterations=iterations, friction=friction, pair_capacity=pair_capacity,
                 collisions=collisions)
    return gpu, ora


def assert_bodies_equal(gpu, ora, what):
    gp, gv = gpu.read_bodies()
    op, ov = ora.bodies()
    for name, g, o in (("poses", gp, op), ("velocities", gv, ov)):
        diff = np.nonzero(g.view(np.uint32) != o.view(np.uint32))
        if diff[0].size:
            i = diff[0][0]
            raise AssertionError(f"{what}: {name} differ at body {i}: gpu {g[i]} oracle {o[i]} "
                                 f"({diff[0].size} words)")


def gpu_array(gpu, buf, nbytes, dtype, width):
    raw = gpu.read_buffer(buf, nbytes)
    return np.frombuffer(raw, dtype=dtype).reshape(-1, width) if width else np.frombuffer(raw, dtype=dtype)


def sort_rows(rows):
    rows = np.asarray(rows)
    keys = rows.view(np.uint32) if rows.dtype == np.float32 else rows
    order = np.lexsort(tuple(keys[:, k] for k in range(keys.shape[1] - 1, -1, -1)))
    return rows[order]


def test_integration_only_300_ticks_bit_exact():
    scene = mh.named_pile("pile_1k")
    gpu, ora = make(scene, collisions_enabled=False)
    for tick in range(1, 301):
        slot = gpu.dispatch_tick(tick)
        ora.step(tick)
        if tick % 50 == 0:
            assert_bodies_equal(gpu, ora, f"tick {tick}")
            snap = gpu.read_snapshot_transforms(slot)
            op, _ = ora.bodies()
            assert np.array_equal(snap.view(np.uint32), op.view(np.uint32))
    assert gpu.read_last_tick().error_flags == 0


def compare_broadphase(gpu, ora, scene):
    c = len(scene.colliders)
    s = max(256, 1 << (c - 1).bit_length())
    keys = gpu_array(gpu, _lib.BUFFER_MORTON_KEYS, s * 8, np.uint64, 0)
    okeys = ora.keys()
    assert np.array_equal(keys, okeys), "sorted Morton entries differ"
    assert np.all(np.diff(keys.astype(np.uint64)) > 0)
    aabbs = gpu_array(gpu, _lib.BUFFER_COLLIDER_AABBS, c * 32, np.float32, 8)
    assert np.array_equal(aabbs.view(np.uint32), ora.aabbs().view(np.uint32)), "collider AABBs differ"
    if c > 1:
        ch = gpu_array(gpu, _lib.BUFFER_NODE_CHILDREN, (c - 1) * 8, np.uint32, 2)
        assert np.array_equal(ch, ora.children()), "LBVH children differ"
        par = gpu_array(gpu, _lib.BUFFER_NODE_PARENTS, (2 * c - 1) * 4, np.uint32, 0)
        assert np.array_equal(par, ora.parents()), "LBVH parents differ"
    nodes = gpu_array(gpu, _lib.BUFFER_NODE_AABBS, (2 * c - 1) * 32, np.float32, 8)
    assert np.array_equal(nodes.view(np.uint32), ora.aabbs(nodes=True).view(np.uint32)), "node AABBs differ"
    diag = np.frombuffer(gpu.read_buffer(_lib.BUFFER_DIAGNOSTICS, 32), np.uint32)
    odiag = ora.diag()
    assert diag[1] == odiag[1], f"pair counts {diag[1]} vs {odiag[1]}"
    pairs = gpu_array(gpu, _lib.BUFFER_PAIRS, int(diag[1]) * 8, np.uint32, 2)
    assert np.array_equal(sort_rows(pairs), sort_rows(ora.pairs())), "pair sets differ"
    return int(diag[1])


@pytest.mark.parametrize("variant", list(BROADPHASE))
@pytest.mark.parametrize("name,ticks", [("pile_1k", 1), ("pile_1k", 90)])
def test_broadphase_bit_exact(name, ticks, variant):
    scene = mh.named_pile(name)
    gpu, ora = make(scene, use_graph=False, tuning=BROADPHASE[variant])
    for tick in range(1, ticks):
        gpu.dispatch_tick(tick)
        ora.step(tick)
    gpu.debug_tick(ticks, 2)
    ora.step_partial(ticks, 2)
    compare_broadphase(gpu, ora, scene)


@pytest.mark.parametrize("variant", list(BROADPHASE))
@pytest.mark.parametrize("shape", [(5, 5, 5), (9, 2, 9), (2, 33, 2)])
def test_broadphase_partial_packets_bit_exact(shape, variant):
    # 125, 162 and 132 overlapping cuboids: the last packet of 64 leaves is
    # partial, and the walkers of a packet split at nodes whose children both
    # overlap it. The keys, tree and pair set must still be the oracle's.
    scene = mh.cuboid_pile(*shape, seed=11, spacing=0.22)
    gpu, ora = make(scene, use_graph=False, tuning=BROADPHASE[variant])
    gpu.debug_tick(1, 2)
    ora.step_partial(1, 2)
    assert compare_broadphase(gpu, ora, scene) > 0


def contact_key(rows):
    u = rows.view(np.uint32)
    return u


def compare_contacts(gpu, ora):
    diag = np.frombuffer(gpu.read_buffer(_lib.BUFFER_DIAGNOSTICS, 32), np.uint32)
    odiag = ora.diag()
    assert diag[2] == odiag[2], f"contact counts {diag[2]} vs {odiag[2]}"
    k = int(diag[2])
    g = gpu_array(gpu, _lib.BUFFER_CONTACTS, k * 64, np.float32, 16)
    o = ora.contacts()
    gs, os_ = sort_rows(g.view(np.uint32)), sort_rows(o.view(np.uint32))
    # Manifold geometry within the 1e-5 m bar, then bit-exact.
    gf, of = gs.view(np.float32), os_.view(np.float32)
    assert np.array_equal(gs[:, :4], os_[:, :4]), "contact bodies/colliders differ"
    assert np.max(np.abs(gf[:, 4:] - of[:, 4:]), initial=0.0) <= 1e-5
    bad = np.nonzero(np.any(gs != os_, axis=1))[0]
    assert bad.size == 0, f"contacts differ in bits: gpu {gf[bad[0]]} oracle {of[bad[0]]}"
    return k


@pytest.mark.parametrize("ticks", [1, 45, 120])
def test_contacts_bit_exact(ticks):
    scene = mh.named_pile("pile_1k")
    gpu, ora = make(scene, use_graph=False)
    for tick in range(1, ticks):
        gpu.dispatch_tick(tick)
        ora.step(tick)
    gpu.debug_tick(ticks, 3)
    ora.step_partial(ticks, 3)
    compare_contacts(gpu, ora)


def test_device_fixed_point_rounding_matches_the_reference_rule():
    """The device conversion of the Q20 velocity deltas (add_velocity_delta,
    collision.wgsl:764-772): i32(round(x)) with round half to even and a
    saturating cast (NaN -> 0, at or above 2^31 -> INT_MAX, below -2^31 ->
    INT_MIN), which the oracle restates; checked on halfway cases, the
    saturation edges, infinities, NaN, signed zeros, subnormals and random
    values over the whole exponent range."""
    import ctypes
    special = np.array([0.0, -0.0, 0.5, 1.5, 2.5, -0.5, -1.5, -2.5, 3.5, -3.5, 8388607.5, -8388607.5,
                        2147483520.0, -2147483520.0, 2147483648.0, -2147483648.0, -2147483904.0, 4294967296.0,
                        1e30, -1e30, np.inf, -np.inf, np.nan, -np.nan, 1e-45, -1e-45, 1e-38, 0.49999997,
                        -0.49999997, 1048576.0 * 0.7, -1048576.0 * 1.3], np.float32)
    rng = np.random.default_rng(7)
    bits = rng.integers(0, 1 << 32, 200_000, dtype=np.uint64).astype(np.uint32)
    values = np.concatenate([special, bits.view(np.float32), (rng.standard_normal(100_000) * 3e6).astype(np.float32)])
    out = np.zeros(values.size, np.int32)
    rc = _lib.load().mh_device_round_i32(values.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), values.size)
    assert rc == 0
    with np.errstate(invalid="ignore"):  # signalling NaNs among the random bit patterns
        r = np.rint(values.astype(np.float64))
    expected = np.where(np.isnan(r), 0, np.clip(np.nan_to_num(r, nan=0.0, posinf=2.0 ** 31, neginf=-2.0 ** 32),
                                                -2.0 ** 31, 2.0 ** 31 - 1)).astype(np.int64).astype(np.int32)
    bad = np.nonzero(out != expected)[0]
    assert bad.size == 0, f"{bad.size} differ, first {values[bad[0]]!r}: device {out[bad[0]]} rule {expected[bad[0]]}"


def test_manifold_rows_follow_their_contacts():
    """MH_BUFFER_MANIFOLDS: each contact's persistent-manifold row of the tick
    (prepare_contacts, collision.wgsl:720-727, with persist_contacts'
    impulse, collision.wgsl:1071-1088): the
    ordered collider pair, the tick and one point, the contact's normal,
    penetration and arm, and a non-negative impulse. A contact the
    preparation culled writes no row this tick."""
    scene = mh.named_pile("pile_1k")
    gpu, _ = make(scene, friction=0.6)
    for tick in range(1, 61):
        gpu.dispatch_tick(tick)
    r = gpu.read_last_tick()
    k = r.contact_count
    c = np.frombuffer(gpu.read_buffer(_lib.BUFFER_CONTACTS, k * 64), np.float32).reshape(-1, 16)
    m = np.frombuffer(gpu.read_buffer(_lib.BUFFER_MANIFOLDS, k * 48), np.float32).reshape(-1, 12)
    cu, mu = c.view(np.uint32), m.view(np.uint32)
    kept = mu[:, 2] == 60
    assert kept.sum() >= r.active_contact_count > 0
    low, high = np.minimum(cu[:, 2], cu[:, 3]), np.maximum(cu[:, 2], cu[:, 3])
    assert np.array_equal(mu[kept, 0], low[kept]) and np.array_equal(mu[kept, 1], high[kept])
    assert np.all(mu[kept, 3] == 1)
    assert np.array_equal(mu[kept, 4:8], cu[kept, 4:8])
    assert np.array_equal(mu[kept, 8:11], cu[kept, 8:11])
    assert np.all(m[kept, 11] >= 0.0) and m[kept, 11].sum() > 0.0


@pytest.mark.parametrize("solver", list(SOLVER))
@pytest.mark.parametrize("friction,iterations", [(0.05, 8), (0.6, 8), (0.6, 3), (0.05, 1)])
def test_full_step_pile_1k_bit_exact(friction, iterations, solver):
    scene = mh.named_pile("pile_1k")
    gpu, ora = make(scene, friction=friction, solver_iterations=iterations, tuning=SOLVER[solver])
    for tick in range(1, 301):
        gpu.dispatch_tick(tick)
        ora.step(tick)
        if tick % 25 == 0 or tick < 4:
            assert_bodies_equal(gpu, ora, f"tick {tick}")
            r = gpu.read_last_tick()
            od = ora.diag()
            assert (r.error_flags, r.pair_count, r.contact_count, r.active_contact_count) == (
                od[0], od[1], od[2], od[5]), f"diagnostics differ at tick {tick}"
    assert gpu.read_last_tick().active_contact_count > 0


def test_graph_and_direct_launches_agree_and_repeat():
    scene = mh.named_pile("pile_1k")
    a = mh.GpuPhysics.new_with_config(scene, mh.GpuPhysicsConfig(use_graph=True))
    b = mh.GpuPhysics.new_with_config(scene, mh.GpuPhysicsConfig(use_graph=False))
    c = mh.GpuPhysics.new_with_config(scene, mh.GpuPhysicsConfig(use_graph=True))
    for tick in range(1, 121):
        a.dispatch_tick(tick)
        b.dispatch_tick(tick)
        c.dispatch_tick(tick)
    pa, va = a.read_bodies()
    pb, vb = b.read_bodies()
    pc, vc = c.read_bodies()
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))
    assert np.array_equal(va.view(np.uint32), vb.view(np.uint32))
    assert np.array_equal(pa.view(np.uint32), pc.view(np.uint32))
    assert np.array_equal(va.view(np.uint32), vc.view(np.uint32))


def test_edge_scenes_single_and_static_bodies():
    # One body: no internal nodes, no pairs; it lands on the ground.
    one = mh.boxes([[0.0, 0.3, 0.0]], [0.125, 0.125, 0.125])
    gpu, ora = make(one)
    for tick in range(1, 61):
        gpu.dispatch_tick(tick)
        ora.step(tick)
    assert_bodies_equal(gpu, ora, "single body")
    # A static body under a dynamic one: the static one never moves.
    two = mh.boxes([[0.0, 0.125, 0.0], [0.02, 0.4, 0.01]], [0.125, 0.125, 0.125],
                   inverse_mass=[0.0, 0.064], inverse_inertia=[0.0, 6.144])
    gpu, ora = make(two)
    for tick in range(1, 91):
        gpu.dispatch_tick(tick)
        ora.step(tick)
    assert_bodies_equal(gpu, ora, "static + dynamic")
    p, _ = gpu.read_bodies()
    assert p[0, 1] == np.float32(0.125)


def test_uneven_counts_and_tiny_pair_capacity_flags_overflow():
    scene = mh.cuboid_pile(7, 9, 5, seed=3, spacing=0.2)  # 315 bodies, overlapping start
    gpu, ora = make(scene, pair_capacity=256)
    for tick in range(1, 4):
        gpu.dispatch_tick(tick)
        ora.step(tick)
    r = gpu.read_last_tick()
    assert r.error_flags & mh.PAIR_OVERFLOW_FLAG
    assert r.pair_count == ora.diag()[1]


def test_non_finite_state_raises_invalid_numeric():
    scene = mh.named_pile("pile_1k")
    gpu, ora = make(scene)
    gpu.dispatch_tick(1)
    ora.step(1)
    p, v = gpu.read_bodies()
    v[5, 0] = np.nan
    gpu.write_bodies(p, v)
    ora.set_bodies(p, v)
    gpu.dispatch_tick(2)
    ora.step(2)
    r = gpu.read_last_tick()
    assert r.error_flags & mh.INVALID_NUMERIC_FLAG
    assert ora.diag()[0] & mh.INVALID_NUMERIC_FLAG


def test_external_impulse_matches_oracle():
    scene = mh.boxes([[0.0, 0.125, 0.0]], [0.5, 0.5, 0.5], inverse_mass=0.001, inverse_inertia=0.006)
    gpu, ora = make(scene)
    gpu.apply_impulse(0, [0.0, 0.4, 0.5], [500.0, 0.0, 0.0])
    ora.apply_impulse(0, [0.0, 0.4, 0.5], [500.0, 0.0, 0.0])
    gpu.dispatch_tick(1)
    ora.step(1)
    assert_bodies_equal(gpu, ora, "impulse")
    p, v = gpu.read_bodies()
    # device.rs:4144-4176: the body moves along +x and turns about -z... here about -y/z sign of the arm.
    assert v[0, 0] > 0.0
    with pytest.raises(mh.GpuImpulseError):
        gpu.apply_impulse(3, [0, 0, 0], [1, 0, 0])
    with pytest.raises(mh.GpuImpulseError):
        gpu.apply_impulse(0, [0, np.inf, 0], [1, 0, 0])


@pytest.mark.parametrize("variant", list(BROADPHASE))
def test_pile_100k_first_tick_broadphase_and_contacts(variant):
    scene = mh.named_pile("pile_100k")
    gpu, ora = make(scene, use_graph=False, tuning=BROADPHASE[variant])
    ora.set_threads(8)
    gpu.debug_tick(1, 3)
    ora.step_partial(1, 3)
    compare_broadphase(gpu, ora, scene)
    compare_contacts(gpu, ora)


@pytest.mark.parametrize("variant", ["radix-scan", "bitonic-lane", "per-body"])
@pytest.mark.parametrize("count", [2, 3, 5, 16, 17, 33])
def test_small_stacks_many_ticks(count, variant):
    # Fewer colliders than the per-tick counters: every counter must still reset.
    centers = [[0.01 * (i % 3), 0.125 + 0.26 * i, -0.01 * (i % 2)] for i in range(count)]
    scene = mh.boxes(centers, [0.125] * 3, inverse_mass=0.064, inverse_inertia=6.144)
    gpu, ora = make(scene, friction=0.6, tuning={**BROADPHASE, **SOLVER}[variant])
    for tick in range(1, 241):
        gpu.dispatch_tick(tick)
        ora.step(tick)
        if tick % 60 == 0:
            assert_bodies_equal(gpu, ora, f"{count} bodies, tick {tick}")


@pytest.mark.parametrize("solver", list(SOLVER))
def test_delta_slots_per_body(solver):
    """MH_BUFFER_BODY_SLOTS after a tick whose contacts are all active and fit
    one 64-contact group: the per-contact passes give every body with a contact
    side one delta slot (its group's sides summed); the per-body passes one
    slot per side."""
    count = 33
    centers = [[0.01 * (i % 3), 0.125 + 0.26 * i, -0.01 * (i % 2)] for i in range(count)]
    scene = mh.boxes(centers, [0.125] * 3, inverse_mass=0.064, inverse_inertia=6.144)
    gpu, _ = make(scene, friction=0.6, tuning=SOLVER[solver])
    for tick in range(1, 121):
        gpu.dispatch_tick(tick)
    r = gpu.read_last_tick()
    assert 0 < r.active_contact_count == r.contact_count <= 64
    rows = np.frombuffer(gpu.read_buffer(_lib.BUFFER_CONTACTS, r.contact_count * 64), np.uint32).reshape(-1, 16)
    sides = np.zeros(count, np.int64)
    for a, b in rows[:, :2]:
        sides[a] += 1
        if b != 0xFFFFFFFF:
            sides[b] += 1
    slots = np.frombuffer(gpu.read_buffer(_lib.BUFFER_BODY_SLOTS, 4 * count), np.uint32)
    expected = (sides > 0).astype(np.int64) if solver == "per-contact" else sides
    assert slots.tolist() == expected.tolist()


@pytest.mark.parametrize("solver", list(SOLVER))
def test_body_with_many_contact_sides(solver):
    """A heavy slab carrying 64 cubes: the slab has 64 cube contacts and one
    ground contact, past the per-contact passes' 31 delta slots (their atomic
    overflow path) and the same rows for the per-body passes."""
    grid = [(-1.75 + 0.5 * i, -1.75 + 0.5 * j) for i in range(8) for j in range(8)]
    centers = [[0.0, 0.1, 0.0]] + [[x, 0.305, z] for x, z in grid]
    half = [[2.0, 0.1, 2.0]] + [[0.1, 0.1, 0.1]] * 64
    inv_m = [1.0 / 400.0] + [1.0] * 64
    inv_i = [0.02] + [150.0] * 64
    scene = mh.boxes(centers, half, inverse_mass=inv_m, inverse_inertia=inv_i)
    gpu, ora = make(scene, friction=0.6, tuning=SOLVER[solver])
    most = 0
    for tick in range(1, 121):
        gpu.dispatch_tick(tick)
        ora.step(tick)
        most = max(most, gpu.read_last_tick().active_contact_count)
        if tick % 30 == 0:
            assert_bodies_equal(gpu, ora, f"slab, tick {tick}")
    assert most >= 65 and gpu.read_last_tick().error_flags == 0


@pytest.mark.parametrize("solver", list(SOLVER))
def test_body_in_more_groups_than_slots(solver):
    """A heavy slab carrying 2,116 cubes: its contact sides fall in about 33
    groups of 64 contacts, so the per-contact passes run out of its 31 delta
    slots and add the rest with atomics (the overflow path); the per-body passes
    lay out all its sides. Both against the oracle, bit for bit."""
    grid = [(-11.25 + 0.5 * i, -11.25 + 0.5 * j) for i in range(46) for j in range(46)]
    centers = [[0.0, 0.1, 0.0]] + [[x, 0.305, z] for x, z in grid]
    half = [[12.0, 0.1, 12.0]] + [[0.1, 0.1, 0.1]] * len(grid)
    inv_m = [1.0 / 4000.0] + [1.0] * len(grid)
    inv_i = [0.0005] + [150.0] * len(grid)
    scene = mh.boxes(centers, half, inverse_mass=inv_m, inverse_inertia=inv_i)
    gpu, ora = make(scene, friction=0.6, tuning=SOLVER[solver])
    for tick in range(1, 41):
        gpu.dispatch_tick(tick)
        ora.step(tick)
    assert_bodies_equal(gpu, ora, "slab with 2,116 cubes, tick 40")
    r = gpu.read_last_tick()
    assert r.active_contact_count > 31 * 64 and r.error_flags == 0
    if solver == "per-contact":
        slots = np.frombuffer(gpu.read_buffer(_lib.BUFFER_BODY_SLOTS, 4 * scene.body_count), np.uint32)
        assert slots[0] > 31  # the slab took a slot in every group, past its 31


def test_driver_publishes_and_stops_on_flags():
    from mechanic_amd import runtime as rt
    scene = mh.named_pile("pile_1k")
    gpu = mh.GpuPhysics.new(scene)
    runtime = rt.PhysicsRuntime()
    runtime.load(scene)
    runtime.start()
    driver = rt.PhysicsDriver(gpu, runtime)
    assert driver.update(250_000_000) == 15
    p = runtime.published_state()
    assert p.current_tick == 15 and p.current_slot == 0
    snap = gpu.read_snapshot_transforms(p.current_slot)
    pose, _ = gpu.read_bodies()
    assert (snap == pose).all()
    poses, vel = gpu.read_bodies()
    vel[3, 1] = float("nan")
    gpu.write_bodies(poses, vel)
    driver.update(100_000_000)
    assert runtime.failure() == rt.InvalidNumericState()


@pytest.mark.parametrize("solver", list(SOLVER))
def test_profiled_ticks_match_dispatched_ticks(solver):
    """A profiled tick (events between kernel groups, clock stamps in the solver
    passes) runs the same tick as a dispatched one. Per-body passes have no
    separate gather kernel."""
    scene = mh.named_pile("pile_1k")
    out = []
    for profiled in (False, True):
        gpu, _ = make(scene, friction=0.6, tuning=SOLVER[solver])
        for tick in range(1, 31):
            if profiled and 10 <= tick < 20:
                t = gpu.profile_tick(tick)
                assert t["total_ms"] > 0 and t["solve_kernel_ms"] > 0
                assert (t["apply_kernel_ms"] > 0) == (solver == "per-contact")
                assert t["solve_ms"] >= 8 * t["solve_kernel_ms"]
            else:
                gpu.dispatch_tick(tick)
        p, v = gpu.read_bodies()
        out.append((p.view(np.uint32).copy(), v.view(np.uint32).copy()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
