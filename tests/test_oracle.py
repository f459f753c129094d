"""Pins the CPU oracle to the reference's own vectors and to independent math.

Each check uses a method different from the oracle's code path: the reference's
SAT known-answer vectors, linear-programming overlap tests, an O(n^2) brute
force, a recursive top-down radix-tree construction, closed-form solver decay,
momentum conservation, and thread-count determinism.
"""
import hashlib
import json
import math
import os

import numpy as np
import pytest
from scipy.optimize import linprog

import mechanic_amd as mh
from tests._oracle import Oracle, morton_code, obb_sat_f64

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def quat_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def boxes_overlap_lp(ca, ra, ha, cb, rb, hb, shrink=0.0):
    """True iff a point lies in both boxes (each shrunk by `shrink`), by LP feasibility."""
    rows, rhs = [], []
    for c, r, h in ((ca, ra, ha), (cb, rb, hb)):
        m = quat_matrix(r)
        for k in range(3):
            axis = m[:, k]
            rows.append(axis)
            rhs.append(h[k] - shrink + axis @ np.asarray(c, float))
            rows.append(-axis)
            rhs.append(h[k] - shrink - axis @ np.asarray(c, float))
    res = linprog(np.zeros(3), A_ub=np.array(rows), b_ub=np.array(rhs), bounds=[(None, None)] * 3, method="highs")
    return res.status == 0


def test_reference_sat_known_answers():
    with open(os.path.join(GOLDEN, "reference_kats.json")) as f:
        kats = json.load(f)
    for case in kats["cases"]:
        (ca, ra, ha), (cb, rb, hb) = case["a"], case["b"]
        normal, pen, point, _ = obb_sat_f64(ca, ra, ha, cb, rb, hb)
        exp = case["expect"]
        # The step keeps a contact when penetration >= -1e-5 (collision.wgsl:579).
        assert (pen < -1e-5) == exp["separated"], case["name"]
        if "normal" in exp:
            assert np.allclose(normal, exp["normal"], atol=exp["tolerance"])
            assert abs(pen - exp["penetration"]) < exp["tolerance"]
        if "point_x" in exp:
            assert abs(point[0] - exp["point_x"]) < exp["tolerance"]


def random_quat(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


def test_sat_penetration_is_the_minimum_translation_by_lp():
    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(150):
        ha, hb = rng.uniform(0.1, 0.6, 3), rng.uniform(0.1, 0.6, 3)
        ra, rb = random_quat(rng), random_quat(rng)
        ca = np.zeros(3)
        cb = rng.uniform(-1.2, 1.2, 3)
        normal, pen, _, _ = obb_sat_f64(ca, ra, ha, cb, rb, hb)
        overlap = boxes_overlap_lp(ca, ra, ha, cb, rb, hb)
        if pen < -1e-9:
            assert not overlap
            continue
        if pen < 1e-6:
            continue
        assert overlap
        assert abs(np.linalg.norm(normal) - 1.0) < 1e-9
        # Pushing B out along the normal by the penetration separates the boxes;
        # pushing by a little less does not.
        moved = cb + normal * (pen + 1e-6)
        assert not boxes_overlap_lp(ca, ra, ha, moved, rb, hb, shrink=1e-9)
        moved = cb + normal * (pen * 0.98)
        assert boxes_overlap_lp(ca, ra, ha, moved, rb, hb)
        checked += 1
    assert checked > 20


def test_sat_axis_aligned_closed_form():
    for d in ([0.3, 0.1, 0.0], [-0.2, 0.55, 0.1], [0.0, 0.0, -0.7]):
        ha, hb = np.array([0.5, 0.4, 0.3]), np.array([0.25, 0.2, 0.45])
        normal, pen, _, _ = obb_sat_f64([0, 0, 0], [0, 0, 0, 1], ha, d, [0, 0, 0, 1], hb)
        overlaps = ha + hb - np.abs(d)
        k = int(np.argmin(overlaps))
        assert abs(pen - overlaps[k]) < 1e-12
        expect = np.zeros(3)
        expect[k] = 1.0 if d[k] >= 0 else -1.0
        assert np.allclose(normal, expect)


def interleave(x, y, z):
    code = 0
    for bit in range(10):
        code |= ((x >> bit) & 1) << (3 * bit)
        code |= ((y >> bit) & 1) << (3 * bit + 1)
        code |= ((z >> bit) & 1) << (3 * bit + 2)
    return code


def test_morton_closed_form():
    assert morton_code(0.0, 0.0, 0.0) == interleave(512, 512, 512)
    assert morton_code(-1000.0, 2000.0, 0.5) == interleave(0, 1023, 512)
    rng = np.random.default_rng(1)
    for _ in range(200):
        c = rng.uniform(-600, 600, 3).astype(np.float32)
        q = [int(min(max(math.floor(float(v) + 512.0), 0), 1023)) for v in c]
        assert morton_code(*[float(v) for v in c]) == interleave(*q)


def build_pile_oracle(nx=6, ny=5, nz=6, seed=2, ticks=40, stop=2, **kw):
    scene = mh.cuboid_pile(nx, ny, nz, seed=seed)
    ora = Oracle(scene, **kw)
    for t in range(1, ticks):
        ora.step(t)
    ora.step_partial(ticks, stop)
    return scene, ora


def test_sorted_keys_and_radix_tree_match_top_down_construction():
    scene, ora = build_pile_oracle(ticks=30)
    keys = ora.keys()
    c = len(scene.colliders)
    assert np.all(np.diff(keys.astype(np.uint64)) > 0)
    assert sorted(int(k) & 0xFFFFFFFF for k in keys[:c]) == list(range(c))
    assert np.all(keys[c:] >> np.uint64(32) == 0xFFFFFFFF)
    k = [int(v) for v in keys[:c]]

    # Top-down: a range splits where its highest differing key bit flips.
    expected = set()

    def build(lo, hi):
        if lo == hi:
            return
        bit = (k[lo] ^ k[hi]).bit_length() - 1
        split = lo
        while (k[split + 1] >> bit) & 1 == 0:
            split += 1
        expected.add((lo, hi, split))
        build(lo, split)
        build(split + 1, hi)

    build(0, c - 1)
    ch = ora.children()
    parents = ora.parents()
    leaf_base = c - 1

    def rng(node, memo={}):
        if node >= leaf_base:
            return node - leaf_base, node - leaf_base
        if node not in memo:
            a, b = rng(int(ch[node][0])), rng(int(ch[node][1]))
            assert a[1] + 1 == b[0]
            memo[node] = (a[0], b[1])
        return memo[node]

    got = set()
    for n in range(c - 1):
        lo, hi = rng(n)
        got.add((lo, hi, rng(int(ch[n][0]))[1]))
        # Karras: internal node n is one end of its range.
        assert n in (lo, hi)
        assert parents[int(ch[n][0])] == n + 1 and parents[int(ch[n][1])] == n + 1
    assert rng(0) == (0, c - 1) and parents[0] == 0
    assert got == expected


def brute_force_pairs(aabbs, bodies):
    mn, mx = aabbs[:, :3].astype(np.float64), aabbs[:, 4:7].astype(np.float64)
    ov = np.all(mn[:, None, :] <= mx[None, :, :], axis=2) & np.all(mn[None, :, :] <= mx[:, None, :], axis=2)
    ov &= bodies[:, None] != bodies[None, :]
    i, j = np.nonzero(np.triu(ov, 1))
    return set(zip(i.tolist(), j.tolist()))


@pytest.mark.parametrize("ticks", [1, 60])
def test_pair_set_equals_brute_force(ticks):
    scene, ora = build_pile_oracle(ticks=ticks)
    bodies = np.arange(len(scene.colliders))
    got = set(map(tuple, ora.pairs().tolist()))
    assert got == brute_force_pairs(ora.aabbs(), bodies)
    assert ora.diag()[1] == len(got)


def test_suppressed_and_same_body_pairs_are_filtered():
    scene = mh.boxes([[0, 2, 0], [0.1, 2, 0], [0.2, 2, 0], [0.3, 2, 0]], [0.125] * 3)
    scene.suppressed_pairs = [(2, 0)]
    ora = Oracle(scene)
    ora.step_partial(1, 2)
    pairs = set(map(tuple, ora.pairs().tolist()))
    assert (0, 2) not in pairs and (0, 1) in pairs and (1, 2) in pairs
    assert pairs == brute_force_pairs(ora.aabbs(), np.arange(4)) - {(0, 2)}


def test_ground_contact_closed_form():
    h = 0.125
    scene = mh.boxes([[0.3, h - 0.01, -0.2]], [h, h, h])
    ora = Oracle(scene)
    ora.step_partial(1, 3)
    c = ora.contacts()
    assert len(c) == 1
    u = c.view(np.uint32)
    assert u[0, 0] == 0 and u[0, 1] == 0xFFFFFFFF
    p, v = ora.bodies()
    bottom = p[0, 1] - h
    assert np.allclose(c[0, 4:7], [0, -1, 0])
    assert abs(c[0, 7] - (-bottom)) < 1e-6
    assert np.allclose(c[0, 8:11], [0.0, -h, 0.0], atol=1e-7)


def test_resting_box_velocity_decays_by_relaxation_per_pass():
    h = 0.125
    dt = np.float32(1.0 / 60.0)
    for iterations in (1, 3, 8):
        scene = mh.boxes([[0.0, h - 0.02, 0.0]], [h, h, h], inverse_mass=0.064, inverse_inertia=6.144)
        ora = Oracle(scene, iterations=iterations)
        p, v = ora.bodies()
        v[0, 1] = -0.7
        ora.set_bodies(p, v)
        ora.step(1)
        p1, v1 = ora.bodies()
        v0 = (-0.7 + (-9.81) * (1.0 / 60.0)) * 0.999
        pen = h - p1[0, 1]
        bias = min(max(pen - 0.001, 0.0) * 0.2 / float(dt), 1.0)
        expect = bias + (1 - 0.125) ** iterations * (v0 - bias)
        assert abs(v1[0, 1] - expect) < 2e-5 * iterations
        assert np.allclose(v1[0, [0, 2, 4, 5, 6]], 0.0, atol=1e-6)


def test_body_contacts_conserve_momentum():
    scene = mh.boxes([[0.0, 3.0, 0.0], [0.2, 3.05, 0.03]], [0.125] * 3,
                     rotations=[[0, 0, 0, 1], [0.0, 0.3826834, 0.0, 0.9238795]],
                     inverse_mass=[0.064, 0.032], inverse_inertia=[6.144, 3.072])
    ora = Oracle(scene, friction=0.6)
    p, v = ora.bodies()
    v[0, :3] = [1.0, 0.0, 0.2]
    v[1, :3] = [-0.5, -0.3, 0.0]
    ora.set_bodies(p, v)
    m = np.array([1 / 0.064, 1 / 0.032])
    before = (m[:, None] * ((v[:, :3] + [0.0, -9.81 / 60.0, 0.0]) * 0.999)).sum(axis=0)
    ora.step(1)
    assert ora.diag()[5] >= 1
    _, v1 = ora.bodies()
    after = (m[:, None] * v1[:, :3]).sum(axis=0)
    assert np.allclose(after, before, atol=1e-3)


def test_thread_count_does_not_change_bits():
    scene = mh.cuboid_pile(8, 6, 8, seed=4)
    a = Oracle(scene, threads=1, friction=0.6)
    b = Oracle(scene, threads=4, friction=0.6)
    for t in range(1, 41):
        a.step(t)
        b.step(t)
    pa, va = a.bodies()
    pb, vb = b.bodies()
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))
    assert np.array_equal(va.view(np.uint32), vb.view(np.uint32))


def test_fp64_oracle_tracks_fp32_over_a_few_ticks():
    scene = mh.cuboid_pile(5, 4, 5, seed=9)
    f32 = Oracle(scene)
    f64 = Oracle(scene, fp64=True)
    for t in range(1, 6):
        f32.step(t)
        f64.step(t)
    p32, _ = f32.bodies()
    p64, _ = f64.bodies_f64()
    assert np.max(np.abs(p32[:, :3] - p64[:, :3])) < 1e-4


def test_pile_checksums_fixture():
    with open(os.path.join(GOLDEN, "oracle_pile_checksums.json")) as f:
        fx = json.load(f)
    scene = mh.cuboid_pile(6, 5, 6, seed=11)
    ora = Oracle(scene, friction=0.6)
    for tick in range(1, 121):
        ora.step(tick)
        if str(tick) in fx["ticks"]:
            p, v = ora.bodies()
            h = hashlib.sha256(np.ascontiguousarray(p).tobytes() + np.ascontiguousarray(v).tobytes()).hexdigest()
            assert h == fx["ticks"][str(tick)]["sha256"], tick
            assert [int(x) for x in ora.diag()] == fx["ticks"][str(tick)]["diag"]


def test_invalid_numeric_flag_on_nan_velocity():
    scene = mh.cuboid_pile(3, 3, 3, seed=1)
    ora = Oracle(scene)
    p, v = ora.bodies()
    v[4, 1] = np.nan
    ora.set_bodies(p, v)
    ora.step(1)
    assert ora.diag()[0] & mh.INVALID_NUMERIC_FLAG
    p1, _ = ora.bodies()
    assert np.array_equal(p1[4].view(np.uint32), p[4].view(np.uint32))


def test_pair_overflow_flag_with_small_capacity():
    scene = mh.cuboid_pile(7, 9, 5, seed=3, spacing=0.2)
    ora = Oracle(scene, pair_capacity=256)
    ora.step(1)
    d = ora.diag()
    assert d[1] > 256 and d[0] & mh.PAIR_OVERFLOW_FLAG


# ---- independent pins of the integration, inertia, cache, friction and
# full-cylinder ground functions --------------------------------------------

DT = np.float32(1.0 / 60.0)
G = np.float32(-9.81)


def set_mass_rows(scene, i, inverse_inertia):
    """Body-frame inverse inertia (3x3, symmetric) of body i, column by column."""
    m = scene.masses[i]
    for k, col in enumerate((m.inverse_inertia_x, m.inverse_inertia_y, m.inverse_inertia_z)):
        col[:] = [float(inverse_inertia[0, k]), float(inverse_inertia[1, k]), float(inverse_inertia[2, k]), 0.0]


def spd(rng, scale):
    a = rng.normal(size=(3, 3))
    return (a @ a.T + 3.0 * np.eye(3)) * scale


def test_world_inverse_inertia_is_r_iinv_rt():
    # update_world_masses (collision.wgsl:144-162) against R I^-1 R^T formed
    # independently in FP64 from the body's quaternion, for rotated bodies with
    # fully anisotropic (non-diagonal) inverse inertia.
    rng = np.random.default_rng(21)
    rots = [random_quat(rng) for _ in range(6)]
    scene = mh.boxes([[3.0 * i, 5.0, 0.0] for i in range(6)], [0.2, 0.1, 0.3], inverse_mass=0.5)
    inv = [spd(rng, 0.7) for _ in range(6)]
    for i in range(6):
        scene.poses[i].rotation[:] = [float(x) for x in rots[i]]
        set_mass_rows(scene, i, inv[i])
    ora = Oracle(scene)
    ora.step_partial(1, 2)
    p, _ = ora.bodies()
    wm = ora.world_masses()
    for i in range(6):
        r = quat_matrix(p[i, 4:8].astype(np.float64))
        expect = r @ inv[i] @ r.T
        got = wm[i].reshape(3, 4)[:, :3].T  # rows of wm are the columns of the matrix
        assert np.allclose(got, expect, rtol=2e-5, atol=2e-6 * np.abs(expect).max()), i
        assert wm[i, 3] == np.float32(0.5)


def inertia_world(q, inverse_inertia):
    r = quat_matrix(q)
    return r @ np.linalg.inv(inverse_inertia) @ r.T


def test_body_contacts_conserve_angular_momentum():
    # Equal and opposite impulses at one world point: the solve changes no
    # angular momentum about the origin, sum m x * v + I_w w, with I_w formed
    # from each body's own rotated inertia. A transposed or wrongly rotated world
    # inverse inertia breaks this.
    rng = np.random.default_rng(8)
    inv = [spd(rng, 2.0), spd(rng, 1.0)]
    scene = mh.boxes([[0.0, 3.0, 0.0], [0.2, 3.05, 0.03]], [0.125] * 3,
                     rotations=[[0, 0, 0, 1], [0.0, 0.3826834, 0.0, 0.9238795]],
                     inverse_mass=[0.064, 0.032])
    q0 = [random_quat(rng), random_quat(rng)]
    for i in range(2):
        set_mass_rows(scene, i, inv[i])
        scene.poses[i].rotation[:] = [float(x) for x in q0[i]]
    ora = Oracle(scene, friction=0.6)
    p, v = ora.bodies()
    v[0, :3] = [1.0, 0.0, 0.2]
    v[1, :3] = [-0.5, -0.3, 0.0]
    v[0, 4:7] = [0.0, 0.0, 0.0]
    v[1, 4:7] = [0.0, 0.0, 0.0]
    ora.set_bodies(p, v)
    # Velocities the solve starts from: integrate's gravity and damping (physics.wgsl:101-108).
    vb = v[:, :3].astype(np.float32).copy()
    vb[:, 1] = (vb[:, 1] + G * DT).astype(np.float32)
    vb = (vb * np.float32(0.999)).astype(np.float32)
    ora.step(1)
    assert ora.diag()[5] >= 1
    p1, v1 = ora.bodies()
    m = np.array([1 / 0.064, 1 / 0.032])

    def momentum(lin, ang):
        total = np.zeros(3)
        for i in range(2):
            total += m[i] * np.cross(p1[i, :3].astype(np.float64), lin[i].astype(np.float64))
            total += inertia_world(p1[i, 4:8].astype(np.float64), inv[i]) @ ang[i].astype(np.float64)
        return total

    before = momentum(vb, np.zeros((2, 3)))
    after = momentum(v1[:, :3], v1[:, 4:7])
    change = np.abs(after - before).max()
    # Something must have turned, or the test pins nothing.
    assert np.abs(v1[:, 4:7]).max() > 0.05
    assert change < 2e-3 * np.abs(before).max(), (before, after)


def test_quaternion_integration_closed_form():
    # integrate (physics.wgsl:109-120): q <- normalize(q + dt/2 (w, 0) q) after
    # w <- 0.98 w. Each step is exactly a world-frame rotation about w's axis by
    # 2 atan(dt |w| / 2), so after K ticks q = r(theta_K, u) q0 with
    # theta_K = 2 sum_k atan(dt |w0| 0.98^k / 2). A body-frame product (q w)
    # or an undamped or halved rate fails.
    rng = np.random.default_rng(3)
    q0 = random_quat(rng)
    u = np.array([0.3, -0.8, 0.52])
    u /= np.linalg.norm(u)
    w0 = 7.0
    scene = mh.boxes([[0.0, 50.0, 0.0]], [0.1, 0.2, 0.3], inverse_mass=1.0)
    scene.poses[0].rotation[:] = [float(x) for x in q0]
    ora = Oracle(scene, collisions=False)
    p, v = ora.bodies()
    v[0, 4:7] = (u * w0).astype(np.float32)
    ora.set_bodies(p, v)
    q0 = p[0, 4:8].astype(np.float64)
    w_exact = np.float64(v[0, 4:7].astype(np.float64) @ u)
    theta = 0.0
    for k in range(1, 61):
        ora.step(k)
        theta += 2.0 * math.atan(0.5 * float(DT) * w_exact * 0.98 ** k)
        if k in (1, 10, 60):
            r = np.array([*(u * math.sin(theta / 2)), math.cos(theta / 2)])
            x1, y1, z1, w1 = r
            x2, y2, z2, w2 = q0
            expect = np.array([w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                               w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2])
            got = ora.bodies()[0][0, 4:8].astype(np.float64)
            assert np.abs(got - expect).max() < 2e-6 * k + 1e-6, (k, got, expect)


def resting_box(vy, y, inverse_mass=0.064, inverse_inertia=6.144):
    scene = mh.boxes([[0.0, y, 0.0]], [0.125] * 3, inverse_mass=inverse_mass, inverse_inertia=inverse_inertia)
    ora = Oracle(scene)
    p, v = ora.bodies()
    v[0, 1] = vy
    ora.set_bodies(p, v)
    return scene, ora


def ground_tick_prediction(p, v, warm, inverse_mass, iterations=8, h=0.125):
    """Normal speed after one tick of a face-down box on the ground: integrate,
    then one contact whose normal denominator is the inverse mass. With
    u0 the integrated speed and b the bias, the first pass gives
    u1 = u0 + warm d + 0.125 (b - u0), every later one u <- u + 0.125 (b - u):
    u_I = b + 0.875^I (u0 - b) + 0.875^(I-1) warm d."""
    u0 = float(np.float32((np.float32(v) + G * DT) * np.float32(0.999)))
    y = float(np.float32(np.float32(p) + np.float32(u0) * DT))
    pen = h - y
    b = min(max(pen - 0.001, 0.0) * 0.2 / float(DT), 1.0)
    return b + 0.875 ** iterations * (u0 - b) + 0.875 ** (iterations - 1) * warm * inverse_mass


def test_cache_warm_start_carries_half_the_previous_impulse():
    # prepare_contacts (collision.wgsl:699-733) + warm_start (collision.wgsl:783):
    # a contact prepared at tick - 1 with the same normal and arm starts from
    # half its previous accumulated impulse.
    h = 0.125
    _, ora = resting_box(-0.6, h - 0.01)
    ora.step(1)
    lam1 = float(ora.contacts()[0, 11])
    assert lam1 > 0.5
    p1, v1 = ora.bodies()
    ora.step(2)
    _, v2 = ora.bodies()
    expect = ground_tick_prediction(p1[0, 1], v1[0, 1], 0.5 * lam1, 0.064)
    cold = ground_tick_prediction(p1[0, 1], v1[0, 1], 0.0, 0.064)
    assert abs(expect - cold) > 0.01
    assert abs(float(v2[0, 1]) - expect) < 2e-5


def cold_start_copy(scene, ora, tick):
    """A fresh oracle (empty manifold cache) in `ora`'s body state, stepped `tick`."""
    fresh = Oracle(scene)
    p, v = ora.bodies()
    fresh.set_bodies(p, v)
    fresh.step(tick)
    return fresh.bodies()


def test_cache_miss_after_a_tick_without_the_contact():
    # The cached entry must come from tick - 1 (collision.wgsl:713): a contact
    # absent for one tick comes back cold.
    h = 0.125
    scene, ora = resting_box(-0.6, h - 0.01)
    ora.step(1)
    p, v = ora.bodies()
    lifted = p.copy()
    lifted[0, 1] = 2.0
    ora.set_bodies(lifted, v)
    ora.step(2)
    assert ora.diag()[2] == 0
    ora.set_bodies(p, v)
    cold_p, cold_v = cold_start_copy(scene, ora, 3)
    ora.step(3)
    p3, v3 = ora.bodies()
    assert ora.diag()[5] == 1
    assert np.array_equal(v3.view(np.uint32), cold_v.view(np.uint32))


def test_cache_miss_when_the_contact_point_moves():
    # |arm - cached arm| <= 0.02 m (collision.wgsl:716-717): tilting the box by
    # two degrees moves its ground support from the face centre to an edge
    # (0.125 m away), so the contact starts cold; a matched contact would not.
    h = 0.125
    scene, ora = resting_box(-0.6, h - 0.01)
    ora.step(1)
    p, v = ora.bodies()
    tilt = p.copy()
    a = math.radians(2.0)
    tilt[0, 4:8] = [0.0, 0.0, math.sin(a / 2), math.cos(a / 2)]
    ora.set_bodies(tilt, v)
    _, cold_v = cold_start_copy(scene, ora, 2)
    ora.step(2)
    _, v2 = ora.bodies()
    assert np.array_equal(v2.view(np.uint32), cold_v.view(np.uint32))
    # The same state without the tilt does take the cached impulse.
    scene, ora = resting_box(-0.6, h - 0.01)
    ora.step(1)
    _, cold_v = cold_start_copy(scene, ora, 2)
    ora.step(2)
    _, v2 = ora.bodies()
    assert abs(float(v2[0, 1]) - float(cold_v[0, 1])) > 0.01


def test_sliding_box_loses_mu_times_normal_impulse():
    # Friction bound (collision.wgsl:818-836, 884-895): mu * lambda on the first
    # pass, mu * |delta lambda| on each later one. While the box slides and the
    # normal impulse only grows, the friction impulses over a tick sum to
    # mu * lambda_final: the linear speed drops by mu lambda / m and the spin
    # about z by h mu lambda / I. A bound of mu * lambda on every pass, or none
    # on the first, gives several times more.
    h, mu, inv_m, inv_i = 0.125, 0.6, 0.064, 6.144
    scene = mh.boxes([[0.0, h - 0.005, 0.0]], [h] * 3, inverse_mass=inv_m, inverse_inertia=inv_i)
    ora = Oracle(scene, friction=mu)
    p, v = ora.bodies()
    v[0, 0] = 3.0
    v[0, 1] = -0.5
    ora.set_bodies(p, v)
    ora.step(1)
    lam = float(ora.contacts()[0, 11])
    _, v1 = ora.bodies()
    assert lam > 1.0
    vx0 = float(np.float32(np.float32(3.0) * np.float32(0.999)))
    assert abs(float(v1[0, 0]) - (vx0 - mu * lam * inv_m)) < 2e-5
    assert abs(float(v1[0, 6]) - (-h * mu * lam * inv_i)) < 2e-5
    assert abs(float(v1[0, 4])) < 1e-6 and abs(float(v1[0, 5])) < 1e-6


def cylinder_oracle(axis_rotation):
    from mechanic_amd import construction as C

    g = C.ConstructionGraph()
    g.spawn(C.Cylinder(1.0, 0.0, 0.5, C.BuildPose((0, 4, 0))))
    scene = g.compile().scene
    scene.poses[0].rotation[:] = [float(x) for x in axis_rotation]
    return scene, Oracle(scene)


def lowest_point(center, axis, radius, half_length):
    """The lowest point of a solid cylinder (FP64 closed form)."""
    down = np.array([0.0, -1.0, 0.0])
    axial = down @ axis
    radial = down - axis * axial
    n = np.linalg.norm(radial)
    point = center + (radial / n) * radius if n > 1e-9 else center.copy()
    # A horizontal axis makes the whole bottom line lowest: its midpoint.
    return point + axis * np.sign(axial) * half_length if abs(axial) > 1e-6 else point


@pytest.mark.parametrize("tilt_deg", [90.0, 60.0, 30.0])
def test_full_cylinder_ground_contact_is_its_lowest_rim_point(tilt_deg):
    # full_cylinder_contact_count / full_cylinder_support_point
    # (collision.wgsl:193-238, 622-664): a tilted full cylinder touches the
    # ground with one analytic contact at its lowest rim point, whatever its
    # sixteen segment colliders' own support points are.
    a = math.radians(tilt_deg)
    q = [math.sin(a / 2), 0.0, 0.0, math.cos(a / 2)]  # about x: the y axis tilts toward z
    scene, ora = cylinder_oracle(q)
    p, v = ora.bodies()
    axis = quat_matrix(np.array(q)) @ np.array([0.0, 1.0, 0.0])
    low = lowest_point(np.zeros(3), axis, 0.5, 0.25)
    p[0, :3] = [0.3, -0.02 - low[1], -0.1]  # 2 cm into the ground
    ora.set_bodies(p, v)
    ora.step_partial(1, 3)
    c = ora.contacts()
    pb, _ = ora.bodies()
    assert len(c) == 1
    u = c.view(np.uint32)
    assert u[0, 1] == 0xFFFFFFFF and c[0, 15] == -1.0
    expect = lowest_point(pb[0, :3].astype(np.float64), quat_matrix(pb[0, 4:8].astype(np.float64)) @ [0, 1, 0], 0.5,
                          0.25)
    got = c[0, 8:11].astype(np.float64) + pb[0, :3]
    assert np.allclose(got, expect, atol=2e-6)
    assert abs(c[0, 7] - (-expect[1])) < 2e-6


def test_flat_full_cylinder_touches_with_four_rim_points():
    # An upright full cylinder (axis within the 0.05 radial threshold of
    # vertical) rests on four analytic rim points of its lower face, roles 1..4
    # at +-ta, +-tb (collision.wgsl:213-231); roles 5..16 add none.
    scene, ora = cylinder_oracle([0.0, 0.0, 0.0, 1.0])
    p, v = ora.bodies()
    p[0, :3] = [0.0, 0.25 - 0.01, 0.0]
    ora.set_bodies(p, v)
    ora.step_partial(1, 3)
    c = ora.contacts()
    pb, _ = ora.bodies()
    assert len(c) == 4
    points = c[:, 8:11].astype(np.float64) + pb[0, :3]
    centre = pb[0, :3].astype(np.float64) - [0.0, 0.25, 0.0]
    expect = [centre + d for d in ([0.5, 0, 0], [-0.5, 0, 0], [0, 0, 0.5], [0, 0, -0.5])]
    for e in expect:
        assert np.min(np.abs(points - e).max(axis=1)) < 2e-6
    assert np.allclose(c[:, 7], 0.25 - float(pb[0, 1]), atol=2e-6) and np.all(c[:, 15] == -1.0)
