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
