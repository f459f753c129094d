"""Articulated scenes: bearing graphs lowered to the rows the step uploads.

`compile_mechanisms` restates the topology part of the reference's compile step
(crates/mechanic-core/src/compile.rs:352-503 and compile_tree_metadata,
compile.rs:504-611) and the mechanism lowering of GpuPhysics
(crates/mechanic-gpu/src/device.rs:500-515, 1684-1808):

* bearings join compounds; every bearing suppresses collisions between its two
  compounds (compile.rs:395);
* the spanning forest takes driven bearings first, then passive ones, in row
  order; an edge inside one tree, or between two trees that each hold a static
  body, becomes a loop-closure edge (compile.rs:418-439);
* tree edges get coordinates in row order (compile.rs:441-475);
* each component is traversed breadth-first from its static bodies, or from its
  lowest body when none is static, visiting neighbours sorted by
  (body, bearing row, direction) (compile.rs:504-585);
* mechanism body rows carry parent, tree-bearing row, direction and root flag,
  with the bind pose relative to the parent (device.rs:1705-1750; compounds are
  compiled with identity root rotation, compile.rs:331).

`articulated_pile` builds the 100,000-body articulated scene: 90,000 cuboids in
the jittered pile lattice plus 10,000 solid or hollow cylinders, about half of
them hinged into zigzag chains of 2-20 links hanging from static roots in
vertical planes, with 500 loop-closure bearings (planar four-joint loops) and 20%
of the bearings driven at +-2 rad/s within +-pi/2 by gentle motors
(0.5 rad/s^2), none inside a loop. Chains are laid out so that none can reach
the ground or another chain; the other cylinders lie free on the ground. Each
cylinder is one OBB collider (half extents r, L/2, r).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .scenes import Scene, cuboid_pile

DENSITY = 1000.0  # kg/m^3 (CUBOID_DENSITY_KG_M3)
INF = float("inf")


@dataclass
class Drive:
    """A coordinate drive row (GpuMechanismDrive, abi.rs:220-238)."""

    mode: int = 0  # 0 passive, 1 speed, 2 angle
    max_acceleration: float = 0.0
    max_speed: float = 0.0
    target_speed: float = 0.0
    target_angle: float = 0.0
    min_angle: float = -INF
    max_angle: float = INF

    def row(self) -> "_lib.MechanismDrive":
        return _lib.MechanismDrive(self.mode, self.max_acceleration, self.max_speed, self.target_speed,
                                   self.target_angle, self.min_angle, self.max_angle, 0.0)


PASSIVE = Drive()


@dataclass
class BearingSpec:
    """A hinge between compounds a and b at a world anchor about a world axis."""

    a: int
    b: int
    anchor: Sequence[float]
    axis: Sequence[float]
    drive: Optional[Drive] = None


class _Dsu:
    def __init__(self, n):
        self.p = list(range(n))

    def find(self, x):
        p = self.p
        while p[x] != x:
            p[x] = p[p[x]]
            x = p[x]
        return x

    def union(self, a, b):
        ra, rb = self.find(a), self.find(b)
        if ra != rb:
            if ra < rb:
                ra, rb = rb, ra
            self.p[ra] = rb
        return rb


@dataclass
class Mechanisms:
    bearings: ctypes.Array
    mechanism_bodies: ctypes.Array
    preorder: np.ndarray
    drives: ctypes.Array
    coordinate_count: int
    closure_count: int
    suppressed_pairs: List[tuple]
    body_components: np.ndarray
    root_flags: np.ndarray
    max_depth: int
    parents: np.ndarray
    depth: np.ndarray
    tree_rows: List[int] = field(default_factory=list)


def compile_mechanisms(positions: np.ndarray, static: np.ndarray, specs: Sequence[BearingSpec]) -> Mechanisms:
    """Lowers bearing specs over compounds at `positions` (centres of mass, (N, 3))."""
    n = len(positions)
    positions = np.asarray(positions, np.float32)
    comp = _Dsu(n)
    suppressed = set()
    for s in specs:
        if s.a == s.b:
            raise ValueError("a bearing must join two compounds")
        comp.union(s.a, s.b)
        suppressed.add((min(s.a, s.b), max(s.a, s.b)))

    # Spanning forest: driven bearings first, then passive ones (compile.rs:418-439).
    forest = _Dsu(n)
    fixed = [bool(x) for x in static]
    tree = set()
    for driven_pass in (True, False):
        for row, s in enumerate(specs):
            if (s.drive is not None and s.drive.mode != 0) != driven_pass:
                continue
            ra, rb = forest.find(s.a), forest.find(s.b)
            if ra == rb or (fixed[ra] and fixed[rb]):
                continue
            has_fixed = fixed[ra] or fixed[rb]
            r = forest.union(ra, rb)
            fixed[r] = has_fixed
            tree.add(row)

    rows = (_lib.Bearing * max(len(specs), 1))()
    coordinate_of_row = {}
    drives: List[Drive] = []
    for row, s in enumerate(specs):
        if row in tree:
            coordinate_of_row[row] = len(drives)
            drives.append(s.drive if s.drive is not None else PASSIVE)
        elif s.drive is not None and s.drive.mode != 0:
            raise ValueError(f"driven bearing {row} is a loop-closure edge")
        anchor = np.asarray(s.anchor, np.float32)
        axis = np.asarray(s.axis, np.float32)
        la = anchor - positions[s.a]
        lb = anchor - positions[s.b]
        r = rows[row]
        r.local_anchor_a[:] = [float(la[0]), float(la[1]), float(la[2]), 0.0]
        r.local_anchor_b[:] = [float(lb[0]), float(lb[1]), float(lb[2]), 0.0]
        r.local_axis_a[:] = [float(axis[0]), float(axis[1]), float(axis[2]), 0.0]
        r.local_axis_b[:] = [float(axis[0]), float(axis[1]), float(axis[2]), 0.0]
        coord = coordinate_of_row.get(row)
        r.metadata[:] = [s.a, s.b, 0xFFFFFFFF if coord is None else coord, 1 if coord is None else 0]

    # Components ordered by their lowest body.
    members = {}
    for b in range(n):
        members.setdefault(comp.find(b), []).append(b)
    components = sorted(members.values(), key=lambda m: m[0])

    adjacency = [[] for _ in range(n)]
    for row, s in enumerate(specs):
        if row in tree:
            adjacency[s.a].append((s.b, row, 0))
            adjacency[s.b].append((s.a, row, 1))
    for nb in adjacency:
        nb.sort()

    parent = np.arange(n, dtype=np.int64)
    tree_row = np.full(n, -1, dtype=np.int64)
    direction = np.zeros(n, dtype=np.uint32)
    depth = np.zeros(n, dtype=np.uint32)
    component_index = np.zeros(n, dtype=np.uint32)
    is_root = np.ones(n, dtype=bool)
    visited = np.zeros(n, dtype=bool)
    preorder: List[int] = []
    for ci, comp_members in enumerate(components):
        fixed_roots = [b for b in comp_members if static[b]]
        roots = fixed_roots if fixed_roots else [comp_members[0]]
        for root in roots:
            if visited[root]:
                continue
            visited[root] = True
            component_index[root] = ci
            queue = [root]
            head = 0
            while head < len(queue):
                p = queue[head]
                head += 1
                preorder.append(p)
                for child, row, dirn in adjacency[p]:
                    if visited[child]:
                        continue
                    visited[child] = True
                    parent[child] = p
                    tree_row[child] = row
                    direction[child] = dirn
                    component_index[child] = ci
                    depth[child] = depth[p] + 1
                    is_root[child] = False
                    queue.append(child)
    preorder_index = np.zeros(n, dtype=np.uint32)
    preorder_index[np.asarray(preorder, dtype=np.int64)] = np.arange(n, dtype=np.uint32)
    post = sorted(preorder, key=lambda b: (-int(depth[b]), b))
    postorder_index = np.zeros(n, dtype=np.uint32)
    postorder_index[np.asarray(post, dtype=np.int64)] = np.arange(n, dtype=np.uint32)

    bodies = (_lib.MechanismBody * n)()
    for b in range(n):
        m = bodies[b]
        m.traversal[:] = [int(component_index[b]), int(depth[b]), int(preorder_index[b]), int(postorder_index[b])]
        if is_root[b]:
            m.metadata[:] = [b, 0xFFFFFFFF, 0, 1]
            m.bind_relative_position[:] = [0.0, 0.0, 0.0, 0.0]
            m.bind_relative_rotation[:] = [0.0, 0.0, 0.0, 1.0]
        else:
            p = int(parent[b])
            rel = positions[b] - positions[p]
            m.metadata[:] = [p, int(tree_row[b]), int(direction[b]), 0]
            m.bind_relative_position[:] = [float(rel[0]), float(rel[1]), float(rel[2]), 0.0]
            m.bind_relative_rotation[:] = [0.0, 0.0, 0.0, 1.0]
    drive_rows = (_lib.MechanismDrive * max(len(drives), 1))(*[d.row() for d in drives])
    return Mechanisms(
        bearings=rows, mechanism_bodies=bodies, preorder=np.asarray(preorder, dtype=np.uint32),
        drives=drive_rows, coordinate_count=len(drives), closure_count=len(specs) - len(tree),
        suppressed_pairs=sorted(suppressed), body_components=component_index, root_flags=is_root.astype(np.uint32),
        max_depth=int(depth.max()) if n else 0, parents=parent, depth=depth, tree_rows=sorted(tree))


def attach(scene: Scene, mech: Mechanisms) -> Scene:
    """A copy of `scene` carrying the mechanism rows."""
    return Scene(scene.body_count, scene.poses, scene.inverse_masses, scene.masses, scene.inertias, scene.colliders,
                 suppressed_pairs=list(mech.suppressed_pairs), body_components=mech.body_components,
                 root_flags=mech.root_flags, bearings=mech.bearings, bearing_count=len(mech.tree_rows) +
                 mech.closure_count, mechanism_bodies=mech.mechanism_bodies, preorder=mech.preorder,
                 drives=mech.drives, coordinate_count=mech.coordinate_count)


# ---- rows for hand-built bodies ------------------------------------------------

def _quat_axis_angle(axis, angle):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    s = math.sin(angle * 0.5)
    return np.array([axis[0] * s, axis[1] * s, axis[2] * s, math.cos(angle * 0.5)])


def _quat_mul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + bw * ax + (ay * bz - az * by), aw * by + bw * ay + (az * bx - ax * bz),
                     aw * bz + bw * az + (ax * by - ay * bx), aw * bw - (ax * bx + ay * by + az * bz)])


def _quat_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def cuboid_mass(half_extents, density=DENSITY):
    hx, hy, hz = (2.0 * float(h) for h in half_extents)
    m = density * hx * hy * hz
    return m, np.diag([m * (hy * hy + hz * hz) / 12.0, m * (hx * hx + hz * hz) / 12.0, m * (hx * hx + hy * hy) / 12.0])


def cylinder_mass(outer_radius, inner_radius, length, density=DENSITY):
    """Full (2 pi) hollow cylinder about its local y axis (compile.rs:853-877)."""
    r2 = outer_radius * outer_radius + inner_radius * inner_radius
    m = density * 2.0 * math.pi * (outer_radius ** 2 - inner_radius ** 2) * length * 0.5
    axial = length * length / 12.0
    return m, np.diag([m * (axial + r2 / 4.0), m * (r2 / 2.0), m * (r2 / 4.0 + axial)])


def body_rows(positions, rotations, half_extents, masses_inertias, static, roles=None) -> Scene:
    """A Scene of single-collider compounds: root rotation identity, the body's
    orientation carried by its collider (compile.rs:331, 889-895)."""
    n = len(positions)
    poses = (_lib.Transform * n)()
    inv = (ctypes.c_float * n)()
    masses = (_lib.Mass * n)()
    inertias = (_lib.SpatialInertia * n)()
    cols = (_lib.Collider * n)()
    for i in range(n):
        p = positions[i]
        poses[i].position[:] = [float(p[0]), float(p[1]), float(p[2]), 0.0]
        poses[i].rotation[:] = [0.0, 0.0, 0.0, 1.0]
        m, inertia_local = masses_inertias[i]
        R = _quat_matrix(rotations[i])
        inertia = R @ inertia_local @ R.T
        inertias[i].mass[:] = [m, 0.0, 0.0, 0.0]
        for k, col in enumerate((inertias[i].inertia_x, inertias[i].inertia_y, inertias[i].inertia_z)):
            col[:] = [float(inertia[0, k]), float(inertia[1, k]), float(inertia[2, k]), 0.0]
        if static[i]:
            inv_m, inv_i = 0.0, np.zeros((3, 3))
        else:
            inv_m, inv_i = 1.0 / m, np.linalg.inv(inertia)
        inv[i] = inv_m
        masses[i].inverse_mass[:] = [inv_m, 0.0, 0.0, 0.0]
        for k, col in enumerate((masses[i].inverse_inertia_x, masses[i].inverse_inertia_y,
                                 masses[i].inverse_inertia_z)):
            col[:] = [float(inv_i[0, k]), float(inv_i[1, k]), float(inv_i[2, k]), 0.0]
        c = cols[i]
        c.local_center[:] = [0.0, 0.0, 0.0, 0.0]
        c.local_rotation[:] = [float(x) for x in rotations[i]]
        h = half_extents[i]
        c.half_extents[:] = [float(h[0]), float(h[1]), float(h[2]), 0.0]
        c.metadata[:] = [i, i, 0, 0 if roles is None else int(roles[i])]
    return Scene(n, poses, inv, masses, inertias, cols)


def _concat(a: Scene, b: Scene) -> Scene:
    """Bodies of a then bodies of b (collider metadata of b re-indexed)."""
    n = a.body_count + b.body_count
    poses = (_lib.Transform * n)(*list(a.poses), *list(b.poses))
    inv = (ctypes.c_float * n)(*list(a.inverse_masses), *list(b.inverse_masses))
    masses = (_lib.Mass * n)(*list(a.masses), *list(b.masses))
    inertias = (_lib.SpatialInertia * n)(*list(a.inertias), *list(b.inertias))
    cols = (_lib.Collider * (len(a.colliders) + len(b.colliders)))(*list(a.colliders), *list(b.colliders))
    for i in range(len(a.colliders), len(cols)):
        body = cols[i].metadata[0] + a.body_count
        cols[i].metadata[0] = body
        cols[i].metadata[1] = body
    return Scene(n, poses, inv, masses, inertias, cols)


def articulated_pile(seed: int = 2, nx: int = 50, ny: int = 36, nz: int = 50, cylinders: int = 10_000,
                     tree_bearings: int = 4_500, closures: int = 500, driven_fraction: float = 0.2,
                     chain_lengths=(2, 20), drive_acceleration: float = 0.5) -> Scene:
    """The articulated configuration (defaults: 90,000 cuboids + 10,000 cylinders,
    5,000 bearings of which 500 close loops, 20% of the tree bearings driven)."""
    rng = np.random.default_rng(seed)
    pile = cuboid_pile(nx, ny, nz, seed=seed)
    # Cylinder dimensions: outer diameter U[0.25, 2], bore U[0, outer - 0.05], length U[0.25, 4].
    outer_d = rng.uniform(0.25, 2.0, cylinders)
    bore = rng.uniform(0.0, 1.0, cylinders) * (outer_d - 0.05)
    length = rng.uniform(0.25, 4.0, cylinders)
    # Chains: consecutive cylinders, lengths uniform in chain_lengths, until the
    # tree-bearing budget is spent.
    chains = []
    used = 0
    bearings_left = tree_bearings
    while bearings_left > 0 and used + 2 <= cylinders:
        L = int(rng.integers(chain_lengths[0], chain_lengths[1] + 1))
        L = min(L, bearings_left + 1, cylinders - used)
        chains.append(list(range(used, used + L)))
        used += L
        bearings_left -= L - 1
    free = list(range(used, cylinders))
    # Links of one chain share one diameter and bore (drawn once per chain, so
    # every cylinder's marginal distribution is unchanged): the Jacobi bearing
    # projection does not converge across the ~3000:1 mass ratios of
    # independently drawn neighbours.
    for chain in chains:
        outer_d[chain] = outer_d[chain[0]]
        bore[chain] = bore[chain[0]]

    # Loop closures: link a to link a + 3 at the centre of a + 3 (a planar
    # four-joint loop, one degree of freedom), every loop slot of the longest
    # chains first (five on a 20-link chain), so that whole chains stay free of
    # loops and carry the drives.
    loops = []
    for ch in sorted((ch for ch in chains if len(ch) >= 4), key=len, reverse=True):
        for start in range(0, len(ch) - 3, 4):
            if len(loops) < closures:
                loops.append((ch, start))
    positions = np.zeros((cylinders, 3))
    rotations = np.zeros((cylinders, 4))
    rotations[:, 3] = 1.0
    static = np.zeros(cylinders, dtype=bool)
    specs: List[BearingSpec] = []
    # Chains hang as zigzags in vertical x-y planes from static roots, hinged
    # about z, so every link stays in its chain's plane. A chain's reach is its
    # total length: roots sit that high above the ground, chains sharing a plane
    # are two reaches apart, and planes are 3 m apart (more than any diameter),
    # so no chain can touch the ground or another chain whatever its joints do.
    z_axis = np.array([0.0, 0.0, 1.0])
    per_row = 5
    x_cursor = {}
    for c, chain in enumerate(chains):
        row = c // per_row
        reach = float(length[chain].sum())
        x0 = x_cursor.get(row, 20.0) + reach + 1.0
        x_cursor[row] = x0 + reach + 1.0
        joint = np.array([x0, reach + 2.0, -40.0 - 3.0 * row])
        for k, b in enumerate(chain):
            a = -math.pi / 2 + math.radians(rng.uniform(20.0, 60.0)) * (1 if k % 2 == 0 else -1)
            d = np.array([math.cos(a), math.sin(a), 0.0])
            positions[b] = joint + d * (length[b] * 0.5)
            # Local y (the cylinder axis) along d: rotate about z by a - pi/2.
            rotations[b] = _quat_axis_angle(z_axis, a - math.pi / 2)
            if k > 0:
                specs.append(BearingSpec(chain[k - 1], b, joint.copy(), z_axis, None))
            joint = joint + d * length[b]
        static[chain[0]] = True
    tree_count = len(specs)
    for ch, start in loops:
        a, b = ch[start], ch[start + 3]
        specs.append(BearingSpec(a, b, positions[b].copy(), z_axis, None))
    # Drives: driven_fraction of all bearings, on the chains that carry no loop.
    # A drive on a looped chain (inside a loop two motors fight over its one
    # degree of freedom; between or below loops it excites them) pumps energy
    # into the loops faster than the reference's velocity scheme removes it.
    looped = {b for ch, _ in loops for b in ch}
    eligible = [i for i in range(tree_count) if specs[i].b not in looped]
    want = min(len(eligible), int(round(driven_fraction * len(specs))))
    for i in sorted(rng.choice(len(eligible), size=want, replace=False)) if want else []:
        specs[eligible[i]].drive = Drive(mode=1, max_acceleration=drive_acceleration, max_speed=2.0,
                                         target_speed=2.0 if rng.uniform() < 0.5 else -2.0,
                                         min_angle=-math.pi / 2, max_angle=math.pi / 2)
    # Free cylinders lie on a grid on the other side of the pile, axis along x,
    # resting 0.3 m up.
    grid = int(math.ceil(math.sqrt(max(len(free), 1))))
    for k, b in enumerate(free):
        gi, gj = divmod(k, grid)
        positions[b] = [-20.0 - gi * 4.5, outer_d[b] * 0.5 + 0.3, (gj - grid / 2) * 4.5]
        rotations[b] = _quat_axis_angle(z_axis, -math.pi / 2)
    half = np.stack([outer_d * 0.5, length * 0.5, outer_d * 0.5], axis=1)
    mi = [cylinder_mass(outer_d[i] * 0.5, bore[i] * 0.5, length[i]) for i in range(cylinders)]
    cyl = body_rows(positions, rotations, half, mi, static)
    scene = _concat(pile, cyl)
    n0 = pile.body_count
    shifted = [BearingSpec(s.a + n0, s.b + n0, s.anchor, s.axis, s.drive) for s in specs]
    all_static = np.concatenate([np.zeros(n0, dtype=bool), static])
    all_pos = np.concatenate([scene.pose_array()[:, :3]])
    mech = compile_mechanisms(all_pos, all_static, shifted)
    return attach(scene, mech)


# ---- small mechanisms (tests, examples) ---------------------------------------

def _cuboids(centers, half_extents, static):
    n = len(centers)
    rot = np.tile([0.0, 0.0, 0.0, 1.0], (n, 1))
    mi = [cuboid_mass(h) for h in half_extents]
    return body_rows(np.asarray(centers, np.float64), rot, half_extents, mi, static)


def hinge_chain(links: int = 4, length: float = 0.5, axis=(0.0, 0.0, 1.0), height: float = 3.0,
                drive: Optional[Drive] = None) -> Scene:
    """A static block with `links` cuboids hanging off it in a horizontal line
    along +x, hinged end to end about `axis`; `drive` drives the first joint."""
    half = [(length * 0.5, 0.05, 0.05)] * (links + 1)
    centers = [(-length * 0.5, height, 0.0)] + [((k + 0.5) * length, height, 0.0) for k in range(links)]
    static = [True] + [False] * links
    scene = _cuboids(centers, half, static)
    specs = [BearingSpec(k, k + 1, (k * length, height, 0.0), axis, drive if k == 0 else None)
             for k in range(links)]
    return attach(scene, compile_mechanisms(np.asarray(centers, np.float32), np.asarray(static), specs))


def four_bar(side: float = 1.0, height: float = 3.0) -> Scene:
    """A planar four-bar of cuboid links about z: a static base from (0, h) to
    (s, h), a crank up to (0, h + s), a coupler to (s, h + s) and a rocker back
    down, closed by a loop-closure bearing between rocker and base at (s, h)."""
    h, s = height, side
    centers = [(s * 0.5, h, 0.0), (0.0, h + s * 0.5, 0.0), (s * 0.5, h + s, 0.0), (s, h + s * 0.5, 0.0)]
    half = [(s * 0.5, 0.05, 0.05), (0.05, s * 0.5, 0.05), (s * 0.5, 0.05, 0.05), (0.05, s * 0.5, 0.05)]
    static = [True, False, False, False]
    scene = _cuboids(centers, half, static)
    z = (0.0, 0.0, 1.0)
    specs = [BearingSpec(0, 1, (0.0, h, 0.0), z), BearingSpec(1, 2, (0.0, h + s, 0.0), z),
             BearingSpec(2, 3, (s, h + s, 0.0), z), BearingSpec(3, 0, (s, h, 0.0), z)]
    return attach(scene, compile_mechanisms(np.asarray(centers, np.float32), np.asarray(static), specs))
