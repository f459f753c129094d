"""Construction-graph lowering: editable parts, welds and bearings to the rows
GpuPhysics uploads.

Restates the reference's compile step and its GPU lowering:

* parts and poses on the quarter-metre grid (crates/mechanic-core/src/geometry.rs:7,
  126-131, 178-190, 447-455, 503-509): cuboids of 1..32 grid units a side,
  control blocks as one-unit cuboids, solid or hollow cylinder sectors about
  their local y axis;
* weld groups by union-find over part-part welds, then rigid links; a weld to
  the ground makes the whole group static (compile.rs:256-343, 936-975);
* mass properties per group about its centre of mass with parallel-axis terms,
  compounds rooted at the centre of mass with identity rotation
  (compile.rs:768-879);
* one collider per cuboid and sixteen radial segment colliders per cylinder
  (compile.rs:881-929), with the full-cylinder ground data of
  device.rs:3008-3045 (segment-centre radius, outer radius, role 1..16);
* bearings: duplicate rows of one physical joint collapse, every bearing
  suppresses its compound pair, the forest takes driven bearings first
  (compile.rs:352-503, through `articulated.compile_mechanisms`);
* each tree coordinate's child-subtree inertia about its axis and the drive
  row of a wired bearing from its limits, first program state and wire
  direction (compile.rs:613-766).

Mass properties and collider rows are computed in FP64 and rounded once to
FP32; the reference computes them in FP32 with glam, so the two agree within
the reference's own known-answer tolerances (compile.rs:1011-1141), not bit
for bit.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .articulated import BearingSpec as _Hinge
from .articulated import Drive, attach, compile_mechanisms
from .scenes import Scene

GRID_UNIT_METERS = 0.25              # geometry.rs:7
CUBOID_DENSITY_KG_M3 = 500.0         # lib.rs:35
CYLINDER_COLLIDER_COUNT = 16         # compile.rs:15
GROUND = -1                          # a weld's second face on the static ground
MAX_DRIVE_SPEED_RAD_S = 25.0         # drive.rs:6
DEFAULT_MAX_SPEED_RAD_S = 3.0        # drive.rs:534
INVALID = 0xFFFFFFFF


class TopologyError(ValueError):
    """TopologyError (compile.rs:218-246)."""

    def __init__(self, kind: str, message: str = ""):
        super().__init__(message or kind)
        self.kind = kind


# ---- grid poses -------------------------------------------------------------

def _quat_mul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return (aw * bx + ax * bw + ay * bz - az * by,
            aw * by - ax * bz + ay * bw + az * bx,
            aw * bz + ax * by - ay * bx + az * bw,
            aw * bw - ax * bx - ay * by - az * bz)


def _quat_rotate(q, v):
    x, y, z, w = q
    vx, vy, vz = v
    # v + 2 w (u x v) + 2 u x (u x v), u = (x, y, z)
    cx, cy, cz = y * vz - z * vy, z * vx - x * vz, x * vy - y * vx
    ccx, ccy, ccz = y * cz - z * cy, z * cx - x * cz, x * cy - y * cx
    return np.array([vx + 2.0 * (w * cx + ccx), vy + 2.0 * (w * cy + ccy), vz + 2.0 * (w * cz + ccz)])


def _quat_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def _axis_quat(axis: int, angle: float):
    s, c = math.sin(angle * 0.5), math.cos(angle * 0.5)
    v = [0.0, 0.0, 0.0]
    v[axis] = s
    return (v[0], v[1], v[2], c)


@dataclass(frozen=True)
class GridRotation:
    """Quarter turns about local x, y, z in XYZ Euler order (geometry.rs:107-131)."""

    x: int = 0
    y: int = 0
    z: int = 0

    def quaternion(self):
        turns = [t % 4 for t in (self.x, self.y, self.z)]
        q = (0.0, 0.0, 0.0, 1.0)
        for axis, t in enumerate(turns):
            q = _quat_mul(q, _axis_quat(axis, t * math.pi * 0.5))
        return q


@dataclass(frozen=True)
class BuildPose:
    """A centre on the grid in quarter-metre units, optionally offset by half a
    unit per axis, and a grid rotation (geometry.rs:136-190)."""

    units: Tuple[int, int, int] = (0, 0, 0)
    rotation: GridRotation = GridRotation()
    half: Tuple[int, int, int] = (0, 0, 0)

    def translation(self) -> np.ndarray:
        return np.array([(2 * u + h) * (GRID_UNIT_METERS * 0.5) for u, h in zip(self.units, self.half)])


@dataclass(frozen=True)
class Cuboid:
    dims: Tuple[int, int, int]  # grid units, 1..32 each
    pose: BuildPose = BuildPose()

    def __post_init__(self):
        if any(not (1 <= int(d) <= 32) for d in self.dims):
            raise ValueError("grid dimensions must be between 1 and 32 quarter-metre units")

    def size(self) -> np.ndarray:
        return np.array([d * GRID_UNIT_METERS for d in self.dims])


@dataclass(frozen=True)
class Controller:
    """A control block: simulated as its fixed one-unit cube (compile.rs:832-837)."""

    pose: BuildPose = BuildPose()

    def cuboid(self) -> Cuboid:
        return Cuboid((1, 1, 1), self.pose)


@dataclass(frozen=True)
class Cylinder:
    """A solid or hollow cylinder (sector) about its local y axis
    (geometry.rs:232-345): diameters in metres, length a whole number of grid
    units, sweep 15..360 degrees in 15-degree steps centred on local +x."""

    outer_diameter: float
    inner_diameter: float
    axial_length: float
    pose: BuildPose = BuildPose()
    sweep_degrees: int = 360

    def __post_init__(self):
        if not (0.05 <= self.outer_diameter <= 8.0):
            raise ValueError("outer diameter out of range")
        if self.inner_diameter < 0.0 or self.inner_diameter > self.outer_diameter - 0.05:
            raise ValueError("inner diameter out of range")
        units = self.axial_length / GRID_UNIT_METERS
        if abs(units - round(units)) > 1e-5 or not (1 <= round(units) <= 32):
            raise ValueError("axial length must be 1..32 grid units")
        if self.sweep_degrees % 15 or not (15 <= self.sweep_degrees <= 360):
            raise ValueError("sweep must be 15..360 degrees in 15-degree steps")

    @property
    def sweep(self) -> float:
        return math.radians(self.sweep_degrees)


def _physical(spec):
    return spec.cuboid() if isinstance(spec, Controller) else spec


# ---- drives -----------------------------------------------------------------

@dataclass
class DriveLimits:
    """Speed, torque and travel envelope of a driven bearing (drive.rs:526-655)."""

    max_speed: float = DEFAULT_MAX_SPEED_RAD_S
    max_torque: float = math.inf
    angle_limits: Optional[Tuple[float, float]] = None

    def __post_init__(self):
        if not (0.0 < self.max_speed <= MAX_DRIVE_SPEED_RAD_S):
            raise ValueError("max speed out of range")
        if not self.max_torque > 0.0:
            raise ValueError("max torque must be positive")
        if self.angle_limits is not None:
            lo, hi = self.angle_limits
            if lo >= hi:
                raise ValueError("inverted angle limits")


@dataclass
class DriveWire:
    """A control block's wire to one bearing (graph.rs:130-165): the first
    program state's target ('speed' or 'angle', value), its limits, and
    whether the wire runs opposite the programmed direction."""

    controller: int
    bearing: int
    target: Tuple[str, float] = ("speed", 0.0)
    limits: DriveLimits = field(default_factory=DriveLimits)
    reversed: bool = False

    def resolved_target(self) -> Tuple[str, float]:
        kind, value = self.target
        return (kind, -value) if self.reversed else (kind, value)


def coordinate_drive(target: Tuple[str, float], limits: DriveLimits, axis_inertia: float) -> Drive:
    """coordinate_drive (compile.rs:707-737)."""
    torque = limits.max_torque
    max_acceleration = math.inf if math.isinf(torque) else torque / axis_inertia
    kind, value = target
    mode, speed, angle = (1, value, 0.0) if kind == "speed" else (2, 0.0, value)
    lo, hi = limits.angle_limits if limits.angle_limits is not None else (-math.inf, math.inf)
    return Drive(mode=mode, max_acceleration=max_acceleration, max_speed=limits.max_speed, target_speed=speed,
                 target_angle=angle, min_angle=lo, max_angle=hi)


# ---- compiled creation -------------------------------------------------------

@dataclass
class MassProperties:
    mass: float
    inverse_mass: float
    center_of_mass: np.ndarray
    inertia: np.ndarray
    inverse_inertia: np.ndarray


@dataclass
class Compound:
    source_parts: List[int]
    root_translation: np.ndarray
    is_static: bool
    mass_properties: MassProperties
    collider_range: Tuple[int, int]


@dataclass
class LocalCollider:
    source_part: int
    compound_index: int
    local_center: np.ndarray
    local_rotation: Tuple[float, float, float, float]
    half_extents: np.ndarray


@dataclass
class CompiledBearing:
    source_bearing: int
    compound_a: int
    compound_b: int
    local_anchor_a: np.ndarray
    local_anchor_b: np.ndarray
    axis: np.ndarray
    coordinate: Optional[int]


@dataclass
class Compiled:
    compounds: List[Compound]
    colliders: List[LocalCollider]
    bearings: List[CompiledBearing]
    suppression: List[Tuple[int, int]]
    part_to_compound: Dict[int, int]
    tree_bearings: List[int]
    closure_bearings: List[int]
    bearing_coordinates: Dict[int, int]
    coordinate_axis_inertia: List[float]
    coordinate_drives: List[Drive]
    scene: Scene
    mechanisms: object = None  # articulated.Mechanisms: parents, roots, traversal (None without bearings)


class _DisjointSet:
    """Union by rank with path compression, in the reference's tie rules
    (compile.rs:936-975): the first argument's root wins equal ranks."""

    def __init__(self, n):
        self.parent = list(range(n))
        self.rank = [0] * n

    def find(self, item):
        root = item
        while self.parent[root] != root:
            root = self.parent[root]
        while self.parent[item] != item:
            nxt = self.parent[item]
            self.parent[item] = root
            item = nxt
        return root

    def union(self, a, b):
        ra, rb = self.find(a), self.find(b)
        if ra == rb:
            return
        if self.rank[ra] < self.rank[rb]:
            ra, rb = rb, ra
        self.parent[rb] = ra
        if self.rank[ra] == self.rank[rb]:
            self.rank[ra] += 1


def part_mass_properties(spec) -> Tuple[float, np.ndarray, np.ndarray]:
    """(mass, local centre, local principal inertia) of one part (compile.rs:839-879)."""
    spec = _physical(spec)
    if isinstance(spec, Cuboid):
        sx, sy, sz = spec.size()
        m = CUBOID_DENSITY_KG_M3 * sx * sy * sz
        return m, np.zeros(3), np.array([m * (sy * sy + sz * sz) / 12.0, m * (sx * sx + sz * sz) / 12.0,
                                         m * (sx * sx + sy * sy) / 12.0])
    outer, inner = spec.outer_diameter * 0.5, spec.inner_diameter * 0.5
    length, sweep = spec.axial_length, spec.sweep
    r2 = outer * outer + inner * inner
    m = CUBOID_DENSITY_KG_M3 * sweep * (outer * outer - inner * inner) * length * 0.5
    cx = 4.0 * math.sin(sweep * 0.5) * (outer ** 3 - inner ** 3) / (3.0 * sweep * (outer * outer - inner * inner))
    parallel = r2 * (sweep + math.sin(sweep)) / (4.0 * sweep)
    perpendicular = r2 * (sweep - math.sin(sweep)) / (4.0 * sweep)
    axial = length * length / 12.0
    return m, np.array([cx, 0.0, 0.0]), np.array([m * (axial + perpendicular),
                                                   m * (parallel + perpendicular - cx * cx),
                                                   m * (parallel + axial - cx * cx)])


def mass_properties(specs: Sequence, is_static: bool, part: int) -> MassProperties:
    """calculate_mass_properties (compile.rs:768-830)."""
    total = 0.0
    weighted = np.zeros(3)
    props = []
    for spec in specs:
        m, lc, li = part_mass_properties(spec)
        q = spec.pose.rotation.quaternion()
        center = spec.pose.translation() + _quat_rotate(q, lc)
        props.append((m, li, q, center))
        total += m
        weighted += center * m
    com = weighted / total
    inertia = np.zeros((3, 3))
    for m, li, q, center in props:
        r = _quat_matrix(q)
        offset = center - com
        inertia += r @ np.diag(li) @ r.T + m * (np.eye(3) * float(offset @ offset) - np.outer(offset, offset))
    det = float(np.linalg.det(inertia))
    if not (math.isfinite(total) and total > 0.0 and np.all(np.isfinite(inertia)) and math.isfinite(det)
            and det > np.finfo(np.float32).eps):
        raise TopologyError("InvalidMassProperties", f"compound containing part {part} has invalid mass properties")
    return MassProperties(total, 0.0 if is_static else 1.0 / total, com, inertia,
                          np.zeros((3, 3)) if is_static else np.linalg.inv(inertia))


def part_colliders(part: int, compound: int, spec, com: np.ndarray) -> List[LocalCollider]:
    """append_part_colliders (compile.rs:881-929)."""
    spec = _physical(spec)
    q = spec.pose.rotation.quaternion()
    if isinstance(spec, Cuboid):
        return [LocalCollider(part, compound, spec.pose.translation() - com, q, spec.size() * 0.5)]
    outer, inner = spec.outer_diameter * 0.5, spec.inner_diameter * 0.5
    half_radial = (outer - inner) * 0.5
    center_radius = (outer + inner) * 0.5
    segment = spec.sweep / 16.0
    half_tangent = outer * math.tan(segment * 0.5)
    start = -segment * 0.5 if spec.sweep_degrees == 360 else -spec.sweep * 0.5
    out = []
    for k in range(CYLINDER_COLLIDER_COUNT):
        angle = start + segment * (k + 0.5)
        radial = np.array([math.cos(angle), 0.0, math.sin(angle)])
        out.append(LocalCollider(part, compound, spec.pose.translation() - com + _quat_rotate(q, radial * center_radius),
                                 _quat_mul(q, _axis_quat(1, -angle)),
                                 np.array([half_radial, spec.axial_length * 0.5, half_tangent])))
    return out


def full_cylinder_ground_data(colliders: Sequence[LocalCollider]) -> List[Tuple[float, float, int]]:
    """full_cylinder_ground_data (device.rs:3015-3045): (segment-centre radius,
    outer radius, role) per collider; roles 1..16 mark the segments of a full
    cylinder (their radial directions sum to zero), 0 everything else."""
    out = [(0.0, 0.0, 0)] * len(colliders)
    start = 0
    while start < len(colliders):
        end = start + 1
        while end < len(colliders) and colliders[end].source_part == colliders[start].source_part:
            end += 1
        group = colliders[start:end]
        if len(group) == CYLINDER_COLLIDER_COUNT:
            radial = sum(_quat_rotate(c.local_rotation, (1.0, 0.0, 0.0)) for c in group)
            if float(radial @ radial) < 1.0e-8:
                center = sum(c.local_center for c in group) * (1.0 / 16.0)
                center_radius = float(np.linalg.norm(group[0].local_center - center))
                outer_radius = center_radius + float(group[0].half_extents[0])
                for k in range(start, end):
                    out[k] = (center_radius, outer_radius, k - start + 1)
        start = end
    return out


class ConstructionGraph:
    """Editable parts, welds, rigid links, bearings and drive wires, compiled to
    upload rows. Part and bearing ids are their insertion indices."""

    def __init__(self):
        self.parts: List[object] = []
        self.welds: List[Tuple[int, int]] = []
        self.rigid_links: List[Tuple[int, int]] = []
        self.bearings: List[Tuple[int, int, np.ndarray, np.ndarray]] = []
        self.wires: List[DriveWire] = []

    def spawn(self, spec) -> int:
        self.parts.append(spec)
        return len(self.parts) - 1

    def weld(self, first: int, second: int = GROUND) -> None:
        self.welds.append((first, second))

    def rigid_link(self, first: int, second: int) -> None:
        self.rigid_links.append((first, second))

    def add_bearing(self, source: int, target: int, anchor, axis) -> int:
        if source == GROUND or target == GROUND:
            raise ValueError("bearings join two parts")
        self.bearings.append((source, target, np.asarray(anchor, np.float32), np.asarray(axis, np.float32)))
        return len(self.bearings) - 1

    def wire(self, wire: DriveWire) -> None:
        self.wires.append(wire)

    def bearing_wire(self, bearing: int) -> Optional[DriveWire]:
        """bearing_drive_link (graph.rs:421-426): the first wire of the bearing."""
        return next((w for w in self.wires if w.bearing == bearing), None)

    def compile(self) -> Compiled:
        return compile_graph(self)


def compile_graph(graph: ConstructionGraph) -> Compiled:
    """compile_graph (compile.rs:256-502) plus the GPU lowering of
    new_with_config (device.rs:424-523)."""
    n_parts = len(graph.parts)
    if n_parts == 0:
        raise TopologyError("EmptyConstruction", "construction contains no parts")
    groups = _DisjointSet(n_parts)
    grounded = [False] * n_parts
    for a, b in graph.welds:
        if a != GROUND and b != GROUND:
            groups.union(a, b)
        elif a != GROUND or b != GROUND:
            grounded[a if a != GROUND else b] = True
    for a, b in graph.rigid_links:
        groups.union(a, b)
    members: Dict[int, List[int]] = {}
    for p in range(n_parts):
        members.setdefault(groups.find(p), []).append(p)

    compounds: List[Compound] = []
    colliders: List[LocalCollider] = []
    part_to_compound: Dict[int, int] = {}
    for root in sorted(members):
        rows = members[root]
        index = len(compounds)
        is_static = any(grounded[p] for p in rows)
        props = mass_properties([graph.parts[p] for p in rows], is_static, rows[0])
        first = len(colliders)
        for p in rows:
            part_to_compound[p] = index
            colliders.extend(part_colliders(p, index, graph.parts[p], props.center_of_mass))
        compounds.append(Compound(rows, props.center_of_mass, is_static, props, (first, len(colliders))))

    # Bearings: collapse rows describing one physical joint; a driven row
    # represents its group (compile.rs:358-399).
    driven = {w.bearing for w in graph.wires}
    physical: List[List[int]] = []  # [representative bearing, compound a, compound b]
    by_key: Dict[tuple, int] = {}
    keys: List[Tuple[int, tuple]] = []
    suppressed = set()
    for bid, (pa, pb, anchor, axis) in enumerate(graph.bearings):
        ca, cb = part_to_compound[pa], part_to_compound[pb]
        if ca == cb:
            raise TopologyError("SelfBearing", f"bearing {bid} connects compound {ca} to itself after weld compilation")
        key = (ca, cb, anchor.astype(np.float32).tobytes(), axis.astype(np.float32).tobytes())
        suppressed.add((min(ca, cb), max(ca, cb)))
        keys.append((bid, key))
        if key not in by_key:
            by_key[key] = len(physical)
            physical.append([bid, ca, cb])
        else:
            row = physical[by_key[key]]
            if bid in driven and row[0] not in driven:
                row[0] = bid

    positions = np.array([c.root_translation for c in compounds], np.float64).astype(np.float32)
    static = np.array([c.is_static for c in compounds])
    # The forest takes driven bearings first (compile.rs:402-432): a driven
    # marker drive routes them through compile_mechanisms' first pass.
    marker = Drive(mode=1)
    hinges = [_Hinge(ca, cb, graph.bearings[bid][2], graph.bearings[bid][3], marker if bid in driven else None)
              for bid, ca, cb in physical]
    try:
        mech = compile_mechanisms(positions, static, hinges) if hinges else None
    except ValueError as e:  # a driven row left on a loop-closure edge
        row = int(str(e).split()[2])
        raise TopologyError("DrivenClosureBearing",
                            f"driven bearing {physical[row][0]} closes a loop and cannot carry a drive") from e
    tree = set(mech.tree_rows) if mech else set()

    compiled_bearings: List[CompiledBearing] = []
    tree_bearings, closure_bearings = [], []
    coordinate_of: Dict[int, int] = {}
    for row, (bid, ca, cb) in enumerate(physical):
        anchor, axis = graph.bearings[bid][2].astype(np.float64), graph.bearings[bid][3].astype(np.float64)
        coord = None
        if row in tree:
            coord = len(tree_bearings)
            tree_bearings.append(bid)
            coordinate_of[bid] = coord
        else:
            closure_bearings.append(bid)
        compiled_bearings.append(CompiledBearing(bid, ca, cb, anchor - compounds[ca].root_translation,
                                                 anchor - compounds[cb].root_translation, axis, coord))
    bearing_coordinates = {}
    for bid, key in keys:
        rep = physical[by_key[key]][0]
        if rep in coordinate_of:
            bearing_coordinates[bid] = coordinate_of[rep]

    axis_inertia = coordinate_axis_inertia(compounds, compiled_bearings, mech)
    drives = []
    for coord, bid in enumerate(tree_bearings):
        w = graph.bearing_wire(bid)
        inertia = axis_inertia[coord]
        if w is None or not math.isfinite(inertia):
            drives.append(Drive())
        else:
            drives.append(coordinate_drive(w.resolved_target(), w.limits, inertia))

    scene = _lower(compounds, colliders, sorted(suppressed))
    if mech is not None:
        # compile_mechanisms numbers coordinates in tree-row order, as here.
        rows = mech.drives
        for q, d in enumerate(drives):
            rows[q] = d.row()
        scene = attach(scene, mech)
    return Compiled(compounds, colliders, compiled_bearings, sorted(suppressed), part_to_compound, tree_bearings,
                    closure_bearings, bearing_coordinates, axis_inertia, drives, scene, mech)


def coordinate_axis_inertia(compounds: List[Compound], bearings: List[CompiledBearing], mech) -> List[float]:
    """compile_coordinate_axis_inertia (compile.rs:613-672): per tree coordinate,
    the child subtree's inertia about the bearing axis through its anchor;
    infinite when the subtree holds a static compound."""
    if mech is None:
        return []
    n = len(compounds)
    children: List[List[int]] = [[] for _ in range(n)]
    root_flags = np.asarray(mech.root_flags)
    child_of_row: Dict[int, int] = {}
    for body in range(n):
        if not root_flags[body]:
            children[int(mech.parents[body])].append(body)
    rows = mech.mechanism_bodies
    for body in range(n):
        if not root_flags[body]:
            child_of_row[int(rows[body].metadata[1])] = body
    out: List[float] = []
    for row, b in enumerate(bearings):
        if b.coordinate is None:
            continue
        child = child_of_row.get(row)
        norm = float(np.linalg.norm(b.axis))
        if child is None or norm == 0.0:
            out.append(math.inf)
            continue
        axis = b.axis / norm
        anchor = compounds[b.compound_a].root_translation + b.local_anchor_a
        total, stack, grounded = 0.0, [child], False
        while stack:
            body = stack.pop()
            c = compounds[body]
            if c.is_static:
                grounded = True
                break
            p = c.mass_properties
            offset = p.center_of_mass - anchor
            radial = offset - axis * float(offset @ axis)
            total += float(axis @ (p.inertia @ axis)) + p.mass * float(radial @ radial)
            stack.extend(children[body])
        out.append(total if not grounded and math.isfinite(total) and total > 0.0 else math.inf)
    return out


def _lower(compounds: List[Compound], colliders: List[LocalCollider], suppressed) -> Scene:
    """Compound and collider rows as new_with_config lowers them (device.rs:424-499)."""
    n, c = len(compounds), len(colliders)
    poses = (_lib.Transform * n)()
    inv = (ctypes.c_float * n)()
    masses = (_lib.Mass * n)()
    inertias = (_lib.SpatialInertia * n)()
    rows = (_lib.Collider * c)()
    for i, comp in enumerate(compounds):
        p = comp.mass_properties
        poses[i].position[:] = [float(v) for v in comp.root_translation] + [0.0]
        poses[i].rotation[:] = [0.0, 0.0, 0.0, 1.0]
        inv[i] = p.inverse_mass
        masses[i].inverse_mass[:] = [p.inverse_mass, 0.0, 0.0, 0.0]
        inertias[i].mass[:] = [p.mass, 0.0, 0.0, 0.0]
        for k, (ic, mc) in enumerate(zip((inertias[i].inertia_x, inertias[i].inertia_y, inertias[i].inertia_z),
                                         (masses[i].inverse_inertia_x, masses[i].inverse_inertia_y,
                                          masses[i].inverse_inertia_z))):
            ic[:] = [float(p.inertia[0, k]), float(p.inertia[1, k]), float(p.inertia[2, k]), 0.0]
            mc[:] = [float(p.inverse_inertia[0, k]), float(p.inverse_inertia[1, k]), float(p.inverse_inertia[2, k]),
                     0.0]
    ground = full_cylinder_ground_data(colliders)
    for k, (col, (center_radius, outer_radius, role)) in enumerate(zip(colliders, ground)):
        r = rows[k]
        r.local_center[:] = [float(v) for v in col.local_center] + [center_radius]
        r.local_rotation[:] = [float(v) for v in col.local_rotation]
        r.half_extents[:] = [float(v) for v in col.half_extents] + [outer_radius]
        r.metadata[:] = [col.compound_index, col.source_part, 0, role]
    return Scene(n, poses, inv, masses, inertias, rows, suppressed_pairs=list(suppressed))
