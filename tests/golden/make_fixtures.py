"""Writes tests/golden/*.json.

reference_kats.json restates, as data, the input/expected-output vectors that
the reference's own CPU SAT tests hold (crates/mechanic-gpu/src/collision.rs:
199-232). oracle_pile_checksums.json pins the oracle's FP32 trajectory of a
small seeded pile (SHA-256 of the body state at fixed ticks), so a change to the
oracle shows up as a fixture diff.

    python tests/golden/make_fixtures.py
"""
import hashlib
import json
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def reference_kats():
    q45 = [0.0, math.sin(math.pi / 8), 0.0, math.cos(math.pi / 8)]
    ident = [0.0, 0.0, 0.0, 1.0]
    half = [0.5, 0.5, 0.5]
    return {
        "source": "crates/mechanic-gpu/src/collision.rs:199-232",
        "cases": [
            {"name": "sat_rejects_separated_boxes", "a": [[0, 0, 0], ident, half], "b": [[1.01, 0, 0], ident, half],
             "expect": {"separated": True}},
            {"name": "sat_finds_minimum_face_axis", "a": [[0, 0, 0], ident, half], "b": [[0.75, 0, 0], ident, half],
             "expect": {"separated": False, "normal": [1, 0, 0], "penetration": 0.25, "tolerance": 1e-6}},
            {"name": "face_contact_retains_four_stable_points", "a": [[0, 0, 0], ident, half],
             "b": [[1.0, 0, 0], ident, half], "expect": {"separated": False, "point_x": 0.5, "tolerance": 1e-6}},
            {"name": "rotated_edge_overlap_is_detected", "a": [[0, 0, 0], ident, half], "b": [[0.8, 0, 0], q45, half],
             "expect": {"separated": False}},
        ],
    }


def pile_checksums():
    import numpy as np

    import mechanic_amd as mh
    from tests._oracle import Oracle

    scene = mh.cuboid_pile(6, 5, 6, seed=11)
    ora = Oracle(scene, friction=0.6)
    out = {}
    for tick in range(1, 121):
        ora.step(tick)
        if tick in (1, 10, 60, 120):
            p, v = ora.bodies()
            h = hashlib.sha256(np.ascontiguousarray(p).tobytes() + np.ascontiguousarray(v).tobytes()).hexdigest()
            out[str(tick)] = {"sha256": h, "diag": [int(x) for x in ora.diag()]}
    return {"scene": "cuboid_pile(6, 5, 6, seed=11), friction 0.6, 8 iterations", "ticks": out}


if __name__ == "__main__":
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(reference_kats(), f, indent=1)
    with open(os.path.join(HERE, "oracle_pile_checksums.json"), "w") as f:
        json.dump(pile_checksums(), f, indent=1)
    print("written")
