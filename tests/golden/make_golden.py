"""Writes the golden vectors under tests/golden/ from the CPU oracle (the reference cannot be
built or imported here, so the oracle, pinned by the closed-form tests in test_oracle.py, is
the source). Each fixture holds, for steps 1, 2, 10 and 100 of a seeded scene, the node
positions, velocities, angular velocities, forces and torques and the real contact pairs.
The scenes: the C1 deposition (2000 spheres at rest in a box of 6 walls, FrictMat,
IdealElPl), a dense variant with contacts from the start, a 216-sphere periodic Hertz box,
and a drum patch with facets and fakeVel.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tests.scenes import build_oracle, deposition, drum, periodic_hertz  # noqa: E402

STEPS = (1, 2, 10, 100)
ENG = ["Collider", "ContactLoop", "Leapfrog"]


def scenes():
    return {
        # C1 as specified: loose random placement at rest, gravity deposition
        "c1_deposition": (deposition(n=2000, seed=1), ENG),
        # the same material and walls on a dense jittered lattice, so that contacts exist at once
        "c1_dense": (deposition(n=1000, seed=2, dense=True, vel_sigma=0.05), ["Leapfrog", "Collider", "ContactLoop"]),
        "hertz_periodic_216": (periodic_hertz(n=216, seed=21, spacing=0.92), ENG),
        "drum_facets": (drum(n=600, seed=22, R=0.025, length=0.016, div=24), ENG),
    }


def record(o, engines):
    out = {}
    done = 0
    for s in STEPS:
        o.run(engines, s - done)
        done = s
        for f in ("pos", "vel", "angVel", "force", "torque"):
            out[f"{f}_{s}"] = o.get(f)
        ids, _, _ = o.contacts(real_only=True)
        out[f"real_{s}"] = ids.astype(np.int32)
    return out


def main():
    for name, (sc, engines) in scenes().items():
        o = build_oracle(sc)
        data = record(o, engines)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, engines=np.array(engines), **data)
        print(name, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
