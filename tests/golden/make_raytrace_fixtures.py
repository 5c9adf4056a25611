"""Write golden evaluation-point vectors from the reference's own ray tracer.

Runs the amdflang-built reference ``evaluation_points`` (eval_points.F90 with
healpix.F90 and heapsort.F90; oracle/_ref/ref_raytrace, built by
`make -C oracle ref` from the reference sources) on three point sets:
a jittered 6^3 cloud (no distance ties), a regular 4x5x6 grid with spacing
1/16 pc (many ties) and the same grid with one ionised point (etype 2, which
stops a ray).  Output: tests/golden/raytrace_ref.npz with the inputs, the
HEALPix ray vectors, and per set the evaluation-point counts, projected
points and coordinates.

Run in the build container (the reference tree and amdflang are needed):
    python tests/golden/make_raytrace_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

MAXPOINTS = 40


def point_sets():
    rng = np.random.default_rng(20251206)
    g6 = np.stack(np.meshgrid(*[np.arange(6)] * 3, indexing="ij"), -1).reshape(-1, 3)[:, ::-1].astype(float)
    jit = g6 * 0.0625 + rng.uniform(-0.02, 0.02, g6.shape)
    iz, iy, ix = np.meshgrid(np.arange(6), np.arange(5), np.arange(4), indexing="ij")
    grid = np.stack([ix, iy, iz], -1).reshape(-1, 3).astype(float) * 0.0625
    et = np.ones(len(grid), dtype=np.int32)
    et[57] = 2
    return {"jitter": (jit, np.ones(len(jit), dtype=np.int32)),
            "grid": (grid, np.ones(len(grid), dtype=np.int32)),
            "grid_ion": (grid, et)}


def main():
    out = {}
    for name, (xyz, et) in point_sets().items():
        vec, epray, proj, epoint = oracle.ref_eval_points(xyz, level=0, theta_crit=1.3, maxpoints=MAXPOINTS,
                                                          etype=et)
        out[f"{name}_xyz"] = xyz
        out[f"{name}_etype"] = et
        out[f"{name}_epray"] = epray
        out[f"{name}_proj"] = proj
        out[f"{name}_epoint"] = epoint
        out["vectors"] = vec
    np.savez_compressed(os.path.join(HERE, "raytrace_ref.npz"), **out)


if __name__ == "__main__":
    main()
