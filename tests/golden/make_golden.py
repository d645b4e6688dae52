"""Writes tests/golden/spect_small.npz: seeded small inputs and the FP64
oracle's outputs for them (projection and back-projection with both rotation
methods, 3 ML-EM and 2x2 OS-EM iterations).  The reference ships no golden
vectors (its tests use unseeded rand); these fixtures pin the oracle against
regressions and are the GPU tests' references.

    python tests/golden/make_golden.py
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def inputs():
    rng = np.random.default_rng(20240601)
    nx, nz, nview, px, pz = 16, 8, 8, 5, 3
    mumap = (0.1 * rng.random((nx, nx, nz))).astype(np.float32)
    p = rng.random((px, pz, nx, nview)).astype(np.float32)
    p = p + p[::-1, ::-1]
    p = p / p.sum(axis=(0, 1), keepdims=True)
    psfs = (0.5 * (p + p[::-1, ::-1])).astype(np.float32)
    image = rng.random((nx, nx, nz)).astype(np.float32)
    views = rng.random((nx, nz, nview)).astype(np.float32)
    twopi = float(np.float32(2 * math.pi))
    angles = np.array([np.float32(k * twopi / nview) for k in range(nview)], np.float32)
    return dict(mumap=mumap, psfs=psfs, image=image, views=views, angles=angles, dy=np.float32(0.48))


def compute(O):
    d = inputs()
    out = dict(d)
    ang = d["angles"].astype(np.float64)
    for m in ("one", "two"):
        P = O.Problem(d["mumap"], d["psfs"], float(d["dy"]), ang, m, np.float64, "direct")
        out[f"proj_{m}"] = P.project(d["image"].astype(np.float64))
        out[f"back_{m}"] = P.backproject(d["views"].astype(np.float64))
    P = O.Problem(d["mumap"], d["psfs"], float(d["dy"]), ang, "two", np.float64, "direct")
    y = out["proj_two"] + 0.1
    b = np.full_like(y, 0.1)
    x0 = np.ones(d["image"].shape)
    out["em_y"], out["em_b"] = y, b
    out["mlem3"] = P.mlem(x0, y, b, 3)
    out["osem_2x2"] = P.osem(x0, y, b, 2, 2)
    return out


if __name__ == "__main__":
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
    import oracle as O

    res = compute(O)
    np.savez_compressed(os.path.join(HERE, "spect_small.npz"), **{k: np.asarray(v) for k, v in res.items()})
    print("wrote", os.path.join(HERE, "spect_small.npz"), sorted(res))
