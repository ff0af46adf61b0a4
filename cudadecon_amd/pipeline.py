"""The large-stack pipeline on one GPU with every buffer resident in HBM:
camera offset subtraction, learned warm start, RL-BA iterations, deskew and
an affine resampling (the 512 x 2048 x 2048 light-sheet workload), with the
time of each stage."""
from __future__ import annotations

import time

import numpy as np


def small_affine(shape_zyx, seed=4, max_deg=5.0, scale=(0.95, 1.05)):
    """Row-major 4x4 pull matrix (output voxel (x, y, z, 1) -> input voxel) of
    a rotation by at most max_deg about a random axis and an isotropic scale
    s, both about the volume centre: input = R^T (out - c) / s + c."""
    rng = np.random.default_rng(seed)
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    ang = np.deg2rad(rng.uniform(0.0, max_deg))
    s = rng.uniform(*scale)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    R = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * (K @ K)
    A = R.T / s
    nz, ny, nx = shape_zyx
    c = np.array([(nx - 1) / 2.0, (ny - 1) / 2.0, (nz - 1) / 2.0])
    M = np.eye(4)
    M[:3, :3] = A
    M[:3, 3] = c - A @ c
    return M.astype(np.float32), dict(axis=axis.round(4).tolist(), degrees=round(float(np.rad2deg(ang)), 3),
                                       scale=round(float(s), 4))


def run(eng, raw_u16, net=None, iters=10, offset=100.0, dz=0.4, dr=0.104, angle=31.5, seed=4):
    """Run the pipeline on an RLEngine of the stack's shape; returns a dict of
    stage times (ms) and the output shape.  The engine keeps the results."""
    t = {}
    t0 = time.perf_counter()
    eng.set_raw_u16(raw_u16, offset)
    t["stage_u16_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    seeded = net is not None
    if seeded:
        t["warm_start_ms"] = round(eng.warm_start(net), 2)
    _, ms = eng.iterate(iters, seeded=seeded)
    t["iterations_ms"] = round(ms, 2)
    t0 = time.perf_counter()
    nx_out = eng.deskew(dz, dr, angle)
    t["deskew_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    M, desc = small_affine(eng.post_shape(), seed)
    t0 = time.perf_counter()
    eng.affine(M)
    t["affine_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    t["output_zyx"] = list(eng.post_shape())
    t["deskewed_nx"] = nx_out
    t["affine"] = desc
    t["device_total_ms"] = round(sum(v for k, v in t.items() if k.endswith("_ms") and k != "stage_u16_ms"), 2)
    return t
