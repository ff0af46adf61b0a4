"""Writes tests/golden/batch_48_digest.npz: the FP64 oracle's RL-BA result
for three stacks of the 48-stack uneven batch (cudadecon_amd.synth
batch_shapes / batch_stack), one per channel PSF, the largest shape drawn
among them, sampled.

Stacks (index: shape, channel sigma_xy / sigma_z):
  11: 256 x 1024 x 768, channel 2 (1.8 / 5.0 px)  - the largest shape drawn
   0: 256 x 384 x 256,  channel 0 (1.2 / 3.0 px)
  13: 64 x 384 x 384,   channel 1 (1.5 / 4.0 px)
Each: R = max(raw - 100, 0), OTF made radialft-style from the channel's
33^3 Gaussian PSF, k-scales as RL_interface_init computes them for
dr 0.104 um, dz 0.2 um, 10 RL-BA iterations.  The digest holds, per stack,
20,000 seeded voxel positions of X with their FP64 values, the position of
max |X|, max |X|, sum X, the acceleration factors and the raw stack's sum.
Takes about 15 minutes on 8 cores (most of it the 201 M-voxel stack).

    python tests/golden/make_batch_digest.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

STACKS = (11, 0, 13)
N_SAMPLES = 20000


def main():
    import oracle as orc
    from cudadecon_amd import synth
    from cudadecon_amd.otf import make_otf
    shapes = synth.batch_shapes()
    out = {}
    for i in STACKS:
        shape = shapes[i]
        c, (sxy, sz) = synth.batch_channel(i)
        t0 = time.time()
        raw = synth.batch_stack(i, shape)
        R = np.maximum(raw.astype(np.float64) - 100.0, 0.0)
        tab = make_otf(synth.gaussian_psf(sxy, sz))
        nz, ny, nx = shape
        g = orc.init_geometry(nx, ny, nz, synth.BATCH_DR, synth.BATCH_DZ, synth.BATCH_DR, synth.BATCH_DZ, 0.0, 0.0,
                              0, False, (2 * tab.shape[1], tab.shape[0], 1))
        H = orc.otf_interp(tab, g["otf_dims"], g["fft_dims"], g["kscale"])
        X, lam = orc.rl_ba(R, H, synth.BATCH_ITERS)
        del H, R
        rng = np.random.default_rng(500 + i)
        idx = rng.choice(X.size, N_SAMPLES, replace=False)
        top = int(np.argmax(np.abs(X)))
        idx = np.concatenate([[top], idx])
        out[f"s{i}_shape"] = np.array(shape)
        out[f"s{i}_channel"] = c
        out[f"s{i}_idx"] = idx
        out[f"s{i}_val"] = X.reshape(-1)[idx]
        out[f"s{i}_xmax"] = np.abs(X).max()
        out[f"s{i}_xsum"] = X.sum()
        out[f"s{i}_lam"] = lam
        out[f"s{i}_raw_sum"] = float(raw.astype(np.float64).sum())
        print(f"stack {i} {shape} channel {c}: {time.time() - t0:.0f} s, lam {lam[2:5]}", flush=True)
        del X, raw
    out["stacks"] = np.array(STACKS)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "batch_48_digest.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path))


if __name__ == "__main__":
    main()
