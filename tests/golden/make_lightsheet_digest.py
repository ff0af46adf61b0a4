"""Writes tests/golden/lightsheet_128x512x512_digest.npz: the FP64 oracle's RL_interface result
for the full light-sheet workload (128 x 512 x 512), sampled.

raw 128 x 512 x 512 uint16 (5,000 beads, PSF sigma 1.3 / 4.0 px), deskew
31.5 deg with dz 0.3 um and dx 0.104 um (128 x 512 x 810), 15 RL-BA
iterations, rotation to coverslip coordinates (192 x 512 x 791).  The digest
holds 40,000 seeded voxel positions of the result with their FP64 values,
max |result| and the acceleration factors.  Takes a few minutes on 8 cores.

    python tests/golden/make_lightsheet_digest.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

N_SAMPLES = 40000


def main():
    import oracle as orc
    from cudadecon_amd.otf import make_otf
    from cudadecon_amd.synth import make_stack
    raw, psf, c = make_stack("lightsheet_128x512x512")
    tab = make_otf(psf)
    nz, ny, nx = raw.shape
    g = orc.init_geometry(nx, ny, nz, c["dr"], c["dz"], c["dr"], c["dz"], c["angle"], c["rot"], 0, False,
                          (2 * tab.shape[1], tab.shape[0], 1))
    t0 = time.time()
    X, _, lam = orc.rl_interface(orc.Session(g, tab), raw, 100.0, False, False, c["iters"], 0)
    print("oracle", X.shape, f"{time.time() - t0:.1f} s")
    rng = np.random.default_rng(123)
    idx = rng.choice(X.size, N_SAMPLES, replace=False)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lightsheet_128x512x512_digest.npz")
    np.savez_compressed(out, shape=np.array(X.shape), idx=idx, val=X.reshape(-1)[idx], xmax=np.abs(X).max(),
                        xsum=X.sum(), lam=lam, raw_sum=float(raw.astype(np.float64).sum()))
    print("wrote", out, os.path.getsize(out))


if __name__ == "__main__":
    main()
