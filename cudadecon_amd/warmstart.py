"""Learned warm start for RL-Biggs-Andrews: training pairs, training on the
GPU, and the comparison with plain RL.

The network (``cudadecon_amd.WarmNet``, kernels in csrc/net_kernels.hip)
maps a preprocessed stack to an estimate of its 50-iteration RL-BA result;
``RLEngine.warm_start`` writes that estimate, clamped at 0, into Y_0 of the
iteration (the FlatStart hook, src/RL-Biggs-Andrews.cpp:305-316).

Training pairs are made by plain RL itself: light-sheet stacks made as the
128 x 512 x 512 workload (PSF sigma_xy 1.3 / sigma_z 4.0 px, bead density
5000 per 128*512*512 voxels, smooth background, Poisson noise, camera offset
100) are generated at 128 x 1024 x 1024, deconvolved for 50 iterations on
the GPU, and both the background-subtracted stack and the result are cut
into 64 x 128 x 128 patches.  Deconvolving whole stacks (not patches) keeps
the circular-convolution wrap of the FFT out of the targets, as it is for
the large stacks the network is used on.

    python -m cudadecon_amd.warmstart train --stacks 8 --out weights.npz
    python -m cudadecon_amd.warmstart eval --weights weights.npz
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import RLEngine, WarmNet, rl_cleanup, rl_init, session_kscale
from .otf import make_otf, write_otf_tiff
from .synth import make_stack

BACKGROUND = 100.0
TARGET_ITERS = 50
TRAIN_STACK = (128, 1024, 1024)
PATCH = (64, 128, 128)
EVAL_STACK = (256, 1024, 1024)
TRAIN_SEED0 = 100000
VALID_SEED = 200000
EVAL_SEED = 3
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "warmstart_lightsheet.npz")


def log(*a):
    print(*a, flush=True)


def kscale_for(shape, tab, dr=0.104, dz=0.3):
    """OTF k-scales of a (nz, ny, nx) grid as RL_interface_init computes them
    (data and PSF sampled alike)."""
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "otf.tif")
        write_otf_tiff(p, tab)
        rl_init(shape, p, dzdata=dz, dxdata=dr, dzpsf=dz, dxpsf=dr)
        ks = session_kscale()
        rl_cleanup()
    return ks


def light_sheet_stack(shape, seed):
    """(background-subtracted float32 stack, PSF) made as the 128x512x512
    workload at another size and seed."""
    raw, psf, _ = make_stack("lightsheet_128x512x512", shape=shape, seed=seed)
    return np.maximum(raw.astype(np.float32) - np.float32(BACKGROUND), np.float32(0.0)), psf


def cut(vol, patch=PATCH):
    nz, ny, nx = vol.shape
    pz, py, px = patch
    v = vol.reshape(nz // pz, pz, ny // py, py, nx // px, px)
    return np.ascontiguousarray(v.transpose(0, 2, 4, 1, 3, 5).reshape(-1, pz, py, px))


def rl_result(eng, R, iters=TARGET_ITERS):
    eng.set_raw(R)
    _, ms = eng.iterate(iters)
    return eng.get(), ms


def build_training_set(net, n_stacks, seed0=TRAIN_SEED0, shape=TRAIN_STACK, target_iters=TARGET_ITERS):
    """Generate n_stacks stacks (CPU, the next one while the GPU works on the
    current one), deconvolve each for target_iters iterations and add the
    patch pairs to the network's device training set.  Returns the count."""
    _, psf = light_sheet_stack((8, 8, 8), 0)
    tab = make_otf(psf)
    eng = RLEngine(shape, tab, kscale_for(shape, tab))
    count = 0
    with ThreadPoolExecutor(1) as ex:
        nxt = ex.submit(light_sheet_stack, shape, seed0)
        for s in range(n_stacks):
            t0 = time.perf_counter()
            R, _ = nxt.result()
            if s + 1 < n_stacks:
                nxt = ex.submit(light_sheet_stack, shape, seed0 + s + 1)
            X, ms = rl_result(eng, R, target_iters)
            net.add_samples(cut(R), cut(X))
            count += cut(R).shape[0]
            log(f"stack {s + 1}/{n_stacks}: RL-{target_iters} {ms:.0f} ms on the GPU, {count} patches, "
                f"{time.perf_counter() - t0:.1f} s")
    eng.close()
    return count


def patch_scores(net, R, X):
    """Relative L2 to the RL target of the raw stack and of the network's
    seed, per patch (the plain RL start is the raw stack)."""
    S = net.forward(R, seed=True)
    n = lambda a: np.sqrt((a.reshape(a.shape[0], -1).astype(np.float64) ** 2).sum(axis=1))
    return n(R - X) / n(X), n(S - X) / n(X)


def train(args):
    net = WarmNet(seed=args.seed)
    t0 = time.perf_counter()
    count = build_training_set(net, args.stacks, target_iters=args.target_iters)
    log(f"training set: {count} patches of {PATCH}, built in {time.perf_counter() - t0:.1f} s")
    # validation pairs from an unseen stack
    _, psf = light_sheet_stack((8, 8, 8), 0)
    tab = make_otf(psf)
    eng = RLEngine(TRAIN_STACK, tab, kscale_for(TRAIN_STACK, tab))
    Rv, _ = light_sheet_stack(TRAIN_STACK, VALID_SEED)
    Xv, _ = rl_result(eng, Rv, args.target_iters)
    eng.close()
    Rv, Xv = cut(Rv)[:16], cut(Xv)[:16]
    raw0, _ = patch_scores(net, Rv, Xv)
    log(f"validation: raw stack vs RL-{args.target_iters}: rel L2 {raw0.mean():.4f}")
    history = []
    total_ms = 0.0
    for lr, steps in zip(args.lr, args.steps):
        done = 0
        while done < steps:
            k = min(args.chunk, steps - done)
            losses, ms = net.train(k, args.batch, lr, seed=args.seed * 7919 + len(history))
            total_ms += ms
            done += k
            history.append((lr, float(losses.mean())))
            _, sc = patch_scores(net, Rv, Xv)
            log(f"lr {lr:g} step {done}/{steps}: loss {losses.mean():.5f} ({ms / k:.2f} ms/step), "
                f"validation seed rel L2 {sc.mean():.4f}")
    p = net.get_params()
    _, sc = patch_scores(net, Rv, Xv)
    meta = dict(arch="unet3-16-32-64", patch=PATCH, patches=count, stacks=args.stacks, batch=args.batch,
                lr=args.lr, steps=args.steps, train_ms=round(total_ms, 1), final_loss=history[-1][1],
                validation_raw_rel=float(raw0.mean()), validation_seed_rel=float(sc.mean()),
                target_iters=args.target_iters)
    np.savez(args.out, params=p, meta=np.array(json.dumps(meta)))
    log("saved", args.out, json.dumps(meta))
    return net


def load_params(path=WEIGHTS):
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32), json.loads(str(z["meta"]))


def first_k(curve, r):
    """Smallest k with curve[k] <= r (None if never)."""
    for k, v in enumerate(curve):
        if v <= r:
            return k
    return None


def evaluate(net, R, tab, ks, iters=TARGET_ITERS, ks_report=(0, 2, 5, 10), warm_iters=None):
    """Residual curves (relative L2 to the plain RL-`iters` result) of plain
    RL (k iterations from the raw stack) and of warm start + k iterations,
    with device times.  Returns a dict."""
    warm_iters = iters if warm_iters is None else warm_iters
    eng = RLEngine(R.shape, tab, ks)
    eng.set_raw(R)
    _, t_plain = eng.iterate(iters)
    t_it = t_plain / iters
    eng.snapshot()
    plain = [1.0]  # k = 0: X_0 = raw
    eng.set_raw(R)
    plain[0] = eng.distance()
    for k in range(iters):
        eng.iterate(1, k0=k)
        plain.append(eng.distance())
    t_net = eng.warm_start(net)
    t_net = min(t_net, eng.warm_start(net))  # second call: buffers allocated
    warm = [eng.distance(seed=True)]
    eng.iterate(1, k0=0, seeded=True)
    warm.append(eng.distance())
    for k in range(1, warm_iters):
        eng.iterate(1, k0=k)
        warm.append(eng.distance())
    eng.close()
    out = {"iteration_ms": round(t_it, 3), "net_ms": round(t_net, 3),
           "plain": {str(k): round(plain[k], 6) for k in ks_report},
           "warm": {str(k): round(warm[k], 6) for k in ks_report if k < len(warm)}}
    best = int(np.argmin(warm))
    out["warm_best"] = {"k": best, "rel_l2": round(float(warm[best]), 6)}
    ttr = {}
    levels = [1e-3, 1e-2, 5e-2, 0.15, 0.2, round(float(warm[0]), 6)]
    for r in levels:
        kp, kw = first_k(plain, r), first_k(warm, r)
        tp = None if kp is None else kp * t_it
        tw = None if kw is None else t_net + kw * t_it
        ttr[f"{r:g}"] = {"plain_k": kp, "plain_ms": None if tp is None else round(tp, 2), "warm_k": kw,
                         "warm_ms": None if tw is None else round(tw, 2),
                         "ratio": None if (tp is None or tw is None or tp == 0) else round(tw / tp, 3)}
    out["time_to_residual"] = ttr
    out["curves"] = {"plain": [round(v, 6) for v in plain], "warm": [round(v, 6) for v in warm]}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--stacks", type=int, default=8)
    t.add_argument("--batch", type=int, default=8)
    t.add_argument("--lr", type=float, nargs="+", default=[1e-3, 3e-4, 1e-4])
    t.add_argument("--steps", type=int, nargs="+", default=[1500, 750, 500])
    t.add_argument("--chunk", type=int, default=250)
    t.add_argument("--seed", type=int, default=1)
    t.add_argument("--out", default=WEIGHTS)
    t.add_argument("--target-iters", type=int, default=TARGET_ITERS,
                   help="train towards the RL result after this many iterations (the evaluation's reference "
                        "stays the 50-iteration result, which the seed then approaches in the remaining ones)")
    i = sub.add_parser("infer", help="time the warm start (network on the device) on a synthetic stack")
    i.add_argument("--weights", default=WEIGHTS)
    i.add_argument("--shape", type=int, nargs=3, default=list(EVAL_STACK))
    i.add_argument("--repeat", type=int, default=3)
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=WEIGHTS)
    e.add_argument("--shape", type=int, nargs=3, default=list(EVAL_STACK))
    e.add_argument("--seed", type=int, default=EVAL_SEED)
    e.add_argument("--out", default=None)
    e.add_argument("--warm-iters", type=int, default=None, help="iterations after the warm start (default 50)")
    a = ap.parse_args(argv)
    if a.cmd == "train":
        if len(a.lr) != len(a.steps):
            ap.error("--lr and --steps need the same number of values")
        train(a)
        return 0
    p, meta = load_params(a.weights)
    net = WarmNet(params=p)
    shape = tuple(a.shape)
    if a.cmd == "infer":
        # a stack of the evaluation kind is not needed for timing: background
        # plus one block of signal, staged on the device
        _, psf = light_sheet_stack((8, 8, 8), 0)
        tab = make_otf(psf)
        eng = RLEngine(shape, tab, kscale_for(shape, tab))
        blk = np.random.default_rng(0).uniform(0, 50, (min(64, shape[0]), 64, 64)).astype(np.float32)
        eng.set_raw_block(blk, (0, 0, 0), background=100.0)
        for r in range(a.repeat):
            log(f"warm start {shape}: {eng.warm_start(net):.3f} ms")
        eng.close()
        return 0
    t0 = time.perf_counter()
    R, psf = light_sheet_stack(shape, a.seed)
    tab = make_otf(psf)
    log(f"evaluation stack {shape} seed {a.seed}: generated in {time.perf_counter() - t0:.1f} s")
    res = evaluate(net, R, tab, kscale_for(shape, tab), warm_iters=a.warm_iters)
    res["weights"] = meta
    js = json.dumps(res)
    log(js)
    if a.out:
        with open(a.out, "w") as f:
            f.write(js + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
