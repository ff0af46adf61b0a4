"""Writes tests/golden/fft_golden.npz: numpy (pocketfft, FP64) transforms used
to pin the oracle's own mixed-radix FFT.

Inputs are regenerated from the stored seeds (numpy PCG64); outputs are
stored.  Covers every axis length of the workloads (64 ... 2048, 810, and
the 3675-wide deskewed stacks) plus small odd/even 3-D shapes.

    python tests/golden/make_golden.py
"""
import os

import numpy as np

LENGTHS_1D = [2, 3, 5, 7, 11, 13, 30, 64, 96, 105, 128, 180, 192, 256, 384, 405, 512, 600, 768, 810, 1024, 2048, 3675]
SHAPES_3D = [(6, 10, 14), (7, 9, 15), (4, 8, 16), (3, 5, 11), (5, 6, 7), (8, 12, 810 // 27)]


def inputs_1d(n):
    rng = np.random.default_rng(1000 + n)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def inputs_3d(shape):
    rng = np.random.default_rng(sum(shape) * 7 + shape[0])
    return rng.standard_normal(shape)


def main():
    out = {}
    for n in LENGTHS_1D:
        x = inputs_1d(n)
        out[f"fft_{n}"] = np.fft.fft(x)
        out[f"ifft_{n}"] = np.fft.ifft(x) * n
    for s in SHAPES_3D:
        v = inputs_3d(s)
        key = "x".join(map(str, s))
        S = np.fft.rfftn(v)
        out[f"rfftn_{key}"] = S
        # irfftn of a non-Hermitian half spectrum (numpy semantics)
        rng = np.random.default_rng(99 + sum(s))
        Z = rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape)
        out[f"irfftn_in_{key}"] = Z
        out[f"irfftn_{key}"] = np.fft.irfftn(Z, s=s, axes=(0, 1, 2)) * np.prod(s)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
