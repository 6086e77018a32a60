"""Writes tests/golden/reference_fft.npz from the reference's own CPU FFT.

The vectors are produced by oracle/_ref/libreffft.so, which oracle/Makefile
compiles from Source/FastFourierTransformCPU/FastFourierTransform.cpp of the
reference checkout (FFT::fft_radix2, FFT::inverse_fft_radix2, FFT::fft_shift,
with the reference's pi = 3.14154965f).  Inputs are seeded; outputs are what
the reference code returned for them.  The filter chain vectors run the
reference's row filter sequence (FBP3D.cpp:399-413) with those same compiled
functions and the gain n * g((float)i/n - 0.5f) of
FastFourierTransformImage.cpp:35-37,45.

    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def ref_gain(n, ftype, pi):
    out = np.empty(n)
    for i in range(n):
        x = float(np.float32(np.float32(i) / np.float32(n)) - np.float32(0.5))
        g = {0: abs(x), 1: np.sin(abs(x)), 2: abs(np.sin(pi * x)) / pi}[ftype]
        out[i] = g * n
    return out


def ref_filter_row(row, npad, ftype):
    pi = oracle.ref_lib().ref_pi()
    z = np.zeros(npad, dtype=np.complex128)
    z[:row.size] = row
    z = oracle.ref_fft_radix2(z)
    z = oracle.ref_fft_shift(z, npad // 2)
    z = z * ref_gain(npad, ftype, pi)
    z = oracle.ref_fft_shift(z, -(npad // 2))
    z = oracle.ref_inverse_fft_radix2(z)
    return z[:row.size].real


def main():
    if oracle.ref_lib() is None:
        raise SystemExit("oracle/_ref/libreffft.so is missing: run make -C oracle ref")
    rng = np.random.default_rng(20240601)
    data = {"reference_pi": np.array(oracle.ref_lib().ref_pi())}
    for n in (8, 64, 512, 2048):
        z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        data[f"fft_in_{n}"] = z
        data[f"fft_out_{n}"] = oracle.ref_fft_radix2(z)
        data[f"ifft_out_{n}"] = oracle.ref_inverse_fft_radix2(z)
        data[f"shift_pos_{n}"] = oracle.ref_fft_shift(z, n // 2)
        data[f"shift_neg_{n}"] = oracle.ref_fft_shift(z, -(n // 2))
    for width, npad, ftype in ((367, 512, 0), (100, 256, 1), (60, 128, 2)):
        row = rng.standard_normal(width).astype(np.float32)
        key = f"filter_{width}_{npad}_{ftype}"
        data[key + "_in"] = row
        data[key + "_out"] = ref_filter_row(row.astype(np.float64), npad, ftype)
    np.savez(os.path.join(HERE, "reference_fft.npz"), **data)
    print("wrote", os.path.join(HERE, "reference_fft.npz"))


if __name__ == "__main__":
    main()
