"""Pins the oracle's FFT and filter restatement.

Checks, chosen so that a dropped term, a wrong sign or a wrong index fails:
* the reference's own compiled FFT (golden vectors in tests/golden, and the
  live library when oracle/_ref was built) -- bit-exact in pi' mode;
* a long-double direct DFT -- true-pi mode;
* an independent numpy formulation of the row filter (fftfreq indexing
  instead of shift indexing);
* the closed-form Ram-Lak impulse response (Kak & Slaney eq. 3.61).
"""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_fft.npz")


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


@pytest.mark.parametrize("n", [8, 64, 512, 2048])
def test_fft_matches_reference_golden(orc, golden, n):
    pi = float(golden["reference_pi"])
    assert pi == orc.reference_pi() == float(np.float32(3.14154965))
    z = golden[f"fft_in_{n}"]
    assert np.array_equal(orc.fft_radix2(z, pi), golden[f"fft_out_{n}"])
    assert np.array_equal(orc.inverse_fft_radix2(z, pi), golden[f"ifft_out_{n}"])
    assert np.array_equal(orc.fft_shift(z, n // 2), golden[f"shift_pos_{n}"])
    assert np.array_equal(orc.fft_shift(z, -(n // 2)), golden[f"shift_neg_{n}"])


@pytest.mark.parametrize("key", ["filter_367_512_0", "filter_100_256_1", "filter_60_128_2"])
def test_filter_chain_matches_reference_golden(orc, golden, key):
    _, width, npad, ftype = key.split("_")
    got = orc.filter_rows(golden[key + "_in"][None, :], int(ftype), int(npad),
                          pi=orc.reference_pi())[0]
    assert np.array_equal(got, golden[key + "_out"])


def test_fft_matches_live_reference_library(orc):
    if orc.ref_lib() is None:
        pytest.skip("oracle/_ref was not built (no reference checkout here)")
    rng = np.random.default_rng(5)
    for n in (1, 2, 4, 16, 128, 1024, 4096):
        z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        assert np.array_equal(orc.fft_radix2(z, orc.reference_pi()), orc.ref_fft_radix2(z))
        assert np.array_equal(orc.inverse_fft_radix2(z, orc.reference_pi()),
                              orc.ref_inverse_fft_radix2(z))
    with pytest.raises(ValueError):
        orc.ref_fft_radix2(np.zeros(12))


@pytest.mark.parametrize("n", [1, 2, 4, 8, 32, 256])
def test_fft_true_pi_matches_direct_dft(orc, n):
    rng = np.random.default_rng(n)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = orc.dft(z, -1)
    got = orc.fft_radix2(z)
    assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())
    back = orc.inverse_fft_radix2(got)
    assert np.abs(back - z).max() <= 1e-13 * max(1.0, np.abs(z).max())
    # conj-FFT-conj / n is the backward DFT / n
    assert np.abs(orc.inverse_fft_radix2(z) - orc.dft(z, +1) / n).max() <= 1e-13 * n


def test_fft_rejects_non_power_of_two(orc):
    for n in (3, 6, 12, 100):
        with pytest.raises(ValueError):
            orc.fft_radix2(np.zeros(n))


@pytest.mark.parametrize("n", [2, 8, 64])
def test_fft_shift_identities(orc, n):
    z = np.arange(n) + 0j
    assert np.array_equal(orc.fft_shift(z).real, np.fft.fftshift(z.real))
    assert np.array_equal(orc.fft_shift(z, -(n // 2)).real, np.fft.ifftshift(z.real))
    assert np.array_equal(orc.fft_shift(orc.fft_shift(z, 3), -3), z)


@pytest.mark.parametrize("ftype", [0, 1, 2])
def test_filter_gain_closed_form(orc, ftype):
    n = 64
    g = orc.filter_gain(n, ftype)
    i = np.arange(n)
    x = i / n - 0.5
    ref = {0: np.abs(x), 1: np.sin(np.abs(x)), 2: np.abs(np.sin(np.pi * x)) / np.pi}[ftype] * n
    assert np.allclose(g, ref, rtol=0, atol=1e-12 * n)
    assert g[n // 2] == 0.0  # DC sits at the centre of the shifted spectrum
    assert np.allclose(g[1:n // 2], g[n - 1:n // 2:-1])  # even about DC


def _numpy_filter(rows, ftype, npad):
    """Independent formulation: unshifted spectrum, gain from fftfreq."""
    f = np.fft.fftfreq(npad)  # [0, 1/n, ..., -1/2, ..., -1/n]
    gain = {0: np.abs(f), 1: np.sin(np.abs(f)), 2: np.abs(np.sin(np.pi * f)) / np.pi}[ftype]
    z = np.zeros((rows.shape[0], npad))
    z[:, :rows.shape[1]] = rows
    return np.fft.ifft(np.fft.fft(z, axis=1) * gain * npad, axis=1).real[:, :rows.shape[1]]


@pytest.mark.parametrize("width,npad,ftype", [(367, 512, 0), (367, 1024, 1), (768, 2048, 2),
                                              (5, 8, 0), (64, 64, 2)])
def test_filter_rows_vs_independent_numpy(orc, width, npad, ftype):
    rng = np.random.default_rng(width)
    rows = rng.standard_normal((6, width)).astype(np.float32)
    got = orc.filter_rows(rows, ftype, npad)
    ref = _numpy_filter(rows.astype(np.float64), ftype, npad)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_filter_default_length_is_next_power_of_two(orc):
    rows = np.random.default_rng(1).standard_normal((2, 367)).astype(np.float32)
    assert np.array_equal(orc.filter_rows(rows, 0), orc.filter_rows(rows, 0, 512))
    with pytest.raises(ValueError):
        orc.filter_rows(rows, 0, 256)   # shorter than the row
    with pytest.raises(ValueError):
        orc.filter_rows(rows, 0, 600)   # not a power of two
    with pytest.raises(ValueError):
        orc.filter_rows(rows, 3, 512)   # unknown filter


def test_ram_lak_impulse_response_closed_form(orc):
    """Unit impulse -> n * h[m] with h[0] = 1/4, h[odd m] -> -1/(pi m)^2,
    h[even m != 0] -> 0 (sampled ramp of unit sample spacing)."""
    n = 4096
    row = np.zeros((1, 2048), dtype=np.float32)
    row[0, 0] = 1.0
    q = orc.filter_rows(row, 0, n)[0] / n
    assert abs(q[0] - 0.25) < 1e-12
    for m in (1, 3, 5, 21):
        assert abs(q[m] + 1.0 / (np.pi * m) ** 2) < 2e-3 / m ** 2
    for m in (2, 4, 10):
        assert abs(q[m]) < 1e-6
