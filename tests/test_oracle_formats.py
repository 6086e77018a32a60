"""Pins the oracle's restatement of the ingest / output formats by closed
forms: partition of unity, block means for integer shrink factors, identity,
exact linear ramps when enlarging, byte order of 16-bit pixels, unorm16 edge
values and rounding."""
import numpy as np
import pytest


@pytest.mark.parametrize("n_dst,n_src", [(512, 2048), (37, 100), (80, 50), (64, 64), (1, 9)])
def test_resample_rows_sum_to_one(orc, n_dst, n_src):
    m = orc.resample_matrix(n_dst, n_src)
    assert np.allclose(m.sum(axis=1), 1.0, atol=1e-12)
    assert (m >= 0).all()


def test_integer_shrink_is_block_mean(orc):
    rng = np.random.default_rng(3)
    img = rng.random((64, 96))
    got = orc.resample_matrix(16, 64) @ img @ orc.resample_matrix(24, 96).T
    ref = img.reshape(16, 4, 24, 4).mean(axis=(1, 3))
    assert np.abs(got - ref).max() < 1e-12


def test_equal_size_is_identity(orc):
    assert np.array_equal(orc.resample_matrix(33, 33), np.eye(33))


def test_enlarge_reproduces_linear_ramp_inside(orc):
    n_src, n_dst = 50, 80
    ramp = 3.0 + 0.25 * np.arange(n_src)
    got = orc.resample_matrix(n_dst, n_src) @ ramp
    x = (np.arange(n_dst) + 0.5) * n_src / n_dst - 0.5
    inside = (x >= 0) & (x <= n_src - 1)
    assert np.abs(got[inside] - (3.0 + 0.25 * x[inside])).max() < 1e-12
    assert got[0] == ramp[0] and got[-1] == ramp[-1]  # clamped ends


def test_non_integer_shrink_preserves_mean(orc):
    rng = np.random.default_rng(5)
    v = rng.random(100)
    m = orc.resample_matrix(37, 100)
    # every source pixel's total weight is n_dst / n_src: the mean is kept
    assert np.allclose(m.sum(axis=0), 37 / 100, atol=1e-12)
    assert abs((m @ v).mean() - v.mean()) < 1e-12


def test_decode_u16_little_endian_channel0(orc):
    # 2 x 1 frame, 3 channels of 2 bytes; channel 0 = 0x1234, 0xffff
    raw = bytes([0x34, 0x12, 9, 9, 9, 9, 0xff, 0xff, 7, 7, 7, 7])
    got = orc.decode_frame(raw, 2, 1, 3, 2, 2, 1)
    assert np.allclose(got, [[0x1234 / 65535.0, 1.0]], atol=0, rtol=1e-15)


def test_decode_u8(orc):
    raw = bytes([0, 255, 51, 102])
    got = orc.decode_frame(raw, 4, 1, 1, 1, 2, 1)
    assert np.allclose(got, [[127.5 / 255.0, 76.5 / 255.0]], rtol=1e-15)


def test_unorm16_edges(orc):
    v = np.array([-1.0, 0.0, 1.0, 2.0, np.nan, 0.5, 1.0 / 65535.0, 0.49999 / 65535.0],
                 dtype=np.float32)
    got = orc.unorm16(v)
    assert got.tolist() == [0, 0, 65535, 65535, 0, 32768, 1, 0]


def test_normalize_range(orc):
    v = np.array([2.0, 4.0, 3.0, np.nan])
    got = orc.normalize_range(v)
    assert got[:3].tolist() == [0.0, 1.0, 0.5]
    assert np.all(orc.normalize_range(np.full(4, 7.0)) == 0.0)
