"""The FP16 LDS-staged backprojector holds a voxel's row x in [0, 128) as the
FP32 number x + 128 and takes two things from that one number: the LDS
address, from the high half of its bits (v_mad_u32_u16 with op_sel selecting
the high half, against a base lowered by 0x4300 rows), and the row weight,
from v_fract_f32.  This pins the IEEE-754 facts the kernel relies on, on the
CPU, for every row of the 72-row buffer and fractions across each cell,
including values that round up onto the next row: the high half is
0x4300 + floor, the address wraps back to floor * pitch + base modulo 2^32
for the LDS pitches in use, and the weight lies in [0, 1) and puts the
sample on the same side of the row boundary as the address does.  The
kernel's rows stay at least 0.05 inside the staged rectangle (its bounds
carry that margin), so the samples here end 0.05 below the last row's end."""
import numpy as np

ROWS = 72                       # 16-bit LDS buffer height
PITCHES_BYTES = (51 * 4, 53 * 4)


def _biased(x):
    # the kernel's rows are FP32 results of FMAs on FP64-derived operands:
    # round the exact value once to FP32
    return (np.asarray(x, dtype=np.float64) + 128.0).astype(np.float32)


def _samples():
    rng = np.random.default_rng(11)
    ints = np.arange(ROWS, dtype=np.float64)
    below = np.nextafter(ints[1:], 0.0)                    # just under a row
    tiny = np.array([0.0, 2.0 ** -20, 2.0 ** -17, 2.0 ** -16])
    grid = (ints[:, None] + np.linspace(0.0, 1.0, 65)[None, :-1]).ravel()
    x = np.concatenate([rng.uniform(0.0, ROWS, 200000), ints, below, tiny, grid])
    return x[x <= ROWS - 0.05]


def test_biased_row_bits():
    xp = _biased(_samples())
    bits = xp.view(np.uint32)
    assert ((bits >> 23) == 134).all()                      # every row in [128, 256)
    row = np.floor(xp.astype(np.float64)) - 128.0
    assert (row >= 0).all() and (row < ROWS).all()
    assert np.array_equal((bits >> 16).astype(np.int64) - 0x4300, row.astype(np.int64))


def test_biased_row_address_and_weight():
    xp = _biased(_samples())
    hi = (xp.view(np.uint32) >> 16).astype(np.uint64)
    row = (np.floor(xp.astype(np.float64)) - 128.0).astype(np.uint64)
    for lp4 in PITCHES_BYTES:
        for base in (0, 2064, 2064 + 2 * ROWS * lp4 - lp4):
            lowered = (base - 0x4300 * lp4) % 2 ** 32
            addr = (hi * lp4 + lowered) % 2 ** 32
            assert np.array_equal(addr, row * lp4 + base)
    w = xp - np.floor(xp)                                   # v_fract_f32, exact here
    assert (w >= 0).all() and (w < 1).all()
    # weight and address come from the same rounded number, so a sample that
    # rounds onto a row boundary takes that row with weight 0
    on_row = xp == np.floor(xp)
    assert (w[on_row] == 0).all()
    # the rounding of the row is at most half of 2^-16
    x = _samples()
    assert np.abs(xp.astype(np.float64) - 128.0 - x).max() <= 2.0 ** -17
