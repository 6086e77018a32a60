"""Pins the oracle functions that the FFT, geometry and format tests leave
open: fan -> parallel rebinning, log-normalisation, the FFT's bit-reversal
and single-stage passes, and the ellipsoid voxeliser.

Each check comes from outside the oracle's own formula -- explicit vector
geometry, analytic sinograms, closed forms, the reference's compiled FFT --
so that a mirrored fan angle, a swapped view/offset mapping, a wrong log
base or a mis-indexed butterfly fails.
"""
import math
import os

import numpy as np
import pytest
from scipy import ndimage

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_fft.npz")


# --- fan -> parallel rebinning (FBP2D::parallelize_fan_beam_sinogram,
#     FBP2D.cpp:31-52) -------------------------------------------------------

def _vector_rebin(fan, sod, odd, det_w, nu_par, par_w):
    """Independent restatement from vectors: for parallel ray (theta_n, s)
    find the source on the circle of radius SOD that lies on the ray (on the
    +e(theta) side), intersect the ray with that source's flat detector, and
    sample the fan sinogram there (bilinear, periodic in views)."""
    n_views, nu = fan.shape
    dth = 2 * math.pi / n_views
    delta = det_w / nu
    pitch = par_w / nu_par
    wrapped = np.vstack([fan, fan[:1]]).astype(np.float64)  # view N == view 0
    out = np.zeros((n_views, nu_par))
    for n in range(n_views):
        th = n * dth
        e = np.array([math.cos(th), math.sin(th)])      # (x, z)
        eu = np.array([-math.sin(th), math.cos(th)])
        for i in range(nu_par):
            s = (i + 0.5 - nu_par / 2) * pitch
            if abs(s) >= sod:
                continue
            S = s * eu + math.sqrt(sod * sod - s * s) * e  # source on the ray
            beta = math.atan2(S[1], S[0]) % (2 * math.pi)
            eb = np.array([math.cos(beta), math.sin(beta)])
            ebu = np.array([-math.sin(beta), math.cos(beta)])
            D0 = -odd * eb
            r = -e                                          # ray direction
            t = -np.dot(S - D0, eb) / np.dot(r, eb)
            H = S + t * r
            u = np.dot(H - D0, ebu)
            fb = beta / dth
            fu = u / delta + nu / 2 - 0.5
            out[n, i] = ndimage.map_coordinates(wrapped, [[fb], [fu]], order=1,
                                                mode="grid-constant", cval=0.0)[0]
    return out


def test_rebin_matches_vector_geometry(orc):
    rng = np.random.default_rng(12)
    n_views, nu, nu_par = 90, 64, 41
    sod, odd, det_w, par_w = 300.0, 200.0, 260.0, 150.0
    fan = rng.standard_normal((n_views, nu)).astype(np.float32)
    got = orc.rebin_fan_to_parallel(fan, sod, odd, det_w, 1.0, nu_par, par_w)
    ref = _vector_rebin(fan, sod, odd, det_w, nu_par, par_w)
    assert np.abs(got - ref).max() < 1e-9


def test_rebin_central_ray_is_the_central_pixel_pair(orc):
    """s = 0 maps to fan angle 0: parallel view n reads fan view n exactly,
    halfway between the two central detector columns."""
    rng = np.random.default_rng(2)
    fan = rng.standard_normal((36, 50)).astype(np.float32)
    got = orc.rebin_fan_to_parallel(fan, 400.0, 100.0, 120.0, 1.0, 31, 60.0)
    centre = 0.5 * (fan[:, 24].astype(np.float64) + fan[:, 25])
    # (n dtheta) / dtheta rounds to within an ulp of n
    assert np.allclose(got[:, 15], centre, rtol=0, atol=1e-12)


def test_rebin_moves_a_point_to_theta_minus_gamma(orc):
    """A single fan sample at source angle beta and detector column c (fan
    angle gamma = atan(u / SDD)) lands at parallel angle beta - gamma and
    offset s = SOD sin(gamma)."""
    n_views, nu, nu_par = 360, 128, 201
    sod, odd, det_w, par_w = 500.0, 500.0, 400.0, 360.0
    for b, c in [(40, 100), (200, 10), (359, 64)]:
        fan = np.zeros((n_views, nu), dtype=np.float32)
        fan[b, c] = 1.0
        par = orc.rebin_fan_to_parallel(fan, sod, odd, det_w, 1.0, nu_par, par_w)
        dth = 2 * math.pi / n_views
        u = (c + 0.5 - nu / 2) * det_w / nu
        gam = math.atan(u / (sod + odd))
        th = (b * dth - gam) % (2 * math.pi)
        s = sod * math.sin(gam)
        n_hit, i_hit = np.unravel_index(np.argmax(par), par.shape)
        dview = abs((n_hit * dth - th + math.pi) % (2 * math.pi) - math.pi)
        assert dview <= dth
        assert abs((i_hit + 0.5 - nu_par / 2) * par_w / nu_par - s) <= par_w / nu_par
        assert par.sum() > 0


def test_rebin_analytic_fan_gives_analytic_parallel(orc):
    """Analytic fan sinogram of off-centre ellipses, rebinned, equals the
    analytic parallel sinogram of the same phantom up to interpolation."""
    ell = np.array([[20.0, 0.0, -35.0, 30.0, 1e6, 18.0, 0.0, 0.4, 0.0, 1.0],
                    [-40.0, 0.0, 25.0, 12.0, 1e6, 25.0, 0.0, 1.9, 0.0, 0.5]])
    n_views, nu, nu_par = 720, 512, 257
    sod, odd, det_w, par_w = 500.0, 500.0, 600.0, 256.0
    fan = orc.analytic_cone(ell, sod, odd, det_w, nu, 1, n_views)[:, 0, :]
    got = orc.rebin_fan_to_parallel(fan.astype(np.float32), sod, odd, det_w, 1.0, nu_par,
                                     par_w)
    ref = orc.analytic_parallel(ell, par_w, nu_par, 1, n_views, 1.0, 0.0)[:, 0, :]
    assert ref.max() > 30.0
    err = np.abs(got - ref)
    assert np.sqrt((err ** 2).mean()) / ref.max() < 5e-3
    # a mirrored fan angle would displace every off-centre shadow
    assert np.corrcoef(got.ravel(), ref.ravel())[0, 1] > 0.999


def test_rebin_rays_outside_the_fan_are_zero(orc):
    fan = np.ones((40, 32), dtype=np.float32)
    # detector 100 wide at SDD 400: fan half-angle atan(50/400); SOD 200 ->
    # rays with |s| > 200 sin(atan(1/8)) ~ 24.8 miss the detector
    par = orc.rebin_fan_to_parallel(fan, 200.0, 200.0, 100.0, 1.0, 101, 101.0)
    s = (np.arange(101) + 0.5 - 50.5)
    assert np.all(par[:, np.abs(s) > 26.0] == 0.0)
    assert np.allclose(par[:, np.abs(s) < 23.0], 1.0, rtol=0, atol=1e-12)


# --- log-normalisation (FBP3D::log_normalize_projections, FBP3D.cpp:333-349)

def test_log_normalize_closed_forms(orc):
    base = 1e5
    p = np.array([0.0, 0.5, 1.0, 3.0, 7.25], dtype=np.float64)
    got = orc.log_normalize(base * np.exp(-p), base)
    assert np.allclose(got, p, rtol=0, atol=1e-12)
    assert orc.log_normalize(np.array([base]), base)[0] == 0.0
    # the floor clamps dark or negative readings
    got = orc.log_normalize(np.array([0.0, -5.0, 0.5, 2.0]), base, floor_value=1.0)
    assert np.allclose(got, [math.log(base)] * 3 + [-math.log(2.0 / base)], rtol=1e-15)


def test_log_normalize_inverts_beer_lambert(orc):
    """Transmission I = I0 exp(-p) of analytic ellipsoid line integrals is
    log-normalised back to those integrals."""
    ell = np.array([[5.0, -3.0, 2.0, 20.0, 14.0, 9.0, 0.2, 0.5, 0.1, 0.02]])
    p = orc.analytic_cone(ell, 200.0, 150.0, 120.0, 24, 20, 6)
    got = orc.log_normalize(1e5 * np.exp(-p), 1e5)
    assert p.max() > 0.3
    assert np.allclose(got, p, rtol=0, atol=1e-12)


# --- FFT passes (FFFT::reverse_bit_swap / fft_single_step,
#     FastFourierTransformGPU.cpp:238-256) -------------------------------------

@pytest.mark.parametrize("n", [1, 2, 8, 64, 1024])
def test_bit_reverse_is_the_index_permutation(orc, n):
    levels = n.bit_length() - 1
    rev = np.array([int(format(i, f"0{levels}b")[::-1], 2) if levels else 0 for i in range(n)])
    z = np.arange(n) + 1j * np.arange(n, 2 * n)
    assert np.array_equal(orc.fft_bit_reverse(z), z[rev])
    with pytest.raises(ValueError):
        orc.fft_bit_reverse(np.zeros(12))


@pytest.mark.parametrize("n", [2, 16, 512, 2048])
def test_stages_after_bit_reversal_are_the_fft(orc, golden_fft, n):
    """Bit reversal then stages 0 .. log2(n) - 1 reproduce the full FFT bit
    for bit, with true pi and with the reference's pi' (where the reference's
    own compiled FFT, via the golden vectors, fixes the result)."""
    rng = np.random.default_rng(n)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for pi in (np.pi, orc.reference_pi()):
        w = orc.fft_bit_reverse(z)
        for s in range(n.bit_length() - 1):
            w = orc.fft_single_stage(w, s, pi)
        assert np.array_equal(w, orc.fft_radix2(z, pi))
    if f"fft_in_{n}" in golden_fft:
        w = orc.fft_bit_reverse(golden_fft[f"fft_in_{n}"])
        for s in range(n.bit_length() - 1):
            w = orc.fft_single_stage(w, s, orc.reference_pi())
        assert np.array_equal(w, golden_fft[f"fft_out_{n}"])


def test_single_stage_closed_form(orc):
    """Stage 0 is x[2m] +- x[2m+1]; the last stage of span n pairs i with
    i + n/2 through the twiddle exp(-2 pi i k / n)."""
    n = 16
    rng = np.random.default_rng(4)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    w = orc.fft_single_stage(z, 0)
    assert np.allclose(w[0::2], z[0::2] + z[1::2], atol=1e-15)
    assert np.allclose(w[1::2], z[0::2] - z[1::2], atol=1e-15)
    w = orc.fft_single_stage(z, 3)
    k = np.arange(n // 2)
    tw = np.exp(-2j * np.pi * k / n)
    assert np.allclose(w[:n // 2], z[:n // 2] + z[n // 2:] * tw, atol=1e-14)
    assert np.allclose(w[n // 2:], z[:n // 2] - z[n // 2:] * tw, atol=1e-14)
    with pytest.raises(ValueError):
        orc.fft_single_stage(z, 4)   # span 32 > n


@pytest.fixture(scope="module")
def golden_fft():
    return dict(np.load(GOLDEN, allow_pickle=False))


# --- ellipsoid voxeliser (FBP3D::render_ellipsoid, FBP3D.cpp:609-628) -------

def test_render_sphere_volume_and_membership(orc):
    X = Y = Z = 48
    r = 15.0
    ball = np.array([[2.0, -3.0, 1.0, r, r, r, 0.7, 0.2, 1.3, 0.25]])
    vol = orc.render_ellipsoids(ball, (X, Y, Z), float(X))
    inside = vol > 0
    assert set(np.unique(vol)) <= {0.0, 0.25}
    assert abs(inside.sum() - 4.0 / 3.0 * math.pi * r ** 3) / (4.0 / 3.0 * math.pi * r ** 3) < 0.02
    # voxel (i, j, k) has centre (i - 23.5, j - 23.5, k - 23.5).  Row j = 20,
    # k = 24 sits 0.5 off the centre in y and z, so it holds the chord
    # |x - 2| <= sqrt(r^2 - 0.5) = 14.98: i = 40 (x = 16.5) is inside,
    # i = 41 (x = 17.5) is not; likewise i = 11 (x = -12.5) and 10 (-13.5)
    row = vol[24, 20, :]
    assert row[40] == 0.25 and row[41] == 0.0
    assert row[11] == 0.25 and row[10] == 0.0
    assert np.count_nonzero(row) == 30


def test_render_rotation_convention(orc):
    """A needle along body x, rotated a quarter turn about z (gamma), lies
    along world y: R = Rz(gamma) Ry(beta) Rx(alpha)."""
    needle = np.array([[0.0, 0.0, 0.0, 10.0, 1.2, 1.2, 0.0, 0.0, math.pi / 2, 1.0]])
    vol = orc.render_ellipsoids(needle, (32, 32, 32), 32.0)
    assert vol[16, :, 16].sum() > 15 and vol[16, 16, :].sum() < 4


# --- the closed-form helpers ----------------------------------------------------

@pytest.mark.parametrize("n", [1, 3, 12, 64, 100])
def test_direct_dft_matches_numpy(orc, n):
    """The long-double DFT that pins the true-pi FFT, against numpy's FFT
    (any n, forward and backward)."""
    rng = np.random.default_rng(n)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.abs(orc.dft(z, -1) - np.fft.fft(z)).max() < 1e-12 * max(1, n)
    assert np.abs(orc.dft(z, +1) - n * np.fft.ifft(z)).max() < 1e-12 * max(1, n)


def test_analytic_cone_sphere_closed_form(orc):
    """analytic_cone's ray for detector pixel (iu, iv) of view n, rebuilt
    from vectors (source SOD e, pixel at -ODD e + u eu + v y): a sphere's
    projection is 2 sqrt(r^2 - h^2) times its density, h the distance of its
    centre from that ray."""
    sod, odd, det_w, nu, nv, N = 300.0, 200.0, 240.0, 24, 18, 5
    c, r, rho = np.array([12.0, -7.0, 5.0]), 30.0, 0.7
    ball = np.array([[*c, r, r, r, 0.3, 0.2, 0.1, rho]])
    p = orc.analytic_cone(ball, sod, odd, det_w, nu, nv, N, 1.0, 0.25)
    delta = det_w / nu
    for n, iv, iu in [(0, 9, 12), (2, 3, 5), (4, 15, 20), (3, 0, 0)]:
        th = 2 * math.pi * n / N + 0.25
        e = np.array([math.cos(th), 0.0, math.sin(th)])
        eu = np.array([-math.sin(th), 0.0, math.cos(th)])
        u, v = (iu + 0.5 - nu / 2) * delta, (iv + 0.5 - nv / 2) * delta
        S = sod * e
        D = -odd * e + u * eu + np.array([0.0, v, 0.0])
        d = (D - S) / np.linalg.norm(D - S)
        w = c - S
        h2 = np.dot(w, w) - np.dot(w, d) ** 2
        ref = rho * 2 * math.sqrt(max(r * r - h2, 0.0))
        assert p[n, iv, iu] == pytest.approx(ref, abs=1e-9)
    assert p.max() > 2 * rho * r * 0.9


# --- per-cell backprojection (the dequantized model of the 16-bit
#     value-and-step texels) ------------------------------------------------

def _plain_as_cells(q):
    """Cells of a plain view stack q (N, nv, nu): cell c = iv + 1 holds row
    iv on top and row iv + 1 below, rows outside the detector zero."""
    n, nv, nu = q.shape
    z = np.zeros((n, 1, nu), dtype=np.float32)
    return np.concatenate([z, q], axis=1), np.concatenate([q, z], axis=1)


def test_cells_backprojection_reduces_to_plain(orc):
    """With each cell's rows taken from one view stack, the per-cell
    backprojection is the pinned plain one (cone and parallel), up to the
    order of the FP64 sum."""
    rng = np.random.default_rng(31)
    q = rng.standard_normal((40, 21, 26)).astype(np.float32)
    top, bot = _plain_as_cells(q)
    args = (150.0, 90.0, 52.0, 20.0, 1.0, 0.3, 0.7, (20, 18, 22))
    a = orc.backproject_cone_cells(top, bot, *args)
    b = orc.backproject_cone(q, *args)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
    pargs = (26.0, 20.0, 0.5, 0.1, 0.7, (20, 18, 22))
    a = orc.backproject_parallel_cells(top, bot, *pargs)
    b = orc.backproject_parallel(q, *pargs)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_cells_take_the_bottom_row_from_bot(orc):
    """A cell's lower row comes from `bot`, not from the next cell's top: with
    bot = top + 1 everywhere (and top from a plain stack), every sample gains
    exactly its vertical weight wv, so the volume gains the backprojection of
    the stack whose cells are all (0, 1).  Adding 1 to top instead gains the
    stack whose cells are all (1, 0); the two gains add up to the all-ones
    stack, which inside the detector is the plain backprojection of ones."""
    rng = np.random.default_rng(32)
    n, nv, nu = 30, 17, 23
    # multiples of 1/64: q + 1 is exact in float32
    q = (np.round(rng.standard_normal((n, nv, nu)) * 64) / 64).astype(np.float32)
    top, bot = _plain_as_cells(q)
    ones = np.ones_like(top)
    args = (150.0, 90.0, 46.0, 16.0, 1.0, 0.2, 0.9, (16, 14, 16))
    base = orc.backproject_cone_cells(top, bot, *args)
    gain = orc.backproject_cone_cells(top, bot + ones, *args) - base
    wv_only = orc.backproject_cone_cells(np.zeros_like(top), ones, *args)
    assert np.abs(gain - wv_only).max() <= 1e-12 * np.abs(wv_only).max()
    assert np.abs(wv_only).max() > 0.05
    swapped = orc.backproject_cone_cells(top + ones, bot, *args) - base
    top_only = orc.backproject_cone_cells(ones, np.zeros_like(top), *args)
    assert np.abs(swapped - top_only).max() <= 1e-12 * np.abs(top_only).max()
    # wv and 1 - wv partition the unit cell: (1, 1) cells give the all-ones
    # stack's backprojection (interior columns and rows)
    full = orc.backproject_cone_cells(ones, ones, *args)
    plain = orc.backproject_cone(np.ones((n, nv, nu), np.float32), *args)
    inner = (slice(4, 12), slice(4, 10), slice(4, 12))
    assert np.allclose((wv_only + top_only)[inner], full[inner], rtol=1e-12, atol=0)
    assert np.allclose(full[inner], plain[inner], rtol=1e-12, atol=0)
