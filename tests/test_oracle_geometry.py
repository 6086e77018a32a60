"""Pins the oracle's geometry: weights, backprojection, forward projection.

The checks restate the scan geometry independently (explicit source point,
detector plane and ray-plane intersection in numpy, scipy's bilinear
sampler), use closed forms (sphere chords, the FDK cosine weight) and
reconstruct analytic phantoms end to end, so that a flipped rotation, a
swapped axis, a half-pixel shift or a wrong scale constant fails.
"""
import math

import numpy as np
import pytest
from scipy import ndimage


def test_fdk_weight_closed_form(orc):
    sod, odd, w, nu, nv = 500.0, 300.0, 400.0, 8, 6
    p = np.ones((1, nv, nu), dtype=np.float32)
    got = orc.fdk_weight(p, sod, odd, w)[0]
    delta = w / nu
    for iv, iu in [(0, 0), (2, 5), (5, 7)]:
        u = (iu + 0.5 - nu / 2) * delta
        v = (iv + 0.5 - nv / 2) * delta
        assert got[iv, iu] == pytest.approx(800.0 / math.sqrt(800.0 ** 2 + u * u + v * v),
                                            rel=1e-15)


def _vector_cone_bp(proj_view, theta, sod, odd, det_w, vol_w, shape, voxels):
    """One view's contribution at the given voxel indices, from vectors."""
    nv, nu = proj_view.shape
    X, Y, Z = shape
    d = vol_w / X
    delta = det_w / nu
    e = np.array([math.cos(theta), 0.0, math.sin(theta)])
    eu = np.array([-math.sin(theta), 0.0, math.cos(theta)])
    ey = np.array([0.0, 1.0, 0.0])
    S = sod * e
    D0 = -odd * e                        # detector centre
    out = []
    for (i, j, k) in voxels:
        P = np.array([(i + 0.5 - X / 2) * d, (j + 0.5 - Y / 2) * d, (k + 0.5 - Z / 2) * d])
        r = P - S
        t = np.dot(D0 - S, -e) / np.dot(r, -e)  # line S + t r meets the plane
        hit = S + t * r
        u = np.dot(hit - D0, eu)
        v = np.dot(hit - D0, ey)
        fu = u / delta + nu / 2 - 0.5
        fv = v / delta + nv / 2 - 0.5
        val = ndimage.map_coordinates(proj_view.astype(np.float64), [[fv], [fu]], order=1,
                                      mode="grid-constant", cval=0.0)[0]
        L = np.dot(r, -e)                  # depth from the source
        out.append((sod / L) ** 2 * val)
    return np.array(out)


@pytest.mark.parametrize("theta", [0.0, 0.7, 2.5, 4.0])
def test_cone_backprojection_vs_vector_geometry(orc, theta):
    rng = np.random.default_rng(int(theta * 10))
    nu, nv = 40, 30
    q = rng.standard_normal((1, nv, nu)).astype(np.float32)
    shape = (20, 16, 24)
    sod, odd, det_w, vol_w = 90.0, 60.0, 80.0, 30.0
    vol = orc.backproject_cone(q, sod, odd, det_w, vol_w, 1.0, theta, 1.0, shape)
    vox = [(rng.integers(0, 20), rng.integers(0, 16), rng.integers(0, 24)) for _ in range(25)]
    vox += [(0, 0, 0), (19, 15, 23), (10, 8, 12)]
    ref = _vector_cone_bp(q[0], theta, sod, odd, det_w, vol_w, shape, vox)
    got = np.array([vol[k, j, i] for (i, j, k) in vox])
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_cone_region_matches_full(orc):
    rng = np.random.default_rng(3)
    q = rng.standard_normal((12, 20, 24)).astype(np.float32)
    args = (100.0, 50.0, 60.0, 30.0, 1.0, 0.2, 0.5)
    full = orc.backproject_cone(q, *args, (16, 12, 14))
    part = orc.backproject_cone_region(q, *args, (16, 12, 14), ((3, 9), (2, 12), (5, 6)))
    assert np.array_equal(part, full[5:6, 2:12, 3:9])


def test_parallel_backprojection_vs_vector_geometry(orc):
    rng = np.random.default_rng(8)
    nu, nv, theta = 30, 12, 1.1
    q = rng.standard_normal((1, nv, nu)).astype(np.float32)
    shape = (18, 10, 22)
    det_w, vol_w = 36.0, 27.0
    vol = orc.backproject_parallel(q, det_w, vol_w, 1.0, theta, 1.0, shape)
    d, delta = vol_w / 18, det_w / nu
    eu = np.array([-math.sin(theta), 0.0, math.cos(theta)])
    for (i, j, k) in [(0, 0, 0), (17, 9, 21), (4, 3, 11), (9, 5, 2)]:
        P = np.array([(i + 0.5 - 9) * d, (j + 0.5 - 5) * d, (k + 0.5 - 11) * d])
        fu = np.dot(P, eu) / delta + nu / 2 - 0.5
        fv = P[1] / delta + nv / 2 - 0.5
        ref = ndimage.map_coordinates(q[0].astype(np.float64), [[fv], [fu]], order=1,
                                      mode="grid-constant", cval=0.0)[0]
        assert vol[k, j, i] == pytest.approx(ref, abs=1e-12)


def test_constant_projection_backprojects_to_weight_sum(orc):
    """q == 1 everywhere inside the detector: each voxel receives
    scale * sum_n (SOD / L_n)^2 with L_n its depth from the source."""
    N, sod, odd = 16, 80.0, 80.0
    q = np.ones((N, 64, 64), dtype=np.float32)
    shape = (6, 5, 7)
    vol = orc.backproject_cone(q, sod, odd, 400.0, 12.0, 1.0, 0.0, 0.5, shape)
    d = 2.0
    for (i, j, k) in [(0, 0, 0), (5, 4, 6), (2, 2, 3)]:
        x, z = (i + 0.5 - 3) * d, (k + 0.5 - 3.5) * d
        ref = sum(0.5 * (sod / (sod - x * math.cos(t) - z * math.sin(t))) ** 2
                  for t in 2 * np.pi * np.arange(N) / N)
        assert vol[k, j, i] == pytest.approx(ref, rel=1e-12)


def test_f32_baseline_loop_matches_fp64(orc):
    rng = np.random.default_rng(4)
    q = rng.standard_normal((45, 40, 48)).astype(np.float32)
    args = (150.0, 100.0, 96.0, 40.0, 1.0, 0.3, 0.01)
    shape = (40, 32, 36)
    ref = orc.backproject_cone(q, *args, shape)
    got = orc.backproject_cone_f32(q, *args, shape)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-4
    part = orc.backproject_cone_f32(q, *args, shape, z_range=(10, 13))
    assert np.array_equal(part, got[10:13])


def test_sphere_chord_closed_form(orc):
    r, rho = 7.0, 0.3
    sph = np.array([[1.0, -2.0, 0.5, r, r, r, 0.3, 1.2, -0.4, rho]])
    for b in (0.0, 2.5, 6.9, 7.5):
        o = np.array([1.0 - 100.0, -2.0 + b, 0.5])
        got = orc.ellipsoid_line_integral(sph, o, np.array([2.0, 0.0, 0.0]))
        ref = rho * 2 * math.sqrt(max(r * r - b * b, 0.0))
        assert got == pytest.approx(ref, abs=1e-12)


def test_ellipsoid_rotation_convention(orc):
    """Semi-axes (a, b, c) along body x, y, z; R = Rz(g) Ry(b) Rx(a).  A
    quarter turn about z maps the body x axis onto world y."""
    e = np.array([[0, 0, 0, 5.0, 2.0, 1.0, 0.0, 0.0, math.pi / 2, 1.0]])
    along_y = orc.ellipsoid_line_integral(e, np.array([0, -50.0, 0]), np.array([0, 1.0, 0]))
    along_x = orc.ellipsoid_line_integral(e, np.array([-50.0, 0, 0]), np.array([1.0, 0, 0]))
    assert along_y == pytest.approx(10.0, abs=1e-9)
    assert along_x == pytest.approx(4.0, abs=1e-9)
    e2 = np.array([[0, 0, 0, 5.0, 2.0, 1.0, 0.0, math.pi / 2, 0.0, 1.0]])  # about y: x -> -z
    along_z = orc.ellipsoid_line_integral(e2, np.array([0, 0, -50.0]), np.array([0, 0, 1.0]))
    assert along_z == pytest.approx(10.0, abs=1e-9)


@pytest.mark.parametrize("kind", ["cone", "parallel"])
def test_forward_projector_vs_analytic(orc, kind):
    """Ray-marched projections of a voxelised off-centre ellipsoid agree with
    its exact line integrals up to the voxelisation error."""
    X = Y = Z = 64
    vol_w = 64.0
    ell = np.array([[6.0, -4.0, 9.0, 14.0, 9.0, 11.0, 0.3, 0.8, 0.2, 1.0]])
    vol = orc.render_ellipsoids(ell, (X, Y, Z), vol_w).astype(np.float32)
    if kind == "cone":
        args = (200.0, 150.0, 160.0, 48, 40, 8)
        fp = orc.forward_cone(vol, vol_w, *args, 1.0, 0.4, 0.5)
        an = orc.analytic_cone(ell, *args, 1.0, 0.4)
    else:
        args = (96.0, 48, 40, 8)
        fp = orc.forward_parallel(vol, vol_w, *args, 0.5, 0.4, 0.5)
        an = orc.analytic_parallel(ell, *args, 0.5, 0.4)
    assert an.max() > 10.0
    err = np.abs(fp - an)
    assert np.sqrt((err ** 2).mean()) / an.max() < 0.02
    # the bulk of the shadow must overlap (a mirrored geometry would not)
    assert np.corrcoef(fp.ravel(), an.ravel())[0, 1] > 0.995


def _analytic_ball_fdk(orc, nfft, n_views=90):
    sod, odd, det_w, nu, nv = 150.0, 150.0, 192.0, 96, 96
    X = Y = Z = 48
    vol_w = 48.0
    ball = np.array([[3.0, 2.0, -4.0, 14.0, 14.0, 14.0, 0, 0, 0, 0.05]])
    p = orc.analytic_cone(ball, sod, odd, det_w, nu, nv, n_views)
    w = orc.fdk_weight(p, sod, odd, det_w).astype(np.float32)
    q = orc.filter_rows(w.reshape(-1, nu), 0, nfft).reshape(w.shape).astype(np.float32)
    scale = math.pi / n_views / (nfft * det_w / nu) * (sod + odd) / sod
    inside = orc.backproject_cone_region(q, sod, odd, det_w, vol_w, 1.0, 0.0, scale,
                                         (X, Y, Z), ((25, 29), (24, 28), (18, 22)))
    outside = orc.backproject_cone_region(q, sod, odd, det_w, vol_w, 1.0, 0.0, scale,
                                          (X, Y, Z), ((2, 6), (22, 26), (40, 44)))
    return inside.mean(), outside.mean()


def test_fdk_scale_recovers_ball_density(orc):
    """Feldkamp scale (pi/N) * (SDD/SOD) / (n * delta) on reference-convention
    filtered projections reproduces the density of an off-centre ball."""
    inside, outside = _analytic_ball_fdk(orc, 512)
    assert inside == pytest.approx(0.05, rel=0.01)
    assert abs(outside) < 0.05 * 0.01


def test_parallel_2d_fbp_recovers_disk(orc):
    """2D = the y = 0 slice: parallel FBP over 180 deg of an analytic disk."""
    N, nu, W = 180, 129, 64
    disk = np.array([[4.0, 0.0, -6.0, 15.0, 1e6, 15.0, 0, 0, 0, 1.0]])
    p = orc.analytic_parallel(disk, float(nu), nu, 1, N, 0.5, 0.0)
    nfft = 512
    q = orc.filter_rows(p.reshape(-1, nu).astype(np.float32), 0, nfft).reshape(p.shape)
    img = orc.backproject_parallel(q.astype(np.float32), float(nu), float(W), 0.5, 0.0,
                                   math.pi / N / nfft, (W, 1, W))[:, 0, :]
    assert img[26 - 2:26 + 2, 36 - 2:36 + 2].mean() == pytest.approx(1.0, rel=0.01)
    assert abs(img[2:6, 2:6].mean()) < 0.01
