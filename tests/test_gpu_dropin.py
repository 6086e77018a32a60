"""The C++ drop-in surface: examples/fbp_cone_fdk_main.cpp drives FBP3D with
the call sequence of the reference's Mains/fbp_cone_fdk_main.cpp; its volume
must equal the same pipeline issued through the C ABI from Python."""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import rel_max, rel_rmse

pytestmark = pytest.mark.gpu

SOD, ODD, DET, VOLW = 730.87, 669.04, 409.60, 213.84


def f32(x):
    return float(np.float32(x))


def shepp_logan_records(size):
    """FBP3D::generate_shepplogan through FBP3D::render_ellipsoid in voxel
    units (centre c/2 * size, semi-axes a/2 * size, angle deg * 3.14/180)."""
    k = np.float32(3.14) / np.float32(180.0)
    rows = [((0, 0, 0), (0.69, 0.9, 0.92), 0, 2.0 / 3.0),
            ((0, 0, 0), (0.6624, 0.88, 0.874), 0, -0.98 / 3.0),
            ((-0.22, -0.25, 0), (0.41, 0.21, 0.16), 108, -0.02 * 75 / 3.0),
            ((0.22, -0.25, 0), (0.31, 0.22, 0.11), 72, -0.02 * 75 / 3.0),
            ((0, -0.25, 0.4), (0.23, 0.46, 0.23), 0, 0.02 * 75 / 3.0),
            ((0, -0.25, 0.1), (0.046, 0.046, 0.046), 0, 0.02 * 75 / 3.0),
            ((-0.8, -0.25, -0.65), (0.046, 0.02, 0.023), 0, 0.01 * 75 / 3.0),
            ((0.06, -0.25, -0.065), (0.046, 0.02, 0.023), 90, 0.01 * 75 / 3.0),
            ((0.06, 0.625, -0.105), (0.56, 0.1, 0.04), 90, 0.02 * 75 / 3.0),
            ((0, -0.625, 0.1), (0.056, 0.1, 0.056), 0, -0.02 * 75 / 3.0)]
    out = []
    for c, a, deg, v in rows:
        c = [f32(x) for x in c]
        a = [f32(x) for x in a]
        out.append([c[0] / 2.0 * size, c[1] / 2.0 * size, c[2] / 2.0 * size,
                    a[0] / 2.0 * size, a[1] / 2.0 * size, a[2] / 2.0 * size,
                    0.0, float(np.float32(np.float32(deg) * k)), 0.0, f32(f32(v))])
    return np.array(out)


@pytest.mark.parametrize("dt", [0, 1, 2])
def test_cpp_fbp3d_matches_c_abi_pipeline(ctr, dev, tmp_path, dt):
    """FP32, FP16 and BF16 staging (`FBP3D::staged_dtype`).  The two pipelines'
    filtered projections agree to FP32 rounding, not bit for bit, so a 16-bit
    texel can round the other way: the 16-bit volumes agree to one 16-bit
    rounding step at the voxels that read such a texel."""
    import torch
    exe = os.path.join(os.path.dirname(ctr.LIB_PATH), "examples", "fbp_cone_fdk_main")
    assert os.path.exists(exe), "examples are built by __graft_entry__.build()"
    size, views = 64, 90
    out = tmp_path / "vol.raw"
    r = subprocess.run([exe, str(size), str(views), str(out), "0", str(dt)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, dtype=np.float32).reshape(size, size, size)

    g = ctr.Geometry(ctr.CONE, nu=size, nv=size, n_views=views, X=size, Y=size, Z=size,
                     detector_width=f32(DET), volume_width=f32(VOLW),
                     source_distance=f32(SOD), detector_distance=f32(ODD))
    vol = torch.zeros((size, size, size), dtype=torch.float32, device=dev)
    ctr.render_ellipsoids(shepp_logan_records(size), vol, float(size))
    proj = ctr.forward_project(vol, g, step=1.0)
    ctr.fdk_weight(proj, g)
    ctr.filter_rows(proj, ctr.SHEPP_LOGAN, 0)
    staged, vs = ctr.stage_projections(proj, g, dt)
    nfft = 64  # the reference default 2^ceil(log2 nu) for nu = 64
    scale = math.pi / views / (nfft * (g.detector_width / size))
    scale *= (g.source_distance + g.detector_distance) / g.source_distance
    ref = ctr.backproject(staged, g, dt, view_scale=vs, scale=scale, clip=True).cpu().numpy()
    assert ref.max() > 0
    assert rel_max(got, ref) < {0: 1e-6, 1: 1e-4, 2: 8e-4}[dt]
    assert rel_rmse(got, ref) < {0: 1e-6, 1: 1e-5, 2: 1e-4}[dt]


def test_cpp_fbp3d_reports_reference_errors(ctr, tmp_path):
    exe = os.path.join(os.path.dirname(ctr.LIB_PATH), "examples", "fbp_cone_fdk_main")
    # 1 voxel per side with the reference geometry is fine; a zero-size volume is not
    r = subprocess.run([exe, "0", "10", str(tmp_path / "x.raw")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1
    assert "[CTAnalyzer Error]" in (r.stdout + r.stderr)


def test_cpp_fbp3d_slice_and_view_copies(ctr, dev, tmp_path):
    """FBP3D's slice and view copies (FBP3D.cpp:639-778) on the drop-in
    class: every load into an exact-size and an odd-size texture, every store
    from both, against numpy index identities.  The copied extent is the min
    of both sizes; the rest of a target keeps its value (-1 for loads)."""
    exe = os.path.join(os.path.dirname(ctr.LIB_PATH), "examples", "fbp3d_copies_main")
    X, Y, Z, nu, nv, N = 20, 12, 16, 24, 10, 9
    rng = np.random.default_rng(4)
    vol = rng.standard_normal((Z, Y, X)).astype(np.float32)
    proj = rng.standard_normal((N, nv, nu)).astype(np.float32)
    sino = rng.standard_normal((nv, N, nu)).astype(np.float32)
    patch = rng.standard_normal(4096).astype(np.float32)
    for name, a in (("volume", vol), ("projections", proj), ("sinograms", sino),
                    ("patch", patch)):
        a.tofile(tmp_path / f"{name}.raw")
    r = subprocess.run([exe, str(tmp_path), *map(str, (X, Y, Z, nu, nv, N))],
                       capture_output=True, text=True, timeout=120)
    # the program ends on the reference's out-of-range error, after its dumps
    assert r.returncode == 1 and "load_projection: index out of range" in r.stdout + r.stderr

    def got(name, h, w):
        return np.fromfile(tmp_path / f"{name}.raw", dtype=np.float32).reshape(h, w)

    sources = {"slice_x": (vol[:, :, X // 3], Y, Z), "slice_y": (vol[:, Y // 2, :], X, Z),
               "slice_z": (vol[Z - 1], X, Y), "projection": (proj[N // 2], nu, nv),
               "sinogram": (sino[nv // 3], nu, N),
               "projection_from_sinograms": (sino[:, 1, :], nu, nv)}
    for name, (src, w, h) in sources.items():
        assert np.array_equal(got(name + "_exact", h, w), src), name
        odd = got(name + "_odd", h - 2, w + 3)
        ref = np.full((h - 2, w + 3), -1.0, dtype=np.float32)
        ref[:, :w] = src[:h - 2, :w]
        assert np.array_equal(odd, ref), name

    def P(h, w):
        return patch[:h * w].reshape(h, w)

    v = vol.copy()
    v[:, :, 1] = P(Z, Y)
    v[:Z, :Y - 2, X - 1] = P(Z + 1, Y - 2)[:Z]
    v[:, 0, :] = P(Z, X)
    v[:Z - 3, Y - 2, :] = P(Z - 3, X + 2)[:, :X]
    v[2] = P(Y, X)
    v[Z // 2, :, :X - 1] = P(Y + 2, X - 1)[:Y]
    assert np.array_equal(got("volume_out", Z * Y, X).reshape(Z, Y, X), v)
    p = proj.copy()
    p[0] = P(nv, nu)
    p[N - 1, :nv - 1] = P(nv - 1, nu + 1)[:, :nu]
    assert np.array_equal(got("projections_out", N * nv, nu).reshape(N, nv, nu), p)
    s = sino.copy()
    s[0] = P(N, nu)
    s[nv - 1, :, :nu - 2] = P(N + 1, nu - 2)[:N]
    s[:, 2, :] = P(nv, nu)
    s[:nv - 2, N - 2, :] = P(nv - 2, nu + 2)[:, :nu]
    assert np.array_equal(got("sinograms_out", nv * N, nu).reshape(nv, N, nu), s)


def test_cpp_ffft_methods(ctr, orc, dev, tmp_path):
    """Every public FFFT method (FastFourierTransformGPU.h:10-66) through the
    drop-in class: the padded fft (2^ceil(log2 W), zero tail) and the
    cropping inverse_fft (:38-91), the shifts in and out of place (:93-142),
    the blits, conjugate / multiply / divide (:144-236), bit reversal plus
    single stages (:238-256), floor/ceil_log2 (:269-280) and the reference's
    precondition errors."""
    exe = os.path.join(os.path.dirname(ctr.LIB_PATH), "examples", "ffft_main")
    W, P, H = 300, 256, 5
    rng = np.random.default_rng(21)
    rows = (rng.standard_normal((H, W)) + 1j * rng.standard_normal((H, W))).astype(np.complex64)
    rows_p2 = (rng.standard_normal((H, P)) + 1j * rng.standard_normal((H, P))).astype(np.complex64)
    rows.tofile(tmp_path / "rows.raw")
    rows_p2.tofile(tmp_path / "rows_p2.raw")
    r = subprocess.run([exe, str(tmp_path), str(W), str(P), str(H)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr

    def got(name, w, complex_=True):
        a = np.fromfile(tmp_path / f"{name}.raw", dtype=np.complex64 if complex_ else np.float32)
        return a.reshape(H, w)

    padded = np.zeros((H, 512), dtype=np.complex128)
    padded[:, :W] = rows
    ref = np.stack([orc.fft_radix2(z) for z in padded])
    assert rel_max(got("fft_padded", 512), ref) < 2e-5
    assert rel_max(got("inverse_cropped", W), rows) < 1e-5
    ref = np.stack([orc.inverse_fft_radix2(z.astype(np.complex128)) for z in rows_p2])
    assert rel_max(got("inverse_same", P), ref) < 1e-5
    assert np.array_equal(got("shift_out", P), np.roll(rows_p2, P // 2, axis=1))
    assert np.array_equal(got("shift_in_place", P), np.roll(rows_p2, P // 2, axis=1))
    assert np.array_equal(got("unshift_in_place", P), rows_p2)
    assert np.array_equal(got("shift_by_3", P), np.roll(rows_p2, 3, axis=1))
    ref = np.stack([orc.fft_radix2(z.astype(np.complex128)) for z in rows_p2])
    assert rel_max(got("stages", P), ref) < 2e-5
    assert np.array_equal(got("real_r", W, False), rows.real)
    assert np.array_equal(got("real_i", W, False), rows.imag)
    assert np.array_equal(got("complex_from_real", W), rows.real.astype(np.complex64))
    c = np.conj(rows)
    ref = ((c.real * np.float32(2.5)) * np.float32(0.25)
           + 1j * ((c.imag * np.float32(2.5)) * np.float32(0.25))).astype(np.complex64)
    assert np.array_equal(got("conj_mul_div", W), ref)
    out = r.stdout
    for n, fl, ce in ((1, 0, 0), (2, 1, 1), (3, 1, 2), (300, 8, 9), (512, 9, 9), (513, 9, 10)):
        assert f"log2 {n}: floor {fl} ceil {ce}" in out
    assert out.count("error as expected") == 2 and "no error" not in out
    assert "[CTAnalyzer Error] FFFT::fft_radix2() is called but texture's width is not a " \
           "power of 2" in out


def test_cpp_fbp3d_detector_and_volume_passes(ctr, orc, dev, tmp_path):
    """FBP3D's log-normalisation, FDK weighting and row filtering of both
    projection layouts and its negative clip (FBP3D.cpp:333-458), through the
    drop-in class, each stage against the oracle applied to the previous
    GPU stage."""
    exe = os.path.join(os.path.dirname(ctr.LIB_PATH), "examples", "fbp3d_passes_main")
    nu, nv, N, X, Y, Z = 100, 24, 7, 20, 12, 16
    base, sod, odd, width = (float(np.float32(v)) for v in (1e5, 400.0, 300.0, 250.0))
    rng = np.random.default_rng(8)
    proj = (base * np.exp(-rng.uniform(0, 5, (N, nv, nu)))).astype(np.float32)
    sino = (base * np.exp(-rng.uniform(0, 5, (nv, N, nu)))).astype(np.float32)
    vol = rng.standard_normal((Z, Y, X)).astype(np.float32)
    for name, a in (("projections", proj), ("sinograms", sino), ("volume", vol)):
        a.tofile(tmp_path / f"{name}.raw")
    r = subprocess.run([exe, str(tmp_path), *map(str, (nu, nv, N, X, Y, Z, base, sod, odd,
                                                       width))],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr

    def got(name, shape):
        return np.fromfile(tmp_path / f"{name}.raw", dtype=np.float32).reshape(shape)

    flt_min = float(np.finfo(np.float32).tiny)
    for layout, data, shape in (("projections", proj, (N, nv, nu)),
                                ("sinograms", sino, (nv, N, nu))):
        log = got(f"{layout}_log", shape)
        assert np.abs(log - orc.log_normalize(data.astype(np.float64), base, flt_min)).max() < 2e-6
        fdk = got(f"{layout}_fdk", shape)
        as_proj = (lambda a: a) if layout == "projections" else (lambda a: a.transpose(1, 0, 2))
        ref = as_proj(orc.fdk_weight(as_proj(log), sod, odd, width))
        assert np.abs(fdk - ref).max() <= 2 ** -24 * np.abs(ref).max()
        ftype = orc.SHEPP_LOGAN if layout == "projections" else orc.COSINE
        filt = got(f"{layout}_filtered", shape)
        ref = orc.filter_rows(fdk.reshape(-1, nu), ftype).reshape(shape)  # default 2^ceil(log2 nu)
        assert rel_max(filt, ref) < 1e-5
    assert np.array_equal(got("volume_clipped", (Z, Y, X)), np.maximum(vol, np.float32(0)))
