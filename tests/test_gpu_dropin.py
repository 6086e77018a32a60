"""The C++ drop-in surface: examples/fbp_cone_fdk_main.cpp drives FBP3D with
the call sequence of the reference's Mains/fbp_cone_fdk_main.cpp; its volume
must equal the same pipeline issued through the C ABI from Python."""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import rel_max

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


def test_cpp_fbp3d_matches_c_abi_pipeline(ctr, dev, tmp_path):
    import torch
    exe = os.path.join(os.path.dirname(ctr.LIB_PATH), "examples", "fbp_cone_fdk_main")
    assert os.path.exists(exe), "examples are built by __graft_entry__.build()"
    size, views = 64, 90
    out = tmp_path / "vol.raw"
    r = subprocess.run([exe, str(size), str(views), str(out)], capture_output=True, text=True,
                       timeout=120)
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
    staged, _ = ctr.stage_projections(proj, g)
    nfft = 64  # the reference default 2^ceil(log2 nu) for nu = 64
    scale = math.pi / views / (nfft * (g.detector_width / size))
    scale *= (g.source_distance + g.detector_distance) / g.source_distance
    ref = ctr.backproject(staged, g, scale=scale, clip=True).cpu().numpy()
    assert ref.max() > 0
    assert rel_max(got, ref) < 1e-6


def test_cpp_fbp3d_reports_reference_errors(ctr, tmp_path):
    exe = os.path.join(os.path.dirname(ctr.LIB_PATH), "examples", "fbp_cone_fdk_main")
    # 1 voxel per side with the reference geometry is fine; a zero-size volume is not
    r = subprocess.run([exe, "0", "10", str(tmp_path / "x.raw")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1
    assert "[CTAnalyzer Error]" in (r.stdout + r.stderr)
