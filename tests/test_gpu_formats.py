"""Formats either side of the path on the device: frame decode + resample
(ctr_decode_frames), min / max / range normalisation and unorm16 slices,
against the oracle's restatement, and the C++ drop-in read_projections /
write_volume / normalize_histogram (examples/fbp_io_main) end to end.
Values are in [0, 1]: the decode bound 2e-6 is FP32 accumulation of the box
filter; unorm16 slices are bit-exact."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _frames(rng, n, w, h, ch, bpc):
    hi = 255 if bpc == 1 else 65535
    px = rng.integers(0, hi + 1, size=(n, h, w, ch), dtype=np.uint32)
    if bpc == 1:
        return px.astype(np.uint8)
    return px.astype('<u2').view(np.uint8).reshape(n, h, w, ch * 2)


@pytest.mark.parametrize("fw,fh,ch,bpc,tw,th", [
    (256, 192, 1, 2, 64, 48),     # integer shrink 4 (the reference main: 2048 -> 512)
    (100, 90, 3, 2, 37, 41),      # non-integer shrink, channel 0 of 3
    (50, 40, 1, 2, 80, 64),       # enlarge (linear)
    (64, 64, 2, 1, 64, 64),       # 8-bit, same size
])
@pytest.mark.parametrize("layout", [0, 1])
def test_decode_frames_vs_oracle(ctr, orc, dev, fw, fh, ch, bpc, tw, th, layout):
    import torch
    rng = np.random.default_rng(fw + fh + ch + layout)
    n = 5
    raw = _frames(rng, n, fw, fh, ch, bpc)
    print(f"[formats] decode {fw}x{fh}x{ch}x{bpc} -> {tw}x{th} layout {layout}", flush=True)
    raw_t = torch.from_numpy(raw.reshape(-1).copy()).to(dev)
    views = n + 2
    shape = (views, th, tw) if layout == 0 else (th, views, tw)
    out = torch.full(shape, -1.0, dtype=torch.float32, device=dev)
    ctr.decode_frames(raw_t, fw, fh, ch, bpc, out, tw, th, first_view=1, layout=layout)
    got = out.cpu().numpy()
    if layout == 1:
        got = got.transpose(1, 0, 2)
    for f in range(n):
        ref = orc.decode_frame(raw[f].tobytes(), fw, fh, ch, bpc, tw, th)
        assert np.abs(got[f + 1] - ref).max() < 2e-6, f
    assert (got[0] == -1.0).all() and (got[-1] == -1.0).all()  # other views untouched


def test_min_max_and_normalize(ctr, orc, dev):
    import torch
    rng = np.random.default_rng(11)
    v = (rng.standard_normal(1 << 20) * 3 - 1).astype(np.float32)
    v[12345] = np.nan
    t = torch.from_numpy(v).to(dev)
    mm = ctr.min_max(t).cpu().numpy()
    assert mm[0] == np.nanmin(v) and mm[1] == np.nanmax(v)
    ctr.normalize_range(t, ctr.min_max(t))
    ref = orc.normalize_range(v)
    got = t.cpu().numpy()
    ok = ~np.isnan(v)
    assert np.abs(got[ok] - ref[ok]).max() < 2e-7
    const = torch.full((1000,), 5.0, device=dev)
    assert (ctr.normalize_range(const, ctr.min_max(const)).cpu().numpy() == 0).all()


def test_slices_unorm16_bit_exact(ctr, orc, dev):
    import torch
    rng = np.random.default_rng(12)
    Z, Y, X = 7, 9, 33
    v = (rng.random((Z, Y, X)) * 1.4 - 0.2).astype(np.float32)
    v[0, 0, :4] = [0.5, 1.0 / 65535.0, np.nan, 1.0]
    got = ctr.slices_unorm16(torch.from_numpy(v).to(dev), 2, 5).cpu().numpy().view(np.uint16)
    ref = orc.unorm16(v[:, 2:7, :].transpose(1, 0, 2))
    assert np.array_equal(got, ref)


def _exe(ctr):
    exe = os.path.join(os.path.dirname(ctr.LIB_PATH), "examples", "fbp_io_main")
    assert os.path.exists(exe), "examples are built by __graft_entry__.build()"
    return exe


@pytest.mark.parametrize("sino", [False, True])
def test_cpp_read_projections(ctr, orc, tmp_path, sino):
    rng = np.random.default_rng(21)
    fw, fh, ch, n, tw, th = 128, 96, 1, 12, 32, 24
    raw = _frames(rng, n, fw, fh, ch, 2)
    d = tmp_path / "frames"
    d.mkdir()
    for i in range(n):  # index in the 4 characters before the extension
        (d / f"scan_{i:04d}.raw").write_bytes(raw[i].tobytes())
    (d / "notes.txt").write_text("not a frame")
    (d / "scan_0099.raw").write_bytes(raw[0].tobytes())  # index beyond the count: ignored
    out = tmp_path / "proj.raw"
    args = [_exe(ctr), "read", str(d), str(fw), str(fh), str(ch), "2", str(tw), str(th), str(n),
            str(out)] + (["sino"] if sino else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "[FBP3D Info] loading from" in r.stdout
    got = np.fromfile(out, dtype=np.float32)
    got = got.reshape(th, n, tw).transpose(1, 0, 2) if sino else got.reshape(n, th, tw)
    for i in range(n):
        ref = orc.decode_frame(raw[i].tobytes(), fw, fh, ch, 2, tw, th)
        assert np.abs(got[i] - ref).max() < 2e-6, i


def test_cpp_read_projections_errors(ctr, tmp_path):
    r = subprocess.run([_exe(ctr), "read", str(tmp_path / "missing"), "8", "8", "1", "2", "8",
                        "8", "2", str(tmp_path / "o.raw")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1
    assert "[CTAnalyzer Error] FBP3D tried to load_projections() but" in r.stdout
    d = tmp_path / "short"
    d.mkdir()
    (d / "f0000.raw").write_bytes(b"\0" * 10)
    (d / "f0001.raw").write_bytes(b"\0" * 128)
    r = subprocess.run([_exe(ctr), "read", str(d), "8", "8", "1", "2", "8", "8", "2",
                        str(tmp_path / "o.raw")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "bytes of a frame" in r.stdout


@pytest.mark.parametrize("mode", ["", "normalize", "textures"])
def test_cpp_write_volume(ctr, orc, tmp_path, mode):
    rng = np.random.default_rng(31)
    X, Y, Z = 24, 10, 16
    v = (rng.random((Z, Y, X)) * 2 - 0.5).astype(np.float32)
    src = tmp_path / "vol.raw"
    v.tofile(src)
    out = tmp_path / "slices"
    args = [_exe(ctr), "write", str(src), str(X), str(Y), str(Z), str(out)] + ([mode] if mode else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    files = sorted(os.listdir(out))
    assert files == [f"CTAnalyzer_Slice#{y:06d}.raw" for y in range(Y)]
    ref_vol = orc.normalize_range(v).astype(np.float32) if mode else v
    for y in range(Y):
        got = np.fromfile(out / files[y], dtype='<u2').reshape(Z, X)
        ref = orc.unorm16(ref_vol[:, y, :])
        # the normalised values are computed in FP32 on the device
        tol = 1 if mode else 0
        assert np.abs(got.astype(int) - ref.astype(int)).max() <= tol, y
