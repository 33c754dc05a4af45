"""HeatPlot (Graph.h:7-109) and the plot members built on it (IMEXRK.h:164-303,
StateVector.h:476-500), through tests/cpp/plot_demo.cpp on a 48x1x64 grid.

The resampling is restated below from Graph.h:34-95 (two periods in x at 50
pixels per unit length, z in [-6, 6), bilinear in x, periodic, and along the
field's own vertical points, first row at the top) and pinned by what the
mathematics fixes: a field linear in z comes out exactly linear in z.  The
PNG files are read back with PIL."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "plot_demo")
N1, N3 = 48, 64

pytestmark = pytest.mark.gpu


def resample(F, pts, L1):
    """Graph.h:34-95 in FP32: F is (N1, N3) nodal values of slice j2."""
    f = np.float32
    L1 = f(L1)
    W, H = int(L1 * f(50)), int(f(2 * 6 * 50))
    img = np.zeros((H, 2 * W), dtype=np.float32)
    for j1 in range(2 * W):
        x = f(j1) * L1 / f(W)
        kl = int(x / (L1 / f(N1)))
        kr = kl + 1
        xl, xr = f(kl) * L1 / f(N1), f(kr) * L1 / f(N1)
        wl, wr = (xr - x) / (xr - xl), (x - xl) / (xr - xl)
        kl, kr = kl % N1, kr % N1
        for j3 in range(H):
            z = f(j3) * f(2) * f(6) / f(H) - f(6)
            ka = 1
            while f(pts[ka]) < z:
                ka += 1
            kb = ka - 1
            zb, za = f(pts[kb]), f(pts[ka])
            wa, wb = (z - zb) / (za - zb), (za - z) / (za - zb)
            img[H - 1 - j3, j1] = (wl * wb * F[kl, kb] + wl * wa * F[kl, ka] + wr * wb * F[kr, kb] +
                                   wr * wa * F[kr, ka])
    return img


def test_heatplot_resampling_and_files(tmp_path):
    from PIL import Image
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    w, h = [int(v) for v in r.stdout.split("size")[1].split()[:2]]
    L1 = 2 * 3.14159265358979 / 0.44
    assert (w, h) == (2 * int(np.float32(L1) * np.float32(50)), 600)
    zf = np.fromfile(str(tmp_path / "zf.bin"), dtype=np.float64)
    lin = np.fromfile(str(tmp_path / "linear.bin"), dtype=np.float32).reshape(h, w)
    wave = np.fromfile(str(tmp_path / "wave.bin"), dtype=np.float32).reshape(h, w)
    # a field linear in z is reproduced exactly along the image's z
    zpix = (np.arange(h)[::-1] * 12.0 / h - 6.0)[:, None]
    assert np.abs(lin - (0.25 + 0.5 * zpix)).max() < 2e-6
    # the general case against the restatement
    x = np.arange(N1) * np.float32(L1) / N1
    F = (np.cos(2 * np.pi * x / np.float32(L1))[:, None] * np.exp(-0.1 * zf[None, :] ** 2) +
         0.01 * np.arange(N3)[None, :]).astype(np.float32)
    ref = resample(F, zf, L1)
    assert np.abs(wave - ref).max() < 1e-5 * np.abs(ref).max()
    # the PNG: the image itself, top row the largest value of the linear field
    im = Image.open(str(tmp_path / "linear.png"))
    assert im.size == (w, h) and im.mode == "RGB"
    px = np.asarray(im)
    assert tuple(px[0, 0]) == (140, 13, 13) and tuple(px[-1, 0]) == (13, 38, 115)
    assert np.all(px[:, 0] == px[:, w // 2])  # no x dependence
    # StateVector::PlotAll and IMEXRK's PlotAll (2D: no u2)
    for name in ("u1", "u3", "b", "vorticity"):
        assert Image.open(str(tmp_path / "state" / (name + ".png"))).size == (w, h)
    assert not (tmp_path / "state" / "u2.png").exists()
    for d in ("pressure", "buoyancy", "u3", "u1", "perturbvorticity", "vorticity"):
        assert Image.open(str(tmp_path / "images" / d / "0-0.png")).size == (w, h), d
    assert (tmp_path / "images" / "u2").is_dir() and not os.listdir(str(tmp_path / "images" / "u2"))
