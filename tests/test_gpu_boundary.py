"""The rest of the drop-in boundary from a reference-style C++ program
(tests/cpp/dropin_boundary.cpp): LensHaloParticles on a borrowed array with
recentring and rotation (particle_halo.h:78-93, 341-417, 634-680), GridMap's
unlensed constructor, ReInitialize, move and copy (gridmap.cpp:122-160,
gridmap.h:139-180) and the uniform pixel maps and FITS files
(gridmap.h:98-116, 506-740), against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from oracle import formats as F
import slsimlib_amd as S

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_dropin_boundary_program(ctx, tmp_path):
    exe = os.path.join(ROOT, "tests", "cpp", "dropin_boundary")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/dropin_boundary is not built (run __graft_entry__.build())")
    out = tmp_path / "boundary.bin"
    base = str(tmp_path / "uniform")
    r = subprocess.run([exe, str(out), base], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, "exit %d: %s%s" % (r.returncode, r.stdout, r.stderr)
    v = np.fromfile(out, dtype=np.float64)
    n = int(v[0])
    at = 1
    xyz = v[at:at + 3 * n].reshape(n, 3)
    at += 3 * n
    cm, dp = v[at:at + 3], v[at + 3:at + 6]
    at += 6
    # the same particles through the oracle's set_up restatement
    x, y = S.gen_nfw(1002, n, 1e13, 8.0, 0.5)
    p = np.stack([x + 1e-3, y - 2e-3, 1e-4 * np.sin(np.arange(n, dtype=np.float64))], 1)
    m = np.full(n, np.float32(1e13 / n), np.float32)
    h = (1e-4 * (1.0 + (np.arange(n) % 7) / 7.0)).astype(np.float32)
    q, mc, dpo, _ = F.particle_setup(p, m, h, recenter=False, theta_rotate=(0.3, -0.2))
    assert np.array_equal(xyz, q)
    assert np.array_equal(cm, mc) and np.array_equal(dp, dpo)
    N = 48
    grid = v[at:at + 8 * N * N].reshape(-1, 8)
    at += 8 * N * N
    L = O.Lens(2.0)
    L.moveinMainHalo(O.Tree(q[:, 0].copy(), q[:, 1].copy(), m, h), 0.5)
    pos = O.gridmap_positions(N, (0.0, 0.0), 30 * S.arcsecTOradians)
    assert np.array_equal(grid[:, :2], pos)
    yr, Ar = L.shoot(pos)
    assert np.abs(grid[:, 2:4] - yr).max() <= 1e-12 * np.abs(yr).max()
    assert np.abs(grid[:, 4:8] - Ar).max() <= 1e-8
    # ReInitialize after ResetSourcePlane(1.5): the rectangular constructor's
    # positions, shot to the new source
    re = v[at:at + 8 * N * N].reshape(-1, 8)
    at += 8 * N * N
    rg = 30 * S.arcsecTOradians
    rp, _ = O.gridmap_rect_positions(N, (0.0, 0.0), rg, rg)
    assert np.array_equal(re[:, :2], rp)
    L2 = O.Lens(1.5)
    L2.moveinMainHalo(O.Tree(q[:, 0].copy(), q[:, 1].copy(), m, h), 0.5)
    yr, Ar = L2.shoot(rp)
    assert np.abs(re[:, 2:4] - yr).max() <= 1e-12 * np.abs(yr).max()
    assert np.abs(re[:, 4:8] - Ar).max() <= 1e-8
    # the uniform alpha1 map: 20 x 13 pixels of range / 19 around (2e-6, -3e-6)
    nx, ny = 20, 13
    umap = v[at:at + nx * ny].reshape(ny, nx)
    assert at + nx * ny == v.size
    A = re[:, 4:8]
    alpha1 = re[:, 0] - re[:, 2]
    ref = F.pixelmap_uniform(re[:, :2], alpha1, (2e-6, -3e-6), nx, ny, rg / (nx - 1))
    assert np.array_equal(umap, ref)
    assert np.count_nonzero(umap) > nx * ny // 2
    # its FITS file (tagged name) holds the same pixels
    cards, img = F.fits_parse(open(base + ".alpha1.fits", "rb").read())
    assert np.array_equal(img.astype(np.float64), umap)
    cards, img = F.fits_parse(open(base + ".invmag_grid.fits", "rb").read())
    invmag = (A[:, 0] * A[:, 3] - A[:, 1] * A[:, 2]).astype(np.float32)
    assert img.shape == (N, N) and np.array_equal(img.reshape(-1), invmag)
