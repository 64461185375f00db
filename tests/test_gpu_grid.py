"""The adaptive Grid (TreeCode_link/grid_maintenance.cpp) on the device
against the oracle's restatement: the same points in the same order (image
positions and cell sizes bit for bit, source positions and magnification
matrices to the FP64 shooter's 1e-9) after rounds of RefineLeaves over the
cells along a critical curve, RefineLeaf at the box's corners and edges, a
refinement whose points leave the source box, and the Einstein area."""
import numpy as np
import pytest

import oracle as O
import slsimlib_amd as S
from slsimlib_amd import workloads as W

pytestmark = pytest.mark.gpu

GRAV = 4.7788e-20


def _same(G, OG):
    x, y, A, gs = G.points()
    ox, oy, oA, ogs = OG.arrays()
    assert len(x) == len(ox)
    assert np.array_equal(x, ox) and np.array_equal(gs, ogs)
    scale = np.abs(oy).max() + np.abs(ox).max()
    assert np.abs(y - oy).max() <= 1e-9 * scale
    assert np.abs(A - oA).max() <= 1e-9 * max(1.0, np.abs(oA).max())
    i = G.info()
    assert list(i["image_p1"]) == OG.ip1 and list(i["image_p2"]) == OG.ip2
    return x, A, gs


def _point_mass(ctx, zl=0.5, zs=2.0, M=1e12):
    lens = S.Lens(ctx, zs)
    x0, m, h = np.array([0.0]), np.array([M], np.float32), np.array([1e-12], np.float32)
    lens.moveinMainHalo(S.LensHaloParticles(ctx, x0, x0, m, h, zl))
    L = O.Lens(zs)
    L.moveinMainHalo(O.Tree(x0, x0, m, h), zl)
    zl_f = float(np.float32(zl))
    Dl, Ds = O.coorDist(0, zl_f), O.coorDist(0, zs)
    th = np.sqrt(4 * GRAV * float(np.float32(M)) * (Ds - Dl) * (1 + zl_f) / (Dl * Ds))
    return lens, L, th


def test_grid_refinement_point_mass(ctx):
    lens, L, th = _point_mass(ctx)
    N, rg, c = 25, 3 * th, (1e-3 * th, -2e-3 * th)
    print("case: Grid point mass N=25", flush=True)
    G = S.Grid(lens, N, c, rg)
    OG = O.Grid(L, N, c, rg)
    x, A, gs = _same(G, OG)
    for rnd in range(3):
        r = np.hypot(x[:, 0], x[:, 1])
        near = np.nonzero(np.abs(r - th) < 1.5 * gs)[0]
        print("case: RefineLeaves round %d, %d cells" % (rnd, len(near)), flush=True)
        got = G.RefineLeaves(near)
        want = OG.refine_leaves(list(near))
        assert list(got) == want
        x, A, gs = _same(G, OG)
    # edges and corners of the image box
    print("case: RefineLeaf / RefineLeaves on the box edges", flush=True)
    for q in (0, N - 1, N * N - 1):
        assert list(G.RefineLeaf(q)) == OG.refine_leaf(q)
    edge = [N // 2, N * (N // 2), N * (N // 2) + N - 1]
    assert list(G.RefineLeaves(edge)) == OG.refine_leaves(edge)
    _same(G, OG)
    assert G.EinsteinArea() == pytest.approx(OG.einstein_area(), rel=1e-12)
    assert G.EinsteinArea() == pytest.approx(np.pi * th * th, rel=0.02)
    with pytest.raises(S.SlsimError):
        G.RefineLeaves([0, 0])
    G.close()
    lens.close()


def test_grid_source_box_and_rect(ctx):
    lens, L, th = _point_mass(ctx)
    # the centre cell of a small grid on the lens: its points map far away
    print("case: Grid source-box drops", flush=True)
    G = S.Grid(lens, 5, (0.0, 0.0), 1e-4 * th)
    OG = O.Grid(L, 5, (0.0, 0.0), 1e-4 * th)
    got, want = G.RefineLeaves([12]), OG.refine_leaves([12])
    assert list(got) == want and len(want) < 8
    assert list(G.RefineLeaf(12)) == OG.refine_leaf(12)
    _same(G, OG)
    G.close()
    # the rectangular constructor on the 65,536-particle NFW halo
    lens2, (px, py, pm, ph) = W.build_c1_lens(ctx, n=65536)
    L2 = O.Lens(W.C1["z_source"])
    L2.moveinMainHalo(O.Tree(px, py, pm, ph), W.C1["z_lens"])
    rg = W.grid_range(30.0)
    print("case: Grid.rect NFW halo 48 x 28", flush=True)
    R = S.Grid.rect(lens2, 48, (1e-6, 0.0), rg, 0.6 * rg)
    OR = O.Grid(L2, 48, (1e-6, 0.0), rg, 0.6 * rg)
    assert R.info()["ngrid_init2"] == 28
    x, A, gs = _same(R, OR)
    det = A[:, 0] * A[:, 3] - A[:, 1] * A[:, 2]
    ids = np.nonzero(np.abs(det) < 0.3)[0]
    assert len(ids) > 0
    assert list(R.RefineLeaves(ids)) == OR.refine_leaves(list(ids))
    _same(R, OR)
    assert R.EinsteinArea() == pytest.approx(OR.einstein_area(), rel=1e-12)
    R.close()
    lens2.close()
    lens.close()


def test_cpp_dropin_grid_program(ctx, tmp_path):
    """tests/cpp/dropin_grid.cpp: ImageFinding::find_crit and
    find_magnification_contour over a GridMap, and a Grid refined along the
    critical curve, from reference-style C++, against the oracle."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "tests", "cpp", "dropin_grid")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/dropin_grid is not built (run __graft_entry__.build())")
    out = tmp_path / "grid.bin"
    print("case: C++ dropin_grid program", flush=True)
    r = subprocess.run([exe, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, "exit %d: %s%s" % (r.returncode, r.stdout, r.stderr)
    v = np.fromfile(out, dtype=np.float64)
    n = 8192
    x, y = S.gen_nfw(1003, n, 1e13, 8.0, 0.5)
    m = np.full(n, np.float32(1e13 / n), np.float32)
    h = np.full(n, np.float32(1e-4), np.float32)
    L = O.Lens(2.0)
    L.moveinMainHalo(O.Tree(x, y, m, h), 0.5)
    rg = 30 * S.arcsecTOradians
    N = int(v[0])
    at = 1
    rays = v[at:at + 8 * N * N].reshape(-1, 8)
    at += 8 * N * N
    pos = O.gridmap_positions(N, (0.0, 0.0), rg)
    assert np.array_equal(rays[:, :2], pos)
    yr, Ar = L.shoot(pos)
    assert np.abs(rays[:, 2:4] - yr).max() <= 1e-12 * np.abs(yr).max()
    gy, gA = rays[:, 2:4].copy(), rays[:, 4:8].copy()
    want = O.image_finding_find_crit(L, N, pos, gy, gA, rg)
    nc = int(v[at])
    at += 1
    assert nc == len(want)
    for w in want:
        typ, hit, k = int(v[at]), bool(v[at + 1]), int(v[at + 2])
        cc, carea = v[at + 3:at + 5], v[at + 5]
        kc, karea, nout = v[at + 6:at + 8], v[at + 8], int(v[at + 9])
        at += 10
        crit = v[at:at + 2 * k].reshape(-1, 2)
        at += 2 * k
        caus = v[at:at + 2 * k].reshape(-1, 2)
        at += 2 * k
        outline = v[at:at + 2 * nout].reshape(-1, 2)
        at += 2 * nout
        assert typ == w["type"] and hit == w["touches_edge"] and k == len(w["critcurve"])
        assert np.array_equal(crit, w["critcurve"]) and np.array_equal(cc, w["critical_center"])
        assert carea == w["critical_area"]
        assert np.abs(caus - w["critcurve_y"]).max() <= 1e-9 * (np.abs(caus).max() + rg)
        assert np.array_equal(outline, O.tightest_hull(caus))
        ccen = np.zeros(2)
        for p in caus:
            ccen += p
        assert np.array_equal(kc, ccen / k)
        assert karea == O.windings(kc, outline)[1]
    c0, h0, _, _ = O.find_magnification_contour(N, pos, gA, rg, 0.0)
    ncont = int(v[at])
    at += 1
    assert ncont == len(c0)
    for i in range(ncont):
        hit, k = bool(v[at]), int(v[at + 1])
        at += 2
        cr = v[at:at + 4 * k].reshape(-1, 4)
        at += 4 * k
        assert hit == h0[i] and np.array_equal(cr[:, :2], c0[i])
        yw, _ = L.shoot(c0[i], zs=np.full(k, 2.0))
        assert np.abs(cr[:, 2:4] - yw).max() <= 1e-9 * (np.abs(yw).max() + rg)
    # the Grid
    OG = O.Grid(L, 25, (0.0, 0.0), rg)
    nref = int(v[at])
    ids = [int(q) for q in v[at + 1:at + 1 + nref]]
    at += 1 + nref
    OG.refine_leaves(ids)
    OG.refine_leaf(0)
    npts = int(v[at])
    at += 1
    P = v[at:at + 9 * npts].reshape(-1, 9)
    at += 9 * npts
    ox, oy, oA, ogs = OG.arrays()
    assert npts == len(ox)
    assert np.array_equal(P[:, :2], ox) and np.array_equal(P[:, 8], ogs)
    assert np.abs(P[:, 2:4] - oy).max() <= 1e-9 * (np.abs(oy).max() + rg)
    assert np.abs(P[:, 4:8] - oA).max() <= 1e-8
    assert v[at] == pytest.approx(OG.einstein_area(), rel=1e-12)
    assert at + 1 == v.size


def test_grids_as_server_clients(ctx):
    """Four grids refining at once from four threads, their rays sent as
    ray-list requests of one server (continuous batching): the same points,
    bit for bit, as the same grids refined alone with direct shoots."""
    import threading
    lens, L, th = _point_mass(ctx)
    centres = [(k * 0.1 * th, -k * 0.07 * th) for k in range(4)]

    def refine(G):
        for _ in range(3):
            x, _, _, gs = G.points()
            r = np.hypot(x[:, 0], x[:, 1])
            G.RefineLeaves(np.nonzero(np.abs(r - th) < 1.5 * gs)[0])
        G.RefineLeaf(0)

    alone = [S.Grid(lens, 21, c, 3 * th) for c in centres]
    for G in alone:
        refine(G)
    srv = S.Server(ctx, hbm_budget=4 << 30)
    served = [S.Grid(lens, 21, c, 3 * th) for c in centres]
    for G in served:
        G.set_server(srv)
    print("case: four Grids as server clients", flush=True)
    th_ = [threading.Thread(target=refine, args=(G,)) for G in served]
    for t in th_:
        t.start()
    for t in th_:
        t.join()
    for a, b in zip(alone, served):
        pa, pb = a.points(), b.points()
        assert all(np.array_equal(u, v) for u, v in zip(pa, pb))
    assert srv.stats()["completed"] >= 4 * 4
    for G in alone + served:
        G.close()
    srv.close()
    lens.close()
