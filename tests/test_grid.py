"""Pins the oracle's restatement of the adaptive Grid
(TreeCode_link/grid_maintenance.cpp): the initial positions (GridMap's
PositionFromIndex grid, pinned elsewhere), the sub-cell geometry of a
refinement (eight points at the parent +- gridsize/3, cell size / 3), the
points a refinement drops at the image box's edges and corners, and the
Einstein area of a point-mass lens converging on pi thetaE^2 as the cells
along the critical curve are refined.  CPU only.
"""
import numpy as np
import pytest

import oracle as O

GRAV = 4.7788e-20


def _point_lens(zl=0.5, zs=2.0, M=1e12):
    L = O.Lens(zs)
    L.moveinMainHalo(O.Tree(np.array([0.0]), np.array([0.0]), np.array([M], np.float32),
                            np.array([1e-12], np.float32)), zl)
    zl_f = float(np.float32(zl))
    Dl, Ds = O.coorDist(0, zl_f), O.coorDist(0, zs)
    thetaE2 = 4 * GRAV * float(np.float32(M)) * (Ds - Dl) * (1 + zl_f) / (Dl * Ds)
    return L, np.sqrt(thetaE2)


def test_initial_grid_and_boxes():
    L, th = _point_lens()
    N, rg, c = 17, 6 * th, (0.3 * th, -0.2 * th)
    G = O.Grid(L, N, c, rg)
    x, y, A, gs = G.arrays()
    assert np.array_equal(x, O.gridmap_positions(N, c, rg))
    assert np.all(gs == rg / (N - 1))
    assert G.ip1 == list(x.min(0)) and G.ip2 == list(x.max(0))
    assert G.sp1[0] == y[:, 0].min() - rg and G.sp2[1] == y[:, 1].max() + rg
    # the rectangular constructor: Ny = (int)(Nx rangeY / rangeX), square cells
    R = O.Grid(L, 16, c, rg, 0.5 * rg)
    xr, _, _, gr = R.arrays()
    assert len(xr) == 16 * 8 and np.all(gr == rg / 15)
    assert np.allclose(np.diff(np.unique(xr[:, 1])), rg / 15, rtol=1e-12)


def test_refinement_geometry():
    L, th = _point_lens()
    N, rg = 9, 6 * th
    G = O.Grid(L, N, (0.0, 0.0), rg)
    q = 4 + 9 * 4 + 1              # an interior point off the centre
    g0 = G.gs[q]
    new = G.refine_leaves([q])
    assert len(new) == 8
    x, _, _, gs = G.arrays()
    d = x[new] - x[q]
    want = np.array([[i % 3 - 1, i // 3 - 1] for i in range(9) if i != 4]) * g0 / 3
    assert np.allclose(d, want, rtol=0, atol=4 * np.spacing(rg))
    assert np.all(gs[new] == g0 * 2 / 3 / 2) and gs[q] == g0 / 3
    # RefineLeaves on an edge and a corner point: the points outside the
    # image box are dropped (5 and 3 stay), order kept
    edge, corner = 4, 0
    new = G.refine_leaves([edge, corner])
    assert len(new) == 5 + 3
    xn = G.arrays()[0][new]
    assert np.all(xn >= np.array(G.ip1)) and np.all(xn <= np.array(G.ip2))
    # RefineLeaf on the opposite corner keeps its three points
    assert len(G.refine_leaf(N * N - 1)) == 3


def test_source_box_drops_points():
    # cells next to the point mass map far outside the source box
    L, th = _point_lens()
    G = O.Grid(L, 5, (0.0, 0.0), 1e-4 * th)
    n0 = G.npoints()
    new = G.refine_leaves([12])           # the centre point, on the lens
    assert len(new) < 8 and G.npoints() == n0 + len(new)
    y = G.arrays()[1][new]
    assert np.all(y >= np.array(G.sp1)) and np.all(y <= np.array(G.sp2))


def test_einstein_area_converges_on_point_mass():
    L, th = _point_lens()
    N, rg = 25, 3 * th
    G = O.Grid(L, N, (0.0, 0.0), rg)
    area = np.pi * th * th
    errs = [abs(G.einstein_area() - area)]
    for _ in range(3):
        x, _, _, gs = G.arrays()
        r = np.hypot(x[:, 0], x[:, 1])
        near = np.nonzero(np.abs(r - th) < 1.5 * gs)[0]
        G.refine_leaves(list(near))
        errs.append(abs(G.einstein_area() - area))
    # the ring's area to a fraction of the coarse cells' band
    assert errs[-1] < 0.25 * errs[0] and errs[-1] < 0.02 * area


def test_refine_rejects_tiny_cells():
    L, th = _point_lens()
    G = O.Grid(L, 3, (1.0, 1.0), 1e-16)
    with pytest.raises(RuntimeError):
        G.refine_leaves([4])
