"""Pins the oracle's restatement of GridMap's critical-curve routines
(find_crit, find_magnification_contour, find_boundaries_of_caustics and the
Utilities contour walk, island fill, convex hull and inCurve under them)
against what fixes them independently: scipy's connected-component labelling
and convex hull, matplotlib's point-in-polygon test, the marching-squares
invariants, and the closed-form critical curves of the singular and cored
isothermal lenses.  CPU only.
"""
import numpy as np
import pytest
from scipy import ndimage
from scipy.spatial import ConvexHull

import oracle as O


def _blobs(n, seed, k=6, holes=True):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:n, 0:n]
    bm = np.zeros((n, n), bool)
    for _ in range(k):
        cx, cy = rng.uniform(3, n - 4, 2)
        a, b = rng.uniform(1.5, n / 6, 2)
        bm |= ((xx - cx) / a) ** 2 + ((yy - cy) / b) ** 2 < 1
    if holes:
        bm &= ~(((xx - n / 2) / 2.2) ** 2 + ((yy - n / 2) / 2.2) ** 2 < 1)
    return bm


def _padded(bm):
    p = bm.copy()
    p[0, :] = p[-1, :] = p[:, 0] = p[:, -1] = False
    return p


@pytest.mark.parametrize("seed", range(6))
def test_boundaries_marching_squares_invariants(seed):
    bm = _blobs(64, seed)
    curves, hits = O.find_boundaries(bm)
    p = _padded(bm)
    # contours = 8-connected components of the inside + 4-connected holes
    _, nin = ndimage.label(p, structure=np.ones((3, 3)))
    out_lab, nout = ndimage.label(~p)
    border = set(np.unique(np.concatenate([out_lab[0], out_lab[-1], out_lab[:, 0],
                                           out_lab[:, -1]])))
    assert len(curves) == nin + (nout - len(border - {0}))
    # every point is the midpoint of one inside/outside pixel pair, and every
    # such pair appears exactly once
    seen = set()
    for c in curves:
        for x, y in c:
            if x % 1 == 0.5:
                a, b = (int(y), int(x - 0.5)), (int(y), int(x + 0.5))
            else:
                a, b = (int(y - 0.5), int(x)), (int(y + 0.5), int(x))
            assert p[a] != p[b]
            assert (a, b) not in seen
            seen.add((a, b))
        # closed: consecutive points are in one cell
        d = np.abs(np.diff(np.vstack([c, c[:1]]), axis=0))
        assert np.all(d.max(1) <= 1.0) and np.all(d.sum(1) > 0)
    pairs = int(np.sum(p[:, 1:] != p[:, :-1]) + np.sum(p[1:, :] != p[:-1, :]))
    assert len(seen) == pairs
    assert hits.dtype == bool and len(hits) == len(curves)


def test_boundaries_hits_edge():
    bm = np.zeros((20, 20), bool)
    bm[5:9, 5:9] = True
    assert list(O.find_boundaries(bm)[1]) == [False]
    bm[5:9, 1:5] = True                     # reaches column 1, next to the padded edge
    assert list(O.find_boundaries(bm)[1]) == [True]
    assert O.find_boundaries(np.ones((10, 10), bool))[1].tolist() == [True]


@pytest.mark.parametrize("seed", range(5))
def test_islands_are_connected_components(seed):
    bm = _blobs(64, 100 + seed, holes=False)
    p = _padded(bm)
    lab, nlab = ndimage.label(p, structure=np.ones((3, 3)))
    _, nout = ndimage.label(~p)
    if nout != 1:
        pytest.skip("blobs enclose a hole")
    isl = O.find_islands(bm)
    curves, _ = O.find_boundaries(bm)
    assert len(isl) == nlab == len(curves)
    got = sorted(tuple(sorted(v)) for v in isl)
    want = sorted(tuple(np.flatnonzero(lab.ravel() == i + 1)) for i in range(nlab))
    assert got == want
    for v in isl:
        assert len(set(v)) == len(v)


def test_convex_hull_matches_scipy():
    rng = np.random.default_rng(3)
    for n in (4, 10, 100, 2000):
        pts = rng.normal(size=(n, 2))
        h = O.convex_hull(pts)
        assert set(h) == set(ConvexHull(pts).vertices)
        # counter-clockwise from the lexicographically smallest point
        q = pts[h]
        assert h[0] == np.lexsort((pts[:, 1], pts[:, 0]))[0]
        e1, e2 = np.roll(q, -1, 0) - q, np.roll(q, -2, 0) - np.roll(q, -1, 0)
        cr = e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]
        assert np.all(cr > 0)
    # points on the edges of a square are not vertices; three points or fewer
    # come back as they are
    g = np.array([[0, 0], [1, 0], [2, 0], [2, 1], [2, 2], [1, 2], [0, 2], [0, 1], [1, 1.0]])
    assert sorted(O.convex_hull(g)) == [0, 2, 4, 6]
    assert list(O.convex_hull(g[:3])) == [0, 1, 2]


def test_in_curve_matches_matplotlib():
    from matplotlib.path import Path
    rng = np.random.default_rng(5)
    t = np.sort(rng.uniform(0, 2 * np.pi, 40))
    r = 1 + 0.4 * np.sin(5 * t)
    poly = np.stack([r * np.cos(t), r * np.sin(t)], 1)
    q = rng.uniform(-1.6, 1.6, size=(3000, 2))
    assert np.array_equal(O.in_curve(poly, q), Path(poly).contains_points(q))
    assert not O.in_curve(poly[:2], q[:5]).any()


# --------------------------------------------------------- lens fields
def _grid(N, rng_):
    x = O.gridmap_positions(N, (0.0, 0.0), rng_)
    return x


def _isothermal(x, te, tc, centre=(0.0, 0.0)):
    """alpha and A = I - d alpha / d x of psi = te sqrt(|x - c|^2 + tc^2)."""
    d = x - np.asarray(centre)
    s = np.sqrt((d ** 2).sum(1) + tc * tc)
    alpha = te * d / s[:, None]
    A = np.zeros((len(x), 4))
    A[:, 0] = 1 - te * (1 / s - d[:, 0] ** 2 / s ** 3)
    A[:, 3] = 1 - te * (1 / s - d[:, 1] ** 2 / s ** 3)
    A[:, 1] = A[:, 2] = te * d[:, 0] * d[:, 1] / s ** 3
    return alpha, A


def _radius(c, centre=(0.0, 0.0)):
    return np.hypot(c[:, 0] - centre[0], c[:, 1] - centre[1])


def test_find_crit_singular_isothermal():
    N, rg, te = 96, 2.0, 0.5
    x = _grid(N, rg)
    a, A = _isothermal(x, te, 0.0)
    curves, hits, types, rep = O.find_crit(N, x, x - a, A, rg)
    res = rg / (N - 1)
    assert rep and list(types) == [O.CRIT_TANGENTIAL, O.CRIT_PSEUDO]
    assert np.all(np.abs(_radius(curves[0]) - te) <= res)
    # the pseudo curve: image positions of the hull of the island's sources,
    # which are the pixels nearest the centre (|y| = te - |x|)
    assert np.all(_radius(curves[1]) < 0.25 * te)
    assert list(hits) == [False, False]


def test_find_crit_cored_isothermal_and_order():
    N, rg = 160, 5.0
    x = _grid(N, rg)
    res = rg / (N - 1)
    # two cored lenses with radial curves, one singular lens (pseudo curve)
    lenses = [((-1.3, 1.2), 0.55, 0.12), ((1.2, -1.2), 0.5, 0.1), ((1.3, 1.4), 0.35, 0.0)]
    alpha = np.zeros_like(x)
    A = np.zeros((len(x), 4))
    A[:, 0] = A[:, 3] = 1.0
    for c, te, tc in lenses:
        a1, A1 = _isothermal(x, te, tc, c)
        alpha += a1
        A += A1
        A[:, 0] -= 1
        A[:, 3] -= 1
    curves, hits, types, rep = O.find_crit(N, x, x - alpha, A, rg)
    T, R, P = O.CRIT_TANGENTIAL, O.CRIT_RADIAL, O.CRIT_PSEUDO
    # the scan meets the lens at y = -1.2 first, then the two at y ~ 1.3
    assert list(types) == [T, R, T, R, T, P]
    order = [lenses[1], lenses[0], lenses[2]]
    for q, (c, te, tc) in enumerate(order):
        rt = np.sqrt(te * te - tc * tc)
        # the other lenses' convergence and shear move the curve by a few pixels
        assert np.median(np.abs(_radius(curves[2 * q], c) - rt)) <= 4 * res
        if tc > 0:
            rr = np.sqrt((te * tc * tc) ** (2 / 3) - tc * tc)
            assert np.max(np.abs(_radius(curves[2 * q + 1], c) - rr)) <= 2 * res
    assert len(hits) == len(curves)


def test_find_crit_without_critical_pixels_keeps_hits():
    N = 16
    x = _grid(N, 1.0)
    A = np.tile([1.0, 0.0, 0.0, 1.0], (N * N, 1))
    curves, hits, types, rep = O.find_crit(N, x, x, A, 1.0)
    assert curves == [] and not rep


def test_magnification_contour_circle():
    N, rg, te = 80, 2.0, 0.4
    x = _grid(N, rg)
    a, A = _isothermal(x, te, 0.0)
    res = rg / (N - 1)
    for thr in (-0.5, 0.0, 0.3):
        curves, hits, _, rep = O.find_magnification_contour(N, x, A, rg, thr)
        # det A = 1 - te / |x| < thr inside |x| = te / (1 - thr)
        assert rep and len(curves) == 1
        assert np.all(np.abs(_radius(curves[0]) - te / (1 - thr)) <= res)


def test_boundaries_of_caustics_isothermal():
    N, rg, te = 48, 2.0, 0.5
    x = _grid(N, rg)
    a, A = _isothermal(x, te, 0.0)
    curves, hits, mult = O.find_boundaries_of_caustics(N, x, x - a, A)
    # an SIS has two images inside |y| = te, one outside (pixel units here)
    res = rg / (N - 1)
    assert len(curves) == 1
    r = np.hypot(curves[0][:, 0] * res - rg / 2, curves[0][:, 1] * res - rg / 2)
    assert np.all(np.abs(r - te) <= 2 * res)
    # (sources whose image falls off the grid have none)
    r = np.hypot(x[:, 0], x[:, 1])
    assert np.all(mult[r > te + 2 * res] <= 1) and np.all(mult[r < te - 3 * res] >= 2)


# ------------------------------------------- the product's host utilities
# Utilities::find_boundaries, convex_hull and inCurve run on the host in the
# product too (the contour walk of find_crit); they equal the oracle's.
def test_product_find_boundaries_equals_oracle():
    import slsimlib_amd as S
    rng = np.random.default_rng(1)
    cases = [_blobs(64, s) for s in range(8)]
    cases += [rng.random((37, 53)) < rng.uniform(0.2, 0.8) for _ in range(12)]   # saddles
    cases += [np.ones((9, 9), bool), np.zeros((5, 7), bool), np.ones((1, 1), bool)]
    for bm in cases:
        a, ha = S.find_boundaries(bm)
        b, hb = O.find_boundaries(bm)
        assert len(a) == len(b) and np.array_equal(ha, hb)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_product_hull_and_in_curve_equal_oracle():
    import slsimlib_amd as S
    rng = np.random.default_rng(2)
    for n in (1, 3, 4, 50, 5000):
        p = rng.normal(size=(n, 2))
        assert np.array_equal(S.convex_hull(p), O.convex_hull(p))
    # ties and collinear points: a coarse lattice with repeats
    p = np.floor(rng.uniform(0, 6, size=(400, 2)))
    assert np.array_equal(S.convex_hull(p), O.convex_hull(p))
    t = np.linspace(0, 2 * np.pi, 30, endpoint=False)
    poly = np.stack([np.cos(t), np.sin(3 * t) * 0.5 + np.sin(t)], 1)
    q = np.concatenate([rng.uniform(-1.5, 1.5, size=(2000, 2)), poly, poly[::2] * [1.0, 0.0]])
    assert np.array_equal(S.in_curve(poly, q), O.in_curve(poly, q))


def test_product_find_boundaries_small_and_thin():
    import slsimlib_amd as S
    rng = np.random.default_rng(4)
    for shape in ((1, 1), (1, 7), (7, 1), (2, 2), (3, 3), (3, 9), (4, 2)):
        for _ in range(5):
            bm = rng.random(shape) < 0.6
            a, ha = S.find_boundaries(bm)
            b, hb = O.find_boundaries(bm)
            assert len(a) == len(b) and np.array_equal(ha, hb)
            for x, y in zip(a, b):
                assert np.array_equal(x, y)
    assert S.find_boundaries(np.zeros((0, 5), bool))[0] == []


def test_rectangular_grid_positions_and_crit():
    # GridMap(lens, Nx, center, rangeX, rangeY): a lattice of spacing
    # rangeX / (Nx - 1) from the lower-left corner, int(Nx rangeY / rangeX) rows
    nx, c, rx, ry = 96, (0.1, -0.05), 2.0, 1.2
    x, ny = O.gridmap_rect_positions(nx, c, rx, ry)
    res = rx / (nx - 1)
    assert ny == 57 and x.shape == (nx * ny, 2)
    assert np.allclose(np.diff(x.reshape(ny, nx, 2)[:, :, 0], axis=1), res, rtol=1e-12)
    assert np.allclose(np.diff(x.reshape(ny, nx, 2)[:, :, 1], axis=0), res, rtol=1e-12)
    assert tuple(x[0]) == (c[0] - 0.5 * rx, c[1] - 0.5 * ry)
    # an SIS in the middle: its tangential curve at the Einstein radius
    te = 0.35
    a, A = _isothermal(x, te, 0.0, centre=(0.11, -0.04))
    curves, hits, types, rep = O.find_crit(nx, x, x - a, A, rx, ny=ny)
    assert rep and list(types) == [O.CRIT_TANGENTIAL, O.CRIT_PSEUDO]
    assert np.all(np.abs(_radius(curves[0], (0.11, -0.04)) - te) <= res)
    cont, _, _, _ = O.find_magnification_contour(nx, x, A, rx, 0.2, ny=ny)
    assert len(cont) == 1
    assert np.all(np.abs(_radius(cont[0], (0.11, -0.04)) - te / 0.8) <= res)


# ------------------------------------------------- TightestHull and windings
def _seg_dist(p, v):
    """Distance of each point p (m, 2) to the closed polygon v (n, 2)."""
    a, b = v, np.roll(v, -1, axis=0)
    d = b - a
    t = np.clip(((p[:, None, :] - a[None]) * d[None]).sum(2) / (d * d).sum(1)[None], 0, 1)
    q = a[None] + t[..., None] * d[None]
    return np.sqrt(((q - p[:, None, :]) ** 2).sum(2)).min(1)


def _shoelace(v):
    x, y = v[:, 0], v[:, 1]
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


def test_windings_closed_forms():
    sq = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    assert O.windings((0.5, 0.5), sq) == (1, 1.0)
    assert O.windings((0.5, 0.5), sq[::-1]) == (-1, 1.0)
    assert O.windings((1.5, 0.5), sq)[0] == 0
    assert O.windings((0.2, 0.2), sq[:3]) == (1, 0.5)
    assert O.windings((0.0, 0.0), sq[:2]) == (0, 0.0)          # fewer than 3 points
    n, r = 37, 2.5
    ph = 2 * np.pi * np.arange(n) / n
    poly = np.stack([r * np.cos(ph) + 1, r * np.sin(ph) - 2], 1)
    wn, area = O.windings((1.0, -2.0), poly)
    assert wn == 1 and area == pytest.approx(0.5 * n * r * r * np.sin(2 * np.pi / n), rel=1e-13)
    # a curve wound twice about the point
    wn, _ = O.windings((1.0, -2.0), np.concatenate([poly, poly]))
    assert wn == 2


def test_tightest_hull_small_and_degenerate():
    p = np.array([[0.0, 0.0], [1.0, 2.0], [3.0, 1.0]])
    assert np.array_equal(O.tightest_hull(p), p)
    line = np.stack([np.full(7, 0.5), np.linspace(0, 1, 7)], 1)
    assert np.array_equal(O.tightest_hull(line), [[0.5, 0.0], [0.5, 1.0]])


@pytest.mark.parametrize("shape", ["circle", "star", "eight"])
def test_tightest_hull_outlines(shape):
    t = np.linspace(0, 2 * np.pi, 400, endpoint=False)
    if shape == "circle":
        v = np.stack([np.cos(t), np.sin(t)], 1) * 3 + [1, -1]
        inner = v
    elif shape == "star":
        r = 1 + 0.45 * np.cos(5 * t)
        v = np.stack([r * np.cos(t), r * np.sin(t)], 1)
        inner = v
    else:
        v = np.stack([np.sin(t), np.sin(t) * np.cos(t)], 1)   # a figure eight
        inner = None
    h = O.tightest_hull(v)
    assert len(h) > 50
    # every outline point lies on the curve
    span = np.ptp(v, 0).min()
    assert _seg_dist(h, v).max() <= 1e-12 * span
    if inner is not None:
        # a simple curve's outline encloses its area, to about a pixel of the
        # 1/100-extent bitmap times its perimeter
        per = np.sqrt((np.diff(np.vstack([v, v[:1]]), axis=0) ** 2).sum(1)).sum()
        assert abs(_shoelace(h) - _shoelace(v)) <= per * span / 100
    else:
        # the eight's outline goes round both lobes
        assert (h[:, 0] > 0.5).any() and (h[:, 0] < -0.5).any()
        assert _shoelace(h) == pytest.approx(4 / 3, rel=0.05)


def test_product_windings_equal_oracle():
    import slsimlib_amd as S
    rng = np.random.default_rng(3)
    for n in (3, 4, 17, 200):
        poly = rng.normal(size=(n, 2))
        for q in rng.normal(size=(40, 2)):
            assert S.windings(q, poly) == O.windings(q, poly)


def test_image_finding_find_crit_point_mass():
    """ImageFinding::find_crit over a GridMap of a point mass: the tangential
    curve is the Einstein ring (area pi thetaE^2 to a pixel times its
    perimeter, centre on the lens), its caustic the point y = 0 (every caustic
    point within the curve's offset from the ring: |y| ~ 2 |r - thetaE|)."""
    GRAV = 4.7788e-20
    zl, zs, M = 0.5, 2.0, 1e12
    L = O.Lens(zs)
    L.moveinMainHalo(O.Tree(np.array([0.0]), np.array([0.0]), np.array([M], np.float32),
                            np.array([1e-12], np.float32)), zl)
    zl_f = float(np.float32(zl))
    Dl, Ds = O.coorDist(0, zl_f), O.coorDist(0, zs)
    te = np.sqrt(4 * GRAV * float(np.float32(M)) * (Ds - Dl) * (1 + zl_f) / (Dl * Ds))
    N, rg, c = 101, 3 * te, (0.01 * te, -0.013 * te)
    pos = O.gridmap_positions(N, c, rg)
    y, A = L.shoot(pos)
    out = O.image_finding_find_crit(L, N, pos, y, A, rg)
    res = rg / (N - 1)
    tang = [o for o in out if o["type"] == O.CRIT_TANGENTIAL]
    assert len(tang) == 1
    t = tang[0]
    r = _radius(t["critcurve"])
    assert np.all(np.abs(r - te) <= res)
    assert abs(t["critical_area"] - np.pi * te * te) <= 2 * np.pi * te * res
    assert np.hypot(*t["critical_center"]) <= res
    yc = np.hypot(t["caustic_curve_intersecting"][:, 0], t["caustic_curve_intersecting"][:, 1])
    assert np.all(yc <= 2.2 * np.abs(r - te) + 1e-6 * te)
    assert np.hypot(*t["caustic_center"]) <= 2 * res
    assert t["caustic_area"] <= np.pi * (2 * res) ** 2
    # the outline is drawn from the caustic points
    assert len(t["caustic_curve_outline"]) >= 1
    assert t["z_source"] == zs
