"""GridMap::find_crit, find_magnification_contour and
find_boundaries_of_caustics on the device against the oracle's restatement
on the same grid outputs: every curve point, curve type, hits_boundary flag
and multiplicity equal.  Cases: the 65,536-particle NFW halo's grid, three
isothermal lenses (tangential, radial and pseudo curves, radial curves moved
after their tangential ones), a field of matrices whose smaller eigenvalue
is zero to the last bit (the eigenvalue rounding), and a source grid with a
non-finite position."""
import numpy as np
import pytest

import oracle as O
import slsimlib_amd as S
from slsimlib_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def shot(ctx):
    lens, _ = W.build_c1_lens(ctx, n=65536)
    N, rg = 256, W.grid_range(30.0)
    g = S.GridMap(lens, N, (0.0, 0.0), rg, S.SLSIM_FP64)
    yield g, g.positions(), g.y(), g.A()
    g.close()


class _Field:
    """Grid outputs uploaded from the host, viewed as a GridMap."""

    def __init__(self, ctx, N, centre, range_, y, A, ny=None, range_y=None):
        self.ctx, self.N, self.center, self.range = ctx, N, tuple(centre), range_
        self.ny = N if ny is None else ny
        self.range_y = range_ if range_y is None else range_y
        self.layout = S._abi.SLSIM_GRID_SQUARE if ny is None else S._abi.SLSIM_GRID_RECT
        self.y_dev = ctx.alloc(16 * N * self.ny)
        self.A_dev = ctx.alloc(32 * N * self.ny)
        ctx.h2d(self.y_dev, np.ascontiguousarray(y, dtype=np.float64))
        ctx.h2d(self.A_dev, np.ascontiguousarray(A, dtype=np.float64))

    _grid = S.GridMap._grid
    find_crit = S.GridMap.find_crit
    find_magnification_contour = S.GridMap.find_magnification_contour
    find_boundaries_of_caustics = S.GridMap.find_boundaries_of_caustics

    def close(self):
        self.ctx.free(self.y_dev)
        self.ctx.free(self.A_dev)


def _same(got, want):
    c1, h1, t1, r1 = got
    c2, h2, t2, r2 = want
    assert r1 == r2
    assert len(c1) == len(c2)
    assert np.array_equal(h1, h2)
    assert np.array_equal(t1, t2)
    for a, b in zip(c1, c2):
        assert np.array_equal(a, b)


def _isothermal(x, lenses):
    alpha = np.zeros_like(x)
    A = np.zeros((len(x), 4))
    A[:, 0] = A[:, 3] = 1.0
    for c, te, tc in lenses:
        d = x - np.asarray(c)
        s = np.sqrt((d ** 2).sum(1) + tc * tc)
        alpha += te * d / s[:, None]
        A[:, 0] -= te * (1 / s - d[:, 0] ** 2 / s ** 3)
        A[:, 3] -= te * (1 / s - d[:, 1] ** 2 / s ** 3)
        A[:, 1] += te * d[:, 0] * d[:, 1] / s ** 3
        A[:, 2] += te * d[:, 0] * d[:, 1] / s ** 3
    return x - alpha, A


def test_find_crit_particle_halo(shot):
    g, x, y, A = shot
    print("case: find_crit NFW particle halo N=256", flush=True)
    got = g.find_crit()
    _same(got, O.find_crit(g.N, x, y, A, g.range))
    assert S.SLSIM_CRIT_TANGENTIAL in got[2]


def test_find_crit_isothermal_lenses(ctx):
    N, rg, centre = 160, 5.0, (0.25, -0.5)
    x = O.gridmap_positions(N, centre, rg)
    lenses = [((-1.05, 0.7), 0.55, 0.12), ((1.45, -1.7), 0.5, 0.1), ((1.55, 0.9), 0.35, 0.0)]
    y, A = _isothermal(x, lenses)
    f = _Field(ctx, N, centre, rg, y, A)
    print("case: find_crit three isothermal lenses N=160", flush=True)
    got = f.find_crit()
    _same(got, O.find_crit(N, x, y, A, rg))
    T, R, P = S.SLSIM_CRIT_TANGENTIAL, S.SLSIM_CRIT_RADIAL, S.SLSIM_CRIT_PSEUDO
    assert list(got[2]) == [T, R, T, R, T, P]
    # no critical pixel: hits_boundary is left to the caller
    one = np.tile([1.0, 0.0, 0.0, 1.0], (N * N, 1))
    f2 = _Field(ctx, N, centre, rg, x, one)
    c, h, t, rep = f2.find_crit()
    assert c == [] and not rep
    f.close()
    f2.close()


def test_find_crit_eigenvalue_rounding(ctx):
    # lambda_1 = m - sqrt(|m^2 - det|) within an ulp of zero on every pixel:
    # the classification follows the rounding of the square root
    N = 256
    rng = np.random.default_rng(9)
    n = N * N
    m = rng.uniform(0.2, 2.0, n)
    d = rng.uniform(-1.0, 1.0, n) * m
    off = np.sqrt(m * m - d * d) * (1 + rng.integers(-3, 4, n) * 2.0 ** -52)
    A = np.stack([m + d, off, off, m - d], 1)
    flip = rng.random(n) < 0.3                       # some pixels radial too
    A[flip] *= -1
    x = O.gridmap_positions(N, (0.0, 0.0), 1.0)
    y = rng.normal(size=(n, 2))
    f = _Field(ctx, N, (0.0, 0.0), 1.0, y, A)
    print("case: find_crit eigenvalue rounding N=256", flush=True)
    got = f.find_crit()
    want = O.find_crit(N, x, y, A, 1.0)
    _same(got, want)
    assert len(got[0]) > 100
    f.close()


@pytest.mark.parametrize("invmag", [-0.3, 0.0, 0.5])
def test_magnification_contour(shot, invmag):
    g, x, y, A = shot
    print("case: find_magnification_contour invmag=%g" % invmag, flush=True)
    _same(g.find_magnification_contour(invmag),
          O.find_magnification_contour(g.N, x, A, g.range, invmag))


def test_boundaries_of_caustics_particle_halo(ctx):
    lens, _ = W.build_c1_lens(ctx, n=65536)
    N, rg = 96, W.grid_range(6.0)            # a 6" field: positions with up to ~13 images
    g = S.GridMap(lens, N, (0.0, 0.0), rg, S.SLSIM_FP64)
    x, y, A = g.positions(), g.y(), g.A()
    print("case: find_boundaries_of_caustics NFW particle halo N=96 6 arcsec", flush=True)
    c, h, mult = g.find_boundaries_of_caustics()
    c0, h0, m0 = O.find_boundaries_of_caustics(N, x, y, A)
    assert np.array_equal(mult, m0)
    assert mult.max() >= 3
    assert np.array_equal(h, h0) and len(c) == len(c0)
    for a, b in zip(c, c0):
        assert np.array_equal(a, b)
    g.close()


def test_boundaries_of_caustics_singular_and_nonfinite(ctx):
    N, rg = 64, 2.0
    x = O.gridmap_positions(N, (0.0, 0.0), rg)
    y, A = _isothermal(x, [((0.013, -0.021), 0.5, 0.0)])
    y[1000] = np.nan                                 # a corner the sign tests see as zeros
    f = _Field(ctx, N, (0.0, 0.0), rg, y, A)
    print("case: find_boundaries_of_caustics SIS with a NaN source position N=64", flush=True)
    c, h, mult = f.find_boundaries_of_caustics()
    c0, h0, m0 = O.find_boundaries_of_caustics(N, x, y, A)
    assert np.array_equal(mult, m0)
    assert np.array_equal(h, h0) and len(c) == len(c0)
    for a, b in zip(c, c0):
        assert np.array_equal(a, b)
    f.close()


def test_find_images_nonfinite_corner_and_spread_sources(ctx):
    # GridMap::find_images with a NaN grid position (its triangles test every
    # source) and sources spread over and beyond the grid, one exactly on a
    # grid vertex: equal to the oracle's all-pairs search
    N, rg = 96, 2.0
    x = O.gridmap_positions(N, (0.0, 0.0), rg)
    y, A = _isothermal(x, [((0.01, -0.02), 0.4, 0.05)])
    y[2 * N + 7] = (np.nan, 0.0)
    f = _Field(ctx, N, (0.0, 0.0), rg, y, A)
    f.find_images = S.GridMap.find_images.__get__(f)
    rng = np.random.default_rng(21)
    ys = np.concatenate([rng.uniform(-1.5, 1.5, size=(300, 2)), y[[500, 4000]],
                         rng.normal(size=(50, 2)) * 0.05])
    print("case: find_images NaN grid corner, 352 sources", flush=True)
    rays, mult, tris = f.find_images(ys, 64)
    r0, m0, t0 = O.find_images(N, x, y, A, ys)
    assert np.array_equal(mult, m0) and np.array_equal(tris, t0)
    assert np.array_equal(rays, r0, equal_nan=True)
    assert mult.max() >= 3
    f.close()


@pytest.mark.parametrize("N", [1, 2, 3])
def test_tiny_grids(ctx, N):
    # grids too small to hold a boundary: no curves; find_crit still reports
    # whether it replaced hits_boundary (any critical pixel, edges included)
    x = O.gridmap_positions(N, (0.0, 0.0), 1.0)
    y, A = _isothermal(x, [((0.01, 0.02), 2.0, 0.0)])
    f = _Field(ctx, N, (0.0, 0.0), 1.0, y, A)
    f.find_images = S.GridMap.find_images.__get__(f)
    print("case: tiny grid N=%d" % N, flush=True)
    _same(f.find_crit(), O.find_crit(N, x, y, A, 1.0))
    _same(f.find_magnification_contour(0.0), O.find_magnification_contour(N, x, A, 1.0, 0.0))
    c, h, mult = f.find_boundaries_of_caustics()
    c0, h0, m0 = O.find_boundaries_of_caustics(N, x, y, A)
    assert np.array_equal(mult, m0) and len(c) == len(c0) == 0
    ys = np.array([[0.0, 0.0], [0.3, -0.2]])
    rays, mult, tris = f.find_images(ys, 16)
    r0, m1, t0 = O.find_images(N, x, y, A, ys)
    assert np.array_equal(mult, m1) and np.array_equal(rays, r0) and np.array_equal(tris, t0)
    f.close()


def test_rectangular_grids(ctx):
    # GridMap(lens, Nx, center, rangeX, rangeY): images, critical curves,
    # contours and caustic boundaries on an Nx x Ny grid equal the oracle's
    nx, c, rx, ry = 120, (0.05, -0.1), 4.0, 2.6
    x, ny = O.gridmap_rect_positions(nx, c, rx, ry)
    lenses = [((-0.9, 0.2), 0.5, 0.1), ((0.9, -0.4), 0.35, 0.0)]
    y, A = _isothermal(x, lenses)
    f = _Field(ctx, nx, c, rx, y, A, ny=ny, range_y=ry)
    f.find_images = S.GridMap.find_images.__get__(f)
    print("case: rectangular grid %d x %d" % (nx, ny), flush=True)
    got = f.find_crit()
    _same(got, O.find_crit(nx, x, y, A, rx, ny=ny))
    T, R, P = S.SLSIM_CRIT_TANGENTIAL, S.SLSIM_CRIT_RADIAL, S.SLSIM_CRIT_PSEUDO
    assert sorted(got[2]) == sorted([T, R, T, P])
    _same(f.find_magnification_contour(0.3), O.find_magnification_contour(nx, x, A, rx, 0.3,
                                                                          ny=ny))
    cb, hb, mb = f.find_boundaries_of_caustics()
    c0, h0, m0 = O.find_boundaries_of_caustics(nx, x, y, A, ny=ny)
    assert np.array_equal(mb, m0) and np.array_equal(hb, h0) and len(cb) == len(c0)
    for a, b in zip(cb, c0):
        assert np.array_equal(a, b)
    rng = np.random.default_rng(8)
    ys = y[rng.integers(0, nx * ny, 300)] + rng.normal(size=(300, 2)) * 1e-3
    rays, mult, tris = f.find_images(ys, 32)
    r0, m1, t0 = O.find_images(nx, x, y, A, ys, ny=ny)
    assert np.array_equal(mult, m1) and np.array_equal(tris, t0) and np.array_equal(rays, r0)
    assert mult.max() >= 3
    f.close()


def test_rectangular_gridmap_shoot(ctx, shot):
    lens = shot[0].lens
    rg = W.grid_range(30.0)
    g = S.GridMap.rect(lens, 64, (1e-6, 0.0), rg, 0.55 * rg)
    x, ny = O.gridmap_rect_positions(64, (1e-6, 0.0), rg, 0.55 * rg)
    assert g.ny == ny and np.array_equal(g.positions(), x)
    # the same rays as a ray list (rayshooterInternal): the grid's outputs bit for bit
    yl, Al = lens.rayshooterInternal(x)
    assert np.array_equal(g.y(), yl) and np.array_equal(g.A(), Al)
    assert g.writePixelMap("KAPPA").shape == (ny, 64)
    got = g.find_crit()
    _same(got, O.find_crit(64, x, g.y(), g.A(), rg, ny=ny))
    g.close()


def _hull_cases():
    rng = np.random.default_rng(21)
    t = np.linspace(0, 2 * np.pi, 500, endpoint=False)
    yield np.stack([np.cos(t), np.sin(t)], 1) * 2e-6 + 1e-5      # radians-sized circle
    r = 1 + 0.45 * np.cos(5 * t)
    yield np.stack([r * np.cos(t), r * np.sin(t)], 1)
    yield np.stack([np.sin(t), 0.5 * np.sin(2 * t)], 1)            # a figure eight
    # an astroid (four cusps, the caustic of an elliptical lens) with noise
    a = np.stack([np.cos(t) ** 3, np.sin(t) ** 3], 1) * 0.3
    yield a + rng.normal(size=a.shape) * 1e-3
    yield rng.normal(size=(60, 2))                                  # a scribble
    yield np.array([[0.0, 0.0], [1.0, 1.0], [2.0, 0.5]])            # three points
    yield np.stack([np.full(9, 0.25), np.linspace(0, 1, 9)], 1)     # zero width


def test_tightest_hull_equals_oracle(ctx):
    for k, v in enumerate(_hull_cases()):
        print("case: tightest hull %d (%d points)" % (k, len(v)), flush=True)
        assert np.array_equal(S.tightest_hull(ctx, v), O.tightest_hull(v)), k


def test_image_finding_find_crit(ctx):
    """ImageFinding::find_crit (find_crit.cpp:1036-1096) on the 65,536-particle
    NFW halo: the critical curves of the device grid and their rays equal the
    oracle's composition on the same grid outputs (curve points, types and
    centres bit for bit, caustic points to the FP64 shooter's 1e-9, the
    outline and areas of the device's caustic points bit for bit)."""
    lens, (px, py, pm, ph) = W.build_c1_lens(ctx, n=65536)
    L = O.Lens(W.C1["z_source"])
    L.moveinMainHalo(O.Tree(px, py, pm, ph), W.C1["z_lens"])
    N, rg = 128, W.grid_range(30.0)
    g = S.GridMap(lens, N, (0.0, 0.0), rg, S.SLSIM_FP64)
    print("case: ImageFinding::find_crit NFW halo N=128", flush=True)
    got = S.ImageFinding.find_crit(lens, g)
    want = O.image_finding_find_crit(L, N, g.positions(), g.y(), g.A(), rg)
    assert len(got) == len(want) > 0
    assert S.SLSIM_CRIT_TANGENTIAL in [c.type for c in got]
    for c, w in zip(got, want):
        assert c.type == w["type"] and c.touches_edge == w["touches_edge"]
        assert c.z_source == w["z_source"]
        assert np.array_equal(c.critcurve_x, w["critcurve"])
        assert np.array_equal(c.critical_center, w["critical_center"])
        assert c.critical_area == w["critical_area"]
        scale = np.abs(w["critcurve_y"]).max() + np.abs(w["critcurve"]).max()
        assert np.abs(c.caustic_curve_intersecting - w["critcurve_y"]).max() <= 1e-9 * scale
        assert np.allclose(c.critcurve_A, w["critcurve_A"], rtol=0, atol=1e-9)
        assert np.array_equal(c.caustic_curve_outline,
                              O.tightest_hull(c.caustic_curve_intersecting))
        assert c.caustic_area == O.windings(c.caustic_center, c.caustic_curve_outline)[1]
        assert c.caustic_intersections == -1
    # the tangential curves enclose the pixels whose smaller eigenvalue is
    # negative (particle-scale islands included): their areas are positive
    # and do not exceed those pixels' area
    res = rg / (N - 1)
    A = g.A()
    m = 0.5 * (A[:, 0] + A[:, 3])
    det = A[:, 0] * A[:, 3] - A[:, 1] * A[:, 2]
    inside = (m - np.sqrt(np.abs(m * m - det)) < 0).sum() * res * res
    tan = [c.critical_area for c in got if c.type == S.SLSIM_CRIT_TANGENTIAL]
    assert min(tan) > 0 and sum(tan) <= inside
    # the magnification contours as rays to the source
    cont, hits = S.ImageFinding.find_magnification_contour(lens, g, 0.0)
    c0, h0, _, _ = O.find_magnification_contour(N, g.positions(), g.A(), rg, 0.0)
    assert len(cont) == len(c0) and np.array_equal(hits, h0)
    for (x, y, Am), xw in zip(cont, c0):
        assert np.array_equal(x, xw)
        yw, Aw = L.shoot(xw, zs=np.full(len(xw), W.C1["z_source"]))
        assert np.abs(y - yw).max() <= 1e-9 * (np.abs(yw).max() + np.abs(xw).max())
    g.close()
    lens.close()
