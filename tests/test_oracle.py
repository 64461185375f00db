"""Pins the CPU oracle (oracle/glamer_oracle.cpp) against closed forms,
invariants, brute force and the scipy distance table.  CPU only.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a sign, a wrong coefficient or index) fails it; none re-types the
oracle's formula.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ARCSEC = 0.000004848136811
GRAV = 4.7788e-20


# ------------------------------------------------------------------ distances
def test_coordist_matches_scipy_table():
    tab = json.load(open(os.path.join(HERE, "golden", "flat_lcdm_distances.json")))
    for r in tab["rows"]:
        v = O.coorDist(r["z1"], r["z2"], tab["Om"], tab["OL"], tab["h"])
        assert abs(v - r["Dc"]) <= 1e-11 * abs(r["Dc"]), r


# ------------------------------------------------------------------ B-spline
def _bs(q, h=1e-3, m=1.0):
    r = q * h
    return O.bspline(np.array([r, 0.0]), r, m, h)


def _sigma(q, h=1e-3):
    return _bs(q, h)[2]          # kappa correction = m * sigma


def _m_enclosed_true(q, h=1e-3):
    from scipy.integrate import quad
    v, _ = quad(lambda qq: 2 * np.pi * qq * h * h * _sigma(qq, h), 0, q, epsabs=1e-14,
                epsrel=1e-12, points=[1.0] if q > 1 else None, limit=200)
    return v


def _m_code(q, h=1e-3):
    # total alpha_x = point mass + correction = -m M(q) / (pi r) in the tree's sign
    r = q * h
    a_corr = _bs(q, h)[0]
    a_pm = -1.0 / (np.pi * r)
    return -(a_pm + a_corr) * np.pi * r


def test_bspline_normalisation():
    # the 2-D cubic spline carries the particle's whole mass inside 2h
    assert abs(_m_enclosed_true(2.0) - 1.0) < 1e-10


def test_bspline_continuity():
    scale = np.abs(_bs(1.0)).max()
    for q0 in (1.0, 2.0):
        lo, hi = _bs(q0 - 1e-9), _bs(q0 + 1e-9)
        assert np.allclose(lo, hi, rtol=0, atol=1e-7 * scale)


def test_bspline_outer_branch_obeys_gauss_law():
    # 1 < q <= 2: enclosed mass implied by alpha equals the integral of kappa
    for q in (1.0, 1.2, 1.5, 1.9, 2.0):
        assert abs(_m_code(q) - _m_enclosed_true(q)) < 1e-9


def test_bspline_inner_branch_carries_reference_term():
    # q <= 1: the reference's M(q) = 10q^2(1 - 3q/4 + 3q^3/10)/7
    # (include/quadTree.h:165) departs from the integral of its own sigma by
    # exactly (15/14) q^3 (q - 1); reproduced on purpose for parity.
    for q in (0.05, 0.2, 0.5, 0.8, 0.99):
        d = _m_code(q) - _m_enclosed_true(q)
        assert abs(d - 15.0 / 14.0 * q ** 3 * (q - 1)) < 1e-9, q


def test_bspline_tangential_shear_identity():
    # gamma_t = mean kappa inside r - kappa(r) for the smoothed particle,
    # with the inner branch's enclosed mass taken from alpha (same quirk)
    h = 1e-3
    for q in (0.3, 0.7, 1.3, 1.8):
        r = q * h
        out = _bs(q, h)
        g0_total = -1.0 / (np.pi * r * r) + out[3]   # point mass + correction
        gt = -g0_total
        expect = _m_code(q, h) / (np.pi * r * r) - out[2]
        assert abs(gt - expect) <= 1e-9 * abs(expect)


# ----------------------------------------------------------------- tree build
def _nfw_like(n, seed, scale=0.05):
    rng = np.random.default_rng(seed)
    r = scale * rng.random(n) ** 2
    ph = 2 * np.pi * rng.random(n)
    return r * np.cos(ph), r * np.sin(ph)


def _children(q):
    """Child lists from the depth-first export (levels)."""
    kids = [[] for _ in range(len(q))]
    stack = []
    for i, lev in enumerate(q[:, 0]):
        while stack and q[stack[-1], 0] >= lev:
            stack.pop()
        if stack:
            kids[stack[-1]].append(i)
        stack.append(i)
    return kids


def test_tree_invariants():
    n = 5000
    x, y = _nfw_like(n, 1)
    m = np.full(n, 1e8, np.float32)
    h = np.full(n, 1e-4, np.float32)
    T = O.Tree(x, y, m, h, theta=0.1, bucket=20)
    d, q = T.export_nodes()
    kids = _children(q)
    leaves = q[:, 3] == 1
    assert q[leaves, 1].sum() == n
    assert np.all(q[leaves, 1] <= 20)
    for i, ks in enumerate(kids):
        if ks:
            assert len(ks) == 4 and q[i, 1] > 20
            assert q[ks, 1].sum() == q[i, 1]
            assert abs(d[ks, 6].sum() - d[i, 6]) <= 1e-12 * d[i, 6]
            # children tile the parent box
            assert np.isclose(d[ks[0], 0], 0.5 * (d[i, 0] + d[i, 2]))
    occ = q[:, 1] > 0
    assert np.all(d[occ, 4] >= d[occ, 0] - 1e-15) and np.all(d[occ, 4] <= d[occ, 2] + 1e-15)
    assert np.all(d[occ, 5] >= d[occ, 1] - 1e-15) and np.all(d[occ, 5] <= d[occ, 3] + 1e-15)
    # every particle is a big particle of exactly one cell here (h < leaf boxes)
    assert q[:, 2].sum() == n


def test_tree_hand_built_kat():
    # bounding box [0,4]x[0,2] -> squared to [0,4]x[0,4]; centre (2,2).
    # 7 particles in (x<=2, y<=2) incl. two on the split lines (they go low),
    # 6 in (x>2, y<=2), 5 in (x<=2, y>2), 3 in (x>2, y>2): 21 > bucket 20.
    pts = [(0.0, 0.0), (2.0, 1.0), (1.0, 2.0), (0.5, 0.5), (1.5, 0.5), (0.5, 1.5), (1.5, 1.5),
           (2.5, 0.5), (3.5, 0.5), (2.5, 1.5), (3.5, 1.5), (3.0, 1.0), (4.0, 0.1),
           (0.5, 2.0001), (1.5, 2.5), (1.0, 2.5), (2.0, 2.2), (0.1, 2.1),
           (2.5, 2.0001), (3.0, 2.0001), (3.5, 2.0001)]
    # shift y so that the maximum is 2 and the minimum 0
    x = np.array([p[0] for p in pts])
    y = np.array([p[1] for p in pts])
    y = y * (2.0 / y.max())
    n = len(pts)
    T = O.Tree(x, y, np.ones(n, np.float32), np.full(n, 1e-6, np.float32), bucket=20)
    d, q = T.export_nodes()
    assert q[0, 1] == 21 and q[0, 3] == 0
    assert tuple(d[0, :4]) == (0.0, 0.0, 4.0, 4.0)
    kids = _children(q)[0]
    # child order 0=(x>,y>) 1=(x<=,y>) 2=(x>,y<=) 3=(x<=,y<=)
    xc, yc = 2.0, 2.0
    expect = [np.sum((x > xc) & (y > yc)), np.sum((x <= xc) & (y > yc)),
              np.sum((x > xc) & (y <= yc)), np.sum((x <= xc) & (y <= yc))]
    assert list(q[kids, 1]) == expect
    assert tuple(d[kids[0], :4]) == (2.0, 2.0, 4.0, 4.0)
    assert tuple(d[kids[3], :4]) == (0.0, 0.0, 2.0, 2.0)
    # one more particle in a leaf bucket -> leaf
    assert all(q[k, 3] == 1 for k in kids)


def test_tree_big_particle_bands():
    # particles larger than half the root box are listed at the root; the
    # rest at the cell whose band box/2 < h <= box contains them, or at the
    # leaf if h < leaf box (include/quadTree.h:379-431)
    n = 400
    x, y = _nfw_like(n, 3, scale=1.0)
    h = np.full(n, 1e-5, np.float32)
    h[:10] = 10.0        # > root box / 2
    h[10:20] = 0.3       # a middle band
    T = O.Tree(x, y, np.ones(n, np.float32), h, bucket=5)
    d, q = T.export_nodes()
    assert q[0, 2] == 10
    assert q[:, 2].sum() == n


def test_walk_equals_direct_sum_when_every_cell_opens():
    n = 3000
    x, y = _nfw_like(n, 5)
    m = (1e8 * (1 + np.random.default_rng(2).random(n))).astype(np.float32)
    h = np.full(n, 2e-6, np.float32)
    rng = np.random.default_rng(9)
    rays = (rng.random((300, 2)) - 0.5) * 0.12
    T = O.Tree(x, y, m, h, theta=1e-12)
    f = T.force(rays)
    dd = O.direct_sum(x, y, m, h, rays)
    scale = np.abs(dd).max(0)
    assert np.all(np.abs(f - dd).max(0) <= 1e-12 * scale + 1e-300)


def test_far_field_error_is_quadrupole_order():
    # The reference's quadrupole deflection uses coefficient 4
    # (include/quadTree.h:1114-1118) where the gradient of its quadrupole
    # potential gives 2, so the far-field error of a lone cell falls as
    # (size/r)^2 rather than (size/r)^3.  Restated, not corrected.
    n = 200
    rng = np.random.default_rng(4)
    x = rng.random(n) * 0.02
    y = rng.random(n) * 0.01
    x[0], y[0] = 0.0, 0.0
    m = np.ones(n, np.float32)
    h = np.full(n, 1e-7, np.float32)
    T = O.Tree(x, y, m, h, theta=0.5, bucket=n)     # one leaf cell
    errs = []
    for R in (2.0, 4.0):
        ang = np.linspace(0, 2 * np.pi, 16, endpoint=False)
        rays = np.stack([0.01 + R * np.cos(ang), 0.005 + R * np.sin(ang)], 1)
        f = T.force(rays)
        dd = O.direct_sum(x, y, m, h, rays)
        errs.append(np.max(np.hypot(*(f[:, :2] - dd[:, :2]).T) / np.hypot(*dd[:, :2].T)))
    slope = np.log(errs[0] / errs[1]) / np.log(2.0)
    assert 1.7 < slope < 2.3, (errs, slope)


def test_gauss_law_on_rings():
    # enclosed mass = (1/2) * flux of the tree alpha through a ring (tree units)
    n = 20000
    x, y = _nfw_like(n, 11, scale=0.2)
    m = np.full(n, 1e8, np.float32)
    h = np.full(n, 1e-7, np.float32)
    T = O.Tree(x, y, m, h, theta=0.1)
    for R in (0.02, 0.08):
        k = 2048
        ang = (np.arange(k) + 0.5) * 2 * np.pi / k
        rays = np.stack([R * np.cos(ang), R * np.sin(ang)], 1)
        f = T.force(rays)
        flux = np.sum(f[:, 0] * np.cos(ang) + f[:, 1] * np.sin(ang)) * R * 2 * np.pi / k
        menc = float(m[0]) * np.sum(np.hypot(x, y) < R)
        assert abs(0.5 * flux - menc) < 5e-3 * menc, (R, 0.5 * flux, menc)


def test_kappa_integrates_to_mass():
    # smoothed particles: sum of kappa over a fine grid recovers the mass
    pts = np.array([[0.0, 0.0], [0.031, 0.017], [-0.02, 0.04]])
    m = np.array([1e10, 2e10, 5e9], np.float32)
    h = np.full(3, 0.01, np.float32)
    T = O.Tree(pts[:, 0], pts[:, 1], m, h, theta=0.1, bucket=20)
    g = np.linspace(-0.1, 0.12, 441)
    dx = g[1] - g[0]
    X, Y = np.meshgrid(g, g)
    f = T.force(np.stack([X.ravel(), Y.ravel()], 1))
    assert abs(f[:, 2].sum() * dx * dx - m.astype(np.float64).sum()) < 2e-3 * m.sum()


# ----------------------------------------------------------------- lens
def _point_lens(zl, zs, M, Om=0.3, OL=0.7, h=0.7):
    L = O.Lens(zs, Om, OL, h)
    T = O.Tree(np.array([0.0]), np.array([0.0]), np.array([M], np.float32),
               np.array([1e-12], np.float32))
    L.moveinMainHalo(T, zl)
    return L


def test_point_mass_lens_closed_form():
    zl, zs, M = 0.5, 2.0, 1e12
    L = _point_lens(zl, zs, M)
    tab = L.tables()
    zl_f = float(np.float32(zl))
    Dl, Ds = O.coorDist(0, zl_f), O.coorDist(0, zs)
    thetaE2 = 4 * GRAV * float(np.float32(M)) * (Ds - Dl) * (1 + zl_f) / (Dl * Ds)
    assert abs(tab["Dl"][0] - Dl) < 1e-9 and abs(tab["Dl"][1] - Ds) < 1e-9
    rng = np.random.default_rng(0)
    th = np.sqrt(thetaE2)
    x = (rng.random((200, 2)) - 0.5) * 6 * th
    y, A = L.shoot(x)
    r2 = (x ** 2).sum(1)
    y_exp = x - thetaE2 * x / r2[:, None]
    assert np.allclose(y, y_exp, rtol=0, atol=1e-10 * th)
    # A = I - thetaE^2 (I/r^2 - 2 x x^T / r^4), row-major
    a0 = 1 - thetaE2 * (1 / r2 - 2 * x[:, 0] ** 2 / r2 ** 2)
    a1 = thetaE2 * 2 * x[:, 0] * x[:, 1] / r2 ** 2
    a3 = 1 - thetaE2 * (1 / r2 - 2 * x[:, 1] ** 2 / r2 ** 2)
    assert np.allclose(A[:, 0], a0, atol=1e-9) and np.allclose(A[:, 3], a3, atol=1e-9)
    assert np.allclose(A[:, 1], a1, atol=1e-9) and np.allclose(A[:, 2], a1, atol=1e-9)


def _multi_lens(nplanes=4, n=3000, zs=3.0):
    zl = np.linspace(0.2, 1.6, nplanes)
    L = O.Lens(zs)
    trees = []
    for j, z in enumerate(zl):
        x, y = _nfw_like(n, 20 + j, scale=0.3)
        m = np.full(n, 3e9, np.float32)
        h = np.full(n, 1e-4, np.float32)
        T = O.Tree(x, y, m, h, theta=0.1)
        T2 = O.Tree(x, y, m, h, theta=0.1)     # an independent copy for the explicit form
        L.moveinMainHalo(T, z)
        trees.append(T2)
    return L, trees


def test_multiplane_recursion_equals_explicit_sum():
    # running-sum recursion (include/lens.h:1003-1066) against the explicit
    # multi-plane lens equation theta_j = theta_1 + sum_{i<j} D_ij/D_j ahat_i,
    # ahat_i = -charge alpha_i (the term SumPrevAlphas accumulates), and
    # A_j = I - sum_{i<j} D_ij/D_j U_i A_i, evaluated plane by plane with
    # comoving D (flat universe, Notes/raytracing.tex:20-27)
    L, trees = _multi_lens()
    tab = L.tables()
    Dl, z, ch = tab["Dl"], tab["z"], tab["charge"]
    k = tab["nplanes"]
    rng = np.random.default_rng(1)
    x = (rng.random((64, 2)) - 0.5) * 60 * ARCSEC
    y, A = L.shoot(x)
    thetas = [x.copy()]
    Amats = [np.tile(np.eye(2), (x.shape[0], 1, 1))]
    ahat, U = [], []
    for j in range(k + 1):
        if j > 0:
            th = x.copy()
            Aj = np.tile(np.eye(2), (x.shape[0], 1, 1))
            for i in range(j):
                w = (Dl[j] - Dl[i]) / Dl[j]
                th = th + w * ahat[i]
                Aj = Aj - w * np.einsum("nab,nbc->nac", U[i], Amats[i])
            thetas.append(th)
            Amats.append(Aj)
        if j == k:
            break
        dphys = Dl[j] / (1 + z[j])
        f = trees[j].force(thetas[j] * dphys)
        ahat.append(-ch * f[:, :2])
        G = np.stack([f[:, 2] + f[:, 3], f[:, 4], f[:, 4], f[:, 2] - f[:, 3]], 1).reshape(-1, 2, 2)
        U.append(G * ch * dphys)
    assert np.allclose(y, thetas[-1], rtol=0, atol=1e-12 * np.abs(x).max())
    assert np.allclose(A, Amats[-1].reshape(-1, 4), rtol=0, atol=1e-10)


def test_lens_plane_rules():
    L = O.Lens(0.4)
    for z in (0.8, 0.3, 0.8):
        T = O.Tree(*_nfw_like(100, 7), np.ones(100, np.float32), np.full(100, 1e-5, np.float32))
        L.moveinMainHalo(T, z)
    tab = L.tables()
    # halos at one redshift share a plane; planes sorted by distance; the
    # source, at or below the last plane, moves to last + 0.1 (lens.cpp:2483-2485)
    assert tab["nplanes"] == 2
    assert np.allclose(tab["z"][:2], [np.float32(0.3), np.float32(0.8)])
    assert abs(tab["z"][2] - (float(np.float32(0.8)) + 0.1)) < 1e-15
    assert np.allclose(np.diff(tab["Dl"]), tab["dDl"][1:])


def test_gridmap_positions():
    N, c, rg = 7, (0.1, -0.2), 3.0
    x = O.gridmap_positions(N, c, rg)
    assert np.allclose(x[0], [c[0] - rg / 2, c[1] - rg / 2])
    assert np.allclose(x[N * N - 1], [c[0] + rg / 2, c[1] + rg / 2])
    assert np.allclose(x[1, 0] - x[0, 0], rg / (N - 1)) and x[1, 1] == x[0, 1]
    assert np.allclose(x[N, 1] - x[0, 1], rg / (N - 1))


def test_same_place_raises():
    with pytest.raises(ValueError):
        O.Tree(np.zeros(5), np.arange(5.0), np.ones(5, np.float32), np.ones(5, np.float32))


# ------------------------------------------------------------ smoothing sizes
def _nn_numpy(x, y, z, nsmooth):
    """The definition: sqrt of the (nsmooth+1)-th smallest squared distance,
    differences in the coordinate type, squares and sums in float64."""
    P = np.stack([x, y, z], 1)
    d = (P[:, None, :] - P[None, :, :]).astype(np.float64)
    r2 = (d[..., 0] ** 2 + d[..., 1] ** 2) + d[..., 2] ** 2
    return np.sqrt(np.sort(r2, 1)[:, nsmooth]).astype(np.float32)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_smoothing_tree_search_equals_definition(dtype):
    rng = np.random.default_rng(3)
    # clustered: two dense clumps and a sparse background
    p = np.concatenate([rng.normal(0, 0.01, (700, 3)), rng.normal(1, 0.1, (700, 3)),
                        rng.uniform(-3, 3, (600, 3))]).astype(dtype)
    x, y, z = p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()
    for k in (1, 8, 40):
        tree = O.smoothing_sizes(x, y, z, k)
        brute = O.smoothing_sizes(x, y, z, k, brute=True)
        assert np.array_equal(tree, brute)
        assert np.array_equal(brute, _nn_numpy(x, y, z, k))


def test_smoothing_known_answers():
    # evenly spaced points on a line: the (k+1)-th nearest of point 0, itself
    # included, is point k
    x = np.arange(50, dtype=np.float64)
    zero = np.zeros(50)
    s = O.smoothing_sizes(x, zero, zero, 7)
    assert s[0] == 7 and s[49] == 7 and s[25] == 4
    # coincident particles are at distance 0 from each other
    s = O.smoothing_sizes(np.r_[np.zeros(10), np.arange(1.0, 11)], np.zeros(20), np.zeros(20), 5)
    assert np.all(s[:10] == 0)
    with pytest.raises(ValueError):
        O.smoothing_sizes(x[:8], zero[:8], zero[:8], 7)


def test_reset_source_plane_choice():
    # nearest: the plane nearest in comoving distance; not nearest: the first
    # plane at or beyond the source, one back when the source sits on a plane
    tab = {"nplanes": 4, "z": np.array([0.2, 0.5, 0.9, 1.4, 3.0])}
    tab["Dl"] = np.array([O.coorDist(0, z) for z in tab["z"]])
    tab["dDl"] = np.diff(np.concatenate([[0.0], tab["Dl"]]))
    for z in np.linspace(0.05, 2.5, 200):
        Ds = O.coorDist(0, z)
        j, zi = O.reset_source_plane(tab, z, False)
        assert j == int(np.sum(tab["Dl"][:4] < Ds)) and zi == z
        j, zi = O.reset_source_plane(tab, z, True)
        if Ds <= tab["Dl"][0] or Ds > tab["Dl"][3] + 0.5 * (tab["Dl"][4] - tab["Dl"][3]):
            continue
        near = int(np.argmin(np.abs(tab["Dl"][:4] - Ds)))
        assert j == near and zi == tab["z"][near], z
    # on a plane: that plane, with or without nearest
    for k in range(4):
        assert O.reset_source_plane(tab, tab["z"][k], False)[0] == k
        assert O.reset_source_plane(tab, tab["z"][k], True) == (k, tab["z"][k])


def _two_clusters(seed):
    """A cluster at the origin and a dense one near (1, 0): the far one lies
    beyond a small maximum range by many times its own cell's size."""
    rng = np.random.default_rng(seed)
    x0, y0 = _nfw_like(4000, seed, scale=0.05)
    x1 = 1.0 + 0.01 * rng.random(3000)
    y1 = 0.01 * rng.random(3000)
    x = np.concatenate([x0, x1])
    y = np.concatenate([y0, y1])
    m = np.full(x.size, 1e8, np.float32)
    h = np.full(x.size, 1e-5, np.float32)
    return x, y, m, h


def test_tree_max_range():
    """TreeQuadParticles' maximum_range (include/quadTree.h:370-380): a cell
    with more than bucket particles stays a leaf only when it lies beyond
    max_range by at least 1.732 box / theta; far beyond every cell the tree is
    the plain one; the walk over the coarse leaves still gives the direct sum
    to the opening angle's accuracy."""
    x, y, m, h = _two_clusters(21)
    T0 = O.Tree(x, y, m, h)
    Tfar = O.Tree(x, y, m, h, max_range=1e3)
    d0, q0 = T0.export_nodes()
    d1, q1 = Tfar.export_nodes()
    assert np.array_equal(d0, d1) and np.array_equal(q0, q1)

    R = 0.02
    T = O.Tree(x, y, m, h, max_range=R)
    d, q = T.export_nodes()
    assert T.nnodes() < T0.nnodes()
    leaves = q[:, 3] == 1
    coarse = leaves & (q[:, 1] > 20)
    assert coarse.any() and q[coarse, 1].sum() >= 2500
    rc = np.hypot((d[:, 0] + d[:, 2]) / 2, (d[:, 1] + d[:, 3]) / 2)
    far = (rc - R) >= (1.732 / 0.1) * (d[:, 2] - d[:, 0]) * (1 - 1e-9)
    assert np.all(far[coarse])                     # coarse leaves only far out
    assert not np.any(leaves & (q[:, 1] > 20) & ~far)   # nearer ones all split
    kids = _children(q)
    for i, ks in enumerate(kids):
        if ks:
            assert q[i, 1] > 20 and not far[i]
    assert q[leaves, 1].sum() == x.size

    rng = np.random.default_rng(3)
    rays = np.concatenate([(rng.random((150, 2)) - 0.5) * 0.1,
                           [1.0, 0.0] + rng.random((150, 2)) * 0.012])
    dd = O.direct_sum(x, y, m, h, rays)
    scale = np.abs(dd).max(0)
    for t in (T0, T):
        assert np.all(np.abs(t.force(rays) - dd).max(0) <= 2e-3 * scale)
