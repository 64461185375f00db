"""Pins the oracle's restatement of LensHaloNFW and TreeQuadHalos (the
halo-of-halos planes of LensHaloHalos) against what the mathematics fixes:
the NFW profile identities, Gauss's law and the tangential-shear identity for
force_halo, the point mass beyond Rmax, and the tree against the sum over
every halo as the opening angle goes to zero (including the reference's
second count of a halo that is big in an internal cell).  CPU only.
"""
import numpy as np
import pytest

import oracle as O


# ------------------------------------------------------------- NFW profile
@pytest.mark.parametrize("x", [0.003, 0.05, 0.4, 0.93, 1.08, 2.5, 17.0, 80.0])
def test_nfw_profile_identities(x):
    g = O.nfw_function(0, x)
    f = O.nfw_function(1, x)
    g2 = O.nfw_function(2, x)
    # projected mass: dM/dx = 2 pi x Sigma, i.e. g'(x) = x f(x)
    h = 1e-3 * x
    dg = (O.nfw_function(0, x + h) - O.nfw_function(0, x - h)) / (2 * h)
    assert abs(dg - x * f) <= 1e-6 * abs(x * f)
    # tangential shear: g2 = 4 g / x^2 - 2 f (mean convergence inside minus local)
    assert abs(g2 - (4 * g / x ** 2 - 2 * f)) <= 1e-9 * max(abs(g2), 1e-12)


def test_nfw_profile_continuous_at_one_and_tables():
    # the closed forms cancel near x = 1 (the reference's own rounding), so
    # continuity is checked 1e-4 away, where the slope moves them by < 1e-4
    for w in range(3):
        a, b, c = (O.nfw_function(w, v) for v in (1 - 1e-4, 1.0, 1 + 1e-4))
        assert abs(a - b) < 2e-4 and abs(c - b) < 2e-4
    # linear interpolation on the 0.01 grid: error <= dx^2/8 max|f''| on the interval
    xs = np.linspace(0.012, 98.5, 400)
    dx = 0.01
    for w in range(3):
        for x in xs:
            exact = O.nfw_function(w, x)
            j = int(x / dx)
            hh = dx / 4
            f2 = max(abs(O.nfw_function(w, t + hh) - 2 * O.nfw_function(w, t) +
                         O.nfw_function(w, t - hh)) / hh ** 2
                     for t in (j * dx + hh, (j + 0.5) * dx, (j + 1) * dx - hh))
            bound = dx * dx / 8 * f2 * 1.5 + 1e-13 * max(1.0, abs(exact))
            assert abs(O.nfw_interpolate(w, x) - exact) <= bound, (w, x)
        # first interval: the function itself
        assert O.nfw_interpolate(w, 0.004) == O.nfw_function(w, 0.004)


# ------------------------------------------------------------ force_halo
def _ring(m, rs, c, r):
    return O.nfw_force(m, rs, c, np.array([[r, 0.0]]))[0]


def test_nfw_force_gauss_law_and_shear_identity():
    from scipy.integrate import quad
    m, rsize, c = 1e12, 0.2, 8.0
    for r in (0.004, 0.03, 0.11, 0.19):
        out = _ring(m, rsize, c, r)
        menc_alpha = -out[0] * np.pi * r            # alpha = -M(<r)/(pi r) xhat (tree sign)
        kap = lambda rr: _ring(m, rsize, c, rr)[2]
        menc_kappa, _ = quad(lambda rr: 2 * np.pi * rr * kap(rr), 1e-9, r, limit=400,
                             epsrel=1e-8)
        # tables interpolate g and f separately: agreement to their interpolation error
        assert abs(menc_alpha - menc_kappa) <= 2e-3 * menc_alpha
        # gamma_t = mean kappa inside - kappa, gamma_t = -gamma0 on the x axis
        gt = -out[3]
        expect = menc_alpha / (np.pi * r * r) - out[2]
        assert abs(gt - expect) <= 1e-3 * abs(expect)
        assert abs(out[1]) == 0.0 and abs(out[4]) == 0.0


def test_nfw_force_point_mass_outside_and_continuity():
    m, rsize, c = 3e11, 0.1, 5.0
    rmax = float(np.float32(rsize))
    for r in (rmax * 1.0001, 0.3, 2.0):
        out = _ring(m, rsize, c, r)
        pm = float(np.float32(m)) / (np.pi * r)          # LensHalo stores the mass as float
        assert abs(-out[0] - pm) <= 1e-14 * pm
        assert out[2] == 0.0
        assert abs(out[3] - (-pm / r)) <= 1e-14 * pm / r     # gamma0 = -M/(pi r^2)
    inside = _ring(m, rsize, c, rmax * (1 - 1e-9))
    outside = _ring(m, rsize, c, rmax * (1 + 1e-9))
    assert abs(inside[0] - outside[0]) <= 1e-6 * abs(outside[0])


# ------------------------------------------------------------------- tree
def _halos(n, seed, rmax_scale=1e-3):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=n) * 0.3
    y = rng.normal(size=n) * 0.3
    m = (10 ** rng.uniform(10, 13, n)).astype(np.float32)
    rs = (rmax_scale * (0.5 + rng.random(n))).astype(np.float32)
    c = (4 + 8 * rng.random(n)).astype(np.float32)
    return x, y, m, rs, c


def test_tree_small_halos_converges_to_direct_sum():
    x, y, m, rs, c = _halos(3000, 1)
    rng = np.random.default_rng(2)
    rays = rng.normal(size=(300, 2)) * 0.3
    rays[:20] = np.stack([x[:20] + 2e-4, y[:20]], 1)       # inside halos
    t = O.HaloTree(x, y, m, rs, c, inv_area=0.0, theta=1e-6)
    d = t.force(rays, direct=True)
    f = t.force(rays)
    scale = np.abs(d[:, :2]).max()
    assert np.abs(f[:, :2] - d[:, :2]).max() <= 1e-11 * scale
    assert np.abs(f[:, 2:] - d[:, 2:]).max() <= 1e-11 * np.abs(d[:, 2:]).max()
    # opening angle 0.5: a multipole approximation, close but not equal
    f5 = O.HaloTree(x, y, m, rs, c, inv_area=0.0, theta=0.5).force(rays)
    err = np.abs(f5[:, :2] - d[:, :2]).max() / scale
    assert 1e-9 < err < 3e-2


def test_tree_counts_a_big_internal_halo_twice():
    # one halo larger than half the root box is on the root's big list and
    # again in its leaf (its cell keeps all halos): with every cell opened
    # the tree gives the direct sum plus that halo once more
    x, y, m, rs, c = _halos(400, 3)
    rs[17] = 0.9 * (max(x.max() - x.min(), y.max() - y.min()))
    t = O.HaloTree(x, y, m, rs, c, inv_area=0.0, theta=1e-6)
    rng = np.random.default_rng(4)
    rays = rng.normal(size=(50, 2)) * 0.3
    d = t.force(rays, direct=True)
    f = t.force(rays)
    one = O.nfw_force(m[17], rs[17], c[17], np.stack([x[17] - rays[:, 0], y[17] - rays[:, 1]], 1))
    assert np.abs(f - (d + one)).max() <= 1e-11 * np.abs(d).max()


def test_tree_background_sheet():
    # inv_area removes a uniform sheet of the halos' total mass: kappa
    # lowered by inv_area * M for every ray whatever the cells
    x, y, m, rs, c = _halos(500, 5)
    ia = 2.5
    rays = np.random.default_rng(6).normal(size=(40, 2)) * 0.3
    for theta in (1e-6, 0.3):
        a = O.HaloTree(x, y, m, rs, c, inv_area=0.0, theta=theta).force(rays)
        b = O.HaloTree(x, y, m, rs, c, inv_area=ia, theta=theta).force(rays)
        dk = b[:, 2] - a[:, 2]
        assert np.allclose(dk, -float(np.float32(ia)) * float(np.sum(m.astype(np.float64))),
                           rtol=1e-9, atol=0)


# ------------------------------------------------ LensPlaneTree field trees
def test_field_tree_without_screening_is_the_halos_tree():
    x, y, m, rs, c = _halos(800, 7)
    rays = np.random.default_rng(8).normal(size=(200, 2)) * 0.3
    a = O.HaloTree(x, y, m, rs, c, inv_area=0.25, theta=0.5).force(rays)
    b = O.FieldTree(x, y, m, rs, c, sigma_background=0.25).force(rays)   # theta 0.5, no buffer
    assert np.array_equal(a, b)


def test_field_tree_periodic_buffer_is_nine_shifted_walks():
    x, y, m, rs, c = _halos(600, 9)
    rays = np.random.default_rng(10).normal(size=(60, 2)) * 0.3
    xl, yl = x.max() - x.min(), y.max() - y.min()
    one = O.FieldTree(x, y, m, rs, c, 0.1, theta=0.3, periodic=False)
    both = O.FieldTree(x, y, m, rs, c, 0.1, theta=0.3, periodic=True).force(rays)
    want = sum(one.force(rays + [i * xl, j * yl]) for i in (-1, 0, 1) for j in (-1, 0, 1))
    assert np.allclose(both, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def test_field_tree_screening_scales_far_cells():
    # two compact clusters 5 apart: a ray at the first sees the second as one
    # far cell, whose monopole and quadrupole are scaled by exp(-r^2 s^2);
    # the near halos' own forces and the background are not screened
    rng = np.random.default_rng(11)
    n = 60
    x = np.concatenate([rng.normal(size=n) * 0.01, 5.0 + rng.normal(size=n) * 1e-3])
    y = np.concatenate([rng.normal(size=n) * 0.01, rng.normal(size=n) * 1e-3])
    m = (10 ** rng.uniform(10, 12, 2 * n)).astype(np.float32)
    rs = np.full(2 * n, 1e-4, np.float32)
    c = np.full(2 * n, 8.0, np.float32)
    ray = np.array([[0.003, -0.002]])
    s = 0.1
    plain = O.FieldTree(x, y, m, rs, c, 0.7, theta=0.5, periodic=False).force(ray)[0]
    scr = O.FieldTree(x, y, m, rs, c, 0.7, inv_screening_scale=s, theta=0.5,
                      periodic=False).force(ray)[0]
    far = O.FieldTree(x[n:], y[n:], m[n:], rs[n:], c[n:], 0.0, theta=1e-6).force(ray, direct=True)[0]
    com = np.array([np.sum(m[n:] * x[n:]), np.sum(m[n:] * y[n:])]) / np.sum(m[n:].astype(float))
    f = np.exp(-np.sum((com - ray[0]) ** 2) * s * s)
    want = plain + (f - 1) * far
    assert np.allclose(scr[:2], want[:2], rtol=1e-6, atol=0)
    assert np.allclose(scr[3:], want[3:], rtol=1e-5, atol=0)
    assert scr[2] == plain[2]                       # kappa: the background only


def test_field_plane_in_lens_tables():
    x, y, m, rs, c = _halos(100, 12)
    L = O.Lens(2.0)
    L.moveinMainHalo(O.HaloTree(x, y, m, rs, c), 0.5)
    L.addFieldPlane(O.FieldTree(x, y, m, rs, c, 0.0, 0.2), 0.31234567)
    L.addFieldPlane(O.FieldTree(x, y, m, rs, c, 0.0, 0.2), 1.2)
    t = L.tables()
    assert t["nplanes"] == 3
    # field redshifts stay doubles; the main plane's is a float
    assert list(t["z"][:3]) == [0.31234567, float(np.float32(0.5)), 1.2]


# ----------------------------------------------------- LensHaloPowerLaw
@pytest.mark.parametrize("beta", [0.3, 1.0, 1.7, -0.5])
def test_powerlaw_gauss_law_and_shear(beta):
    from scipy.integrate import quad
    m, R = 2e12, 0.15
    Rf = float(np.float32(R))
    mf = float(np.float32(m))
    for r in (1e-3, 0.02, 0.09, 0.149):
        out = O.powerlaw_force(m, R, beta, np.array([[r, 0.0]]))[0]
        # M(<r) = m (r/R)^(2-beta) inside, alpha = -M(<r)/(pi r) x-hat
        menc = mf * (r / Rf) ** (2 - beta)
        assert abs(-out[0] * np.pi * r - menc) <= 1e-12 * menc
        kap = lambda rr: O.powerlaw_force(m, R, beta, np.array([[rr, 0.0]]))[0][2]
        mk, _ = quad(lambda rr: 2 * np.pi * rr * kap(rr), 0, r, limit=200, epsrel=1e-10)
        assert abs(mk - menc) <= 1e-6 * menc
        gt = -out[3]
        assert abs(gt - (menc / (np.pi * r * r) - out[2])) <= 1e-10 * abs(gt)
    # beyond Rmax: the point mass
    out = O.powerlaw_force(m, R, beta, np.array([[0.4, 0.0]]))[0]
    assert abs(-out[0] - mf / (np.pi * 0.4)) <= 1e-14 * mf / 0.4 and out[2] == 0.0


def test_powerlaw_rejects_beta_two():
    with pytest.raises(ValueError):
        O.MixedTree([0.0, 1.0], [0.0, 1.0], [1e12, 1e12], [0.1, 0.1], [1, 1], [2.0, 1.0])


def test_mixed_tree_converges_to_direct_sum():
    x, y, m, rs, c = _halos(1200, 21)
    rng = np.random.default_rng(22)
    types = (rng.random(x.size) < 0.4).astype(np.int32)
    params = np.where(types == 1, rng.uniform(0.2, 1.8, x.size), c.astype(np.float64))
    rays = rng.normal(size=(200, 2)) * 0.3
    rays[:10] = np.stack([x[:10] + 2e-4, y[:10]], 1)
    t = O.MixedTree(x, y, m, rs, types, params, inv_area=0.3, theta=1e-6)
    d = t.force(rays, direct=True)
    f = t.force(rays)
    assert np.abs(f - d).max() <= 1e-11 * np.abs(d).max()
    # NFW-only mixed trees are the NFW tree, bit for bit
    t0 = O.MixedTree(x, y, m, rs, np.zeros(x.size, np.int32), c.astype(np.float64),
                     inv_area=float(np.float32(0.3)), theta=float(np.float32(0.2)))
    assert np.array_equal(t0.force(rays), O.HaloTree(x, y, m, rs, c, 0.3, 0.2).force(rays))


# ---------------------------------------- moveinMainHalo(addplanes=false)
def _tree(seed):
    rng = np.random.default_rng(seed)
    n = 600
    x = rng.normal(0, 0.05, n)
    y = rng.normal(0, 0.05, n)
    m = np.full(n, np.float32(1e10), np.float32)
    h = np.full(n, np.float32(1e-3), np.float32)
    return x, y, m, h


def test_nearest_plane_insertion_reduces_to_addplanes():
    """addMainHaloToNearestPlane (MultiPlane/lens.cpp:822-840): a halo at
    z = 0.42 next to main planes at 0.3 and 0.5 deflects as if it were at the
    plane nearer in comoving distance; on an empty lens it deflects twice."""
    rays = np.random.default_rng(5).uniform(-1e-4, 1e-4, (300, 2))
    a, b, c = _tree(1), _tree(2), _tree(3)
    d03, d05, d = O.coorDist(0, 0.3), O.coorDist(0, 0.5), O.coorDist(0, np.float32(0.42))
    znear = 0.5 if abs(d05 - d) < abs(d03 - d) else 0.3

    L = O.Lens(2.0)
    L.moveinMainHalo(O.Tree(*a), 0.3)
    L.moveinMainHalo(O.Tree(*b), 0.5)
    L.moveinMainHalo(O.Tree(*c), 0.42, addplanes=False)
    R = O.Lens(2.0)
    R.moveinMainHalo(O.Tree(*a), 0.3)
    R.moveinMainHalo(O.Tree(*b), 0.5)
    R.moveinMainHalo(O.Tree(*c), znear)
    assert L.tables()["nplanes"] == 2
    for got, want in zip(L.shoot(rays), R.shoot(rays)):
        assert np.array_equal(got, want)
    # the other plane is not the nearest: moving the halo there changes y
    F = O.Lens(2.0)
    F.moveinMainHalo(O.Tree(*a), 0.3)
    F.moveinMainHalo(O.Tree(*b), 0.5)
    F.moveinMainHalo(O.Tree(*c), 0.8 - znear)
    assert not np.allclose(L.shoot(rays)[0], F.shoot(rays)[0], rtol=0, atol=1e-12)

    E = O.Lens(2.0)
    E.moveinMainHalo(O.Tree(*a), 0.5, addplanes=False)
    T = O.Lens(2.0)
    T.moveinMainHalo(O.Tree(*a), 0.5)
    T.moveinMainHalo(O.Tree(*a), 0.5)
    assert E.tables()["nplanes"] == T.tables()["nplanes"] == 1
    for got, want in zip(E.shoot(rays), T.shoot(rays)):
        assert np.array_equal(got, want)
