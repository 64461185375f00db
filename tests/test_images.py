"""Pins the oracle's restatement of GridMap::find_images (oracle/
images_oracle.cpp) with lens maps whose images are known: the identity
map (one image, in the triangle that holds the source) and a point mass
(two images at (b +- sqrt(b^2 + 4 tE^2)) / 2 along the source direction).
CPU only."""
import numpy as np

import oracle as O


def _grid(N, half):
    g = (np.arange(N) / (N - 1) - 0.5) * 2 * half
    X, Y = np.meshgrid(g, g)            # x fastest, as GridMap's
    return np.stack([X.ravel(), Y.ravel()], 1)


def test_identity_map_one_image_in_its_triangle():
    N, half = 33, 1.0
    x = _grid(N, half)
    A = np.tile([1.0, 0.0, 0.0, 1.0], (N * N, 1))
    rng = np.random.default_rng(1)
    ys = (rng.random((500, 2)) - 0.5) * 1.9 * half
    rays, mult, tris = O.find_images(N, x, x, A, ys)
    assert np.all(mult == 1)
    cell = 2 * half / (N - 1)
    assert np.all(np.abs(rays[:, :2] - ys).max(axis=1) <= cell)
    assert np.array_equal(rays[:, 2:4], ys)
    assert np.array_equal(rays[:, 4:], A[:rays.shape[0]])
    # the source lies in the reported triangle
    def cross(u, v):
        return u[0] * v[1] - u[1] * v[0]
    for r, t in zip(ys, tris):
        a, b, c = x[t[0]], x[t[1]], x[t[2]]
        s = [cross(b - a, r - a), cross(c - b, r - b), cross(a - c, r - c)]
        assert (min(s) >= 0) or (max(s) <= 0)


def test_point_mass_two_images():
    N, half, tE = 512, 2.0, 0.6
    x = _grid(N, half)
    r2 = (x ** 2).sum(1)
    y = x - tE ** 2 * x / r2[:, None]
    A = np.tile([1.0, 0.0, 0.0, 1.0], (N * N, 1))
    rng = np.random.default_rng(2)
    b = 0.25 + 0.5 * rng.random(60)
    ph = 2 * np.pi * rng.random(60)
    ys = np.stack([b * np.cos(ph), b * np.sin(ph)], 1)
    rays, mult, tris = O.find_images(N, x, y, A, ys)
    cell = 2 * half / (N - 1)
    k = 0
    for s in range(len(ys)):
        found = rays[k:k + mult[s], :2]
        k += mult[s]
        assert mult[s] in (2, 3)          # a third can sit on the singular centre cells
        u = ys[s] / b[s]
        for th in ((b[s] + np.sqrt(b[s] ** 2 + 4 * tE ** 2)) / 2,
                   (b[s] - np.sqrt(b[s] ** 2 + 4 * tE ** 2)) / 2):
            d = np.sqrt(((found - th * u) ** 2).sum(1)).min()
            assert d <= 1.5 * cell, (s, th, d)
