"""GridMap::find_images on the device against the oracle's restatement on the
same grid outputs: every image, triangle and multiplicity equal, for
sources around a lens's caustics, sources placed exactly on grid vertices
and edges (the reference's edge and vertex rules), and the per-source cap."""
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


def _check(g, x, y, A, ys, cap=32):
    rays, mult, tris = g.find_images(ys, cap)
    r0, m0, t0 = O.find_images(g.N, x, y, A, ys)
    assert np.array_equal(mult, m0)
    assert np.array_equal(tris, t0)
    assert np.array_equal(rays, r0)
    return mult


def test_images_match_oracle_around_caustics(shot):
    g, x, y, A = shot
    rng = np.random.default_rng(3)
    # sources drawn from the grid's own source positions near the centre
    # (the multiply imaged region) plus random ones
    centre = np.argsort((x ** 2).sum(1))[:4000]
    ys = y[rng.choice(centre, 300)] + rng.normal(size=(300, 2)) * 1e-7
    ys = np.concatenate([ys, (rng.random((200, 2)) - 0.5) * np.ptp(y, axis=0)])
    print("case: find_images 500 sources", flush=True)
    mult = _check(g, x, y, A, ys)
    assert mult.max() >= 3                 # some sources are multiply imaged


def test_images_on_vertices_and_edges(shot):
    g, x, y, A = shot
    rng = np.random.default_rng(4)
    k = rng.integers(0, g.N * g.N, 200)
    on_vertex = y[k]
    mid = 0.5 * (y[k] + y[np.minimum(k + 1, g.N * g.N - 1)])
    _check(g, x, y, A, np.concatenate([on_vertex, mid]))


def test_images_cap(shot):
    g, x, y, A = shot
    centre = np.argsort((x ** 2).sum(1))[:10]
    with pytest.raises(S.SlsimError) as e:
        g.find_images(y[centre], max_per_source=1)
    assert e.value.code == 1
