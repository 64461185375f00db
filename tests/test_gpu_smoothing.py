"""GPU smoothing sizes (LensHaloParticles::calculate_smoothing,
include/particle_halo.h:683-730) against the oracle's restatement of the
reference's tree search (include/simpleTree.h:483-620) and its brute-force
definition.  The size is the (Nsmooth+1)-th smallest of numbers that the GPU
and the oracle compute with the same arithmetic, so they must be equal."""
import numpy as np
import pytest

import oracle as O
import slsimlib_amd as S
from slsimlib_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nsmooth", [8, 32, 64])
def test_smoothing_matches_oracle(ctx, dtype, nsmooth):
    x, y, z = W.particles_3d(7, 20000, dtype=dtype)
    g = S.LensHaloParticles.calculate_smoothing(ctx, nsmooth, x, y, z)
    o = O.smoothing_sizes(x, y, z, nsmooth)
    assert np.array_equal(g, o), np.abs(g - o).max()


def test_smoothing_large_subsample(ctx):
    x, y, z = W.particles_3d(8, 300000)
    g = S.LensHaloParticles.calculate_smoothing(ctx, 64, x, y, z)
    which = np.random.default_rng(2).choice(x.size, 3000, replace=False)
    o = O.smoothing_sizes(x, y, z, 64, brute=True, which=which)
    assert np.array_equal(g[which], o)
    assert np.all(g > 0)


@pytest.mark.parametrize("case", ["minimal", "dupes", "line", "grid"])
def test_smoothing_edge_cases(ctx, case):
    rng = np.random.default_rng(4)
    k = 16
    if case == "minimal":                  # n = Nsmooth + 2
        x, y, z = rng.normal(size=(3, k + 2))
    elif case == "dupes":                  # many particles in one place
        x, y, z = rng.normal(size=(3, 3000))
        x[:500], y[:500], z[:500] = x[0], y[0], z[0]
    elif case == "line":                   # a degenerate box in two dimensions
        x = rng.normal(size=2000)
        y = np.zeros(2000)
        z = np.full(2000, 3.0)
    else:                                  # ties on a lattice
        g = np.arange(12, dtype=np.float64)
        X, Y, Z = np.meshgrid(g, g, g)
        x, y, z = X.ravel(), Y.ravel(), Z.ravel()
    g = S.LensHaloParticles.calculate_smoothing(ctx, k, x, y, z)
    o = O.smoothing_sizes(x, y, z, k, brute=True)
    assert np.array_equal(g, o)
    if case == "dupes":
        assert np.all(g[:500] == 0)


def test_smoothing_errors(ctx):
    x, y, z = np.random.default_rng(5).normal(size=(3, 17))
    with pytest.raises(S.SlsimError) as e:       # n <= Nsmooth + 1
        S.LensHaloParticles.calculate_smoothing(ctx, 16, x, y, z)
    assert e.value.code == 1
    with pytest.raises(S.SlsimError):
        S.LensHaloParticles.calculate_smoothing(ctx, 120, *np.random.default_rng(6).normal(size=(3, 500)))


def test_smoothing_dense_cluster_with_outliers(ctx):
    """A dense cluster with sparse outliers: an outlier's neighbour sphere
    takes in the whole cluster's subtree, so its box stack outgrows the
    in-LDS one (about 520 entries at this size) and the query is done again
    with a stack in global memory.  Outliers and cluster particles against
    the brute-force definition."""
    rng = np.random.default_rng(11)
    nc, no = 400000, 300
    x, y, z = np.concatenate([rng.normal(0, 0.01, (3, nc)), rng.uniform(-1, 1, (3, no))], 1)
    g = S.LensHaloParticles.calculate_smoothing(ctx, 64, x, y, z)
    which = np.concatenate([np.arange(nc, nc + no), rng.choice(nc, 200, replace=False)])
    o = O.smoothing_sizes(x, y, z, 64, brute=True, which=which)
    assert np.array_equal(g[which], o)
