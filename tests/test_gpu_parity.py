"""GPU parity of the HIP ray shooter against the CPU oracle, through the C ABI.

Tolerances (stated, per ray):
  FP64 path vs oracle:   relative |d alpha| <= 1e-4, |d kappa|, |d gamma| <= 1e-3
                         (same algorithm in the same precision; observed ~1e-12)
  FP32C path vs FP64:    relative |d alpha| <= 1e-5, |d kappa|, |d gamma| <= 1e-3
                         (multi-plane kappa/gamma: at the 99.9th percentile, the worst
                         ray bounded by the FP64 path's own conditioning; see
                         test_multiplane_fp32c_vs_fp64)
Relative error of a vector quantity v: |dv| / max(|v_ref|, 1e-3 rms|v_ref|).
Tree structure and interaction counts must match the oracle exactly.
"""
import numpy as np
import pytest

import oracle as O
import slsimlib_amd as S
from slsimlib_amd import workloads as W

pytestmark = pytest.mark.gpu

ALPHA_TOL = 1e-4
KG_TOL = 1e-3
FP32C_ALPHA_TOL = 1e-5
# worst-ray kappa and gamma of FP32C against FP64 on a composite multi-plane
# lens (conditioning-limited; the 99.9th percentile is held to KG_TOL)
KG_MAX_FP32C = 5e-2


def rel_err(d, ref):
    """|d| / max(|ref|, 1e-3 rms(|ref|)) per row."""
    d = np.asarray(d).reshape(len(d), -1)
    ref = np.asarray(ref).reshape(len(ref), -1)
    nr = np.linalg.norm(ref, axis=1)
    nd = np.linalg.norm(d, axis=1)
    floor = 1e-3 * np.sqrt(np.mean(nr ** 2))
    return nd / np.maximum(nr, floor if floor > 0 else 1e-300)


def lensing_fields(x, y, A):
    alpha = x - y
    kappa = 1 - 0.5 * (A[:, 0] + A[:, 3])
    gamma = np.stack([0.5 * (A[:, 0] - A[:, 3]), 0.5 * (A[:, 1] + A[:, 2]),
                      0.5 * (A[:, 2] - A[:, 1])], 1)
    return alpha, kappa, gamma


def assert_fields_close(x, y, A, y_ref, A_ref, atol_alpha, atol_kg):
    a, k, g = lensing_fields(x, y, A)
    ar, kr, gr = lensing_fields(x, y_ref, A_ref)
    ea = rel_err(a - ar, ar).max()
    ek = rel_err(k - kr, kr).max()
    eg = rel_err(g - gr, gr).max()
    assert ea <= atol_alpha, ("alpha", ea)
    assert ek <= atol_kg, ("kappa", ek)
    assert eg <= atol_kg, ("gamma", eg)
    return ea, ek, eg


# ----------------------------------------------------------------- tree build
def _compare_trees(dev, orc):
    d, q = dev.export_nodes()
    od, oq = orc.export_nodes()
    assert d.shape == od.shape, (d.shape, od.shape)
    assert np.array_equal(q[:, :4], oq[:, :4])          # level, nparticles, nbig, leaf
    assert np.array_equal(d[:, :4], od[:, :4])          # boxes, bit-exact
    occ = q[:, 1] > 0
    box = (d[:, 2] - d[:, 0])
    assert np.all(np.abs(d[occ, 6] - od[occ, 6]) <= 1e-13 * np.abs(od[occ, 6]))
    assert np.all(np.abs(d[occ, 4:6] - od[occ, 4:6]) <= 1e-13 * box[occ, None])
    qs = np.abs(od[occ, 6]) * box[occ] ** 2
    assert np.all(np.abs(d[occ, 7:10] - od[occ, 7:10]) <= 1e-11 * qs[:, None])
    assert np.all(np.abs(d[occ, 10] - od[occ, 10]) <= 1e-11 * od[occ, 10])


def test_tree_matches_oracle_c1(ctx):
    x, y, m, h = W.c1_particles()
    dev = S.LensHaloParticles(ctx, x, y, m, h, 0.5)
    _compare_trees(dev, O.Tree(x, y, m, h))
    info = dev.info()
    assert info["n_particles"] == x.size and info["n_big_internal"] == 0


@pytest.mark.parametrize("case", ["single", "bucket", "ties", "bands", "float32", "dupes"])
def test_tree_edge_cases(ctx, case):
    rng = np.random.default_rng(5)
    if case == "single":
        x, y = np.array([0.3]), np.array([-0.1])
    elif case == "bucket":
        x, y = rng.random(20), rng.random(20)
    elif case == "ties":                       # particles on every split line
        g = np.linspace(0, 1, 17)
        X, Y = np.meshgrid(g, g)
        x, y = X.ravel(), Y.ravel()
    elif case == "dupes":
        x, y = rng.random(3000), rng.random(3000)
        x[100:110] = x[5]
        y[100:110] = y[5]
    else:
        x, y = rng.random(4000) ** 3, rng.random(4000) ** 3
    n = x.size
    m = (1e9 * (0.5 + rng.random(n))).astype(np.float32)
    h = np.full(n, 1e-6, np.float32)
    if case == "bands":
        h = (10.0 ** rng.uniform(-6, 0.2, n)).astype(np.float32)
    if case == "float32":
        x, y = x.astype(np.float32), y.astype(np.float32)
    dev = S.LensHaloParticles(ctx, x, y, m, h, 0.5)
    _compare_trees(dev, O.Tree(x, y, m, h))
    if case == "bands":
        assert dev.info()["n_big_internal"] > 0


def test_tree_same_place_error(ctx):
    with pytest.raises(S.SlsimError) as e:
        S.LensHaloParticles(ctx, np.zeros(10), np.arange(10.0), np.ones(10, np.float32),
                            np.ones(10, np.float32), 0.5)
    assert e.value.code == 4


# ----------------------------------------------------------------- one plane
def _c1_rays(N=256):
    z = W.C1["z_lens"]
    DA = S.coorDist(0, np.float32(z)) / (1 + np.float32(z))
    pos = O.gridmap_positions(N, (0.0, 0.0), W.grid_range(W.C1["range_arcsec"]))
    return pos * DA


def test_plane_force_fp64_matches_oracle(ctx):
    x, y, m, h = W.c1_particles()
    dev = S.LensHaloParticles(ctx, x, y, m, h, 0.5)
    T = O.Tree(x, y, m, h)
    rays = _c1_rays(128)
    f = dev.force(rays, S.SLSIM_FP64)
    o = T.force(rays)
    assert rel_err(f[:, :2] - o[:, :2], o[:, :2]).max() < 1e-11
    assert rel_err(f[:, 2:] - o[:, 2:], o[:, 2:]).max() < 1e-10


def test_plane_force_fp32c_close_to_fp64(ctx):
    x, y, m, h = W.c1_particles()
    dev = S.LensHaloParticles(ctx, x, y, m, h, 0.5)
    rays = _c1_rays(256)
    f64 = dev.force(rays, S.SLSIM_FP64)
    f32 = dev.force(rays, S.SLSIM_FP32C)
    assert rel_err(f32[:, :2] - f64[:, :2], f64[:, :2]).max() <= FP32C_ALPHA_TOL
    assert rel_err(f32[:, 2:] - f64[:, 2:], f64[:, 2:]).max() <= KG_TOL


def test_plane_force_mass_sheet(ctx):
    """inv_area != 0 (the compensating negative mass sheet): FP64 against the
    oracle, FP32C (which adds the sheet in closed form) against FP64."""
    x, y, m, h = W.c1_particles()
    span = max(x.max() - x.min(), y.max() - y.min())
    ia = 1.0 / (0.25 * span * span)
    dev = S.LensHaloParticles(ctx, x, y, m, h, 0.5, inv_area=ia)
    rays = _c1_rays(128)
    f64 = dev.force(rays, S.SLSIM_FP64)
    o = O.Tree(x, y, m, h, inv_area=ia).force(rays)
    assert rel_err(f64[:, :2] - o[:, :2], o[:, :2]).max() < 1e-11
    assert rel_err(f64[:, 2:] - o[:, 2:], o[:, 2:]).max() < 1e-10
    f32 = dev.force(rays, S.SLSIM_FP32C)
    assert rel_err(f32[:, :2] - f64[:, :2], f64[:, :2]).max() <= FP32C_ALPHA_TOL
    assert rel_err(f32[:, 2:3] - f64[:, 2:3], f64[:, 2:3]).max() <= KG_TOL
    assert rel_err(f32[:, 3:] - f64[:, 3:], f64[:, 3:]).max() <= KG_TOL
    # the sheet moves kappa by -inv_area * M everywhere
    f0 = S.LensHaloParticles(ctx, x, y, m, h, 0.5).force(rays, S.SLSIM_FP64)
    assert np.allclose(f64[:, 2] - f0[:, 2], -ia * m.astype(np.float64).sum(), rtol=1e-9)


def test_plane_force_mixed_sign_masses_with_sheet(ctx):
    """Masses of both signs (a compensated field: every tenth particle carries
    a negative mass) with inv_area != 0: cells whose masses nearly cancel
    have centres of mass far outside them, so the sheet's per-cell
    correction uses the signed dipole.  FP64 against the oracle, FP32C
    against FP64."""
    x, y, m, h = W.c1_particles()
    m = m.copy()
    m[::10] *= np.float32(-0.8)
    span = max(x.max() - x.min(), y.max() - y.min())
    ia = 1.0 / (0.25 * span * span)
    dev = S.LensHaloParticles(ctx, x, y, m, h, 0.5, inv_area=ia)
    rays = _c1_rays(128)
    f64 = dev.force(rays, S.SLSIM_FP64)
    o = O.Tree(x, y, m, h, inv_area=ia).force(rays)
    assert rel_err(f64[:, :2] - o[:, :2], o[:, :2]).max() < 1e-11
    assert rel_err(f64[:, 2:] - o[:, 2:], o[:, 2:]).max() < 1e-10
    f32 = dev.force(rays, S.SLSIM_FP32C)
    assert rel_err(f32[:, :2] - f64[:, :2], f64[:, :2]).max() <= FP32C_ALPHA_TOL
    assert rel_err(f32[:, 2:3] - f64[:, 2:3], f64[:, 2:3]).max() <= KG_TOL
    assert rel_err(f32[:, 3:] - f64[:, 3:], f64[:, 3:]).max() <= KG_TOL


def test_direct_sum_matches_oracle(ctx):
    x, y, m, h = W.c1_particles()
    dev = S.LensHaloParticles(ctx, x, y, m, h, 0.5)
    rays = _c1_rays(64)[::7]
    g = dev.direct_sum(rays)
    o = O.direct_sum(x, y, m, h, rays)
    assert rel_err(g[:, :2] - o[:, :2], o[:, :2]).max() < 1e-11
    assert rel_err(g[:, 2:] - o[:, 2:], o[:, 2:]).max() < 1e-9


# ----------------------------------------------------------- whole ray shoot
@pytest.fixture(scope="module")
def c1(ctx):
    lens, parts = W.build_c1_lens(ctx)
    x, y, m, h = parts
    L = O.Lens(W.C1["z_source"])
    L.moveinMainHalo(O.Tree(x, y, m, h), W.C1["z_lens"])
    N = W.C1["N"]
    rg = W.grid_range(W.C1["range_arcsec"])
    pos = O.gridmap_positions(N, (0.0, 0.0), rg)
    y_ref, A_ref, cnt = L.shoot(pos, counts=True)
    return {"lens": lens, "oracle": L, "pos": pos, "y": y_ref, "A": A_ref, "cnt": cnt,
            "N": N, "range": rg}


def test_lens_tables_match_oracle(c1):
    a, b = c1["lens"].tables(), c1["oracle"].tables()
    assert a["nplanes"] == b["nplanes"] == 1
    assert np.array_equal(a["z"], b["z"])
    assert np.allclose(a["Dl"], b["Dl"], rtol=1e-13) and a["charge"] == b["charge"]


def test_gridmap_fp64_matches_oracle_c1(c1):
    g = S.GridMap(c1["lens"], c1["N"], (0.0, 0.0), c1["range"], S.SLSIM_FP64)
    assert np.array_equal(g.positions(), c1["pos"])
    ea, ek, eg = assert_fields_close(c1["pos"], g.y(), g.A(), c1["y"], c1["A"], ALPHA_TOL, KG_TOL)
    assert ea < 1e-10 and ek < 1e-8 and eg < 1e-8     # same arithmetic: rounding only


def test_interaction_counts_match_oracle(c1):
    g = S.GridMap(c1["lens"], c1["N"], (0.0, 0.0), c1["range"], S.SLSIM_FP64, shoot=False)
    cnt = g.counts()
    assert np.array_equal(cnt, c1["cnt"])


def test_gridmap_fp32c_vs_fp64_c1(c1):
    g64 = S.GridMap(c1["lens"], c1["N"], (0.0, 0.0), c1["range"], S.SLSIM_FP64)
    g32 = S.GridMap(c1["lens"], c1["N"], (0.0, 0.0), c1["range"], S.SLSIM_FP32C)
    ea, ek, eg = assert_fields_close(c1["pos"], g32.y(), g32.A(), g64.y(), g64.A(),
                                     FP32C_ALPHA_TOL, KG_TOL)
    print("fp32c vs fp64: alpha %.2e kappa %.2e gamma %.2e" % (ea, ek, eg))


def test_array_rays_bitwise_equal_grid_rays(c1):
    # the array path Morton-sorts rays into different bundles: each ray's
    # result must not depend on its bundle mates
    rng = np.random.default_rng(3)
    perm = rng.permutation(c1["pos"].shape[0])
    for prec in (S.SLSIM_FP64, S.SLSIM_FP32C):
        g = S.GridMap(c1["lens"], c1["N"], (0.0, 0.0), c1["range"], prec)
        y, A = c1["lens"].rayshooterInternal(c1["pos"][perm], precision=prec)
        assert np.array_equal(y, g.y()[perm]) and np.array_equal(A, g.A()[perm])


def test_gridmap_repeatable(c1):
    g1 = S.GridMap(c1["lens"], c1["N"], (0.0, 0.0), c1["range"], S.SLSIM_FP32C)
    g2 = S.GridMap(c1["lens"], c1["N"], (0.0, 0.0), c1["range"], S.SLSIM_FP32C)
    assert np.array_equal(g1.y(), g2.y()) and np.array_equal(g1.A(), g2.A())


# ------------------------------------------------------------- multi-plane
@pytest.fixture(scope="module")
def field(ctx):
    # 10 planes of 2^16 particles with the particle mass of the 2^20-particle
    # configuration (halo masses scaled by 1/16)
    n = 1 << 16
    lens, parts, zl = W.build_field_lens(ctx, config=2, nplanes=10, n=n, keep_particles=True,
                                         mass_scale=1.0 / 16)
    L = O.Lens(3.0)
    for (x, y, m, h, z) in parts:
        L.moveinMainHalo(O.Tree(x, y, m, h), z)
    return {"lens": lens, "oracle": L, "parts": parts}


def test_multiplane_tables(field):
    a, b = field["lens"].tables(), field["oracle"].tables()
    assert a["nplanes"] == b["nplanes"] == 10
    assert np.array_equal(a["z"], b["z"])
    assert np.allclose(a["Dl"], b["Dl"], rtol=1e-13)


def test_multiplane_gridmap_matches_oracle(field):
    N, rg = 128, W.grid_range(300.0)
    pos = O.gridmap_positions(N, (0.0, 0.0), rg)
    y_ref, A_ref, cnt = field["oracle"].shoot(pos, counts=True)
    g = S.GridMap(field["lens"], N, (0.0, 0.0), rg, S.SLSIM_FP64)
    ea, ek, eg = assert_fields_close(pos, g.y(), g.A(), y_ref, A_ref, ALPHA_TOL, KG_TOL)
    assert ea < 1e-9
    assert np.array_equal(g.counts(), cnt)


def test_multiplane_fp32c_vs_fp64(field):
    """FP32C deflections within FP32C_ALPHA_TOL of FP64 at every ray.  Kappa and
    gamma of the composite lens amplify any change of a ray's path (rays pass
    within a softening length of particles on the later planes), so the
    FP32-rounded path moves them by more than the per-plane FP32 error: they
    are held to KG_TOL at the 99.9th percentile and to KG_MAX_FP32C at the
    worst ray, the scale of the change the FP64 path itself shows when its
    input rays move by 1e-7 of the rms deflection (about one FP32 rounding).
    Per plane, where nothing amplifies, test_field_plane_force_fp32c holds
    every ray to KG_TOL."""
    N, rg = 128, W.grid_range(300.0)
    pos = O.gridmap_positions(N, (0.0, 0.0), rg)
    g = S.GridMap(field["lens"], N, (0.0, 0.0), rg, S.SLSIM_FP64)
    g32 = S.GridMap(field["lens"], N, (0.0, 0.0), rg, S.SLSIM_FP32C)
    y64, A64 = g.y(), g.A()
    a, k, gm = lensing_fields(pos, g32.y(), g32.A())
    ar, kr, gr = lensing_fields(pos, y64, A64)
    assert rel_err(a - ar, ar).max() <= FP32C_ALPHA_TOL
    ek, eg = rel_err(k - kr, kr), rel_err(gm - gr, gr)
    assert np.quantile(ek, 0.999) <= KG_TOL and np.quantile(eg, 0.999) <= KG_TOL
    amp = 1e-7 * np.sqrt(np.mean(np.sum(ar ** 2, 1)))
    phi = np.random.default_rng(3).uniform(0, 2 * np.pi, pos.shape[0])
    xn = pos + amp * np.stack([np.cos(phi), np.sin(phi)], 1)
    yn, An = field["lens"].rayshooterInternal(xn, precision=S.SLSIM_FP64)
    _, kn, gn = lensing_fields(xn, yn, An)
    nk, ng = rel_err(kn - kr, kr), rel_err(gn - gr, gr)
    print("fp32c kappa max %.3e p99.9 %.3e | gamma max %.3e p99.9 %.3e | fp64 nudge kappa "
          "max %.3e gamma max %.3e" % (ek.max(), np.quantile(ek, 0.999), eg.max(),
                                       np.quantile(eg, 0.999), nk.max(), ng.max()))
    # fixed bounds for the worst ray: the conditioning of the composite lens
    # (the FP64 walk's own nudge response, printed above) is where they come from
    assert ek.max() <= KG_MAX_FP32C and eg.max() <= KG_MAX_FP32C


def test_field_plane_force_fp32c(field):
    """One plane of the field lens, FP32C against FP64, every ray."""
    x, y, m, h, z = field["parts"][5]
    dev = S.LensHaloParticles(field["lens"].ctx, x, y, m, h, z)
    span = max(x.max() - x.min(), y.max() - y.min())
    g = (np.arange(256) + 0.5) / 256 - 0.5
    gx, gy = np.meshgrid(g * span, g * span)
    rays = np.stack([gx.ravel() + x.mean(), gy.ravel() + y.mean()], 1)
    f64 = dev.force(rays, S.SLSIM_FP64)
    f32 = dev.force(rays, S.SLSIM_FP32C)
    assert rel_err(f32[:, :2] - f64[:, :2], f64[:, :2]).max() <= FP32C_ALPHA_TOL
    assert rel_err(f32[:, 2:3] - f64[:, 2:3], f64[:, 2:3]).max() <= KG_TOL
    assert rel_err(f32[:, 3:] - f64[:, 3:], f64[:, 3:]).max() <= KG_TOL


def test_per_ray_source_redshifts_match_oracle(field):
    rng = np.random.default_rng(11)
    n = 3000
    x = (rng.random((n, 2)) - 0.5) * W.grid_range(300.0)
    zs = rng.uniform(0.05, 4.0, n)          # includes sources in front of planes
    y, A = field["lens"].rayshooterInternal(x, zs=zs, precision=S.SLSIM_FP64)
    y_ref, A_ref = field["oracle"].shoot(x, zs=zs)
    assert_fields_close(x, y, A, y_ref, A_ref, ALPHA_TOL, KG_TOL)
    front = zs <= np.float32(0.1)
    assert np.array_equal(y[front], x[front])


def test_reset_source_plane(field):
    lens = field["lens"]
    N, rg = 64, W.grid_range(300.0)
    pos = O.gridmap_positions(N, (0.0, 0.0), rg)
    j, zi = lens.ResetSourcePlane(1.2)
    g = S.GridMap(lens, N, (0.0, 0.0), rg, S.SLSIM_FP64)
    lens.RevertSourcePlane()
    assert zi == 1.2
    y_ref, A_ref = field["oracle"].shoot(pos, zs=np.full(pos.shape[0], 1.2))
    assert_fields_close(pos, g.y(), g.A(), y_ref, A_ref, 1e-9, 1e-8)
    # the plane index and, with nearest, the implanted redshift as the
    # oracle's restatement of ResetSourcePlane gives them
    t = field["oracle"].tables()
    zp = t["z"]
    cases = [0.05, zp[0], 0.5 * (zp[0] + zp[1]), zp[3], zp[3] + 1e-9, 1.2,
             zp[-2], 0.5 * (zp[-2] + zp[-1]), zp[-1] + 0.5]
    for z in cases:
        for nearest in (False, True):
            got = lens.ResetSourcePlane(float(z), nearest)
            want = O.reset_source_plane(t, float(z), nearest)
            assert got[0] == want[0] and abs(got[1] - want[1]) <= 1e-15 * want[1], (z, nearest)
    lens.RevertSourcePlane()


# ------------------------------------------------------------ small cases
def test_no_plane_lens_is_identity(ctx):
    lens = S.Lens(ctx, 1.0)
    g = S.GridMap(lens, 16, (0.01, -0.02), 1e-3)
    assert np.array_equal(g.y(), g.positions())
    assert np.array_equal(g.A(), np.tile([1.0, 0, 0, 1.0], (256, 1)))


def test_ragged_grid_sizes(c1):
    for N in (1, 7, 63, 65):
        pos = O.gridmap_positions(N, (1e-5, 0.0), c1["range"])
        y_ref, A_ref = c1["oracle"].shoot(pos)
        g = S.GridMap(c1["lens"], N, (1e-5, 0.0), c1["range"], S.SLSIM_FP64)
        if N == 1:
            # one ray: its kappa is zero to rounding, so the rms floor of the
            # relative measure is no scale; A and y directly
            assert np.abs(g.A() - A_ref).max() <= 1e-12
            assert np.abs(g.y() - y_ref).max() <= 1e-12 * np.abs(y_ref).max()
        else:
            assert_fields_close(pos, g.y(), g.A(), y_ref, A_ref, 1e-9, 1e-8)


def test_error_paths(ctx):
    with pytest.raises(S.SlsimError) as e:
        S.Lens(ctx, 2.0, 0.3, 0.6)
    assert e.value.code == 6
    small = S.Context(0, hbm_budget=1 << 20)
    x, y, m, h = W.c1_particles()
    with pytest.raises(S.SlsimError) as e:
        S.LensHaloParticles(small, x, y, m, h, 0.5)
    assert e.value.code == 3
    assert small.memory()[0] == 0          # nothing leaked on the failed build
    small.close()
    lens = S.Lens(ctx, 2.0)
    with pytest.raises(S.SlsimError):
        S.GridMap(lens, 0, (0, 0), 1e-3)


def test_multi_halo_plane_matches_oracle(ctx):
    """Several particle halos on one plane (LensPlaneSingular sums them,
    MultiPlane/planes.cpp:114-161): three halos at z=0.5, one with a mass
    sheet, and one halo at z=1.  FP64 against the oracle, per-ray counts
    exactly; FP32C (one traversal loop over the plane's trees) against FP64."""
    rng = np.random.default_rng(12)
    lens = S.Lens(ctx, 2.0)
    L = O.Lens(2.0)
    specs = [(0.5, 0.0, (0.0, 0.0)), (0.5, 0.0, (0.02, -0.01)), (0.5, 40.0, (-0.015, 0.02)),
             (1.0, 0.0, (0.01, 0.01))]
    for k, (z, ia, c) in enumerate(specs):
        n = 8000 + 1000 * k
        x, y = S.gen_nfw(500 + k, n, 3e12, 6.0, z, center=c)
        m = np.full(n, np.float32(3e12 / n), np.float32)
        h = np.full(n, np.float32(2e-4), np.float32)
        lens.moveinMainHalo(S.LensHaloParticles(ctx, x, y, m, h, z, inv_area=ia))
        L.moveinMainHalo(O.Tree(x, y, m, h, inv_area=ia), z)
    assert lens.tables()["nplanes"] == L.tables()["nplanes"] == 2
    N, rg = 128, W.grid_range(40.0)
    pos = O.gridmap_positions(N, (0.0, 0.0), rg)
    y_ref, A_ref, cnt = L.shoot(pos, counts=True)
    g = S.GridMap(lens, N, (0.0, 0.0), rg, S.SLSIM_FP64)
    ea, _, _ = assert_fields_close(pos, g.y(), g.A(), y_ref, A_ref, ALPHA_TOL, KG_TOL)
    assert ea < 1e-9
    assert np.array_equal(g.counts(), cnt)
    g32 = S.GridMap(lens, N, (0.0, 0.0), rg, S.SLSIM_FP32C)
    assert_fields_close(pos, g32.y(), g32.A(), g.y(), g.A(), FP32C_ALPHA_TOL, KG_TOL)


def test_tree_max_range_matches_oracle(ctx):
    """LensHaloParticles' max_range (include/particle_halo.h:58-75, 416;
    include/quadTree.h:370-380): the device tree's cells equal the oracle's
    bit for bit (coarse leaves far beyond the range included), for float64
    and float32 positions, and a lens of it shoots as the oracle's does."""
    rng = np.random.default_rng(21)
    r = 0.05 * rng.random(4000) ** 2
    ph = 2 * np.pi * rng.random(4000)
    x = np.concatenate([r * np.cos(ph), 1.0 + 0.01 * rng.random(3000)])
    y = np.concatenate([r * np.sin(ph), 0.01 * rng.random(3000)])
    m = np.full(x.size, 1e8, np.float32)
    h = np.full(x.size, 1e-5, np.float32)
    for R in (0.02, 0.3):
        dev = S.LensHaloParticles(ctx, x, y, m, h, 0.5, max_range=R)
        orc = O.Tree(x, y, m, h, max_range=R)
        _compare_trees(dev, orc)
        q = orc.export_nodes()[1]
        assert np.any((q[:, 3] == 1) & (q[:, 1] > 20))       # coarse leaves exist
        xf, yf = x.astype(np.float32), y.astype(np.float32)
        _compare_trees(S.LensHaloParticles(ctx, xf, yf, m, h, 0.5, max_range=R),
                       O.Tree(xf, yf, m, h, max_range=R))
    lens = S.Lens(ctx, 2.0)
    lens.moveinMainHalo(S.LensHaloParticles(ctx, x, y, m, h, 0.5, max_range=0.02))
    L = O.Lens(2.0)
    L.moveinMainHalo(O.Tree(x, y, m, h, max_range=0.02), 0.5)
    N, rg = 48, W.grid_range(60.0)
    g = S.GridMap(lens, N, (0.0, 0.0), rg, S.SLSIM_FP64)
    yr, Ar = L.shoot(O.gridmap_positions(N, (0.0, 0.0), rg))
    assert np.abs(g.y() - yr).max() <= 1e-12 * np.abs(yr).max()
    assert np.abs(g.A() - Ar).max() <= 1e-8


def test_full_size_c2_fp64_matches_oracle_on_subsample(ctx):
    """The multi-plane configuration at full size (10 planes x 2^20 particles,
    2048^2 rays over 300"): the FP64 GridMap against the oracle on 1,024 rays
    spread over the grid (every 4096th ray), deflection, kappa and gamma to
    ALPHA_TOL / KG_TOL and within 1e-9 in alpha, per-ray counts exactly."""
    print("case: full-size C2 build", flush=True)
    lens, parts, _ = W.build_field_lens(ctx, config=2, nplanes=10, n=1 << 20, keep_particles=True)
    L = O.Lens(3.0)
    for (x, y, m, h, z) in parts:
        L.moveinMainHalo(O.Tree(x, y, m, h), z)
    N, rg = 2048, W.grid_range(300.0)
    print("case: full-size C2 FP64 GridMap", flush=True)
    g = S.GridMap(lens, N, (0.0, 0.0), rg, S.SLSIM_FP64)
    sub = np.arange(0, N * N, 4096)
    pos = O.gridmap_positions(N, (0.0, 0.0), rg)[sub]
    assert np.array_equal(g.positions()[sub], pos)
    y_ref, A_ref, cnt = L.shoot(pos, counts=True)
    ea, _, _ = assert_fields_close(pos, g.y()[sub], g.A()[sub], y_ref, A_ref, ALPHA_TOL, KG_TOL)
    assert ea < 1e-9
    assert np.array_equal(g.counts()[sub], cnt)
    g.close()
    lens.close()
