"""Halo-of-halos planes (LensHaloHalos<LensHaloNFW>) on the device: the
TreeQuadHalos walk and the sum over every halo against the oracle, a
light-cone that mixes particle halos and NFW halos (one plane holding both)
through GridMap against the oracle's Lens, and batches and the server with
such a lens giving each request its lone GridMap's bits.
"""
import threading

import numpy as np
import pytest

import oracle as O
import slsimlib_amd as S
from slsimlib_amd import workloads as W

pytestmark = pytest.mark.gpu


def _halo_set(n, seed, spread=0.4, rmax=(0.01, 0.25)):
    """n NFW halos: masses log-uniform 1e10-1e13, Rsize in rmax (a few
    larger than their cells), concentrations 3-15, Gaussian positions."""
    rng = np.random.default_rng(seed)
    x = rng.normal(size=n) * spread
    y = rng.normal(size=n) * spread
    m = (10 ** rng.uniform(10, 13, n)).astype(np.float32)
    rs = (10 ** rng.uniform(np.log10(rmax[0]), np.log10(rmax[1]), n)).astype(np.float32)
    c = rng.uniform(3, 15, n).astype(np.float32)
    return x, y, m, rs, c


def _rays(x, y, rs, c, n, seed):
    rng = np.random.default_rng(seed)
    r = rng.normal(size=(n, 2)) * 0.5
    k = min(64, x.size)
    r[:k, 0] = x[:k] + (rs[:k] / c[:k]) * 0.004      # x < 0.01: the closed-form interval
    r[:k, 1] = y[:k]
    r[k, 0], r[k, 1] = x[k], y[k]                     # on a halo centre (r^2 clamp)
    return r


def _close(got, want, near=65):
    for k in range(5):
        scale = np.abs(want[near:, k]).max()
        assert np.abs(got[near:, k] - want[near:, k]).max() <= 1e-11 * scale, k
        assert np.all(np.abs(got[:near, k] - want[:near, k]) <=
                      1e-9 * np.abs(want[:near, k]) + 1e-11 * scale), k


@pytest.mark.parametrize("theta,ia", [(0.1, 0.0), (0.5, 1.5e-3), (1e-6, 0.0)])
def test_halo_plane_force_matches_oracle(ctx, theta, ia):
    x, y, m, rs, c = _halo_set(2000, 11)
    print("case: halo plane theta=%g inv_area=%g" % (theta, ia), flush=True)
    h = S.LensHaloHalos(ctx, x, y, m, rs, c, 0.4, inv_area=ia, theta=theta)
    t = O.HaloTree(x, y, m, rs, c, inv_area=ia, theta=theta)
    rays = _rays(x, y, rs, c, 4096, 12)
    got = h.force(rays)
    want = t.force(rays)
    # rays within 0.01 rscale of a centre take the profile's closed forms,
    # whose cancellation (the reference's own) magnifies the last-bit
    # differences between the device's and the host's log/atanh: per ray
    _close(got, want)
    _close(h.direct_sum(rays[:512]), t.force(rays[:512], direct=True))
    info = h.info()
    assert info["n_particles"] == 2000 and info["n_nodes"] > 1


def test_halo_plane_errors(ctx):
    x, y, m, rs, c = _halo_set(10, 13)
    c[3] = np.float32(150.0)                   # Rsize/rscale beyond the tables' end
    with pytest.raises(S.SlsimError) as e:
        S.LensHaloHalos(ctx, x, y, m, rs, c, 0.4)
    assert e.value.code == 1
    with pytest.raises(S.SlsimError) as e:
        S.LensHaloHalos(ctx, np.zeros(8), np.zeros(8), m[:8], rs[:8], np.full(8, 5, np.float32),
                        0.4)
    assert e.value.code == 4                   # particles in same place


def _mixed_lens(ctx, oracle=False):
    n = 16384
    px, py = S.gen_nfw(77, n, 1e13, 8.0, 0.5)
    pm = np.full(n, np.float32(1e13 / n), np.float32)
    ph = np.full(n, np.float32(2e-4), np.float32)
    ax, ay, am, ars, ac = _halo_set(600, 14, spread=0.15, rmax=(0.005, 0.1))
    bx, by, bm, brs, bc = _halo_set(300, 15, spread=0.2, rmax=(0.005, 0.05))
    if oracle:
        L = O.Lens(2.0)
        L.moveinMainHalo(O.Tree(px, py, pm, ph), 0.5)
        L.moveinMainHalo(O.HaloTree(ax, ay, am, ars, ac, inv_area=0.0), 0.3)
        L.moveinMainHalo(O.HaloTree(bx, by, bm, brs, bc, inv_area=0.0), 0.5)
        return L
    lens = S.Lens(ctx, 2.0)
    lens.moveinMainHalo(S.LensHaloParticles(ctx, px, py, pm, ph, 0.5))
    lens.moveinMainHalo(S.LensHaloHalos(ctx, ax, ay, am, ars, ac, 0.3))
    lens.moveinMainHalo(S.LensHaloHalos(ctx, bx, by, bm, brs, bc, 0.5))
    return lens


def test_mixed_lightcone_gridmap_matches_oracle(ctx):
    lens = _mixed_lens(ctx)
    L = _mixed_lens(ctx, oracle=True)
    assert L.tables()["nplanes"] == 2
    N, rg = 64, W.grid_range(60.0)
    g = S.GridMap(lens, N, (0.0, 0.0), rg, S.SLSIM_FP64)
    y, A = g.y(), g.A()
    yr, Ar = L.shoot(O.gridmap_positions(N, (0.0, 0.0), rg))
    assert np.abs(y - yr).max() <= 1e-12 * np.abs(yr).max()
    assert np.abs(A - Ar).max() <= 1e-8
    # FP32C: the particle walk in FP32 (compensated), the halo trees in FP64
    g32 = S.GridMap(lens, N, (0.0, 0.0), rg, S.SLSIM_FP32C)
    x = g.positions()
    a64, a32 = x - y, x - g32.y()
    assert np.abs(a32 - a64).max() <= 1e-5 * np.abs(a64).max()


def test_batch_and_server_with_halo_planes(ctx):
    mixed = _mixed_lens(ctx)
    plain, _ = W.build_c1_lens(ctx, n=16384)
    rg = W.grid_range(60.0)
    reqs = [(mixed, 48, (0.0, 0.0), rg, 0.0), (plain, 40, (1e-6, 0.0), W.grid_range(30.0), 0.0),
            (mixed, 33, (2e-5, -1e-5), rg, 1.2)]
    for prec in (S.SLSIM_FP64, S.SLSIM_FP32C):
        ref = []
        for lens, N, c, r, zs in reqs:
            g = S.GridMap(lens, N, c, r, prec, zs=zs)
            ref.append((g.y(), g.A()))
            g.close()
        b = S.Batch(ctx, reqs, prec)
        for sched in ({}, {"tile_log2": 2, "queue": 1, "order": 1, "graph": 1},
                      {"fuse": 1, "queue": 1, "graph": 1}):
            b.set_schedule(**sched)
            b.run()
            for i in range(len(reqs)):
                y, A = b.output(i)
                assert np.array_equal(y, ref[i][0]) and np.array_equal(A, ref[i][1]), (prec, i)
    srv = S.Server(ctx, hbm_budget=4 << 30)
    out = {}

    def client(k):
        r = reqs[k % len(reqs)]
        rid = srv.submit(r[0], r[1], r[2], r[3], r[4], S.SLSIM_FP32C)
        out[k] = srv.wait(rid, r[1])

    th = [threading.Thread(target=client, args=(k,)) for k in range(6)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for k, (y, A, info) in out.items():
        lens, N, c, r, zs = reqs[k % len(reqs)]
        g = S.GridMap(lens, N, c, r, S.SLSIM_FP32C, zs=zs)
        assert info["status"] == 0
        assert np.array_equal(y, g.y()) and np.array_equal(A, g.A()), k
        g.close()
    srv.close()


def test_cpp_dropin_halos_program(ctx, tmp_path):
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests",
                       "cpp", "dropin_halos")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/dropin_halos is not built (run __graft_entry__.build())")
    out = tmp_path / "halos.bin"
    img = tmp_path / "images.bin"
    crit = tmp_path / "crit.bin"
    fpl = tmp_path / "field.bin"
    r = subprocess.run([exe, str(out), str(img), str(crit), str(fpl)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    v = np.fromfile(out, dtype=np.float64).reshape(-1, 8)
    # the same light-cone through the Python view: LensHaloNFW positions are
    # theta * D_A(zlens)
    n = 16384
    px, py = S.gen_nfw(2024, n, 5e12, 8.0, 0.5)
    pm = np.full(n, np.float32(5e12 / n), np.float32)
    ph = np.full(n, np.float32(2e-4), np.float32)
    i = np.arange(200)
    m = np.array([np.float32(10.0 ** (10.5 + 2.0 * ((k * 37) % 100) / 100.0)) for k in i])
    rs = np.array([np.float32(0.01 + 0.002 * (k % 17)) for k in i])
    c = np.array([np.float32(4.0 + (k % 9)) for k in i])
    dist = S.coorDist(0.0, 0.3) / (1 + 0.3)
    hx = ((i % 20) - 9.5) * 3e-6 * dist
    hy = ((i // 20) - 4.5) * 5e-6 * dist
    lens = S.Lens(ctx, 2.0)
    lens.moveinMainHalo(S.LensHaloParticles(ctx, px, py, pm, ph, 0.5))
    lens.moveinMainHalo(S.LensHaloHalos(ctx, hx, hy, m, rs, c, 0.3))
    g = S.GridMap(lens, 64, (0.0, 0.0), 60 * S.arcsecTOradians)
    assert np.array_equal(v[:, 2:4], g.y()) and np.array_equal(v[:, 4:8], g.A())
    # GridMap::find_images from C++ equals the Python view's
    w = np.fromfile(img, dtype=np.float64)
    ys = g.y()[[1000, 2080, 2112, 3000]]
    rays, mult, _ = g.find_images(ys)
    assert np.array_equal(w[:4], mult) and np.array_equal(w[4:].reshape(-1, 8), rays)
    # GridMap::find_crit and find_magnification_contour from C++ equal the
    # Python view's
    q = np.fromfile(crit, dtype=np.float64)
    curves, hits, types, _ = g.find_crit()
    contours, chits, _, _ = g.find_magnification_contour(0.5)
    assert (q[0], q[1]) == (len(curves), len(contours)) and len(curves) > 0
    at = 2
    for cv, hb, ty in list(zip(curves, hits, types)) + list(zip(contours, chits,
                                                                [0] * len(chits))):
        assert (q[at], q[at + 1], q[at + 2]) == (ty, hb, len(cv))
        assert np.array_equal(q[at + 3:at + 3 + 2 * len(cv)].reshape(-1, 2), cv)
        at += 3 + 2 * len(cv)
    assert at == len(q)
    # LensPlaneTree from C++: its force and a lens of it equal the Python view's
    f = np.fromfile(fpl, dtype=np.float64)
    j = np.arange(30)
    pm_ = np.array([np.float32(10.0 ** (10.0 + 0.05 * k)) for k in j])
    pr_ = np.array([np.float32(0.01 + 0.001 * k) for k in j])
    pb_ = np.array([-0.4 + 0.07 * k for k in j])
    px_ = ((j % 6) - 2.5) * 7e-6 * dist
    py_ = ((j // 6) - 2.0) * 6e-6 * dist
    fp = S.LensPlaneTree.mixed(
        ctx, 0.9, np.concatenate([hx, px_]), np.concatenate([hy, py_]),
        np.concatenate([m, pm_]), np.concatenate([rs, pr_]),
        np.concatenate([np.zeros(hx.size, np.int32), np.ones(30, np.int32)]),
        np.concatenate([c.astype(np.float64), pb_]), 100.0, 2.0)
    assert np.array_equal(f[:5], fp.force(np.array([[1e-3, -2e-3]]))[0])
    lens2 = S.Lens(ctx, 2.0)
    lens2.addFieldPlane(fp)
    g2 = S.GridMap(lens2, 32, (0.0, 0.0), 60 * S.arcsecTOradians)
    v2 = f[5:].reshape(-1, 8)
    assert np.array_equal(v2[:, 2:4], g2.y()) and np.array_equal(v2[:, 4:8], g2.A())
    g2.close()
    L = O.Lens(2.0)
    L.moveinMainHalo(O.Tree(px, py, pm, ph), 0.5)
    L.moveinMainHalo(O.HaloTree(hx, hy, m, rs, c), 0.3)
    yr, Ar = L.shoot(v[:, :2])
    assert np.abs(v[:, 2:4] - yr).max() <= 1e-12 * np.abs(yr).max()
    assert np.abs(v[:, 4:8] - Ar).max() <= 1e-8
    g.close()


# --------------------------------------------------- LensPlaneTree field planes
@pytest.mark.parametrize("sigma,scr", [(0.0, 0.0), (2.0e3, 0.0), (1.0e3, 2.0)])
def test_field_plane_force_matches_oracle(ctx, sigma, scr):
    x, y, m, rs, c = _halo_set(1500, 31)
    print("case: field plane sigma=%g screening=%g" % (sigma, scr), flush=True)
    p = S.LensPlaneTree(ctx, 0.7, x, y, m, rs, c, sigma, scr)
    t = O.FieldTree(x, y, m, rs, c, sigma, scr)
    rays = _rays(x, y, rs, c, 2048, 32)
    _close(p.force(rays), t.force(rays))
    _close(p.direct_sum(rays[:256]), t.force(rays[:256], direct=True))


def _field_lens(ctx, oracle=False):
    n = 16384
    px, py = S.gen_nfw(91, n, 1e13, 8.0, 0.5)
    pm = np.full(n, np.float32(1e13 / n), np.float32)
    ph = np.full(n, np.float32(2e-4), np.float32)
    ax, ay, am, ars, ac = _halo_set(500, 33, spread=0.15, rmax=(0.005, 0.08))
    bx, by, bm, brs, bc = _halo_set(400, 34, spread=0.2, rmax=(0.005, 0.05))
    za, zb = 0.2713, 1.1372
    if oracle:
        L = O.Lens(2.5)
        L.moveinMainHalo(O.Tree(px, py, pm, ph), 0.5)
        L.addFieldPlane(O.FieldTree(ax, ay, am, ars, ac, 50.0, 0.0), za)
        L.addFieldPlane(O.FieldTree(bx, by, bm, brs, bc, 80.0, 3.0), zb)
        return L
    lens = S.Lens(ctx, 2.5)
    lens.moveinMainHalo(S.LensHaloParticles(ctx, px, py, pm, ph, 0.5))
    lens.addFieldPlane(S.LensPlaneTree(ctx, za, ax, ay, am, ars, ac, 50.0, 0.0))
    lens.addFieldPlane(S.LensPlaneTree(ctx, zb, bx, by, bm, brs, bc, 80.0, 3.0))
    return lens


def test_field_planes_lightcone_matches_oracle(ctx):
    lens = _field_lens(ctx)
    L = _field_lens(ctx, oracle=True)
    t = L.tables()
    assert t["nplanes"] == 3
    print("case: light-cone with two field planes, GridMap 64", flush=True)
    N, rg = 64, W.grid_range(60.0)
    g = S.GridMap(lens, N, (0.0, 0.0), rg, S.SLSIM_FP64)
    y, A = g.y(), g.A()
    yr, Ar = L.shoot(O.gridmap_positions(N, (0.0, 0.0), rg))
    assert np.abs(y - yr).max() <= 1e-12 * np.abs(yr).max()
    assert np.abs(A - Ar).max() <= 1e-8
    g32 = S.GridMap(lens, N, (0.0, 0.0), rg, S.SLSIM_FP32C)
    x = g.positions()
    assert np.abs((x - g32.y()) - (x - y)).max() <= 1e-5 * np.abs(x - y).max()
    # batches and the server give each request its lone GridMap's bits
    reqs = [(lens, 40, (0.0, 0.0), rg, 0.0), (lens, 33, (2e-5, 1e-5), rg, 1.4)]
    ref = []
    for lz, n_, c_, r_, zs in reqs:
        gg = S.GridMap(lz, n_, c_, r_, S.SLSIM_FP32C, zs=zs)
        ref.append((gg.y(), gg.A()))
        gg.close()
    b = S.Batch(ctx, reqs, S.SLSIM_FP32C)
    b.run()
    for i in range(len(reqs)):
        yb, Ab = b.output(i)
        assert np.array_equal(yb, ref[i][0]) and np.array_equal(Ab, ref[i][1]), i
    g.close()
    g32.close()


# ------------------------------------------- power-law members (LensHaloPowerLaw)
def _mixed_set(n, seed, frac=0.4):
    x, y, m, rs, c = _halo_set(n, seed)
    rng = np.random.default_rng(seed + 1000)
    kinds = (rng.random(n) < frac).astype(np.int32)
    params = np.where(kinds == 1, rng.uniform(-0.5, 1.8, n), c.astype(np.float64))
    return x, y, m, rs, c, kinds, params


@pytest.mark.parametrize("ia", [0.0, 1.5e-3])
def test_mixed_halo_plane_force_matches_oracle(ctx, ia):
    x, y, m, rs, c, kinds, params = _mixed_set(1800, 41)
    print("case: mixed NFW / power-law halo plane inv_area=%g" % ia, flush=True)
    h = S.LensHaloHalos.mixed(ctx, x, y, m, rs, kinds, params, 0.4, inv_area=ia, theta=0.1)
    t = O.MixedTree(x, y, m, rs, kinds, params, inv_area=float(np.float32(ia)),
                    theta=float(np.float32(0.1)))
    rays = _rays(x, y, rs, c, 4096, 42)
    _close(h.force(rays), t.force(rays))
    _close(h.direct_sum(rays[:512]), t.force(rays[:512], direct=True))
    with pytest.raises(S.SlsimError) as e:
        bad = params.copy()
        bad[np.flatnonzero(kinds == 1)[0]] = 2.0        # "bad gamma"
        S.LensHaloHalos.mixed(ctx, x, y, m, rs, kinds, bad, 0.4)
    assert e.value.code == 1


def test_mixed_field_plane_lightcone_matches_oracle(ctx):
    x, y, m, rs, c, kinds, params = _mixed_set(700, 43)
    n = 16384
    px, py = S.gen_nfw(92, n, 1e13, 8.0, 0.5)
    pm = np.full(n, np.float32(1e13 / n), np.float32)
    ph = np.full(n, np.float32(2e-4), np.float32)
    lens = S.Lens(ctx, 2.0)
    lens.moveinMainHalo(S.LensHaloParticles(ctx, px, py, pm, ph, 0.5))
    lens.addFieldPlane(S.LensPlaneTree.mixed(ctx, 0.8123, x, y, m, rs, kinds, params, 60.0, 2.5))
    L = O.Lens(2.0)
    L.moveinMainHalo(O.Tree(px, py, pm, ph), 0.5)
    L.addFieldPlane(O.MixedTree(x, y, m, rs, kinds, params, 60.0, 0.1, 2.5, True), 0.8123)
    print("case: light-cone with a mixed screened field plane, GridMap 64", flush=True)
    N, rg = 64, W.grid_range(60.0)
    g = S.GridMap(lens, N, (0.0, 0.0), rg, S.SLSIM_FP64)
    yr, Ar = L.shoot(O.gridmap_positions(N, (0.0, 0.0), rg))
    assert np.abs(g.y() - yr).max() <= 1e-12 * np.abs(yr).max()
    assert np.abs(g.A() - Ar).max() <= 1e-8
    g.close()


def test_clear_main_halos_keeps_field_planes(ctx):
    lens = _field_lens(ctx)
    lens.clearMainHalos()
    ax, ay, am, ars, ac = _halo_set(500, 33, spread=0.15, rmax=(0.005, 0.08))
    bx, by, bm, brs, bc = _halo_set(400, 34, spread=0.2, rmax=(0.005, 0.05))
    L = O.Lens(2.5)
    L.addFieldPlane(O.FieldTree(ax, ay, am, ars, ac, 50.0, 0.0), 0.2713)
    L.addFieldPlane(O.FieldTree(bx, by, bm, brs, bc, 80.0, 3.0), 1.1372)
    print("case: clearMainHalos with two field planes", flush=True)
    N, rg = 32, W.grid_range(60.0)
    g = S.GridMap(lens, N, (0.0, 0.0), rg, S.SLSIM_FP64)
    yr, Ar = L.shoot(O.gridmap_positions(N, (0.0, 0.0), rg))
    assert np.abs(g.y() - yr).max() <= 1e-12 * np.abs(yr).max()
    assert np.abs(g.A() - Ar).max() <= 1e-8
    g.close()


@pytest.mark.gpu
def test_nearest_plane_insertion_matches_oracle(ctx):
    """moveinMainHalo(halo, addplanes=false) (include/lens.h:172-185): the
    halo joins the nearest main plane; on an empty lens the reference adds it
    to its own new plane twice."""
    n = 4096
    px, py = S.gen_nfw(78, n, 1e13, 8.0, 0.5)
    pm = np.full(n, np.float32(1e13 / n), np.float32)
    ph = np.full(n, np.float32(2e-4), np.float32)
    hx, hy, m, rs, c = _halo_set(200, 16, spread=0.15, rmax=(0.005, 0.05))
    qx, qy = S.gen_nfw(79, n, 5e12, 6.0, 0.3)
    N, rg = 48, W.grid_range(60.0)
    pos = O.gridmap_positions(N, (0.0, 0.0), rg)

    lens = S.Lens(ctx, 2.0)
    lens.moveinMainHalo(S.LensHaloParticles(ctx, px, py, pm, ph, 0.5))
    lens.moveinMainHalo(S.LensHaloParticles(ctx, qx, qy, pm, ph, 0.3))
    lens.moveinMainHalo(S.LensHaloHalos(ctx, hx, hy, m, rs, c, 0.35), addplanes=False)
    L = O.Lens(2.0)
    L.moveinMainHalo(O.Tree(px, py, pm, ph), 0.5)
    L.moveinMainHalo(O.Tree(qx, qy, pm, ph), 0.3)
    L.moveinMainHalo(O.HaloTree(hx, hy, m, rs, c, inv_area=0.0), 0.35, addplanes=False)
    assert lens.tables()["nplanes"] == L.tables()["nplanes"] == 2
    g = S.GridMap(lens, N, (0.0, 0.0), rg, S.SLSIM_FP64)
    yr, Ar = L.shoot(pos)
    assert np.abs(g.y() - yr).max() <= 1e-12 * np.abs(yr).max()
    assert np.abs(g.A() - Ar).max() <= 1e-8

    one = S.Lens(ctx, 2.0)
    one.moveinMainHalo(S.LensHaloParticles(ctx, px, py, pm, ph, 0.5), addplanes=False)
    E = O.Lens(2.0)
    E.moveinMainHalo(O.Tree(px, py, pm, ph), 0.5, addplanes=False)
    g1 = S.GridMap(one, N, (0.0, 0.0), rg, S.SLSIM_FP64)
    yr, Ar = E.shoot(pos)
    assert np.abs(g1.y() - yr).max() <= 1e-12 * np.abs(yr).max()
    assert np.abs(g1.A() - Ar).max() <= 1e-8


@pytest.mark.gpu
def test_replace_main_halos_equals_a_fresh_lens(ctx):
    """Lens::replaceMainHalos (include/lens.h:244-263): after it the lens
    shoots as a new lens holding only the new halos."""
    n = 4096
    px, py = S.gen_nfw(80, n, 1e13, 8.0, 0.5)
    qx, qy = S.gen_nfw(81, n, 4e12, 6.0, 0.4)
    pm = np.full(n, np.float32(1e13 / n), np.float32)
    ph = np.full(n, np.float32(2e-4), np.float32)
    lens = S.Lens(ctx, 2.0)
    lens.moveinMainHalos([S.LensHaloParticles(ctx, px, py, pm, ph, 0.5)])
    lens.replaceMainHalos([S.LensHaloParticles(ctx, qx, qy, pm, ph, 0.4),
                           S.LensHaloParticles(ctx, px, py, pm, ph, 0.7)])
    assert lens.getNMainHalos() == 2 and lens.tables()["nplanes"] == 2
    fresh = S.Lens(ctx, 2.0)
    fresh.moveinMainHalo(S.LensHaloParticles(ctx, qx, qy, pm, ph, 0.4))
    fresh.moveinMainHalo(S.LensHaloParticles(ctx, px, py, pm, ph, 0.7))
    N, rg = 32, W.grid_range(40.0)
    a = S.GridMap(lens, N, (0.0, 0.0), rg, S.SLSIM_FP64)
    b = S.GridMap(fresh, N, (0.0, 0.0), rg, S.SLSIM_FP64)
    assert np.array_equal(a.y(), b.y()) and np.array_equal(a.A(), b.A())
