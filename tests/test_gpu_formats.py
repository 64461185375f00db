"""Particle files through the device: MakeParticleLenses with smoothing sizes
from the device k-NN (csv4, Gadget-2), the ASCII LensHaloParticles and its
sizes cache, CreateHalos into a Lens, and GridMap::writeFits.  Sizes are
checked bit for bit against the oracle's restatement of TreeSimple's search;
files against the oracle's format restatements.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import slsimlib_amd as S
from oracle import formats as F

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _blobs(rng, n, scale=1.0):
    """n particles in three Gaussian clumps (comoving Mpc/h)."""
    c = np.array([[0, 0, 0], [0.8, -0.3, 0.2], [-0.5, 0.6, -0.4]]) * scale
    k = rng.integers(0, 3, n)
    return (c[k] + rng.normal(size=(n, 3)) * 0.15 * scale).astype(np.float32)


def _csv4(path, xyz, mass):
    with open(path, "w") as f:
        f.write("# x,y,z,mass\n")
        for (a, b, c), m in zip(xyz, mass):
            f.write("%r,%r,%r,%r\n" % (float(a), float(b), float(c), float(m)))
    return str(path)


def test_csv4_sizes_cache_and_halos(ctx, tmp_path):
    rng = np.random.default_rng(21)
    n = 20000
    xyz = _blobs(rng, n)
    mass = np.full(n, np.float32(1e9))
    fn = _csv4(tmp_path / "p.csv", xyz, mass)
    print("case: csv4 n=%d nsmooth=32" % n, flush=True)
    m = S.MakeParticleLenses(ctx, fn, S.SLSIM_FMT_CSV4, 32)
    d = m.data.copy()
    assert np.array_equal(d["x"], xyz) and np.array_equal(d["Mass"], mass)
    want = O.smoothing_sizes(xyz[:, 0], xyz[:, 1], xyz[:, 2], 32)
    assert np.array_equal(d["Size"], want)
    assert open(fn + "_S32.glmb", "rb").read() == F.glmb_bytes(d, 32, [n, 0, 0, 0, 0, 0], -1.0)
    # CreateHalos: units, then one halo whose plane is the one built from the same arrays
    halos = m.CreateHalos(0.7, 0.5, 0.0)
    assert len(halos) == 1
    du = m.data.copy()
    assert du.tobytes() == F.physical_units(d, 0.7).tobytes()
    direct = S.LensHaloParticles(ctx, np.ascontiguousarray(du["x"][:, 0]),
                                 np.ascontiguousarray(du["x"][:, 1]), du["Mass"], du["Size"], 0.5)
    outs = []
    for h in (halos[0], direct):
        lens = S.Lens(ctx, 2.0)
        lens.moveinMainHalo(h)
        g = S.GridMap(lens, 64, (0.0, 0.0), 120 * S.arcsecTOradians)
        outs.append((g.y(), g.A()))
        g.close()
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("ignore_type", [False, True])
def test_gadget2_pipeline(ctx, tmp_path, ignore_type):
    rng = np.random.default_rng(22)
    npart = [5000, 7000, 0, 0, 300, 0]
    masstab = [0.0, 0.0125, 0.0, 0.0, 0.0, 0.0]
    pos = (_blobs(rng, sum(npart), 1000.0) + 5000.0).astype(np.float32)   # comoving kpc/h
    mb = (rng.random(5300) * 0.01 + 0.001).astype(np.float32)
    blob = F.gadget_bytes(npart, masstab, pos, mb, redshift=0.3, big_endian=True)
    fn = str(tmp_path / "snap_050")
    with open(fn, "wb") as f:
        f.write(blob)
    print("case: gadget2 ignore_type=%s" % ignore_type, flush=True)
    m = S.MakeParticleLenses(ctx, fn, S.SLSIM_FMT_GADGET2, 24,
                             ignore_type_in_smoothing=ignore_type)
    d = m.data.copy()
    hd, p0 = F.gadget_parse(blob)
    p0 = F.gadget_units(p0, hd["redshift"])
    assert m.getnp() == npart and m.sim_redshift() == 0.3
    assert m.getmp() == [float(np.float32(v)) for v in masstab]
    assert np.all(np.diff(d["type"]) >= 0)

    def rows(a):
        return np.sort(np.concatenate([a["x"], a["Mass"][:, None],
                                       a["type"][:, None].astype(np.float32)], 1).view(
            np.dtype((np.void, 20))).ravel())
    assert np.array_equal(rows(d), rows(p0))                   # same particles, grouped by type
    if ignore_type:
        want = O.smoothing_sizes(d["x"][:, 0], d["x"][:, 1], d["x"][:, 2], 24)
        assert np.array_equal(d["Size"], want)
    else:
        skip = 0
        for t, k in enumerate(npart):
            if k:
                seg = d[skip:skip + k]
                want = O.smoothing_sizes(seg["x"][:, 0], seg["x"][:, 1], seg["x"][:, 2], 24)
                assert np.array_equal(seg["Size"], want), t
            skip += k
    cache = fn + "_S24.glmb"
    q, ns, nb, z = F.glmb_parse(open(cache, "rb").read())
    assert q.tobytes() == d.tobytes() and nb == npart and z == 0.3


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ascii_halo_sizes_computed_then_cached(ctx, tmp_path, dtype):
    rng = np.random.default_rng(23)
    n = 6000
    xyz = _blobs(rng, n).astype(np.float64) * 3
    m = rng.random(n) * 1e10 + 1e9
    fn = str(tmp_path / "halo.txt")
    with open(fn, "w") as f:
        f.write("# nparticles %d\n" % n)
        for (a, b, c), w in zip(xyz, m):
            f.write("%.12g %.12g %.12g %.8g\n" % (a, b, c, w))
    print("case: ascii dtype=%s" % np.dtype(dtype).name, flush=True)
    halo = S.LensHaloParticles.from_ascii(ctx, fn, 0.4, 16, recenter=True, MinPSize=0.02,
                                          dtype=dtype)
    xyz0, m0 = F.read_ascii(fn, True, dtype)
    raw = O.smoothing_sizes(xyz0[:, 0], xyz0[:, 1], xyz0[:, 2], 16)   # comoving positions
    assert open(fn + ".16sizes").read() == F.sizes_text(raw, 16)
    size0 = np.where(raw < 0.02, np.float32(0.02), raw)
    assert np.array_equal(halo.size, size0)
    xyz1, m1, info1 = F.ascii_halo(xyz0, m0, size0, 0.4, True, True)
    assert np.array_equal(halo.xyz, xyz1) and np.array_equal(halo.mass, m1)
    assert halo.halo_info["mass"] == info1["mass"] and halo.halo_info["rsize"] == info1["rsize"]
    # the second construction reads the sizes file (six significant digits)
    again = S.LensHaloParticles.from_ascii(ctx, fn, 0.4, 16, recenter=True, MinPSize=0.02,
                                           dtype=dtype)
    s6 = np.array([F._strtof("%.6g" % v) for v in raw], np.float32)
    assert np.array_equal(again.size, np.where(s6 < 0.02, np.float32(0.02), s6))


def test_gridmap_writefits(ctx, tmp_path):
    x, y = S.gen_nfw(31, 8192, 1e13, 8.0, 0.5)
    m = np.full(8192, np.float32(1e13 / 8192))
    h = np.full(8192, np.float32(2e-4))
    lens = S.Lens(ctx, 2.0)
    lens.moveinMainHalo(S.LensHaloParticles(ctx, x, y, m, h, 0.5))
    g = S.GridMap(lens, 96, (1e-5, -2e-5), 40 * S.arcsecTOradians)
    for var, dt in (("KAPPA", np.float64), ("GAMMA1", np.float32)):
        fn = str(tmp_path / ("%s.fits" % var))
        g.writeFits(var, fn, dt)
        cards, img = F.fits_parse(open(fn, "rb").read())
        assert np.array_equal(img, g.writePixelMap(var).astype(dt))
        assert F.fits_value(cards["CD2_2"]) == float("%.15G" % (180 * (g.range / 95) / np.pi))
    g.close()


def test_cpp_dropin_formats_program(ctx, tmp_path):
    exe = os.path.join(ROOT, "tests", "cpp", "dropin_formats")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/dropin_formats is not built (run __graft_entry__.build())")
    rng = np.random.default_rng(24)
    n = 12000
    xyz = _blobs(rng, n)
    mass = np.full(n, np.float32(5e8))
    fn = _csv4(tmp_path / "p.csv", xyz, mass)
    out = str(tmp_path / "kappa.fits")
    r = subprocess.run([exe, fn, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "1 halos" in r.stdout
    _, img = F.fits_parse(open(out, "rb").read())
    # the same chain through the Python view (the cache written by the program is read back)
    m = S.MakeParticleLenses(ctx, fn, S.SLSIM_FMT_CSV4, 32)
    halos = m.CreateHalos(0.7, 0.5, 0.0)
    lens = S.Lens(ctx, 2.0)
    lens.moveinMainHalo(halos[0])
    g = S.GridMap(lens, 128, (0.0, 0.0), 60 * S.arcsecTOradians)
    assert np.array_equal(img, g.writePixelMap("KAPPA"))
    g.close()
