"""Particle files in and pixel maps out (MakeParticleLenses' readers, the glmb
and sizes caches, LensHaloParticles' ASCII file, PixelMap::printFITS): the
library against the oracle's restatement, and the oracle against values the
formats fix.  CPU only: no call here computes smoothing sizes (those paths
are in test_gpu_formats.py).
"""
import os
import struct

import numpy as np
import pytest

import slsimlib_amd as S
from oracle import formats as F


def _write(path, text, mode="w"):
    with open(path, mode) as f:
        f.write(text)
    return str(path)


def _particles(rng, n, ntypes=1):
    p = np.zeros(n, S.PARTICLE_DTYPE)
    p["x"] = rng.normal(size=(n, 3)).astype(np.float32)
    p["Mass"] = (rng.random(n) + 0.5).astype(np.float32)
    p["Size"] = (rng.random(n) * 0.1 + 0.01).astype(np.float32)
    p["type"] = np.sort(rng.integers(0, ntypes, n)).astype(np.int32)
    return p


def _csv_text(p, columns, header="# particles\n#x,y,z\n", crlf=False, fmt=repr):
    rows = []
    for r in p:
        v = [r["x"][0], r["x"][1], r["x"][2], r["Mass"], r["Size"]][:columns]
        rows.append(",".join(fmt(float(np.float32(t))) for t in v))
    nl = "\r\n" if crlf else "\n"
    return header + nl.join(rows) + nl


# ------------------------------------------------------------ oracle pins
def test_oracle_csv_known_values(tmp_path):
    fn = _write(tmp_path / "k.csv", "# c\n0.5,1.25,-2,3e2,0.125\n 1.5,0x1p-2,+7, 1e-3 ,2\r\n")
    rc, p, _ = F.read_csv(fn, 5)
    assert rc == 0 and p.size == 2
    assert p["x"][0].tolist() == [0.5, 1.25, -2.0]
    assert p["Mass"][0] == 300.0 and p["Size"][0] == 0.125
    assert p["x"][1].tolist() == [1.5, 0.25, 7.0]          # stof: leading blanks, hex, '+'
    assert p["Mass"][1] == np.float32(1e-3) and p["Size"][1] == 2.0
    rc, p, _ = F.read_csv(fn, 3)
    assert rc == 0 and p["Mass"].tolist() == [1.0, 1.0]     # csv3: unit masses
    assert F.read_csv(_write(tmp_path / "few.csv", "1,2\n3,4\n"), 3)[0] == 1
    assert F.read_csv(_write(tmp_path / "e.csv", "1,2,3\n1,,3\n"), 3)[:3:2] == (2, 1)
    assert F.read_csv(_write(tmp_path / "r.csv", "1,2,3\n1,2,1e-50\n"), 3)[:3:2] == (3, 1)
    assert F.read_csv(_write(tmp_path / "s.csv", "1,2,3\n1,2\n"), 3)[:3:2] == (4, 1)


def test_oracle_glmb_layout():
    rng = np.random.default_rng(3)
    p = _particles(rng, 5, 2)
    blob = F.glmb_bytes(p, 17, [3, 2, 0, 0, 0, 0], 0.25)
    assert len(blob) == 68 + 24 * 5
    assert struct.unpack_from("<i", blob, 0)[0] == 17
    assert struct.unpack_from("<Q", blob, 4)[0] == 5
    assert struct.unpack_from("<6Q", blob, 12) == (3, 2, 0, 0, 0, 0)
    assert struct.unpack_from("<d", blob, 60)[0] == 0.25
    assert struct.unpack_from("<3ffi", blob, 68 + 24)[:3] == tuple(p["x"][1].tolist())
    q, ns, nb, z = F.glmb_parse(blob)
    assert (ns, nb, z) == (17, [3, 2, 0, 0, 0, 0], 0.25)
    assert q.tobytes() == p.tobytes()


@pytest.mark.parametrize("big", [False, True])
def test_oracle_gadget_roundtrip(big):
    rng = np.random.default_rng(4)
    npart = [3, 4, 0, 0, 2, 0]
    masstab = [0.0, 2.5, 0.0, 0.0, 0.0, 0.0]
    pos = rng.normal(size=(9, 3)).astype(np.float32)
    mb = (rng.random(5) + 1).astype(np.float32)
    blob = F.gadget_bytes(npart, masstab, pos, mb, redshift=0.75, big_endian=big)
    hd, p = F.gadget_parse(blob)
    assert hd["npart"] == npart and hd["masstab"] == masstab and hd["redshift"] == 0.75
    assert p["type"].tolist() == [0, 0, 0, 1, 1, 1, 1, 4, 4]
    assert np.array_equal(p["x"], pos)
    assert p["Mass"].tolist() == mb[:3].tolist() + [2.5] * 4 + mb[3:].tolist()


def test_oracle_fits_parse_handmade():
    cards = ["SIMPLE  =                    T", "BITPIX  =                  -32",
             "NAXIS   =                    2", "NAXIS1  =                    3",
             "NAXIS2  =                    2", "KEY     = 'ab''c   '",
             "HIERARCH long key = 1.5 / c", "END"]
    head = "".join(c.ljust(80) for c in cards).ljust(2880).encode()
    data = np.arange(6, dtype=">f4").tobytes().ljust(2880, b"\0")
    c, img = F.fits_parse(head + data)
    assert img.tolist() == [[0, 1, 2], [3, 4, 5]]
    assert F.fits_value(c["KEY"]) == "ab'c" and F.fits_value(c["long key"]) == 1.5


def test_oracle_mass_centre_small():
    p = np.zeros(3, S.PARTICLE_DTYPE)
    p["x"] = [[1, 2, 3], [-1, 0, 1], [0.5, 0.5, -4]]
    p["Mass"] = [1, 3, 4]
    cm, ll, ur = F.mass_centre(p)
    assert np.allclose(cm, [(1 - 3 + 2) / 8, (2 + 0 + 2) / 8, (3 + 3 - 16) / 8], rtol=0, atol=0)
    assert ll.tolist() == [-1, 0, -4] and ur.tolist() == [1, 2, 3]
    ll0 = F.mass_centre(p, box_from_first=False)[1]
    assert ll0.tolist() == [-1, 0, -4]
    q = p.copy()
    q["x"] += 10
    assert F.mass_centre(q, box_from_first=False)[1].tolist() == [0, 0, 0]


# ---------------------------------------------------- csv against oracle
@pytest.mark.parametrize("crlf", [False, True])
def test_csv5_matches_oracle_and_caches_glmb(tmp_path, crlf):
    rng = np.random.default_rng(5)
    p = _particles(rng, 50000)
    fn = _write(tmp_path / "p.csv", _csv_text(p, 5, crlf=crlf, fmt=lambda v: "%.9e" % v))
    m = S.MakeParticleLenses(None, fn, S.SLSIM_FMT_CSV5, 24)
    rc, q, _ = F.read_csv(fn, 5)
    assert rc == 0
    d = m.data
    assert d.tobytes() == q.tobytes()
    assert m.getnp() == [50000, 0, 0, 0, 0, 0] and m.sim_redshift() == -1.0
    cm, ll, ur = F.mass_centre(q)
    assert m.getCenterOfMass() == tuple(cm)
    assert m.getBoundingBox() == (tuple(ll), tuple(ur))
    cache = fn + "_S24.glmb"
    assert open(cache, "rb").read() == F.glmb_bytes(q, 24, [50000, 0, 0, 0, 0, 0], -1.0)
    # a second open reads the cache, not the table
    _write(fn, "garbage\n")
    m2 = S.MakeParticleLenses(None, fn, S.SLSIM_FMT_CSV5, 24)
    assert m2.data.tobytes() == q.tobytes()


def test_csv_recenter(tmp_path):
    rng = np.random.default_rng(6)
    p = _particles(rng, 3000)
    p["x"] += 5
    fn = _write(tmp_path / "p.csv", _csv_text(p, 5))
    m = S.MakeParticleLenses(None, fn, S.SLSIM_FMT_CSV5, 8, recenter=True)
    rc, q, _ = F.read_csv(fn, 5)
    cm, ll, ur = F.mass_centre(q)
    assert m.data.tobytes() == F.recenter(q, cm).tobytes()
    assert m.getCenterOfMass() == (0.0, 0.0, 0.0)
    assert m.getBoundingBox() == (tuple(ll - cm), tuple(ur - cm))


def test_csv_errors_follow_reference(tmp_path):
    def err(fn, fmt=S.SLSIM_FMT_CSV5):
        with pytest.raises(S.SlsimError) as e:
            S.MakeParticleLenses(None, fn, fmt, 8)
        return e.value
    e = err(_write(tmp_path / "a.csv", "1,2,3,4\n"))
    assert e.code == S._abi.SLSIM_ERR_ARG and "enough columns" in str(e)
    e = err(_write(tmp_path / "b.csv", "1,2,3,4,5\n1,2,x,4,5\n"))
    assert e.code == S._abi.SLSIM_ERR_ARG and "data line 2" in str(e)
    e = err(_write(tmp_path / "c.csv", "1,2,3,4,5\n1,2,3,4,1e-60\n"))
    assert e.code == S._abi.SLSIM_ERR_ARG and "out of range" in str(e)
    e = err(_write(tmp_path / "d.csv", "1,2,3,4,5\n1,2,3,4,5\n\n"))      # blank line: stof("")
    assert "data line 3" in str(e)
    e = err(_write(tmp_path / "f.csv", "1,2,3,4,5\n1,2,3\n"))
    assert e.code == S._abi.SLSIM_ERR_IO
    e = err(_write(tmp_path / "g.csv", "1,2,3,4,5,0\n"), S.SLSIM_FMT_CSV6)
    assert e.code == S._abi.SLSIM_ERR_ARG and "not supported" in str(e)
    e = err(_write(tmp_path / "h.csv", "1,2,3,4,5\n"), S.SLSIM_FMT_ASCII)
    assert e.code == S._abi.SLSIM_ERR_ARG
    e = err(str(tmp_path / "missing.csv"))
    assert e.code == S._abi.SLSIM_ERR_ARG
    # no file was cached by any failed read
    assert not [f for f in os.listdir(tmp_path) if f.endswith(".glmb")]


def test_csv_parse_does_not_depend_on_threads(tmp_path):
    rng = np.random.default_rng(7)
    p = _particles(rng, 200000)
    txt = _csv_text(p, 5, fmt=lambda v: "%.8g" % v)
    fn = _write(tmp_path / "big.csv", txt)
    m = S.MakeParticleLenses(None, fn, S.SLSIM_FMT_CSV5, 8)
    rc, q, _ = F.read_csv(fn, 5)
    assert m.data.tobytes() == q.tobytes()


# ----------------------------------------------------------------- glmb
def test_glmb_roundtrip_both_ways(tmp_path):
    rng = np.random.default_rng(8)
    p = _particles(rng, 777, 3)
    fn = str(tmp_path / "a.glmb")
    S.glmb_write(fn, p, 33, [300, 400, 77, 0, 0, 0], 1.5)
    assert open(fn, "rb").read() == F.glmb_bytes(p, 33, [300, 400, 77, 0, 0, 0], 1.5)
    fn2 = _write(tmp_path / "b.glmb", F.glmb_bytes(p[::-1], 9, [1, 2, 3, 4, 5, 762], -1.0), "wb")
    q, ns, nb, z = S.glmb_read(fn2)
    assert (ns, nb, z) == (9, [1, 2, 3, 4, 5, 762], -1.0)
    assert q.tobytes() == p[::-1].tobytes()
    with open(fn2, "r+b") as f:
        f.truncate(68 + 24 * 100)
    with pytest.raises(S.SlsimError) as e:
        S.glmb_read(fn2)
    assert e.value.code == S._abi.SLSIM_ERR_IO


def test_glmb_constructors(tmp_path):
    rng = np.random.default_rng(9)
    p = _particles(rng, 2000, 2)
    p["x"] = np.abs(p["x"]) + 1          # all positive: the glmb-only box starts at 0
    nb = [int((p["type"] == t).sum()) for t in range(6)]
    fn = _write(tmp_path / "c.glmb", F.glmb_bytes(p, 16, nb, 0.3), "wb")
    m = S.MakeParticleLenses(None, fn)                       # MakeParticleLenses(filename, recenter)
    cm, ll, ur = F.mass_centre(p, box_from_first=False)
    assert m.getBoundingBox()[0] == (0.0, 0.0, 0.0)
    assert m.getBoundingBox() == (tuple(ll), tuple(ur))
    assert m.getCenterOfMass() == tuple(cm) and m.nsmooth() == 16 and m.getnp() == nb
    assert m.sim_redshift() == 0.3
    m2 = S.MakeParticleLenses(None, fn, S.SLSIM_FMT_GLMB, 99)   # format glmb: read as is
    cm2, ll2, ur2 = F.mass_centre(p)
    assert m2.getBoundingBox() == (tuple(ll2), tuple(ur2)) and m2.nsmooth() == 16
    m3 = S.MakeParticleLenses(None, fn, recenter=True)
    assert m3.data.tobytes() == F.recenter(p, cm).tobytes()
    assert m3.getCenterOfMass() == (0.0, 0.0, 0.0)
    assert m3.getBoundingBox() == (tuple(ll - cm), tuple(ur - cm))
    with pytest.raises(S.SlsimError):
        S.MakeParticleLenses(None, str(tmp_path / "none.glmb"), S.SLSIM_FMT_GLMB, 8)


def test_cuts_units_densest(tmp_path):
    rng = np.random.default_rng(10)
    p = _particles(rng, 5000, 3)
    nb = [int((p["type"] == t).sum()) for t in range(6)]
    fn = _write(tmp_path / "d.glmb", F.glmb_bytes(p, 8, nb, 0.0), "wb")

    def per_type(a):
        return [np.sort(a[a["type"] == t].view(np.void)) for t in range(6)]

    m = S.MakeParticleLenses(None, fn)
    keep = F.keep_radial(p, (0.1, -0.2, 0.3), 1.2)
    m.radialCut((0.1, -0.2, 0.3), 1.2)
    d = m.data
    assert d.size == keep.sum()
    assert np.all(np.diff(d["type"]) >= 0)                  # sorted by type again
    assert all(np.array_equal(a, b) for a, b in zip(per_type(d), per_type(p[keep])))
    assert m.getnp() == [int((p[keep]["type"] == t).sum()) for t in range(6)]
    m = S.MakeParticleLenses(None, fn)
    keep = F.keep_cylindrical(p, (0.3, 0.1), 0.9)
    m.cylindricalCut((0.3, 0.1), 0.9)
    assert all(np.array_equal(a, b) for a, b in zip(per_type(m.data), per_type(p[keep])))
    m = S.MakeParticleLenses(None, fn)
    assert m.densest_particle() == F.densest(p)
    # the unit conversion of CreateHalos (the halos themselves need a device)
    assert S.lib.slsim_mpl_physical_units(m.h, 0.7) == 0
    assert m.data.tobytes() == F.physical_units(p, 0.7).tobytes()
    m.Recenter((1.0, 2.0, 3.0))
    assert m.data.tobytes() == F.recenter(F.physical_units(p, 0.7), (1, 2, 3)).tobytes()


# --------------------------------------------------------------- Gadget-2
@pytest.mark.parametrize("big", [False, True])
def test_gadget_raw_matches_oracle(tmp_path, big):
    rng = np.random.default_rng(11)
    npart = [700, 500, 0, 0, 30, 0]
    masstab = [0.0, 0.031, 0.0, 0.0, 0.0, 0.0]
    pos = (rng.random((1230, 3)) * 1e5).astype(np.float32)
    mb = (rng.random(730) * 0.01).astype(np.float32)
    vel = np.zeros((1230, 3), np.float32).tobytes()
    blob = F.gadget_bytes(npart, masstab, pos, mb, redshift=0.4, big_endian=big,
                          extra_blocks=[("VEL ", vel)], order=("HEAD", "MASS", "POS "))
    fn = _write(tmp_path / "snap", blob, "wb")
    hd, p = S.gadget2_read(fn)
    hd0, p0 = F.gadget_parse(blob)
    assert p.tobytes() == p0.tobytes()
    for k in ("npart", "masstab", "redshift", "time", "omega", "lambda", "hubble"):
        assert hd[k] == hd0[k], k


def test_gadget_errors(tmp_path):
    rng = np.random.default_rng(12)
    pos = rng.random((10, 3)).astype(np.float32)
    # all masses in the table and no MASS block: the reference hits end of file and exits
    blob = F.gadget_bytes([10, 0, 0, 0, 0, 0], [1.0, 0, 0, 0, 0, 0], pos, [],
                          order=("HEAD", "POS "))
    with pytest.raises(S.SlsimError) as e:
        S.gadget2_read(_write(tmp_path / "nomass", blob, "wb"))
    assert e.value.code == S._abi.SLSIM_ERR_IO and "MASS" in str(e.value)
    # SnapFormat 1 (no label records)
    with pytest.raises(S.SlsimError) as e:
        S.gadget2_read(_write(tmp_path / "fmt1", struct.pack("<i", 256) + b"\0" * 260, "wb"))
    assert "incorrect format" in str(e.value)
    # a snapshot split over files
    _write(tmp_path / "split.0", blob, "wb")
    with pytest.raises(S.SlsimError) as e:
        S.gadget2_read(str(tmp_path / "split"))
    assert "split" in str(e.value)


# --------------------------------------------------------------- ASCII halo
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("recenter,theta", [(False, (0.0, 0.0)), (True, (0.3, -0.7))])
def test_ascii_halo_with_sizes_file(tmp_path, dtype, recenter, theta):
    rng = np.random.default_rng(13)
    n = 4000
    xyz = rng.normal(size=(n, 3)) * 2 + 0.5
    m = rng.random(n) * 1e10 + 1e9
    lines = ["# nparticles %d" % n, "# a comment line"]
    lines += ["%.10g %.10g %.10g %.7g" % (a, b, c, w) for (a, b, c), w in zip(xyz, m)]
    fn = _write(tmp_path / "halo.txt", "# nparticles %d\n" % n + "\n".join(lines[1:]) + "\n")
    sizes = (rng.random(n) * 0.05 + 0.001).astype(np.float32)
    _write(tmp_path / "halo.txt.32sizes", F.sizes_text(sizes, 32))
    X, M, H, info = S.LensHaloParticles.load_ascii(None, fn, 0.5, 32, theta, recenter, True,
                                                   MinPSize=0.01, rescale_mass=2.0, dtype=dtype)
    xyz0, m0 = F.read_ascii(fn, True, dtype)
    size0 = np.array([F._strtof("%.6g" % s) for s in sizes], np.float32)
    size0 = np.where(size0 < 0.01, np.float32(0.01), size0)
    xyz1, m1, info1 = F.ascii_halo(xyz0, m0, size0, 0.5, True, recenter, 2.0, theta)
    assert X.dtype == dtype and np.array_equal(X, xyz1)
    assert np.array_equal(M, m1) and M[0] == m0[0]          # the first mass is not rescaled
    assert np.array_equal(H, size0)
    for k in ("mass", "rsize"):
        assert info[k] == info1[k], k
    assert np.allclose(info["mcenter"], info1["mcenter"], rtol=1e-15, atol=1e-300)
    assert np.allclose(info["densest_point"], info1["densest_point"], rtol=1e-15, atol=1e-300)


def test_ascii_header_and_size_file_errors(tmp_path):
    fn = _write(tmp_path / "h.txt", "1 2 3 4\n")
    with pytest.raises(S.SlsimError) as e:
        S.LensHaloParticles.load_ascii(None, fn, 0.5, 8)
    assert "nparticles" in str(e.value)
    fn = _write(tmp_path / "g.txt", "# nparticles 3\n1 2 3 4\n5 6 7 8\n")
    with pytest.raises(S.SlsimError) as e:
        S.LensHaloParticles.load_ascii(None, fn, 0.5, 8)
    assert "does not match" in str(e.value)
    fn = _write(tmp_path / "k.txt", "# nparticles 2\n1 2 3 4\n5 6 7 8\n")
    _write(tmp_path / "k.txt.8sizes", "# nparticles 3\n# nsmooth 8\n1\n2\n3\n")
    with pytest.raises(S.SlsimError) as e:
        S.LensHaloParticles.load_ascii(None, fn, 0.5, 8)
    assert "does not match" in str(e.value)


# ------------------------------------------------------------------- FITS
@pytest.mark.parametrize("dtype,flip,ctype", [(np.float64, False, False), (np.float32, True, True)])
def test_fits_pixelmap(tmp_path, dtype, flip, ctype):
    rng = np.random.default_rng(14)
    img = rng.normal(size=(37, 53)).astype(dtype)
    res, center = 2.5e-6, (1.0e-4, -3.0e-4)
    fn = str(tmp_path / "m.fits")
    S.write_fits_pixelmap("!" + fn, img, res, center, units=4, flipX=flip, ctype=ctype)
    blob = open(fn, "rb").read()
    assert len(blob) % 2880 == 0
    cards, got = F.fits_parse(blob)
    want = img[:, ::-1] if flip else img
    assert got.dtype.itemsize == np.dtype(dtype).itemsize and np.array_equal(got, want)
    v = {k: F.fits_value(c) for k, c in cards.items()}
    assert v["SIMPLE"] is True and v["BITPIX"] == (-64 if dtype == np.float64 else -32)
    assert v["NAXIS1"] == 53 and v["NAXIS2"] == 37 and v["NX"] == 53 and v["NY"] == 37
    assert v["CRPIX1"] == 27.0 and v["CRPIX2"] == 19.0 and v["WCSAXES"] == 2
    # real values carry 15 significant digits, as cfitsio writes TDOUBLE keys

    def g15(x):
        return float("%.15G" % x)
    assert v["CD1_1"] == g15(-180 * res / np.pi) and v["CD2_2"] == g15(180 * res / np.pi)
    assert v["range x"] == g15((center[0] + 53 * res / 2) - (center[0] - 53 * res / 2))
    assert v["CENTER_X"] == g15(center[0]) and v["CENTER_Y"] == g15(center[1])
    assert v["CRVAL1"] == g15(center[0] / 0.01745329251994)
    assert v["UNITS"] == "mass" and v["RA_global"] == 0.0
    assert ("CTYPE1" in v) == ctype
    if ctype:
        assert v["CTYPE1"] == "RA---TAN" and v["RADESYS"] == "ICRS"
    with pytest.raises(S.SlsimError):
        S.write_fits_pixelmap(fn, img, res, center, units=9)


def test_particle_setup_oracle():
    """LensHaloParticles::set_up's restatement: the centre of mass and
    densest point of a hand case, the recentring quirk (only the first
    particle moves), and the rotation: its x and y rows are those of the
    rotation R_x(tx) R_y(-ty) (scipy), its z row keeps the reference's cos(tx)
    where a rotation has cos(tx) sin(ty)."""
    from scipy.spatial.transform import Rotation
    xyz = np.array([[1.0, 2.0, 3.0], [-1.0, 0.5, 0.0], [4.0, -2.0, 1.0]])
    m = np.array([1.0, 2.0, 1.0], np.float32)
    h = np.array([0.3, 0.1, 0.1], np.float32)
    q, mc, dp, mt = F.particle_setup(xyz, m, h)
    assert mt == 4.0
    assert np.allclose(mc, [(1 - 2 + 4) / 4, (2 + 1 - 2) / 4, (3 + 0 + 1) / 4], rtol=0, atol=1e-15)
    assert np.array_equal(dp, xyz[1]) and np.array_equal(q, xyz)
    q, mc2, dp2, _ = F.particle_setup(xyz, m, h, recenter=True)
    assert np.allclose(q[0], xyz[0] - mc) and np.array_equal(q[1:], xyz[1:])
    assert np.array_equal(mc2, np.zeros(3)) and np.allclose(dp2, xyz[1] - mc)
    tx, ty = 0.3, -0.2
    rng = np.random.default_rng(5)
    p = rng.normal(size=(50, 3))
    q, _, _, _ = F.particle_setup(p, np.ones(50, np.float32), np.ones(50, np.float32),
                                  theta_rotate=(tx, ty))
    R = (Rotation.from_euler("x", tx) * Rotation.from_euler("y", -ty)).as_matrix()
    assert np.allclose(q[:, :2], p @ R[:2].T, rtol=0, atol=1e-14)
    e = np.eye(3)
    qz, _, _, _ = F.particle_setup(e, np.ones(3, np.float32), np.ones(3, np.float32),
                                   theta_rotate=(tx, ty))
    assert np.allclose(qz[:, 2], [np.cos(tx), np.sin(tx), np.cos(ty) * np.cos(tx)])
    assert not np.isclose(qz[0, 2], R[2, 0])


def test_pixelmap_uniform_oracle():
    """writePixelMapUniform's restatement: on the grid's own geometry (N x N
    pixels of range / (N - 1) around the grid centre) pixel i holds point i;
    on a map of twice the pixel size each pixel keeps the last of its four
    points; points outside the map are dropped."""
    N, rg, c = 8, 3.0, (0.25, -0.5)
    i = np.arange(N * N)
    x = np.stack([c[0] + rg * ((i % N) / (N - 1) - 0.5), c[1] + rg * ((i // N) / (N - 1) - 0.5)], 1)
    vals = np.arange(N * N, dtype=np.float64) + 1
    assert np.array_equal(F.pixelmap_uniform(x, vals, c, N, N, rg / (N - 1)).reshape(-1), vals)
    res = 2 * rg / (N - 1)
    got = F.pixelmap_uniform(x, vals, c, N // 2, N // 2, res)
    for py in range(N // 2):
        for px in range(N // 2):
            pts = [(2 * py + b) * N + 2 * px + a for b in (0, 1) for a in (0, 1)]
            assert got[py, px] == vals[max(pts)]
    small = F.pixelmap_uniform(x, vals, c, 2, 2, rg / (N - 1))
    inside = [k for k in range(N * N) if abs(x[k, 0] - c[0]) < rg / (N - 1)
              and abs(x[k, 1] - c[1]) < rg / (N - 1)]
    assert len(inside) == 4 and sorted(small.reshape(-1)) == sorted(vals[inside])
