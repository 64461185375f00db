"""CPU restatement of GLAMER's particle-file readers and FITS output.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's cpu_baseline leg
as the checker.  The product never uses it.

Each function cites the reference code it follows (paths relative to the
reference root).  Arithmetic keeps the reference's types: float32 products of
float32 operands, float32 = float32 op float64 rounded once, sequential
float64 sums (np.cumsum, not the pairwise np.sum).

Parity unpinned: the exact header-card text cfitsio would write (cfitsio is
absent, so no file written by it exists to compare with); the keys, values
and pixel data are pinned.
"""
import ctypes
import struct

import numpy as np

from . import _lib

PARTICLE_DTYPE = np.dtype([("x", "<f4", (3,)), ("Mass", "<f4"), ("Size", "<f4"),
                           ("type", "<i4")])

_lib.orc_csv_read.restype = ctypes.c_int
_lib.orc_csv_read.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                              ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t),
                              ctypes.POINTER(ctypes.c_size_t)]

_libc = ctypes.CDLL(None, use_errno=True)
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)]


def read_csv(path, columns):
    """MakeParticleLenses::readCSV (particle_lens.cpp:211-342), 3-5 columns,
    by the C++ restatement in formats_oracle.cpp.  Returns (status,
    particles, bad_line): status 0 ok, 1 too few columns, 2 stof invalid,
    3 stof out of range, 4 a later line with too few fields."""
    n = ctypes.c_size_t()
    bad = ctypes.c_size_t()
    fn = path.encode() if isinstance(path, str) else path
    rc = _lib.orc_csv_read(fn, int(columns), None, 0, ctypes.byref(n), ctypes.byref(bad))
    if rc:
        return rc, None, None
    buf = np.zeros((n.value, 5), np.float32)
    rc = _lib.orc_csv_read(fn, int(columns), buf.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                           n.value, ctypes.byref(n), ctypes.byref(bad))
    if rc:
        return rc, None, bad.value
    p = np.zeros(n.value, PARTICLE_DTYPE)
    p["x"] = buf[:, :3]
    p["Mass"] = buf[:, 3]
    p["Size"] = buf[:, 4]
    return 0, p, None


def glmb_bytes(particles, nsmooth, numbytype, redshift):
    """writeSizesB (particle_halo.h:1037-1054): int Nsmooth, size_t ntot,
    size_t numbytype[6], double redshift, then the ParticleType<float>
    records, little endian."""
    p = np.ascontiguousarray(particles, PARTICLE_DTYPE)
    return (struct.pack("<iQ6Qd", int(nsmooth), p.size, *[int(v) for v in numbytype],
                        float(redshift)) + p.tobytes())


def glmb_parse(blob):
    """readSizesB (particle_halo.h:1056-1073): (particles, nsmooth,
    numbytype, redshift)."""
    ns, n, *rest = struct.unpack_from("<iQ6Qd", blob, 0)
    nb, z = rest[:6], rest[6]
    p = np.frombuffer(blob, PARTICLE_DTYPE, count=n, offset=68).copy()
    return p, ns, list(nb), z


def mass_centre(p, box_from_first=True):
    """Centre of mass and bounding box of MakeParticleLenses' constructors
    (particle_lens.cpp:76-93, 120-131): cm[i] += p[i]*Mass in float32 then
    summed in float64 in order; the glmb-only constructor starts its box at
    the origin."""
    x = p["x"]
    prod = (x * p["Mass"][:, None]).astype(np.float32)
    cm = np.array([np.cumsum(prod[:, i].astype(np.float64))[-1] for i in range(3)])
    m = np.cumsum(p["Mass"].astype(np.float64))[-1]
    cm = cm / m
    if box_from_first:
        ll = x.min(axis=0).astype(np.float64)
        ur = x.max(axis=0).astype(np.float64)
    else:
        ll = np.minimum(x.min(axis=0).astype(np.float64), 0.0)
        ur = np.maximum(x.max(axis=0).astype(np.float64), 0.0)
    return cm, ll, ur


def recenter(p, c):
    """p[i] -= c[i] for float32 positions and a float64 centre."""
    q = p.copy()
    q["x"] = (q["x"].astype(np.float64) - np.asarray(c, np.float64)).astype(np.float32)
    return q


def physical_units(p, hubble):
    """The unit conversion of CreateHalos (particle_lens.cpp:153-170)."""
    u = 1.0 / np.float64(np.float32(hubble))
    q = p.copy()
    q["x"] = (q["x"].astype(np.float64) * u).astype(np.float32)
    q["Size"] = (q["Size"].astype(np.float64) * u).astype(np.float32)
    q["Mass"] = (q["Mass"].astype(np.float64) * u).astype(np.float32)
    return q


def keep_radial(p, center, radius):
    """Particles radialCut keeps (particle_lens.cpp:447-478): r2 in float64."""
    d = p["x"].astype(np.float64) - np.asarray(center, np.float64)
    r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    return r2 <= radius * radius


def keep_cylindrical(p, center, radius):
    """Particles cylindricalCut keeps (particle_lens.cpp:497-528)."""
    d = p["x"][:, :2].astype(np.float64) - np.asarray(center, np.float64)
    r2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]
    return r2 <= radius * radius


def densest(p):
    """densest_particle (particle_lens.cpp:480-495): the first maximum of
    Mass/Size/Size/Size in float32."""
    s = p["Size"]
    d = ((p["Mass"] / s).astype(np.float32) / s).astype(np.float32) / s
    return tuple(p["x"][int(np.argmax(d.astype(np.float32)))].astype(np.float64))


def gadget_bytes(npart, masstab, pos, mass_block, redshift=0.5, time=None, omega=0.3,
                 lam=0.7, hubble=0.7, big_endian=False, extra_blocks=(), order=("HEAD", "POS ",
                                                                                 "MASS")):
    """A SnapFormat 2 Gadget-2 snapshot (Gadget-2 user guide; the reader is
    include/gadget.hh): each block is an 8-byte label record (label, size of
    the data record + 8) followed by the Fortran data record.  The header is
    256 bytes: npart[6], massarr[6], time, redshift, flag_sfr, flag_feedback,
    npartTotal[6], flag_cooling, num_files, BoxSize, Omega0, OmegaLambda,
    HubbleParam, padding."""
    e = ">" if big_endian else "<"
    head = struct.pack(e + "6i6dddii6iiidddd", *[int(v) for v in npart],
                       *[float(v) for v in masstab], float(time if time is not None else
                                                           1 / (1 + redshift)), float(redshift),
                       0, 0, *[int(v) for v in npart], 0, 1, 100.0, omega, lam, hubble)
    head += b"\0" * (256 - len(head))
    blocks = {"HEAD": head, "POS ": np.asarray(pos, e + "f4").tobytes(),
              "MASS": np.asarray(mass_block, e + "f4").tobytes()}
    for lab, data in extra_blocks:
        blocks[lab] = data
    out = b""
    for lab in list(order) + [lab for lab, _ in extra_blocks]:
        if lab not in blocks:
            continue
        data = blocks[lab]
        out += struct.pack(e + "i", 8) + lab.encode() + struct.pack(e + "i", len(data) + 8)
        out += struct.pack(e + "i", 8)
        out += struct.pack(e + "i", len(data)) + data + struct.pack(e + "i", len(data))
    return out


def gadget_parse(blob):
    """GadgetFile's header, POS and MASS (gadget.hh:296-330, 374-418,
    567-610): (header dict, particles in file units with types in block
    order and masses from masstab or the MASS block)."""
    e = "<" if struct.unpack_from("<i", blob, 0)[0] == 8 else ">"
    blocks = {}
    pos = 0
    while pos + 16 <= len(blob):
        assert struct.unpack_from(e + "i", blob, pos)[0] == 8
        lab = blob[pos + 4:pos + 8].decode()
        size = struct.unpack_from(e + "i", blob, pos + 8)[0]
        blocks[lab] = blob[pos + 20:pos + 16 + size - 4]
        pos += 16 + size
    h = blocks["HEAD"]
    npart = struct.unpack_from(e + "6i", h, 0)
    masstab = struct.unpack_from(e + "6d", h, 24)
    time, z = struct.unpack_from(e + "2d", h, 72)
    om, la, hub = struct.unpack_from(e + "3d", h, 136)
    n = sum(npart)
    p = np.zeros(n, PARTICLE_DTYPE)
    p["type"] = np.repeat(np.arange(6), npart)
    p["x"] = np.frombuffer(blocks["POS "], e + "f4").reshape(n, 3)
    mb = np.frombuffer(blocks["MASS"], e + "f4")
    tab = np.asarray(masstab)[p["type"]]
    ind = tab == 0
    p["Mass"][~ind] = tab[~ind].astype(np.float32)
    p["Mass"][ind] = mb
    hd = {"npart": list(npart), "masstab": list(masstab), "time": time, "redshift": z,
          "omega": om, "lambda": la, "hubble": hub}
    return hd, p


def gadget_units(p, z):
    """readGadget2's conversion (particle_lens.cpp:355-362): positions times
    1e-3/(1+z), masses times 1e10, float32 = float32 * double."""
    a = 1.0e-3 / (1 + z)
    q = p.copy()
    q["x"] = (q["x"].astype(np.float64) * a).astype(np.float32)
    q["Mass"] = (q["Mass"].astype(np.float64) * 1.0e10).astype(np.float32)
    return q


def _strtof(tok):
    end = ctypes.c_char_p()
    ctypes.set_errno(0)
    v = _libc.strtof(tok.encode(), ctypes.byref(end))
    return v


def read_ascii(path, multimass, dtype=np.float64):
    """readPositionFileASCII (particle_halo.h:453-555) for files whose rows
    are plain decimal numbers: the "# nparticles N" header, then the first N
    lines not starting with '#', three (four with multimass) columns; mass 0
    without multimass (the reference never stores the header mass)."""
    text = open(path).read()
    lines = text.split("\n")
    if text.endswith("\n"):
        lines.pop()
    i = 0
    npoints = None
    count = 0
    label = ""
    while i < len(lines):
        toks = lines[i].split()
        i += 1
        if toks:
            label = toks[0]
        if label == "#":
            label = toks[1] if len(toks) > 1 else label
            if label == "nparticles":
                npoints = int(toks[2])
                count += 1
            if not multimass and label == "mass":
                count += 1
        else:
            break
        if multimass and count == 1:
            break
        if not multimass and count == 2:
            break
    rows = [ln for ln in lines[i:] if not ln.startswith("#")][:npoints]
    ncol = 4 if multimass else 3
    xyz = np.zeros((npoints, 3), dtype)
    m = np.zeros(npoints, np.float32)
    for r, ln in enumerate(rows):
        t = ln.split()[:ncol]
        if dtype == np.float32:
            xyz[r] = [_strtof(v) for v in t[:3]]
        else:
            xyz[r] = [float(v) for v in t[:3]]
        if multimass:
            m[r] = _strtof(t[3])
    return xyz, m


def sizes_text(sizes, nsmooth):
    """writeSizes (particle_halo.h:734-762): the header and one size per
    line as `ostream << float` writes it (%g, six significant digits)."""
    out = "# nparticles %d\n# nsmooth %d\n" % (len(sizes), nsmooth)
    return out + "".join("%.6g\n" % float(s) for s in np.asarray(sizes, np.float32))


def ascii_halo(xyz, mass, size, redshift, multimass, recenter, rescale_mass=1.0,
               theta_rotate=(0.0, 0.0)):
    """The rest of LensHaloParticles' file constructor (particle_halo.h:
    275-338) on positions of type xyz.dtype: rescale_mass for all but the
    first particle, comoving -> physical, centre of mass, recentring,
    rotation.  Returns (xyz, mass, info)."""
    T = xyz.dtype.type
    xyz = xyz.copy()
    mass = mass.copy()
    mass[1:] = (mass[1:].astype(np.float64) * rescale_mass).astype(np.float32)
    smallest = 0
    for i in range(1, len(size)):
        if size[i] < size[smallest]:
            smallest = i
    dp = xyz[smallest].astype(np.float64)
    sf = 1 / (1 + redshift)
    xyz = (xyz.astype(np.float64) * sf).astype(T)
    mi = mass if multimass else np.full_like(mass, mass[0])
    prod = (xyz * mi.astype(T)[:, None]).astype(T).astype(np.float64)
    mc = np.array([np.cumsum(prod[:, k])[-1] for k in range(3)])
    mtot = np.cumsum(mi.astype(np.float64))[-1]
    mc = mc / mtot
    rsize = 1.0e3
    if recenter:
        xyz = (xyz.astype(np.float64) - mc).astype(T)
        r2 = (xyz[:, 0] * xyz[:, 0] + xyz[:, 1] * xyz[:, 1]) + xyz[:, 2] * xyz[:, 2]
        rsize = float(np.sqrt(np.float64(max(0.0, float(r2.max())))))
        dp = dp - mc
        mc = np.zeros(3)
    tx, ty = theta_rotate
    if not (tx == 0.0 and ty == 0.0):
        cx, sx, cy, sy = np.cos(tx), np.sin(tx), np.cos(ty), np.sin(ty)
        c = np.array([[cy, 0, -sy], [-sy * sx, cx, -cy * sx], [cx, sx, cy * cx]])

        def rot(v):
            v = np.asarray(v, np.float64)
            t = np.zeros(v.shape)
            for j in range(3):
                for r in range(3):
                    t[..., r] += c[r, j] * v[..., j]
            return t
        xyz = rot(xyz).astype(T)
        mc = rot(mc)
        dp = rot(dp)
    return xyz, mass, {"mass": mtot, "mcenter": tuple(mc), "densest_point": tuple(dp),
                       "rsize": rsize}


def fits_parse(blob):
    """A FITS primary image (FITS standard 4.0): header cards of 80
    characters in 2880-byte records up to END, then big-endian data.
    Returns (cards {key: value-text}, image (NAXIS2, NAXIS1))."""
    assert len(blob) % 2880 == 0
    cards = {}
    off = 0
    while True:
        c = blob[off:off + 80].decode("ascii")
        off += 80
        if c.startswith("END") and c[3:].strip() == "":
            break
        if c.startswith("HIERARCH "):
            k, v = c[9:].split("=", 1)
            cards[k.strip()] = v
        elif c[8:10] == "= ":
            cards[c[:8].strip()] = c[10:]
    data_off = ((off + 2879) // 2880) * 2880
    bitpix = int(cards["BITPIX"].split("/")[0])
    nx = int(cards["NAXIS1"].split("/")[0])
    ny = int(cards["NAXIS2"].split("/")[0])
    dt = ">f8" if bitpix == -64 else ">f4"
    img = np.frombuffer(blob, dt, count=nx * ny, offset=data_off).reshape(ny, nx)
    return cards, img


def fits_value(text):
    """The value of a card: a number, T/F or a quoted string."""
    t = text.strip()
    if t.startswith("'"):
        end = t.index("'", 1)
        while end + 1 < len(t) and t[end + 1] == "'":
            end = t.index("'", end + 2)
        return t[1:end].replace("''", "'").rstrip()
    v = t.split("/")[0].strip()
    if v in ("T", "F"):
        return v == "T"
    try:
        return int(v)
    except ValueError:
        return float(v)


def _rotation(tx, ty):
    """The matrix of LensHaloParticles::rotate_particles
    (include/particle_halo.h:634-648): rows (cy, 0, -sy), (-sy sx, cx, -cy sx),
    (cx, sx, cy cx)."""
    cx, sx, cy, sy = np.cos(tx), np.sin(tx), np.cos(ty), np.sin(ty)
    return np.array([[cy, 0, -sy], [-sy * sx, cx, -cy * sx], [cx, sx, cy * cx]])


def _rotate(c, v):
    """t[r] = sum over j of c[r][j] v[j], j ascending (particle_halo.h:653-661)."""
    v = np.asarray(v, np.float64)
    t = np.zeros(v.shape)
    for j in range(3):
        for r in range(3):
            t[..., r] += c[r, j] * v[..., j]
    return t


def particle_setup(xyz, mass, size, recenter=False, theta_rotate=(0.0, 0.0)):
    """LensHaloParticles::set_up (include/particle_halo.h:341-417) on the
    particles of the vector and pointer constructors (:58-93): the densest
    point is the first particle of smallest Size; the centre of mass sums
    x * mass() (position type times float32, in double) and the mass in
    double, particle order; recentring subtracts the centre inside the
    particle loop after which the centre is zeroed, so only the first particle
    moves (reference); then the particles, the centre and the densest point
    are rotated (:634-680).  Returns (xyz, mcenter, densest_point, mass)."""
    T = xyz.dtype.type
    xyz = xyz.copy()
    smallest = 0
    for i in range(1, len(size)):
        if size[i] < size[smallest]:
            smallest = i
    dp = xyz[smallest].astype(np.float64)
    prod = xyz.astype(np.float64) * mass.astype(np.float64)[:, None]
    if T is np.float32:
        prod = (xyz * mass[:, None]).astype(np.float64)
    mc = np.array([np.cumsum(prod[:, k])[-1] for k in range(3)])
    mtot = np.cumsum(mass.astype(np.float64))[-1]
    mc = mc / mtot
    if recenter and len(xyz):
        xyz[0] = (xyz[0].astype(np.float64) - mc).astype(T)
        dp = dp - mc
        mc = np.zeros(3)
    tx, ty = theta_rotate
    if not (tx == 0.0 and ty == 0.0):
        c = _rotation(tx, ty)
        xyz = _rotate(c, xyz).astype(T)
        mc = _rotate(c, mc)
        dp = _rotate(c, dp)
    return xyz, mc, dp, mtot


def pixelmap_uniform(x, values, center, nx, ny, resolution):
    """GridMap::writePixelMapUniform (include/gridmap.h:506-524, 605-681) into
    a PixelMap(center, nx, ny, resolution) (ImageProcessing/pixelmap.cpp:
    111-130): the map is cleaned, then each grid point's value goes to the
    pixel PixelMap::find_index gives for its image position
    (pixelmap.cpp:2227-2242: (long)((x - p1) / resolution) per axis,
    p1 = center - (N resolution) / 2, -1 outside).  Where points share a
    pixel the reference's threads keep any one; here the last in index order.
    Returns the map (ny, nx)."""
    out = np.zeros(nx * ny)
    p1x = center[0] - (nx * resolution) / 2.
    p1y = center[1] - (ny * resolution) / 2.
    for i in range(x.shape[0]):
        ix = int((x[i, 0] - p1x) / resolution)
        iy = int((x[i, 1] - p1y) / resolution)
        if 0 <= ix < nx and 0 <= iy < ny:
            out[ix + nx * iy] = values[i]
    return out.reshape(ny, nx)
