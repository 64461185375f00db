"""CPU oracle of GLAMER's particle ray-shooting path (ctypes view of
oracle/liboracle.so, built from oracle/glamer_oracle.cpp by oracle/Makefile).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never uses it.

Parity unpinned (nothing the reference ships fixes them): the tie rule of the
nearest-plane insertion for a halo exactly halfway between two planes
(Utilities::closest is not in the reference tree; the first plane at the
least |dDl| is taken), and the cosmological distances' integrator (the
closed-form flat-LCDM integral stands in for CosmoLib's, which is absent).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)
    return LIB_PATH


if not os.path.exists(LIB_PATH):
    build()

_lib = ctypes.CDLL(LIB_PATH)
P = ctypes.POINTER
c_d, c_f, c_sz, c_int, c_vp = ctypes.c_double, ctypes.c_float, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p


class Counters(ctypes.Structure):
    _fields_ = [("visits", ctypes.c_uint64), ("far", ctypes.c_uint64), ("leafp", ctypes.c_uint64),
                ("bigtest", ctypes.c_uint64), ("bspline", ctypes.c_uint64)]


_sigs = {
    "orc_coordist": (c_d, [c_d, c_d, c_d, c_d, c_d]),
    "orc_tree_build": (c_vp, [P(c_d), P(c_d), P(c_f), P(c_f), c_sz, c_d, c_int, c_d, P(c_int)]),
    "orc_tree_build_range": (c_vp, [P(c_d), P(c_d), P(c_f), P(c_f), c_sz, c_d, c_int, c_d, c_d,
                                    P(c_int)]),
    "orc_tree_free": (None, [c_vp]),
    "orc_tree_nnodes": (c_sz, [c_vp]),
    "orc_tree_export": (None, [c_vp, P(c_d), P(ctypes.c_int64)]),
    "orc_tree_force": (None, [c_vp, P(c_d), c_sz, P(c_d), c_vp, c_int]),
    "orc_direct_plane": (None, [P(c_d), P(c_d), P(c_f), P(c_f), c_sz, c_d, P(c_d), c_sz, P(c_d),
                                c_int]),
    "orc_smoothing_f64": (c_int, [P(c_d), P(c_d), P(c_d), c_sz, c_int, c_int, P(c_f), c_int,
                                  P(ctypes.c_int64), c_sz]),
    "orc_smoothing_f32": (c_int, [P(c_f), P(c_f), P(c_f), c_sz, c_int, c_int, P(c_f), c_int,
                                  P(ctypes.c_int64), c_sz]),
    "orc_bspline": (None, [P(c_d), c_d, c_d, c_d, P(c_d)]),
    "orc_lens_create": (c_vp, [c_d, c_d, c_d, c_d]),
    "orc_lens_free": (None, [c_vp]),
    "orc_lens_add_halo": (None, [c_vp, c_vp, c_f]),
    "orc_halotree_build": (c_vp, [P(c_d), P(c_d), P(c_f), P(c_f), P(c_f), c_sz, c_f, c_f,
                                  P(c_int)]),
    "orc_halotree_build_field": (c_vp, [P(c_d), P(c_d), P(c_f), P(c_f), P(c_f), c_sz, c_d, c_d,
                                         c_d, c_int, P(c_int)]),
    "orc_lens_add_field_plane": (None, [c_vp, c_vp, c_d]),
    "orc_halotree_build_mixed": (c_vp, [P(c_d), P(c_d), P(c_f), P(c_f), P(c_int), P(c_d), c_sz,
                                        c_d, c_d, c_d, c_int, P(c_int)]),
    "orc_powerlaw_force": (None, [c_f, c_f, c_d, P(c_d), c_sz, P(c_d)]),
    "orc_halotree_free": (None, [c_vp]),
    "orc_halotree_force": (None, [c_vp, P(c_d), c_sz, c_int, P(c_d)]),
    "orc_nfw_force": (None, [c_f, c_f, c_f, P(c_d), c_sz, P(c_d)]),
    "orc_nfw_function": (c_d, [c_int, c_d]),
    "orc_nfw_interpolate": (c_d, [c_int, c_d]),
    "orc_lens_add_halotree": (None, [c_vp, c_vp, c_f]),
    "orc_lens_add_nearest": (None, [c_vp, c_vp, c_f, c_int]),
    "orc_find_images": (c_sz, [c_sz, P(c_d), P(c_d), P(c_d), P(c_d), c_sz, P(ctypes.c_int32),
                               P(c_d), P(ctypes.c_int64), c_sz]),
    "orc_lens_nplanes": (c_int, [c_vp]),
    "orc_lens_tables": (None, [c_vp, P(c_d), P(c_d), P(c_d), P(c_d)]),
    "orc_lens_shoot": (None, [c_vp, c_sz, P(c_d), P(c_d), P(c_d), P(c_d), c_int, c_vp]),
    "orc_gridmap_positions": (None, [c_sz, c_d, c_d, c_d, P(c_d)]),
    "orc_find_images2": (c_sz, [c_sz, c_sz, P(c_d), P(c_d), P(c_d), P(c_d), c_sz,
                                P(ctypes.c_int32), P(c_d), P(ctypes.c_int64), c_sz]),
    "orc_find_crit2": (c_sz, [c_int, c_sz, c_sz, P(c_d), P(c_d), P(c_d), P(ctypes.c_int32), c_d,
                              c_d, c_sz, c_sz, P(ctypes.c_int64), P(c_d), P(ctypes.c_int32),
                              P(ctypes.c_int32), P(c_sz), P(c_sz), P(c_int)]),
    "orc_find_crit": (c_sz, [c_int, c_sz, P(c_d), P(c_d), P(c_d), P(ctypes.c_int32), c_d, c_d,
                             c_sz, c_sz, P(ctypes.c_int64), P(c_d), P(ctypes.c_int32),
                             P(ctypes.c_int32), P(c_sz), P(c_sz), P(c_int)]),
    "orc_find_boundaries": (c_sz, [P(ctypes.c_uint8), c_sz, c_sz, c_sz, c_sz, P(ctypes.c_int64),
                                   P(c_d), P(ctypes.c_int32), P(c_sz)]),
    "orc_find_islands": (c_sz, [P(ctypes.c_uint8), c_sz, c_sz, c_sz, c_sz, P(ctypes.c_int64),
                                P(ctypes.c_int64), P(c_sz)]),
    "orc_convex_hull": (c_sz, [P(c_d), c_sz, P(ctypes.c_int64)]),
    "orc_in_curve": (None, [P(c_d), c_sz, P(c_d), c_sz, P(ctypes.c_int32)]),
    "orc_tightest_hull": (c_sz, [P(c_d), c_sz, c_sz, P(c_d)]),
    "orc_windings": (c_int, [P(c_d), P(c_d), c_sz, P(c_d)]),
}
for _n, (_r, _a) in _sigs.items():
    getattr(_lib, _n).restype = _r
    getattr(_lib, _n).argtypes = _a


def _p(a, t=c_d):
    return a.ctypes.data_as(P(t))


def coorDist(z1, z2, Om=0.3, OL=0.7, h=0.7):
    return _lib.orc_coordist(Om, OL, h, z1, z2)


def _counts(arr):
    return np.array([[c.visits, c.far, c.leafp, c.bigtest, c.bspline] for c in arr], np.uint64)


class Tree:
    """TreeQuadParticles over (x, y, m, h); positions promote to float64."""

    def __init__(self, x, y, m, h, theta=0.1, bucket=20, inv_area=0.0, max_range=-1.0):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.m = np.ascontiguousarray(m, dtype=np.float32)
        self.hh = np.ascontiguousarray(h, dtype=np.float32)
        self.inv_area = inv_area
        err = c_int(0)
        self.t = _lib.orc_tree_build_range(_p(self.x), _p(self.y), _p(self.m, c_f),
                                           _p(self.hh, c_f), self.m.size, inv_area, bucket,
                                           theta, max_range, ctypes.byref(err))
        if not self.t:
            raise ValueError("particles in same place.")
        self.owned = True

    def nnodes(self):
        return _lib.orc_tree_nnodes(self.t)

    def export_nodes(self):
        nn = self.nnodes()
        d = np.zeros((nn, 13))
        q = np.zeros((nn, 6), np.int64)
        _lib.orc_tree_export(self.t, _p(d), q.ctypes.data_as(P(ctypes.c_int64)))
        return d, q

    def force(self, rays, nthreads=8, counts=False):
        r = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 2)
        out = np.zeros((r.shape[0], 5))
        cnt = (Counters * r.shape[0])() if counts else None
        _lib.orc_tree_force(self.t, _p(r), r.shape[0], _p(out),
                            ctypes.cast(cnt, c_vp) if counts else None, nthreads)
        return (out, _counts(cnt)) if counts else out

    def __del__(self):
        if getattr(self, "owned", False) and self.t:
            _lib.orc_tree_free(self.t)


def direct_sum(x, y, m, h, rays, inv_area=0.0, nthreads=8):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    m = np.ascontiguousarray(m, dtype=np.float32)
    h = np.ascontiguousarray(h, dtype=np.float32)
    r = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 2)
    out = np.zeros((r.shape[0], 5))
    _lib.orc_direct_plane(_p(x), _p(y), _p(m, c_f), _p(h, c_f), m.size, inv_area, _p(r),
                          r.shape[0], _p(out), nthreads)
    return out


def smoothing_sizes(x, y, z, nsmooth, brute=False, which=None, nthreads=8):
    """LensHaloParticles::calculate_smoothing: the distance from each particle
    to its (nsmooth+1)-th nearest particle in 3-D (itself included), with the
    coordinate type of x (float32 or float64).  which: optional particle
    indices to compute (all by default).  brute: by selection over all
    particles instead of the reference's tree search."""
    t = np.float32 if np.asarray(x).dtype == np.float32 else np.float64
    x, y, z = (np.ascontiguousarray(a, dtype=t) for a in (x, y, z))
    ct = c_f if t == np.float32 else c_d
    w = None if which is None else np.ascontiguousarray(which, dtype=np.int64)
    nq = x.size if w is None else w.size
    out = np.zeros(nq, np.float32)
    fn = _lib.orc_smoothing_f32 if t == np.float32 else _lib.orc_smoothing_f64
    rc = fn(_p(x, ct), _p(y, ct), _p(z, ct), x.size, int(nsmooth), int(bool(brute)),
            _p(out, c_f), nthreads, None if w is None else _p(w, ctypes.c_int64),
            0 if w is None else w.size)
    if rc:
        raise ValueError("number of neighbors > total number of particles")
    return out


def bspline(xcm, r, mass, size):
    xcm = np.ascontiguousarray(xcm, dtype=np.float64)
    out = np.zeros(5)
    _lib.orc_bspline(_p(xcm), r, mass, size, _p(out))
    return out


class HaloTree:
    """LensHaloHalos<LensHaloNFW>'s TreeQuadHalos (include/quadTreeHalos.h):
    NFW halos (float mass, Rsize, concentration) at physical positions."""

    def __init__(self, x, y, m, rsize, conc, inv_area=0.0, theta=0.1):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.m = np.ascontiguousarray(m, dtype=np.float32)
        self.rsize = np.ascontiguousarray(rsize, dtype=np.float32)
        self.conc = np.ascontiguousarray(conc, dtype=np.float32)
        st = c_int()
        self.t = _lib.orc_halotree_build(_p(self.x), _p(self.y), _p(self.m, c_f),
                                         _p(self.rsize, c_f), _p(self.conc, c_f), self.x.size,
                                         np.float32(inv_area), np.float32(theta),
                                         ctypes.byref(st))
        if st.value:
            raise ValueError("halo tree build failed (halos in one place or out of the tables)")
        self.owned = True

    def force(self, rays, direct=False):
        rays = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 2)
        out = np.zeros((rays.shape[0], 5))
        _lib.orc_halotree_force(self.t, _p(rays), rays.shape[0], int(bool(direct)), _p(out))
        return out

    def __del__(self):
        if getattr(self, "owned", False) and getattr(self, "t", None):
            _lib.orc_halotree_free(self.t)


class FieldTree(HaloTree):
    """LensPlaneTree's TreeQuadHalos (MultiPlane/planes.cpp:13-19): NFW halos
    with sigma_background; with inv_screening_scale != 0 bucket 5, theta 0.1,
    the periodic buffer and screening, else theta 0.5 (theta and periodic may
    be given to test the parts apart)."""

    def __init__(self, x, y, m, rsize, conc, sigma_background=0.0, inv_screening_scale=0.0,
                 theta=None, periodic=None):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.m = np.ascontiguousarray(m, dtype=np.float32)
        self.rsize = np.ascontiguousarray(rsize, dtype=np.float32)
        self.conc = np.ascontiguousarray(conc, dtype=np.float32)
        scr = inv_screening_scale != 0
        theta = (0.1 if scr else 0.5) if theta is None else theta
        periodic = scr if periodic is None else periodic
        st = c_int()
        self.t = _lib.orc_halotree_build_field(
            _p(self.x), _p(self.y), _p(self.m, c_f), _p(self.rsize, c_f), _p(self.conc, c_f),
            self.x.size, float(sigma_background), float(theta), float(inv_screening_scale),
            int(bool(periodic)), ctypes.byref(st))
        if st.value:
            raise ValueError("halo tree build failed (halos in one place or out of the tables)")
        self.owned = True


class MixedTree(HaloTree):
    """A TreeQuadHalos of NFW (type 0, param = concentration) and power-law
    (type 1, param = beta) halos; theta, screening and buffer as given."""

    def __init__(self, x, y, m, rsize, types, params, inv_area=0.0, theta=0.1,
                 inv_screening_scale=0.0, periodic=False):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.m = np.ascontiguousarray(m, dtype=np.float32)
        self.rsize = np.ascontiguousarray(rsize, dtype=np.float32)
        self.types = np.ascontiguousarray(types, dtype=np.int32)
        self.params = np.ascontiguousarray(params, dtype=np.float64)
        st = c_int()
        self.t = _lib.orc_halotree_build_mixed(
            _p(self.x), _p(self.y), _p(self.m, c_f), _p(self.rsize, c_f), _p(self.types, c_int),
            _p(self.params), self.x.size, float(inv_area), float(theta),
            float(inv_screening_scale), int(bool(periodic)), ctypes.byref(st))
        if st.value:
            raise ValueError("halo tree build failed")
        self.owned = True


def powerlaw_force(m, rsize, beta, xcm):
    """LensHaloPowerLaw::force_halo (no point subtraction) at positions
    relative to the halo: (n, 5) = alpha0, alpha1, kappa, gamma0, gamma1."""
    xcm = np.ascontiguousarray(xcm, dtype=np.float64).reshape(-1, 2)
    out = np.zeros((xcm.shape[0], 5))
    _lib.orc_powerlaw_force(np.float32(m), np.float32(rsize), float(beta), _p(xcm),
                            xcm.shape[0], _p(out))
    return out


def nfw_force(m, rsize, conc, xcm):
    """One LensHaloNFW's force_halo at positions relative to it: (n, 5)."""
    xcm = np.ascontiguousarray(xcm, dtype=np.float64).reshape(-1, 2)
    out = np.zeros((xcm.shape[0], 5))
    _lib.orc_nfw_force(np.float32(m), np.float32(rsize), np.float32(conc), _p(xcm),
                       xcm.shape[0], _p(out))
    return out


def nfw_function(which, x):
    """LensHaloNFW's gfunction (0), ffunction (1), g2function (2)."""
    return _lib.orc_nfw_function(int(which), float(x))


def nfw_interpolate(which, x):
    """InterpolateFromTable of the g (0), f (1), g2 (2) tables."""
    return _lib.orc_nfw_interpolate(int(which), float(x))


class Lens:
    """Lens with moveinMainHalo'd particle halos; shoot = rayshooterInternal."""

    def __init__(self, zsource, Om=0.3, OL=0.7, h=0.7):
        self.L = _lib.orc_lens_create(zsource, Om, OL, h)

    def addFieldPlane(self, tree, z):
        """A LensPlaneTree at redshift z (double) made of a FieldTree."""
        tree.owned = False
        _lib.orc_lens_add_field_plane(self.L, tree.t, float(z))

    def moveinMainHalo(self, tree, z, addplanes=True):
        tree.owned = False          # the lens owns it now
        if not addplanes:
            _lib.orc_lens_add_nearest(self.L, tree.t, np.float32(z),
                                      int(isinstance(tree, HaloTree)))
        elif isinstance(tree, HaloTree):
            _lib.orc_lens_add_halotree(self.L, tree.t, np.float32(z))
        else:
            _lib.orc_lens_add_halo(self.L, tree.t, np.float32(z))

    def tables(self):
        n = _lib.orc_lens_nplanes(self.L)
        z, Dl, dDl = (np.zeros(n + 1) for _ in range(3))
        ch = c_d()
        _lib.orc_lens_tables(self.L, _p(z), _p(Dl), _p(dDl), ctypes.byref(ch))
        return {"nplanes": n, "z": z, "Dl": Dl, "dDl": dDl, "charge": ch.value}

    def shoot(self, x, zs=None, nthreads=8, counts=False):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 2)
        n = x.shape[0]
        y = np.zeros((n, 2))
        A = np.zeros((n, 4))
        zp = None
        if zs is not None:
            zs = np.ascontiguousarray(zs, dtype=np.float64)
            zp = _p(zs)
        cnt = (Counters * n)() if counts else None
        _lib.orc_lens_shoot(self.L, n, _p(x), zp, _p(y), _p(A), nthreads,
                            ctypes.cast(cnt, c_vp) if counts else None)
        return (y, A, _counts(cnt)) if counts else (y, A)

    def __del__(self):
        if getattr(self, "L", None):
            _lib.orc_lens_free(self.L)


def gridmap_positions(N, center, range_):
    x = np.zeros((N * N, 2))
    _lib.orc_gridmap_positions(N, center[0], center[1], range_, _p(x))
    return x


def gridmap_rect_positions(nx, center, range_x, range_y):
    """GridMap(lens, Nx, center, rangeX, rangeY)'s points
    (TreeCode_link/gridmap.cpp:26-91): ny = int(nx rangeY / rangeX) rows,
    point ii + jj nx at (x0 + ii res, y0 + jj res), res = rangeX / (nx - 1),
    (x0, y0) the lower-left corner.  Returns (positions (nx*ny, 2), ny)."""
    ny = int(nx * (range_y / range_x))
    xo = center[0] - 0.5 * range_x
    yo = center[1] - 0.5 * range_y
    res = range_x / (nx - 1)
    k = np.arange(nx * ny)
    return np.stack([xo + (k % nx) * res, yo + (k // nx) * res], 1), ny


def find_images(N, x, y, A, ys, cap=None, ny=None):
    """GridMap::find_images over a grid (x, y: (N*ny, 2), A: (N*ny, 4); ny
    rows of N points, ny = N by default) for source positions ys (n, 2):
    (rays (m, 8) = x, y, A; multiplicity; triangles (m, 3)), sources in
    order."""
    ny = N if ny is None else ny
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    A = np.ascontiguousarray(A, dtype=np.float64)
    ys = np.ascontiguousarray(ys, dtype=np.float64).reshape(-1, 2)
    ns = ys.shape[0]
    cap = cap or 64 * max(ns, 1)
    mult = np.zeros(ns, np.int32)
    rays = np.zeros((cap, 8))
    tris = np.zeros((cap, 3), np.int64)
    n = _lib.orc_find_images2(N, ny, _p(x), _p(y), _p(A), _p(ys), ns,
                             mult.ctypes.data_as(P(ctypes.c_int32)), _p(rays),
                             tris.ctypes.data_as(P(ctypes.c_int64)), cap)
    if n > cap:
        raise ValueError("more images than cap")
    return rays[:n], mult, tris[:n]


# ------------------------------------------------------- critical curves
CRIT_TANGENTIAL, CRIT_RADIAL, CRIT_PSEUDO = 1, 2, 3


def _split(off, pts, nc):
    return [pts[off[i]:off[i + 1]].copy() for i in range(nc)]


def find_boundaries(bitmap):
    """Utilities::find_boundaries(bitmap (ny, nx) bool): (curves in pixel
    units, hits_edge)."""
    bm = np.ascontiguousarray(bitmap, dtype=np.uint8)
    ny, nx = bm.shape
    cc, cp = 16, 4 * bm.size + 16
    while True:
        off = np.zeros(cc + 1, np.int64)
        pts = np.zeros((cp, 2))
        hits = np.zeros(cc, np.int32)
        npt = c_sz()
        nc = _lib.orc_find_boundaries(_p(bm, ctypes.c_uint8), bm.size, nx, cc, cp,
                                      _p(off, ctypes.c_int64), _p(pts), _p(hits, ctypes.c_int32),
                                      ctypes.byref(npt))
        if nc <= cc:
            return _split(off, pts, nc), hits[:nc].astype(bool)
        cc = nc


def find_islands(bitmap):
    """Utilities::find_islands(bitmap (ny, nx) bool): pixel index lists."""
    bm = np.ascontiguousarray(bitmap, dtype=np.uint8)
    ny, nx = bm.shape
    ci = 16
    while True:
        off = np.zeros(ci + 1, np.int64)
        pix = np.zeros(bm.size + 1, np.int64)
        npx = c_sz()
        ni = _lib.orc_find_islands(_p(bm, ctypes.c_uint8), bm.size, nx, ci, bm.size + 1,
                                   _p(off, ctypes.c_int64), _p(pix, ctypes.c_int64),
                                   ctypes.byref(npx))
        if ni <= ci:
            return _split(off, pix, ni)
        ci = ni


def convex_hull(points):
    """Utilities::convex_hull(P, hull): indices of the hull, counter-clockwise."""
    p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
    h = np.zeros(2 * len(p) + 1, np.int64)
    n = _lib.orc_convex_hull(_p(p), len(p), _p(h, ctypes.c_int64))
    return h[:n]


def in_curve(poly, points):
    """Utilities::inCurve(x, H) for each of the points."""
    h = np.ascontiguousarray(poly, dtype=np.float64).reshape(-1, 2)
    q = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
    out = np.zeros(len(q), np.int32)
    _lib.orc_in_curve(_p(h), len(h), _p(q), len(q), _p(out, ctypes.c_int32))
    return out.astype(bool)


def _crit(mode, N, x, y, A, mult, range_, invmag, ny=None):
    ny = N if ny is None else ny
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y if y is not None else x, dtype=np.float64)
    A = np.ascontiguousarray(A if A is not None else np.zeros((N * ny, 4)), dtype=np.float64)
    m = np.ascontiguousarray(mult if mult is not None else np.zeros(N * ny), dtype=np.int32)
    cc, cp = 16, 4 * N * ny + 16
    while True:
        off = np.zeros(cc + 1, np.int64)
        pts = np.zeros((cp, 2))
        types = np.zeros(cc, np.int32)
        hits = np.zeros(cc, np.int32)
        nh, npt, rep = c_sz(), c_sz(), c_int()
        nc = _lib.orc_find_crit2(mode, N, ny, _p(x), _p(y), _p(A), _p(m, ctypes.c_int32), range_,
                                invmag, cc, cp, _p(off, ctypes.c_int64), _p(pts),
                                _p(types, ctypes.c_int32), _p(hits, ctypes.c_int32),
                                ctypes.byref(nh), ctypes.byref(npt), ctypes.byref(rep))
        if nc <= cc and npt.value <= cp:
            return (_split(off, pts, nc), hits[:nh.value].astype(bool), types[:nc],
                    bool(rep.value))
        cc, cp = max(cc, nc), max(cp, npt.value)


def find_crit(N, x, y, A, range_, ny=None):
    """GridMap::find_crit over a grid (x, y (N*ny, 2), A (N*ny, 4); range_
    the x range): (curves, hits_boundary, crit types, whether hits_boundary
    was replaced)."""
    return _crit(0, N, x, y, A, None, range_, 0.0, ny)


def find_magnification_contour(N, x, A, range_, invmag, ny=None):
    """GridMap::find_magnification_contour: (curves, hits_boundary, -, replaced)."""
    return _crit(1, N, x, None, A, None, range_, invmag, ny)


def find_boundaries_of_caustics(N, x, y, A, ny=None):
    """GridMap::find_boundaries_of_caustics: the boundaries (pixel units) of
    the grid positions with more than one image, and hits_edge."""
    ny = N if ny is None else ny
    _, mult, _ = find_images(N, x, y, A, x, cap=64 * N * ny, ny=ny)
    c, h, _, _ = _crit(2, N, x, y, A, mult, 0.0, 0.0, ny)
    return c, h, mult


def tightest_hull(points):
    """Utilities::TightestHull(v) (TreeCode_link/curve_routines.cpp:3519-3592)."""
    p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
    cap = 4096
    while True:
        out = np.zeros((cap, 2))
        n = _lib.orc_tightest_hull(_p(p), len(p), cap, _p(out))
        if n <= cap:
            return out[:n].copy()
        cap = n


def windings(x, points):
    """Utilities::windings(x, points, &area) (curve_routines.cpp:1709-1736):
    (winding number, absolute area)."""
    p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
    q = np.ascontiguousarray(x, dtype=np.float64).reshape(2)
    area = c_d()
    wn = _lib.orc_windings(_p(q), _p(p), len(p), ctypes.byref(area))
    return wn, area.value


def image_finding_find_crit(L, N, x, y, A, range_, ny=None):
    """ImageFinding::find_crit(lens, gridmap, critcurves)
    (TreeCode_link/find_crit.cpp:1036-1096) over a shot grid (x, y, A) and the
    oracle lens L: per curve a dict with the reference's fields (critcurve
    points and their rays, centres, areas, caustic points and outline)."""
    zs = float(L.tables()["z"][-1])
    curves, hits, types, replaced = find_crit(N, x, y, A, range_, ny)
    out = []
    for i, c in enumerate(curves):
        n = len(c)
        center = np.zeros(2)
        for j in range(n):
            center += c[j]
        center = center / n
        cy, cA = L.shoot(c, zs=np.full(n, zs))
        ccen = np.zeros(2)
        for j in range(n):
            ccen += cy[j]
        ccen = ccen / n
        outline = tightest_hull(cy)
        out.append({"type": int(types[i]), "touches_edge": bool(hits[i]) if replaced else False,
                    "z_source": zs, "critcurve": c, "critcurve_y": cy, "critcurve_A": cA,
                    "critical_center": center, "critical_area": windings(center, c)[1],
                    "caustic_curve_intersecting": cy, "caustic_center": ccen,
                    "caustic_curve_outline": outline,
                    "caustic_area": windings(ccen, outline)[1]})
    return out


# ------------------------------------------------------------ Grid
def _position_from_index(i, N, range_, center):
    """Utilities::PositionFromIndex (TreeCode_link/utilities.cpp:149-158)."""
    if N == 1:
        return [center[0], center[1]]
    return [center[0] + range_ * (1.0 * (i % N) / (N - 1) - 0.5),
            center[1] + range_ * (1.0 * (i // N) / (N - 1) - 0.5)]


def _xygridpoints(range_, center, N, remove_center):
    """Grid::xygridpoints (TreeCode_link/image_finder.cpp:790-824): positions
    and gridsize; the centre point left out when remove_center and N odd."""
    pts = []
    if remove_center and N % 2 == 1:
        j = 0
        for i in range(N * N):
            if (2 * (i // N)) // (N - 1) == 1 and i % N == N // 2 + 1:
                j = 1
            p = _position_from_index(i, N, range_, center)
            if i - j < len(pts):
                pts[i - j] = p
            else:
                pts.append(p)
        pts = pts[:N * N - 1]
    else:
        pts = [_position_from_index(i, N, range_, center) for i in range(N * N)]
    return pts, range_ / (N - 1)


def _inbox(x, p1, p2):
    return (x[0] >= p1[0]) and (x[0] <= p2[0]) and (x[1] >= p1[1]) and (x[1] <= p2[1])


class Grid:
    """Grid (TreeCode_link/grid_maintenance.cpp) restated over the oracle
    lens: the point lists in the reference's order (x, y, A, gridsize), the
    two trees' root boxes, RefineLeaves (:681-982), RefineLeaf (:543-679)
    and EinsteinArea (:368-378).  The trees themselves are left out: only
    their root boxes enter the refinement."""

    def __init__(self, L, N, center, range_, range_y=None):
        self.L = L
        self.Ngrid_block = 3
        if range_y is None:                        # grid_maintenance.cpp:22-59
            pts, gs = _xygridpoints(range_, center, N, 0)
            self.x = [list(p) for p in pts]
            self.gs = [gs] * len(pts)
            buffer = range_
        else:                                      # :67-129
            axisratio = range_y / range_
            N2 = int(N * axisratio)
            self.x = [None] * (N * N2)
            self.gs = [0.0] * (N * N2)
            for ii in range(N):
                for jj in range(N2):
                    i = ii + jj * N
                    self.x[i] = [center[0] + range_ * ii / (N - 1) - 0.5 * range_,
                                 center[1] + range_ * jj / (N - 1) - 0.5 * range_y]
                    self.gs[i] = range_ / (N - 1)
            buffer = max(range_, range_y)
        y, A = L.shoot(np.array(self.x))
        self.y = [list(v) for v in y]
        self.A = [list(v) for v in A]
        # TreeStruct root boxes (tree_maintenance.cpp:96-124)
        xs = np.array(self.x)
        self.ip1, self.ip2 = list(xs.min(0)), list(xs.max(0))
        ys = np.array(self.y)
        self.sp1 = [ys[:, 0].min() - buffer, ys[:, 1].min() - buffer]
        self.sp2 = [ys[:, 0].max() + buffer, ys[:, 1].max() + buffer]

    def npoints(self):
        return len(self.x)

    def refine_leaves(self, ids):
        """The new points' ids (appended in the reference's order)."""
        nb = self.Ngrid_block
        ncell = nb * nb - 1
        i_points, i_gs = [], []
        Nadded = 0
        Nout_tot = 0
        addedtocell = [0] * len(ids)
        for ii, q in enumerate(ids):
            if self.gs[q] < 10.0 ** -15 * max(abs(self.x[q][0]), abs(self.x[q][1])):
                raise RuntimeError("Cell size is too small for double precision")
            pts, g = _xygridpoints(self.gs[q] * (nb - 1) / nb, self.x[q], nb, 1)
            i_points[Nadded:] = pts
            i_gs[Nadded:] = [g] * len(pts)
            self.gs[q] /= nb
            # points outside the initial image grid, taken out by shifting
            Nout = 0
            for kk in range(ncell):
                k = Nadded + kk - Nout
                p = i_points[k]
                if (not _inbox(p, self.ip1, self.ip2)
                        or i_gs[k] < 1.0e-10 * max(abs(p[0]), abs(p[1]))):
                    del i_points[k]
                    del i_gs[k]
                    Nout += 1
            Nout_tot += Nout
            Nadded += ncell - Nout
            addedtocell[ii] = ncell - Nout
            if addedtocell[ii] == 0:
                self.gs[q] *= nb
        if Nadded == 0:
            return []
        y, A = self.L.shoot(np.array(i_points))
        s_points = [list(v) for v in y]
        s_A = [list(v) for v in A]
        # points outside the initial source grid, taken out by shifting
        kk = 0
        Nout = 0
        for ii, q in enumerate(ids):
            if addedtocell[ii] == 0:
                continue
            Noutcell = 0
            for j in range(addedtocell[ii]):
                k = kk - Nout
                if not _inbox(s_points[k], self.sp1, self.sp2):
                    del i_points[k], i_gs[k], s_points[k], s_A[k]
                    Nout += 1
                    Noutcell += 1
                kk += 1
            addedtocell[ii] -= Noutcell
            if addedtocell[ii] == 0:
                self.gs[q] *= nb
        first = len(self.x)
        self.x += i_points
        self.y += s_points
        self.A += s_A
        self.gs += i_gs
        return list(range(first, len(self.x)))

    def refine_leaf(self, q):
        nb = self.Ngrid_block
        ncell = nb * nb - 1
        if not self.gs[q] > 10.0 ** -15:
            raise RuntimeError("cell too small")
        i_points, g = _xygridpoints(self.gs[q] * (nb - 1) / nb, self.x[q], nb, 1)
        i_gs = [g] * ncell
        self.gs[q] /= nb
        Nout = 0
        xq = self.x[q]
        if (xq[0] == self.ip1[0] or xq[0] == self.ip2[0] or xq[1] == self.ip1[1]
                or xq[1] == self.ip2[1]):
            for kk in range(ncell):
                a, b = kk - Nout, ncell - 1 - Nout
                if not _inbox(i_points[a], self.ip1, self.ip2):
                    i_points[a], i_points[b] = i_points[b], i_points[a]
                    i_gs[a], i_gs[b] = i_gs[b], i_gs[a]
                    Nout += 1
        Ntemp = ncell - Nout
        i_points, i_gs = i_points[:Ntemp], i_gs[:Ntemp]
        if Ntemp > 0:
            y, A = self.L.shoot(np.array(i_points))
            s_points, s_A = [list(v) for v in y], [list(v) for v in A]
        else:
            s_points, s_A = [], []
        Nout = 0
        for kk in range(Ntemp):
            a, b = kk - Nout, Ntemp - 1 - Nout
            if not _inbox(s_points[a], self.sp1, self.sp2):
                for arr in (i_points, i_gs, s_points, s_A):
                    arr[a], arr[b] = arr[b], arr[a]
                Nout += 1
        n = Ntemp - Nout
        if n <= 0:
            self.gs[q] *= nb
            return []
        first = len(self.x)
        self.x += i_points[:n]
        self.y += s_points[:n]
        self.A += s_A[:n]
        self.gs += i_gs[:n]
        return list(range(first, len(self.x)))

    def einstein_area(self):
        total = 0.0
        for a, g in zip(self.A, self.gs):
            if a[0] * a[3] - a[1] * a[2] < 0:
                total += g * g
        return total

    def arrays(self):
        return (np.array(self.x), np.array(self.y), np.array(self.A), np.array(self.gs))


# ------------------------------------------------ Lens::ResetSourcePlane
def _nr_locate(xx, x):
    """Numerical Recipes' locate (Numerical Recipes in C, 2nd ed., sec. 3.4)
    on the 0-based table xx, returning the book's 1-based j: xx[j-1] <= x <
    xx[j], 1 at x == xx[0], n - 1 at x == xx[n-1]."""
    n = len(xx)
    jl, ju = 0, n + 1
    ascnd = xx[n - 1] >= xx[0]
    while ju - jl > 1:
        jm = (ju + jl) >> 1
        if (x >= xx[jm - 1]) == ascnd:
            jl = jm
        else:
            ju = jm
    if x == xx[0]:
        return 1
    if x == xx[n - 1]:
        return n - 1
    return jl


def inv_coordist(D, Om=0.3, OL=0.7, h=0.7):
    """The redshift at comoving distance D (bisection on coorDist)."""
    lo, hi = 0.0, 1.0
    while coorDist(0.0, hi, Om, OL, h) < D and hi < 1e4:
        hi *= 2
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if mid in (lo, hi):
            break
        if coorDist(0.0, mid, Om, OL, h) < D:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def reset_source_plane(tables, z, nearest, Om=0.3, OL=0.7, h=0.7):
    """Lens::ResetSourcePlane(z, nearest) (MultiPlane/lens.cpp:2609-2657) on
    a lens's plane tables (Lens.tables()): (plane index, implanted source
    redshift)."""
    zp, Dl, dDl, n = tables["z"], tables["Dl"], tables["dDl"], tables["nplanes"]
    Ds = coorDist(0.0, z, Om, OL, h)
    j = _nr_locate(Dl[:n], Ds) if len(Dl) > 1 and n > 0 else 0
    if j > 0:
        if Dl[j - 1] == Ds:
            j -= 1
        elif nearest:
            if z < inv_coordist(Dl[j] - 0.5 * dDl[j], Om, OL, h):
                j -= 1
    return j, (float(zp[j]) if nearest and j < n else z)
