"""MI355X multi-plane particle ray shooter (GLAMER's Lens / GridMap /
LensHaloParticles particle path), Python view over the C ABI in
include/slsim_amd.h.

The compute path is the HIP library libslsim_amd.so built in-tree; importing
this package fails loudly if it is missing (there is no CPU fallback).
"""
import ctypes
import os
import weakref

import numpy as np

from . import _abi
from ._abi import SLSIM_DIRECT64, SLSIM_FP32C, SLSIM_FP64, SlsimError, lib
from ._abi import SLSIM_CRIT_ND, SLSIM_CRIT_PSEUDO, SLSIM_CRIT_RADIAL, SLSIM_CRIT_TANGENTIAL
from ._abi import SLSIM_HALO_NFW, SLSIM_HALO_POWERLAW
from ._abi import (SLSIM_FMT_ASCII, SLSIM_FMT_CSV3, SLSIM_FMT_CSV4, SLSIM_FMT_CSV5,
                   SLSIM_FMT_CSV6, SLSIM_FMT_GADGET2, SLSIM_FMT_GLMB)

arcsecTOradians = 0.000004848136811   # include/standard.h:51
Grav = 4.7788e-20                      # include/standard.h:43

__all__ = [
    "Context", "LensHaloParticles", "LensHaloHalos", "Lens", "GridMap", "Batch", "Server", "coorDist",
    "invCoorDist",
    "nfw_r200", "gen_nfw", "gen_halo_field", "SLSIM_FP64", "SLSIM_FP32C", "SLSIM_DIRECT64",
    "SlsimError", "MakeParticleLenses", "PARTICLE_DTYPE", "glmb_write", "glmb_read",
    "gadget2_read",
    "write_fits_pixelmap", "SLSIM_FMT_GLMB", "SLSIM_FMT_CSV3", "SLSIM_FMT_CSV4",
    "SLSIM_FMT_CSV5", "SLSIM_FMT_CSV6", "SLSIM_FMT_GADGET2", "SLSIM_FMT_ASCII",
    "arcsecTOradians", "Grav", "lib",
]

# ParticleType<float> (include/particle_types.h:46-58), 24 bytes
PARTICLE_DTYPE = np.dtype([("x", "<f4", (3,)), ("Mass", "<f4"), ("Size", "<f4"),
                           ("type", "<i4")])


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def _check(rc):
    _abi.check(rc)


class Context:
    """Device context with an HBM budget (0 = 90 % of free HBM)."""

    def __init__(self, device=0, hbm_budget=0):
        self.h = _abi.slsim_ctx()
        self._children = []    # weak references to the objects made on it, oldest first
        _check(lib.slsim_ctx_create(int(device), ctypes.c_size_t(int(hbm_budget)),
                                    ctypes.byref(self.h)))

    def _adopt(self, obj):
        """Registers an object that holds memory or handles of this context,
        so that close() releases it first (newest first)."""
        if len(self._children) >= 256:
            self._children = [r for r in self._children if r() is not None]
        self._children.append(weakref.ref(obj))

    def memory(self):
        used, bud = ctypes.c_size_t(), ctypes.c_size_t()
        _check(lib.slsim_ctx_memory(self.h, ctypes.byref(used), ctypes.byref(bud)))
        return used.value, bud.value

    def synchronize(self):
        _check(lib.slsim_ctx_synchronize(self.h))

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        _check(lib.slsim_device_alloc(self.h, ctypes.c_size_t(int(nbytes)), ctypes.byref(p)))
        return p

    def free(self, p):
        _check(lib.slsim_device_free(self.h, p))

    def d2h(self, host, dev):
        _check(lib.slsim_memcpy_d2h(self.h, host.ctypes.data_as(ctypes.c_void_p), dev,
                                    ctypes.c_size_t(host.nbytes)))

    def h2d(self, dev, host):
        _check(lib.slsim_memcpy_h2d(self.h, dev, host.ctypes.data_as(ctypes.c_void_p),
                                    ctypes.c_size_t(host.nbytes)))

    def d2d(self, dst, src, nbytes):
        _check(lib.slsim_memcpy_d2d(self.h, dst, src, ctypes.c_size_t(nbytes)))

    def close(self):
        """Closes every live object made on the context (newest first), then
        the context; objects still referenced afterwards are inert."""
        if self.h:
            for r in reversed(self._children):
                o = r()
                if o is not None:
                    try:
                        o.close()
                    except Exception:
                        pass
            self._children = []
            lib.slsim_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LensHaloParticles:
    """Particle halo with its device quadtree (particle_halo.h:58-75).

    x, y: positions in physical Mpc (float64 -> ParticleType<double>,
    float32 -> ParticleType<float>); m: masses (Msun, stored float32);
    h: sizes (Mpc, float32); max_range > 0: the tree's maximum range
    (particle_halo.h:58-75, 416), beyond which cells stay unsplit."""

    def __init__(self, ctx, x, y, m, h, z, theta=0.1, bucket=20, inv_area=0.0, max_range=-1.0):
        self.ctx = ctx
        ctx._adopt(self)
        self.z = np.float32(z)
        self.owned_by_lens = False
        m = np.ascontiguousarray(m, dtype=np.float32)
        h = np.ascontiguousarray(h, dtype=np.float32)
        n = m.size
        self.h = _abi.slsim_plane()
        if np.asarray(x).dtype == np.float32:
            xs = np.ascontiguousarray(x, dtype=np.float32)
            ys = np.ascontiguousarray(y, dtype=np.float32)
            fn = lib.slsim_plane_create_range_f32
            ct = ctypes.c_float
        else:
            xs = np.ascontiguousarray(x, dtype=np.float64)
            ys = np.ascontiguousarray(y, dtype=np.float64)
            fn = lib.slsim_plane_create_range_f64
            ct = ctypes.c_double
        _check(fn(ctx.h, _ptr(xs, ct), _ptr(ys, ct), _ptr(m, ctypes.c_float),
                  _ptr(h, ctypes.c_float), ctypes.c_size_t(n), float(theta), int(bucket),
                  float(inv_area), float(max_range), ctypes.byref(self.h)))

    @staticmethod
    def calculate_smoothing(ctx, Nsmooth, x, y, z):
        """LensHaloParticles::calculate_smoothing (particle_halo.h:683-730) on
        the device: for each particle the distance to its (Nsmooth+1)-th
        nearest particle in 3-D (itself included), in the positions' type
        (float64 or float32).  Returns the sizes (float32), which the
        reference stores in ParticleType::Size."""
        t = np.float32 if np.asarray(x).dtype == np.float32 else np.float64
        xs, ys, zs = (np.ascontiguousarray(a, dtype=t) for a in (x, y, z))
        ct = ctypes.c_float if t == np.float32 else ctypes.c_double
        fn = lib.slsim_smoothing_sizes_f32 if t == np.float32 else lib.slsim_smoothing_sizes_f64
        out = np.zeros(xs.size, np.float32)
        _check(fn(ctx.h, _ptr(xs, ct), _ptr(ys, ct), _ptr(zs, ct), ctypes.c_size_t(xs.size),
                  int(Nsmooth), _ptr(out, ctypes.c_float)))
        return out

    @classmethod
    def from_ascii(cls, ctx, filename, redshift, Nsmooth, theta_rotate=(0.0, 0.0),
                   recenter=False, multimass=True, inv_area=0.0, MinPSize=0.0,
                   rescale_mass=1.0, dtype=np.float64):
        """LensHaloParticles<PType>(simulation_filename, SimFileFormat::ascii,
        redshift, Nsmooth, cosmo, theta_rotate, recenter, multimass, inv_area,
        MinPSize, rescale_mass) (particle_halo.h:226-340): positions from the
        ASCII file, sizes from "<file>.<Nsmooth>sizes" or computed on the
        device and written there, then the tree (theta 0.1, bucket 20).
        dtype float64 -> ParticleType<double>, float32 -> ParticleType<float>.
        The halo keeps .xyz, .mass, .size, .halo_info."""
        xyz, m, h, info = cls.load_ascii(ctx, filename, redshift, Nsmooth, theta_rotate, recenter,
                                         multimass, MinPSize, rescale_mass, dtype)
        halo = cls(ctx, np.ascontiguousarray(xyz[:, 0]), np.ascontiguousarray(xyz[:, 1]), m, h,
                   redshift, theta=0.1, bucket=20, inv_area=inv_area)
        halo.xyz, halo.mass, halo.size, halo.halo_info = xyz, m, h, info
        return halo

    @staticmethod
    def load_ascii(ctx, filename, redshift, Nsmooth, theta_rotate=(0.0, 0.0), recenter=False,
                   multimass=True, MinPSize=0.0, rescale_mass=1.0, dtype=np.float64):
        """The particle set of from_ascii without building a tree:
        (xyz, mass, size, info)."""
        fn = os.fsencode(filename)
        n = ctypes.c_size_t()
        _check(lib.slsim_ascii_halo_count(fn, int(bool(multimass)), ctypes.byref(n)))
        n = n.value
        t = np.float32 if np.dtype(dtype) == np.float32 else np.float64
        xyz = np.zeros((n, 3), t)
        m = np.zeros(n, np.float32)
        h = np.zeros(n, np.float32)
        th = np.array([theta_rotate[0], theta_rotate[1]], np.float64)
        info = _abi.slsim_ascii_halo_info()
        fn_load = lib.slsim_ascii_halo_load_f32 if t == np.float32 else lib.slsim_ascii_halo_load_f64
        ct = ctypes.c_float if t == np.float32 else ctypes.c_double
        _check(fn_load(ctx.h if ctx is not None else None, fn, float(redshift), int(Nsmooth),
                       _ptr(th, ctypes.c_double), int(bool(recenter)), int(bool(multimass)),
                       float(MinPSize), float(rescale_mass), ctypes.c_size_t(n), _ptr(xyz, ct),
                       _ptr(m, ctypes.c_float), _ptr(h, ctypes.c_float), ctypes.byref(info)))
        return xyz, m, h, {"mass": info.mass, "mcenter": tuple(info.mcenter),
                           "densest_point": tuple(info.densest_point), "rsize": info.rsize}

    def info(self):
        i = _abi.slsim_plane_info()
        _check(lib.slsim_plane_get_info(self.h, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in i._fields_ if k != "pad"}

    def export_nodes(self):
        nn = self.info()["n_nodes"]
        d = np.zeros((nn, 13), np.float64)
        q = np.zeros((nn, 6), np.int64)
        _check(lib.slsim_plane_export_nodes(self.h, _ptr(d, ctypes.c_double),
                                            _ptr(q, ctypes.c_int64)))
        return d, q

    def force(self, rays, precision=SLSIM_FP64):
        """Tree walk at physical positions rays (n, 2): raw force2D_recur
        outputs (alpha0, alpha1, kappa, gamma0, gamma1)."""
        r = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 2)
        out = np.zeros((r.shape[0], 5), np.float64)
        _check(lib.slsim_plane_force(self.h, ctypes.c_size_t(r.shape[0]),
                                     _ptr(r, ctypes.c_double), _ptr(out, ctypes.c_double),
                                     int(precision)))
        return out

    def direct_sum(self, rays):
        r = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 2)
        out = np.zeros((r.shape[0], 5), np.float64)
        _check(lib.slsim_plane_direct_sum(self.h, ctypes.c_size_t(r.shape[0]),
                                          _ptr(r, ctypes.c_double), _ptr(out, ctypes.c_double)))
        return out

    def close(self):
        if self.h and not self.owned_by_lens and self.ctx.h:
            lib.slsim_plane_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _halos_plane(ctx, x, y, mass, rsize, kinds, params, inv_area, theta, scr, periodic):
    xs = np.ascontiguousarray(x, dtype=np.float64)
    ys = np.ascontiguousarray(y, dtype=np.float64)
    m = np.ascontiguousarray(mass, dtype=np.float32)
    r = np.ascontiguousarray(rsize, dtype=np.float32)
    k = np.ascontiguousarray(kinds, dtype=np.int32)
    p = np.ascontiguousarray(params, dtype=np.float64)
    h = _abi.slsim_plane()
    _check(lib.slsim_halos_create(ctx.h, _ptr(xs, ctypes.c_double), _ptr(ys, ctypes.c_double),
                                  _ptr(m, ctypes.c_float), _ptr(r, ctypes.c_float),
                                  _ptr(k, ctypes.c_int32), _ptr(p, ctypes.c_double),
                                  ctypes.c_size_t(xs.size), float(inv_area), float(theta),
                                  float(scr), int(periodic), ctypes.byref(h)))
    return h


class LensHaloHalos:
    """LensHaloHalos<LensHaloNFW>(halos, redshift, cosmo, inv_area,
    theta_force) (include/particle_halo.h:1092-1227): a main halo made of
    symmetric NFW halos, LensHaloNFW(mass, Rsize, zlens, concentration)
    each, at physical positions (x, y) (Mpc), on a TreeQuadHalos quadtree
    (bucket 5).  Goes into a Lens with moveinMainHalo like a particle halo."""

    def __init__(self, ctx, x, y, mass, rsize, concentration, z, inv_area=0.0, theta=0.1):
        self.ctx = ctx
        ctx._adopt(self)
        self.z = np.float32(z)
        self.owned_by_lens = False
        xs = np.ascontiguousarray(x, dtype=np.float64)
        ys = np.ascontiguousarray(y, dtype=np.float64)
        m = np.ascontiguousarray(mass, dtype=np.float32)
        r = np.ascontiguousarray(rsize, dtype=np.float32)
        c = np.ascontiguousarray(concentration, dtype=np.float32)
        self.h = _abi.slsim_plane()
        _check(lib.slsim_halos_create_nfw(ctx.h, _ptr(xs, ctypes.c_double), _ptr(ys, ctypes.c_double),
                                          _ptr(m, ctypes.c_float), _ptr(r, ctypes.c_float),
                                          _ptr(c, ctypes.c_float), ctypes.c_size_t(xs.size),
                                          ctypes.c_float(inv_area), ctypes.c_float(theta),
                                          ctypes.byref(self.h)))

    @classmethod
    def mixed(cls, ctx, x, y, mass, rsize, kinds, params, z, inv_area=0.0, theta=0.1):
        """LensHaloHalos of symmetric halos of mixed kinds: SLSIM_HALO_NFW
        (param = concentration) or SLSIM_HALO_POWERLAW (param = beta)."""
        self = cls.__new__(cls)
        self.ctx = ctx
        ctx._adopt(self)
        self.z = np.float32(z)
        self.owned_by_lens = False
        self.h = _halos_plane(ctx, x, y, mass, rsize, kinds, params,
                              float(np.float32(inv_area)), float(np.float32(theta)), 0.0, 0)
        return self

    # the plane handle's methods (a halo-of-halos plane is a slsim_plane)
    info = LensHaloParticles.info
    force = LensHaloParticles.force
    direct_sum = LensHaloParticles.direct_sum
    close = LensHaloParticles.close
    __del__ = LensHaloParticles.__del__


class LensPlaneTree:
    """LensPlaneTree(z, halos, Nhalos, sigma_background, inv_screening_scale)
    (MultiPlane/planes.cpp:13-19): a field plane of symmetric NFW halos at
    physical positions (x, y) on a TreeQuadHalos with sigma_background as its
    background; inv_screening_scale != 0 screens its far cells and adds the
    periodic buffer (theta 0.1), else theta 0.5.  z is kept as a double.  Goes
    into a Lens with addFieldPlane; force() is LensPlaneTree::force's tree
    part (TreeQuadHalos::force2D)."""

    def __init__(self, ctx, z, x, y, mass, rsize, concentration, sigma_background=0.0,
                 inv_screening_scale=0.0):
        self.ctx = ctx
        ctx._adopt(self)
        self.z = float(z)
        self.owned_by_lens = False
        xs = np.ascontiguousarray(x, dtype=np.float64)
        ys = np.ascontiguousarray(y, dtype=np.float64)
        m = np.ascontiguousarray(mass, dtype=np.float32)
        r = np.ascontiguousarray(rsize, dtype=np.float32)
        c = np.ascontiguousarray(concentration, dtype=np.float32)
        self.h = _abi.slsim_plane()
        _check(lib.slsim_plane_tree_create_nfw(
            ctx.h, _ptr(xs, ctypes.c_double), _ptr(ys, ctypes.c_double), _ptr(m, ctypes.c_float),
            _ptr(r, ctypes.c_float), _ptr(c, ctypes.c_float), ctypes.c_size_t(xs.size),
            float(sigma_background), float(inv_screening_scale), ctypes.byref(self.h)))

    @classmethod
    def mixed(cls, ctx, z, x, y, mass, rsize, kinds, params, sigma_background=0.0,
              inv_screening_scale=0.0):
        """A LensPlaneTree of symmetric halos of mixed kinds (SLSIM_HALO_NFW:
        param = concentration, SLSIM_HALO_POWERLAW: param = beta)."""
        self = cls.__new__(cls)
        self.ctx = ctx
        ctx._adopt(self)
        self.z = float(z)
        self.owned_by_lens = False
        scr = float(inv_screening_scale) != 0.0
        self.h = _halos_plane(ctx, x, y, mass, rsize, kinds, params, sigma_background,
                              0.1 if scr else 0.5, inv_screening_scale, int(scr))
        return self

    info = LensHaloParticles.info
    force = LensHaloParticles.force
    direct_sum = LensHaloParticles.direct_sum
    close = LensHaloParticles.close
    __del__ = LensHaloParticles.__del__


class Lens:
    """Lens(seed, zs, cosmo) with flat LCDM (MultiPlane/lens.cpp:59-77)."""

    def __init__(self, ctx, zsource, Om=0.3, OL=0.7, h=0.7):
        self.ctx = ctx
        ctx._adopt(self)
        self.cosmo = (Om, OL, h)
        self.h = _abi.slsim_lens()
        self.halos = []
        self._implant = None
        _check(lib.slsim_lens_create(ctx.h, float(zsource), float(Om), float(OL), float(h),
                                     ctypes.byref(self.h)))

    def clearMainHalos(self):
        """Lens::clearMainHalos (MultiPlane/lens.cpp:1025-1036)."""
        _check(lib.slsim_lens_clear_main_halos(self.h))
        keep = []
        for hh in self.halos:
            if isinstance(hh, LensPlaneTree):
                keep.append(hh)
            else:
                hh.h = None           # released by the lens
        self.halos = keep

    def moveinMainHalos(self, halos, addplanes=True):
        """Lens::insertMainHalos (include/lens.h:219-237) for halos the lens
        takes (their device planes move in)."""
        for hh in halos:
            self.moveinMainHalo(hh, addplanes)

    def replaceMainHalos(self, halos):
        """Lens::replaceMainHalos (include/lens.h:244-263): the halos replace
        the main halos, one plane per redshift."""
        self.clearMainHalos()
        self.moveinMainHalos(halos, True)

    def getNMainHalos(self):
        """Lens::getNMainHalos (include/lens.h:303, 887-891)."""
        return sum(1 for hh in self.halos if not isinstance(hh, LensPlaneTree))

    def addFieldPlane(self, plane):
        """A LensPlaneTree as a field plane at its (double) redshift; the
        lens sorts it with the main planes (Lens::combinePlanes)."""
        _check(lib.slsim_lens_add_field_plane(self.h, plane.h, float(plane.z)))
        plane.owned_by_lens = True
        self.halos.append(plane)

    def moveinMainHalo(self, halo, addplanes=True):
        """Lens::moveinMainHalo(halo, addplanes) (include/lens.h:172-185):
        addplanes=False puts the halo on the nearest main plane."""
        f = lib.slsim_lens_add_halo if addplanes else lib.slsim_lens_add_halo_nearest
        _check(f(self.h, halo.h, ctypes.c_float(float(halo.z))))
        halo.owned_by_lens = True
        self.halos.append(halo)

    def ResetSourcePlane(self, z, nearest=False):
        """Lens::ResetSourcePlane(z, nearest) (MultiPlane/lens.cpp:2609-2657):
        returns (the reference's plane index, the implanted source redshift).
        z <= 0 turns the implanted source on and changes nothing else, as the
        reference does; RevertSourcePlane goes back to the lens source."""
        j = ctypes.c_int32()
        zi = ctypes.c_double()
        _check(lib.slsim_lens_reset_source_plane(self.h, float(z), int(bool(nearest)),
                                                 ctypes.byref(j), ctypes.byref(zi)))
        self._implant = zi.value
        return j.value, zi.value

    def RevertSourcePlane(self):
        _check(lib.slsim_lens_reset_source(self.h, -1.0))
        self._implant = None

    def getSourceZ(self):
        """Lens::getSourceZ (include/lens.h:107): the implanted source's
        redshift after ResetSourcePlane, else the lens's source plane."""
        return self._implant if self._implant is not None else float(self.tables()["z"][-1])

    def tables(self):
        n = ctypes.c_int()
        _check(lib.slsim_lens_tables(self.h, ctypes.byref(n), None, None, None, None))
        k = n.value + 1
        z, Dl, dDl = (np.zeros(k) for _ in range(3))
        ch = ctypes.c_double()
        _check(lib.slsim_lens_tables(self.h, ctypes.byref(n), _ptr(z, ctypes.c_double),
                                     _ptr(Dl, ctypes.c_double), _ptr(dDl, ctypes.c_double),
                                     ctypes.byref(ch)))
        return {"nplanes": n.value, "z": z, "Dl": Dl, "dDl": dDl, "charge": ch.value}

    def rayshooterInternal(self, x, zs=None, precision=SLSIM_FP64, stats=None):
        """x: (n, 2) image positions (radians) -> (y (n, 2), A (n, 4))."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 2)
        n = x.shape[0]
        y = np.zeros((n, 2))
        A = np.zeros((n, 4))
        zp = None
        if zs is not None:
            zs = np.ascontiguousarray(zs, dtype=np.float64).reshape(-1)
            if zs.size != n:
                raise ValueError("missing redshifts")
            zp = _ptr(zs, ctypes.c_double)
        st = _abi.slsim_shoot_stats()
        _check(lib.slsim_shoot(self.h, ctypes.c_size_t(n), _ptr(x, ctypes.c_double), zp,
                               _ptr(y, ctypes.c_double), _ptr(A, ctypes.c_double),
                               int(precision), ctypes.byref(st)))
        if stats is not None:
            stats.update({k: getattr(st, k) for k, _ in st._fields_ if k != "pad"})
        return y, A

    def shoot_into(self, x, y, A, zs=None, precision=SLSIM_FP64):
        """rayshooterInternal into caller-owned host arrays: x (n, 2) in,
        y (n, 2) and A (n, 4) out, all C-contiguous float64."""
        n = x.shape[0]
        for a, k in ((x, 2), (y, 2), (A, 4)):
            if a.dtype != np.float64 or not a.flags.c_contiguous or a.shape != (n, k):
                raise ValueError("x, y, A must be C-contiguous float64 (n,2), (n,2), (n,4)")
        zp = None
        if zs is not None:
            if zs.dtype != np.float64 or not zs.flags.c_contiguous or zs.size != n:
                raise ValueError("zs must be C-contiguous float64 of n values")
            zp = _ptr(zs, ctypes.c_double)
        _check(lib.slsim_shoot(self.h, ctypes.c_size_t(n), _ptr(x, ctypes.c_double), zp,
                               _ptr(y, ctypes.c_double), _ptr(A, ctypes.c_double),
                               int(precision), None))

    def close(self):
        if self.h:
            if self.ctx.h:
                lib.slsim_lens_destroy(self.h)
            for hl in self.halos:
                hl.h = None
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _read_curves(h):
    try:
        info = _abi.slsim_curves_info()
        _check(lib.slsim_curves_get_info(h, ctypes.byref(info)))
        nc, npt = info.n_curves, info.n_points
        off = np.zeros(nc + 1, np.int64)
        pts = np.zeros((max(npt, 1), 2))
        types = np.zeros(max(nc, 1), np.int32)
        hits = np.zeros(max(nc, 1), np.int32)
        _check(lib.slsim_curves_read(h, _ptr(off, ctypes.c_int64), _ptr(pts, ctypes.c_double),
                                     _ptr(types, ctypes.c_int32), _ptr(hits, ctypes.c_int32)))
        curves = [pts[off[i]:off[i + 1]].copy() for i in range(nc)]
        return curves, hits[:nc].astype(bool), types[:nc], bool(info.hits_replaced)
    finally:
        lib.slsim_curves_destroy(h)


def tightest_hull(ctx, points):
    """Utilities::TightestHull(v) (TreeCode_link/curve_routines.cpp:3519-3592)
    of a closed curve of (n, 2) points, the distance passes on ctx's device:
    (m, 2) points on the curve tracing its outer boundary."""
    p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
    h = _abi.slsim_curves()
    _check(lib.slsim_tightest_hull(ctx.h, _ptr(p, ctypes.c_double), len(p), ctypes.byref(h)))
    c, _, _, _ = _read_curves(h)
    return c[0] if c else np.zeros((0, 2))


def windings(x, points):
    """Utilities::windings(x, points, &area) (TreeCode_link/curve_routines.cpp:
    1709-1736): (winding number about x, absolute area) of a closed curve."""
    p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
    q = np.ascontiguousarray(x, dtype=np.float64).reshape(2)
    wn = ctypes.c_int32()
    area = ctypes.c_double()
    _check(lib.slsim_windings(_ptr(q, ctypes.c_double), _ptr(p, ctypes.c_double), len(p),
                              ctypes.byref(wn), ctypes.byref(area)))
    return wn.value, area.value


class CriticalCurve:
    """ImageFinding::CriticalCurve (include/grid_maintenance.h:245-330): the
    critical curve as rays (x, y, A at z_source), its centre and area, the
    caustic points, centre, outline (TightestHull) and area."""

    def __init__(self):
        self.critcurve_x = np.zeros((0, 2))
        self.critcurve_y = np.zeros((0, 2))
        self.critcurve_A = np.zeros((0, 4))
        self.z_source = 0.0
        self.type = SLSIM_CRIT_ND
        self.touches_edge = False
        self.critical_center = np.zeros(2)
        self.critical_area = 0.0
        self.caustic_curve_intersecting = np.zeros((0, 2))
        self.caustic_center = np.zeros(2)
        self.caustic_curve_outline = np.zeros((0, 2))
        self.caustic_area = 0.0
        self.caustic_intersections = -1


class ImageFinding:
    """The GridMap-based ImageFinding routines (TreeCode_link/find_crit.cpp)."""

    @staticmethod
    def find_crit(lens, gridmap, precision=SLSIM_FP64):
        """ImageFinding::find_crit(lens, gridmap, critcurves)
        (find_crit.cpp:1036-1096): the grid's critical curves (device
        classification), each one's points shot to the lens's source in one
        batch, centres as point means, areas by windings, caustic outlines by
        tightest_hull."""
        zs = lens.getSourceZ()
        curves, hits, types, replaced = gridmap.find_crit()
        out = []
        for i, c in enumerate(curves):
            cc = CriticalCurve()
            cc.type = int(types[i])
            cc.touches_edge = bool(hits[i]) if replaced else False
            cc.z_source = zs
            n = len(c)
            cc.critcurve_x = c
            center = np.zeros(2)
            for j in range(n):            # the reference's running sum
                center += c[j]
            cc.critical_center = center / n
            cc.critical_area = windings(cc.critical_center, c)[1]
            y, A = lens.rayshooterInternal(c, zs=np.full(n, zs), precision=precision)
            cc.critcurve_y, cc.critcurve_A = y, A
            cc.caustic_curve_intersecting = y.copy()
            center = np.zeros(2)
            for j in range(n):
                center += y[j]
            cc.caustic_center = center / n
            cc.caustic_curve_outline = tightest_hull(lens.ctx, y)
            cc.caustic_area = windings(cc.caustic_center, cc.caustic_curve_outline)[1]
            out.append(cc)
        return out

    @staticmethod
    def find_magnification_contour(lens, gridmap, invmag, precision=SLSIM_FP64):
        """ImageFinding::find_magnification_contour(lens, gridmap, invmag,
        contour, hits_boundary) (find_crit.cpp:1098-1119): each contour of
        det A = invmag shot to the lens's source: a list of (x, y, A) and
        hits_boundary."""
        curves, hits, _, _ = gridmap.find_magnification_contour(invmag)
        zs = lens.getSourceZ()
        out = []
        for c in curves:
            y, A = lens.rayshooterInternal(c, zs=np.full(len(c), zs), precision=precision)
            out.append((c, y, A))
        return out, hits


def find_boundaries(bitmap):
    """Utilities::find_boundaries(bitmap, nx, points, hits_edge)
    (include/concave_hull.h:200-494) of a (ny, nx) bitmap: (curves in pixel
    units, hits_edge)."""
    bm = np.ascontiguousarray(bitmap, dtype=np.uint8)
    ny, nx = bm.shape
    h = _abi.slsim_curves()
    _check(lib.slsim_find_boundaries(_ptr(bm, ctypes.c_uint8), nx, ny, ctypes.byref(h)))
    c, hits, _, _ = _read_curves(h)
    return c, hits


def convex_hull(points):
    """Utilities::convex_hull(P, hull) (include/concave_hull.h:560-609):
    hull indices into points, counter-clockwise."""
    p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
    hull = np.zeros(2 * len(p) + 1, np.int64)
    n = ctypes.c_size_t()
    _check(lib.slsim_convex_hull(_ptr(p, ctypes.c_double), len(p), _ptr(hull, ctypes.c_int64),
                                 ctypes.byref(n)))
    return hull[:n.value]


def in_curve(poly, points):
    """Utilities::inCurve(x, H) (include/concave_hull.h:1022-1063) for each
    point."""
    h = np.ascontiguousarray(poly, dtype=np.float64).reshape(-1, 2)
    q = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
    out = np.zeros(max(len(q), 1), np.int32)
    _check(lib.slsim_in_curve(_ptr(h, ctypes.c_double), len(h), _ptr(q, ctypes.c_double), len(q),
                              _ptr(out, ctypes.c_int32)))
    return out[:len(q)].astype(bool)


class Grid:
    """Grid(lens, N1d, center, range) (TreeCode_link/grid_maintenance.cpp:
    22-59), or Grid.rect(lens, Nx, center, rangeX, rangeY) (:67-129): an
    image-plane grid refined cell by cell.  Its points (image and source
    positions, magnification matrix, cell size) stay in device memory in the
    reference's order; point ids index that order.  The reference's kd-trees
    of the points are not built."""

    def __init__(self, lens, N, center, range_, precision=SLSIM_FP64, range_y=None):
        self.lens = lens
        self.ctx = lens.ctx
        self.ctx._adopt(self)
        self.h = _abi.slsim_agrid()
        c = np.ascontiguousarray(center, dtype=np.float64).reshape(2)
        square = range_y is None
        ry = float(range_) if square else float(range_y)
        _check(lib.slsim_agrid_create(lens.h, int(N), _ptr(c, ctypes.c_double), float(range_), ry,
                                      int(square), int(precision), ctypes.byref(self.h)))

    @classmethod
    def rect(cls, lens, Nx, center, range_x, range_y, precision=SLSIM_FP64):
        return cls(lens, Nx, center, range_x, precision, range_y=range_y)

    def info(self):
        i = _abi.slsim_agrid_info()
        _check(lib.slsim_agrid_get_info(self.h, ctypes.byref(i)))
        return {"npoints": i.npoints, "ngrid_init": i.ngrid_init, "ngrid_init2": i.ngrid_init2,
                "ngrid_block": i.ngrid_block, "image_p1": np.array(i.image_p1[:]),
                "image_p2": np.array(i.image_p2[:]), "source_p1": np.array(i.source_p1[:]),
                "source_p2": np.array(i.source_p2[:])}

    def getNumberOfPoints(self):
        return self.info()["npoints"]

    def getInitNgrid(self):
        return self.info()["ngrid_init"]

    def getNgrid_block(self):
        return self.info()["ngrid_block"]

    def getInitRange(self):
        """include/grid_maintenance.h:85: the image box's x width."""
        i = self.info()
        return i["image_p2"][0] - i["image_p1"][0]

    def getInitCenter(self):
        """TreeCode_link/image_finder.cpp:826-831: the image box's centre."""
        i = self.info()
        return (i["image_p2"] + i["image_p1"]) / 2

    def points(self, first=0, count=None):
        """(x (n, 2), y (n, 2), A (n, 4), gridsize (n,)) of points first ..."""
        if count is None:
            count = self.getNumberOfPoints() - first
        x, y = np.zeros((count, 2)), np.zeros((count, 2))
        A, gs = np.zeros((count, 4)), np.zeros(count)
        _check(lib.slsim_agrid_read(self.h, int(first), int(count), _ptr(x, ctypes.c_double),
                                    _ptr(y, ctypes.c_double), _ptr(A, ctypes.c_double),
                                    _ptr(gs, ctypes.c_double)))
        return x, y, A, gs

    def RefineLeaves(self, ids):
        """Grid::RefineLeaves (grid_maintenance.cpp:681-982): the ids of the
        new points."""
        ids = np.ascontiguousarray(ids, dtype=np.int64).reshape(-1)
        first, count = ctypes.c_int64(), ctypes.c_int64()
        _check(lib.slsim_agrid_refine(self.h, ids.size, _ptr(ids, ctypes.c_int64),
                                      ctypes.byref(first), ctypes.byref(count)))
        return np.arange(first.value, first.value + count.value)

    def RefineLeaf(self, i):
        """Grid::RefineLeaf (grid_maintenance.cpp:543-679)."""
        first, count = ctypes.c_int64(), ctypes.c_int64()
        _check(lib.slsim_agrid_refine_one(self.h, int(i), ctypes.byref(first),
                                          ctypes.byref(count)))
        return np.arange(first.value, first.value + count.value)

    def set_server(self, server):
        """Each later refinement's rays as one ray-list request of `server`
        (continuous batching with other clients), or None: the same points."""
        self._server = server
        _check(lib.slsim_agrid_set_server(self.h, server.h if server is not None else None))

    def EinsteinArea(self):
        """Grid::EinsteinArea (grid_maintenance.cpp:368-378)."""
        a = ctypes.c_double()
        _check(lib.slsim_agrid_einstein_area(self.h, ctypes.byref(a)))
        return a.value

    def close(self):
        if self.h:
            if self.ctx.h:
                lib.slsim_agrid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GridMap:
    """GridMap(lens, N, center, range) (TreeCode_link/gridmap.cpp:93-120):
    shoots the N x N grid on the device; outputs stay in HBM until read."""

    def __init__(self, lens, N, center, range_, precision=SLSIM_FP64, zs=0.0, shoot=True):
        self.lens = lens
        self.ctx = lens.ctx
        self.ctx._adopt(self)
        self.N = int(N)
        self.center = (float(center[0]), float(center[1]))
        self.range = float(range_)
        self.precision = precision
        self.zs = float(zs)
        self.ny = self.N
        self.range_y = self.range
        self.layout = _abi.SLSIM_GRID_SQUARE
        n = self.N * self.N
        self.y_dev = self.ctx.alloc(16 * n)
        self.A_dev = self.ctx.alloc(32 * n)
        self.stats = {}
        if shoot:
            self.shoot()

    @classmethod
    def rect(cls, lens, Nx, center, range_x, range_y, precision=SLSIM_FP64):
        """GridMap(lens, Nx, center, rangeX, rangeY) (gridmap.cpp:26-91):
        ny = int(Nx rangeY / rangeX) rows of Nx points, square cells of
        rangeX / (Nx - 1) from the lower-left corner; the points are shot as
        a ray list and the outputs stay in HBM."""
        self = cls.__new__(cls)
        self.lens, self.ctx = lens, lens.ctx
        self.ctx._adopt(self)
        self.N = int(Nx)
        self.center = (float(center[0]), float(center[1]))
        self.range, self.range_y = float(range_x), float(range_y)
        self.precision, self.zs = precision, 0.0
        self.ny = int(self.N * (self.range_y / self.range))
        self.layout = _abi.SLSIM_GRID_RECT
        n = self.N * self.ny
        self.y_dev = self.ctx.alloc(16 * max(n, 1))
        self.A_dev = self.ctx.alloc(32 * max(n, 1))
        self.stats = {}
        x = np.ascontiguousarray(self.positions())
        xd = self.ctx.alloc(16 * max(n, 1))
        try:
            self.ctx.h2d(xd, x)
            st = _abi.slsim_shoot_stats()
            _check(lib.slsim_shoot_device(lens.h, ctypes.c_size_t(n),
                                          ctypes.cast(xd, ctypes.POINTER(ctypes.c_double)), None,
                                          ctypes.cast(self.y_dev, ctypes.POINTER(ctypes.c_double)),
                                          ctypes.cast(self.A_dev, ctypes.POINTER(ctypes.c_double)),
                                          int(precision), ctypes.byref(st)))
        finally:
            self.ctx.free(xd)
        return self

    def _grid(self):
        g = _abi.slsim_grid()
        g.nx, g.ny = self.N, self.ny
        g.center[0], g.center[1] = self.center
        g.range_x, g.range_y = self.range, self.range_y
        g.layout = self.layout
        return g

    def shoot(self):
        st = _abi.slsim_shoot_stats()
        _check(lib.slsim_gridmap_shoot(self.lens.h, ctypes.c_size_t(self.N), self.center[0],
                                       self.center[1], self.range, self.zs,
                                       ctypes.cast(self.y_dev, ctypes.POINTER(ctypes.c_double)),
                                       ctypes.cast(self.A_dev, ctypes.POINTER(ctypes.c_double)),
                                       int(self.precision), ctypes.byref(st)))
        self.stats = {k: getattr(st, k) for k, _ in st._fields_ if k != "pad"}
        return self.stats

    def positions(self):
        """Image positions of the grid (x fastest): PositionFromIndex, or the
        rectangular constructor's lattice."""
        N = self.N
        if self.layout == _abi.SLSIM_GRID_RECT:
            k = np.arange(N * self.ny)
            res = self.range / (N - 1)
            return np.stack([(self.center[0] - 0.5 * self.range) + (k % N) * res,
                             (self.center[1] - 0.5 * self.range_y) + (k // N) * res], 1)
        i = np.arange(N * N)
        if N == 1:
            return np.array([[self.center[0], self.center[1]]])
        x0 = self.center[0] + self.range * (1.0 * (i % N) / (N - 1) - 0.5)
        x1 = self.center[1] + self.range * (1.0 * (i // N) / (N - 1) - 0.5)
        return np.stack([x0, x1], axis=1)

    def y(self):
        out = np.zeros((self.N * self.ny, 2))
        self.ctx.d2h(out, self.y_dev)
        return out

    def A(self):
        out = np.zeros((self.N * self.ny, 4))
        self.ctx.d2h(out, self.A_dev)
        return out

    def counts(self):
        n = self.N * self.N
        out = np.zeros((n, 5), np.uint64)
        _check(lib.slsim_gridmap_count(self.lens.h, ctypes.c_size_t(self.N), self.center[0],
                                       self.center[1], self.range, self.zs,
                                       _ptr(out, ctypes.c_uint64)))
        return out

    def bundle_stats(self):
        """Lane utilisation of the 64-ray bundles (visits, leaf particles)."""
        out = np.zeros(4, np.uint64)
        _check(lib.slsim_gridmap_bundle_stats(self.lens.h, ctypes.c_size_t(self.N), self.center[0],
                                              self.center[1], self.range, self.zs,
                                              _ptr(out, ctypes.c_uint64)))
        v, wv, p, wp = (float(t) for t in out)
        return {"visits": v, "wave_visits": wv, "visit_utilisation": v / max(64 * wv, 1),
                "leaf_particles": p, "wave_leaf_iterations": wp,
                "leaf_utilisation": p / max(64 * wp, 1)}

    def writePixelMap(self, var):
        """LensingVariable maps of gridmap.h:533-603 as (N, N) arrays."""
        A = self.A()
        y = self.y()
        x = self.positions()
        v = var.upper()
        if v == "ALPHA1":
            r = x[:, 0] - y[:, 0]
        elif v == "ALPHA2":
            r = x[:, 1] - y[:, 1]
        elif v == "ALPHA":
            r = np.hypot(x[:, 0] - y[:, 0], x[:, 1] - y[:, 1])
        elif v == "KAPPA":
            r = 1 - (A[:, 0] + A[:, 3]) / 2
        elif v == "GAMMA1":
            r = (A[:, 0] - A[:, 3]) / 2
        elif v == "GAMMA2":
            r = (A[:, 1] + A[:, 2]) / 2
        elif v == "GAMMA3":
            r = (A[:, 2] - A[:, 1]) / 2
        elif v == "GAMMA":
            r = np.hypot((A[:, 0] - A[:, 3]) / 2, (A[:, 1] + A[:, 2]) / 2)
        elif v == "INVMAG":
            r = A[:, 0] * A[:, 3] - A[:, 1] * A[:, 2]
        elif v == "DELAYT":
            r = np.zeros(self.N * self.ny)
        else:
            raise ValueError("GridMap::writePixelMap() does not work for the input LensingVariable")
        return r.reshape(self.ny, self.N)

    def find_images(self, ys, max_per_source=32):
        """GridMap::find_images(ys, multiplicity) (gridmap.cpp:927-1032,
        1326-1348) on the device over this grid's outputs: (rays (n, 8) =
        image x, source y, A the mean over the image's triangle;
        multiplicity per source; triangles (n, 3) grid indices), sources in
        order, each source's images in the reference's order."""
        ys = np.ascontiguousarray(ys, dtype=np.float64).reshape(-1, 2)
        ns = ys.shape[0]
        cap = int(max_per_source)
        mult = np.zeros(ns, np.int32)
        rays = np.zeros((max(ns * cap, 1), 8))
        tris = np.zeros((max(ns * cap, 1), 3), np.int64)
        n = ctypes.c_size_t()
        g = self._grid()
        _check(lib.slsim_grid_find_images(
            self.ctx.h, ctypes.byref(g),
            self.y_dev, self.A_dev, ctypes.c_size_t(ns), _ptr(ys, ctypes.c_double),
            ctypes.c_size_t(cap), mult.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            _ptr(rays, ctypes.c_double), tris.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            ctypes.byref(n)))
        return rays[:n.value], mult, tris[:n.value]

    def find_crit(self):
        """GridMap::find_crit(points, hits_boundary, crit_type)
        (gridmap.cpp:525-647) on the device over this grid's A: (curves, a
        list of (n, 2) arrays in the grid's units; hits_boundary (bool);
        crit types SLSIM_CRIT_*; whether hits_boundary was replaced, False
        when no pixel is critical)."""
        h = _abi.slsim_curves()
        g = self._grid()
        _check(lib.slsim_grid_find_crit(self.ctx.h, ctypes.byref(g), self.y_dev, self.A_dev,
                                        ctypes.byref(h)))
        return _read_curves(h)

    def find_magnification_contour(self, invmag):
        """GridMap::find_magnification_contour(curves, hits_boundary, invmag)
        (gridmap.cpp:493-522): (curves, hits_boundary, -, replaced)."""
        h = _abi.slsim_curves()
        g = self._grid()
        _check(lib.slsim_grid_find_magnification_contour(self.ctx.h, ctypes.byref(g), self.A_dev,
                                                         float(invmag), ctypes.byref(h)))
        return _read_curves(h)

    def find_boundaries_of_caustics(self):
        """GridMap::find_boundaries_of_caustics(boundaries, hits_edge)
        (gridmap.h:278-296): (boundaries in pixel units, hits_edge, the
        multiplicity of every grid position taken as a source)."""
        h = _abi.slsim_curves()
        mult = np.zeros(self.N * self.ny, np.int32)
        g = self._grid()
        _check(lib.slsim_grid_find_boundaries_of_caustics(self.ctx.h, ctypes.byref(g), self.y_dev,
                                                          _ptr(mult, ctypes.c_int32),
                                                          ctypes.byref(h)))
        c, hits, _, _ = _read_curves(h)
        return c, hits, mult

    def writeFits(self, var, filename, dtype=np.float64):
        """GridMap::writeFits<T>(lensvar, filename) (gridmap.h:524-531): the
        map of writePixelMap written by PixelMap::printFITS, resolution
        range / (N - 1)."""
        write_fits_pixelmap(filename, np.asarray(self.writePixelMap(var), dtype),
                            self.range / (self.N - 1), self.center)

    def close(self):
        if self.y_dev:
            if self.ctx.h:
                self.ctx.free(self.y_dev)
                self.ctx.free(self.A_dev)
            self.y_dev = self.A_dev = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many GridMap requests shot together (one launch per plane step).

    requests: list of (lens, N, (c0, c1), range, zs) with zs <= 0 for the
    lens source."""

    def __init__(self, ctx, requests, precision=SLSIM_FP64):
        self.ctx = ctx
        ctx._adopt(self)
        self.requests = list(requests)
        n = len(self.requests)
        lenses = (_abi.slsim_lens * n)(*[r[0].h for r in self.requests])
        N = np.array([r[1] for r in self.requests], np.int64)
        cen = np.array([c for r in self.requests for c in r[2]], np.float64)
        rg = np.array([r[3] for r in self.requests], np.float64)
        zs = np.array([r[4] for r in self.requests], np.float64)
        self._keep = (lenses, N, cen, rg, zs)
        self.h = _abi.slsim_batch()
        _check(lib.slsim_batch_create(ctx.h, n, lenses, _ptr(N, ctypes.c_int64),
                                      _ptr(cen, ctypes.c_double), _ptr(rg, ctypes.c_double),
                                      _ptr(zs, ctypes.c_double), int(precision),
                                      ctypes.byref(self.h)))

    def set_schedule(self, tile_log2=3, queue=0, order=0, graph=0, blocks_per_cu=4, fuse=0):
        s = _abi.slsim_schedule(tile_log2, queue, order, graph, blocks_per_cu, fuse)
        _check(lib.slsim_batch_set_schedule(self.h, ctypes.byref(s)))

    def schedule(self):
        s = _abi.slsim_schedule()
        _check(lib.slsim_batch_get_schedule(self.h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_ if k != "pad"}

    def run(self):
        st = _abi.slsim_shoot_stats()
        _check(lib.slsim_batch_run(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_ if k != "pad"}

    def autotune(self, repeats=1):
        s = _abi.slsim_schedule()
        b, t = ctypes.c_double(), ctypes.c_double()
        _check(lib.slsim_batch_autotune(self.h, int(repeats), ctypes.byref(s), ctypes.byref(b),
                                        ctypes.byref(t)))
        return {k: getattr(s, k) for k, _ in s._fields_ if k != "pad"}, b.value, t.value

    def output(self, i):
        y = ctypes.POINTER(ctypes.c_double)()
        A = ctypes.POINTER(ctypes.c_double)()
        _check(lib.slsim_batch_output(self.h, int(i), ctypes.byref(y), ctypes.byref(A)))
        N = self.requests[i][1]
        yh = np.zeros((N * N, 2))
        Ah = np.zeros((N * N, 4))
        self.ctx.d2h(yh, ctypes.cast(y, ctypes.c_void_p))
        self.ctx.d2h(Ah, ctypes.cast(A, ctypes.c_void_p))
        return yh, Ah

    def close(self):
        if self.h:
            if self.ctx.h:
                lib.slsim_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop with continuous batching over resident lenses."""

    def __init__(self, ctx, hbm_budget=0):
        self.ctx = ctx
        ctx._adopt(self)
        self.h = _abi.slsim_server()
        _check(lib.slsim_server_create(ctx.h, ctypes.c_size_t(int(hbm_budget)),
                                       ctypes.byref(self.h)))

    def submit(self, lens, N, center, range_, zs=0.0, precision=SLSIM_FP32C):
        rid = ctypes.c_uint64()
        _check(lib.slsim_server_submit(self.h, lens.h, int(N), float(center[0]), float(center[1]),
                                       float(range_), float(zs), int(precision),
                                       ctypes.byref(rid)))
        return rid.value

    def submit_rays(self, lens, x, zs=None, precision=SLSIM_FP32C):
        """rayshooterInternal(Npoints, Point*) as a served request: rays at
        x (n, 2), per-ray source redshifts zs (n,) or the lens's source.
        Collect with wait_rays(rid, n)."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 2)
        zp = None
        if zs is not None:
            zs = np.ascontiguousarray(zs, dtype=np.float64)
            zp = _ptr(zs, ctypes.c_double)
        rid = ctypes.c_uint64()
        _check(lib.slsim_server_submit_rays(self.h, lens.h, ctypes.c_size_t(x.shape[0]),
                                            _ptr(x, ctypes.c_double), zp, int(precision),
                                            ctypes.byref(rid)))
        return rid.value

    def wait_rays(self, rid, n):
        info = _abi.slsim_request_info()
        y = np.zeros((n, 2))
        A = np.zeros((n, 4))
        _check(lib.slsim_server_wait(self.h, ctypes.c_uint64(rid), _ptr(y, ctypes.c_double),
                                     _ptr(A, ctypes.c_double), ctypes.byref(info)))
        return y, A, {k: getattr(info, k) for k, _ in info._fields_}

    def poll(self, rid):
        """True once request rid has finished (non-blocking)."""
        done = ctypes.c_int(0)
        _check(lib.slsim_server_poll(self.h, ctypes.c_uint64(rid), ctypes.byref(done)))
        return bool(done.value)

    def wait(self, rid, N=None, fetch=True):
        info = _abi.slsim_request_info()
        y = A = None
        if fetch:
            y = np.zeros((N * N, 2))
            A = np.zeros((N * N, 4))
            rc = lib.slsim_server_wait(self.h, ctypes.c_uint64(rid), _ptr(y, ctypes.c_double),
                                       _ptr(A, ctypes.c_double), ctypes.byref(info))
        else:
            rc = lib.slsim_server_wait(self.h, ctypes.c_uint64(rid), None, None,
                                       ctypes.byref(info))
        _check(rc)
        return y, A, {k: getattr(info, k) for k, _ in info._fields_}

    def stats(self):
        st = _abi.slsim_server_stats()
        _check(lib.slsim_server_get_stats(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if self.h:
            if self.ctx.h:
                lib.slsim_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MakeParticleLenses:
    """MakeParticleLenses (particle_halo.h:964-1080, MultiPlane/
    particle_lens.cpp): the particles of a simulation file grouped by type.

    MakeParticleLenses(ctx, filename, format, Nsmooth, recenter,
    ignore_type_in_smoothing) reads csv3/csv4/csv5, Gadget-2 or glmb files
    (smoothing sizes on the device, cached in "<filename>_S<Nsmooth>.glmb");
    MakeParticleLenses(ctx, filename, recenter=...) with format=None opens a
    glmb file.  ctx may be None when no sizes have to be computed."""

    def __init__(self, ctx, filename, format=None, Nsmooth=64, recenter=False,
                 ignore_type_in_smoothing=False):
        self.ctx = ctx
        self.h = ctypes.c_void_p()
        ch = ctx.h if ctx is not None else None
        fn = os.fsencode(filename)
        if format is None:
            _check(lib.slsim_mpl_create_glmb(ch, fn, int(bool(recenter)), ctypes.byref(self.h)))
        else:
            _check(lib.slsim_mpl_create(ch, fn, int(format), int(Nsmooth), int(bool(recenter)),
                                        int(bool(ignore_type_in_smoothing)),
                                        ctypes.byref(self.h)))
        self.halos = []

    @property
    def data(self):
        """The library's particle array as a structured numpy view
        (PARTICLE_DTYPE), valid until the particle count changes."""
        p = ctypes.POINTER(_abi.slsim_particle_f32)()
        n = ctypes.c_size_t()
        _check(lib.slsim_mpl_data(self.h, ctypes.byref(p), ctypes.byref(n)))
        if n.value == 0:
            return np.zeros(0, PARTICLE_DTYPE)
        buf = (ctypes.c_char * (24 * n.value)).from_address(ctypes.addressof(p.contents))
        return np.frombuffer(buf, PARTICLE_DTYPE)

    def _info(self):
        i = _abi.slsim_mpl_info()
        _check(lib.slsim_mpl_get_info(self.h, ctypes.byref(i)))
        return i

    def getnp(self):
        return [int(v) for v in self._info().nparticles]

    def getmp(self):
        return [float(v) for v in self._info().masses]

    def sim_redshift(self):
        return self._info().z_original

    getZoriginal = sim_redshift

    def getCenterOfMass(self):
        return tuple(self._info().cm)

    def getBoundingBox(self):
        i = self._info()
        return tuple(i.bbox_ll), tuple(i.bbox_ur)

    def nsmooth(self):
        return self._info().nsmooth

    def Recenter(self, x):
        c = np.asarray(x, np.float64)
        _check(lib.slsim_mpl_recenter(self.h, _ptr(c, ctypes.c_double)))
        self.halos = []

    def radialCut(self, center, radius):
        c = np.asarray(center, np.float64)
        _check(lib.slsim_mpl_radial_cut(self.h, _ptr(c, ctypes.c_double), float(radius)))

    def cylindricalCut(self, center, radius):
        c = np.asarray(center, np.float64)
        _check(lib.slsim_mpl_cylindrical_cut(self.h, _ptr(c, ctypes.c_double), float(radius)))

    def densest_particle(self):
        x = np.zeros(3, np.float64)
        _check(lib.slsim_mpl_densest_particle(self.h, _ptr(x, ctypes.c_double)))
        return tuple(x)

    def CreateHalos(self, hubble, redshift, inv_area):
        """CreateHalos(cosmo, redshift, inv_area) (particle_lens.cpp:153-209):
        units times 1/h, then one LensHaloParticles (theta 0.1, bucket 20,
        no recentring or rotation) per particle type with particles."""
        _check(lib.slsim_mpl_physical_units(self.h, ctypes.c_float(hubble)))
        d = self.data
        self.halos = []
        skip = 0
        for n in self.getnp():
            if n > 0:
                part = d[skip:skip + n]
                self.halos.append(LensHaloParticles(
                    self.ctx, np.ascontiguousarray(part["x"][:, 0]),
                    np.ascontiguousarray(part["x"][:, 1]), part["Mass"], part["Size"],
                    redshift, theta=0.1, bucket=20, inv_area=inv_area))
            skip += n
        return self.halos

    def close(self):
        if self.h:
            lib.slsim_mpl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def glmb_write(filename, particles, nsmooth, numbytype, redshift):
    """MakeParticleLenses::writeSizesB (particle_halo.h:1037-1054)."""
    p = np.ascontiguousarray(particles, PARTICLE_DTYPE)
    nb = np.asarray(numbytype, np.uint64).astype(np.uintp)
    _check(lib.slsim_glmb_write(os.fsencode(filename), p.ctypes.data_as(ctypes.c_void_p),
                                ctypes.c_size_t(p.size), int(nsmooth),
                                nb.ctypes.data_as(ctypes.POINTER(ctypes.c_size_t)),
                                float(redshift)))


def glmb_read(filename):
    """MakeParticleLenses::readSizesB (particle_halo.h:1056-1073):
    (particles, nsmooth, numbytype, redshift)."""
    ns = ctypes.c_int()
    n = ctypes.c_size_t()
    nb = (ctypes.c_size_t * 6)()
    z = ctypes.c_double()
    fn = os.fsencode(filename)
    _check(lib.slsim_glmb_read_header(fn, ctypes.byref(ns), ctypes.byref(n), nb, ctypes.byref(z)))
    p = np.zeros(n.value, PARTICLE_DTYPE)
    _check(lib.slsim_glmb_read(fn, p.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(n.value)))
    return p, ns.value, [int(v) for v in nb], z.value


def gadget2_read(filename):
    """GadgetFile<ParticleType<float>> header + readBlock("POS"/"MASS")
    (include/gadget.hh): (header dict, particles in file units)."""
    npart = (ctypes.c_int32 * 6)()
    mt = (ctypes.c_double * 6)()
    v = [ctypes.c_double() for _ in range(5)]
    fn = os.fsencode(filename)
    _check(lib.slsim_gadget2_header(fn, npart, mt, *[ctypes.byref(x) for x in v]))
    n = sum(npart)
    p = np.zeros(n, PARTICLE_DTYPE)
    _check(lib.slsim_gadget2_read(fn, p.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(n)))
    hd = {"npart": list(npart), "masstab": list(mt), "time": v[0].value, "redshift": v[1].value,
          "omega": v[2].value, "lambda": v[3].value, "hubble": v[4].value}
    return hd, p


def write_fits_pixelmap(filename, image, resolution, center, ra=0.0, dec=0.0, units=0,
                        flipX=False, ctype=False):
    """PixelMap<T>::printFITS(filename, flipX, ctype) (ImageProcessing/
    pixelmap.cpp:1318-1388) for an (Ny, Nx) float64 or float32 image."""
    im = np.ascontiguousarray(image)
    if im.dtype not in (np.float64, np.float32) or im.ndim != 2:
        raise ValueError("image must be a 2-D float64 or float32 array")
    c = np.asarray(center, np.float64)
    _check(lib.slsim_fits_write_pixelmap(os.fsencode(filename), im.ctypes.data_as(ctypes.c_void_p),
                                         int(im.dtype == np.float64), ctypes.c_size_t(im.shape[1]),
                                         ctypes.c_size_t(im.shape[0]), float(resolution),
                                         _ptr(c, ctypes.c_double), float(ra), float(dec),
                                         int(units), int(bool(flipX)), int(bool(ctype))))


def now():
    return lib.slsim_now()


def coorDist(z1, z2, Om=0.3, OL=0.7, h=0.7):
    return lib.slsim_coordist(float(Om), float(OL), float(h), float(z1), float(z2))


def invCoorDist(D, Om=0.3, OL=0.7, h=0.7):
    return lib.slsim_inv_coordist(float(Om), float(OL), float(h), float(D))


def nfw_r200(mass, z, Om=0.3, OL=0.7, h=0.7):
    return lib.slsim_nfw_r200(float(mass), float(z), float(Om), float(OL), float(h))


def gen_nfw(seed, n, mass, conc, z, Om=0.3, OL=0.7, h=0.7, center=(0.0, 0.0)):
    x = np.zeros(n)
    y = np.zeros(n)
    _check(lib.slsim_gen_nfw(ctypes.c_uint64(seed), ctypes.c_size_t(n), float(mass), float(conc),
                             float(z), float(Om), float(OL), float(h), float(center[0]),
                             float(center[1]), _ptr(x, ctypes.c_double), _ptr(y, ctypes.c_double)))
    return x, y


def gen_halo_field(seed, n, z, n_halos=50, Mmin=1e11, Mmax=1e14, field_arcsec=300.0,
                   softening=5e-4, Om=0.3, OL=0.7, h=0.7):
    x = np.zeros(n)
    y = np.zeros(n)
    m = np.zeros(n, np.float32)
    s = np.zeros(n, np.float32)
    _check(lib.slsim_gen_halo_field(ctypes.c_uint64(seed), ctypes.c_size_t(n), int(n_halos),
                                    float(Mmin), float(Mmax), float(field_arcsec), float(z),
                                    float(Om), float(OL), float(h), ctypes.c_float(softening),
                                    _ptr(x, ctypes.c_double), _ptr(y, ctypes.c_double),
                                    _ptr(m, ctypes.c_float), _ptr(s, ctypes.c_float)))
    return x, y, m, s
