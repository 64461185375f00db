"""Synthetic workloads of the benchmark configurations.

Seeds follow seed = 1000 * config + 16 * realization + plane.  Cosmology is
flat LCDM (Om=0.3, OL=0.7, h=0.7); plane redshifts are stored as float32 as
the reference's LensHaloParticles constructor does (particle_halo.h:59).
"""
import numpy as np

from . import (Context, GridMap, Lens, LensHaloParticles, SLSIM_FP64, arcsecTOradians,
               coorDist, gen_halo_field, gen_nfw, invCoorDist)

OM, OL, H = 0.3, 0.7, 0.7


def seed(config, realization=0, plane=0):
    return 1000 * config + 16 * realization + plane


def c1_particles(n=65536, mass=1e13, conc=8.0, z=0.5, soft=1e-4):
    """Config 1: one projected NFW halo of n equal-mass particles at z=0.5."""
    x, y = gen_nfw(seed(1), n, mass, conc, z, OM, OL, H)
    m = np.full(n, np.float32(mass / n), np.float32)
    h = np.full(n, np.float32(soft), np.float32)
    return x, y, m, h


C1 = {"z_lens": 0.5, "z_source": 2.0, "N": 256, "range_arcsec": 30.0, "theta": 0.1}


def planes_evenly_in_Dc(z_lo, z_hi, k):
    d0, d1 = coorDist(0, z_lo, OM, OL, H), coorDist(0, z_hi, OM, OL, H)
    if k == 1:
        return [z_lo]
    return [invCoorDist(d0 + (d1 - d0) * j / (k - 1), OM, OL, H) for j in range(k)]


def field_plane(config, j, z, n, soft=5e-4, realization=0, dtype=np.float64, mass_scale=1.0):
    x, y, m, h = gen_halo_field(seed(config, realization, j), n, z, n_halos=50,
                                Mmin=1e11 * mass_scale, Mmax=1e14 * mass_scale,
                                field_arcsec=300.0, softening=soft, Om=OM, OL=OL, h=H)
    if dtype == np.float32:
        x, y = x.astype(np.float32), y.astype(np.float32)
    return x, y, m, h


C2 = {"nplanes": 10, "z_lo": 0.1, "z_hi": 2.0, "z_source": 3.0, "n": 1 << 20, "N": 2048,
      "range_arcsec": 300.0, "theta": 0.1, "soft": 5e-4}


def build_c1_lens(ctx, theta=0.1, n=65536):
    x, y, m, h = c1_particles(n)
    lens = Lens(ctx, C1["z_source"], OM, OL, H)
    halo = LensHaloParticles(ctx, x, y, m, h, C1["z_lens"], theta=theta)
    lens.moveinMainHalo(halo)
    return lens, (x, y, m, h)


def build_field_lens(ctx, config=2, nplanes=10, n=1 << 20, z_lo=0.1, z_hi=2.0, zs=3.0,
                     theta=0.1, soft=5e-4, dtype=np.float64, keep_particles=False,
                     mass_scale=1.0, realization=0):
    """Multi-plane field lens; mass_scale < 1 with fewer particles keeps the
    particle mass of the full-size configuration."""
    zl = planes_evenly_in_Dc(z_lo, z_hi, nplanes)
    lens = Lens(ctx, zs, OM, OL, H)
    parts = []
    for j, z in enumerate(zl):
        x, y, m, h = field_plane(config, j, z, n, soft=soft, dtype=dtype, mass_scale=mass_scale,
                                 realization=realization)
        halo = LensHaloParticles(ctx, x, y, m, h, z, theta=theta)
        lens.moveinMainHalo(halo)
        if keep_particles:
            parts.append((x, y, m, h, np.float32(z)))
    return lens, parts, zl


def grid_range(arcsec):
    return arcsec * arcsecTOradians


def particles_3d(seed_, n, n_halos=20, conc=8.0, rmax=1.0, dtype=np.float64):
    """3-D positions (Mpc) for the smoothing-size workload: n particles in
    n_halos NFW spheres (scale radius rmax / conc, truncated at rmax, masses
    log-uniform over two decades, centres uniform in a 10 Mpc box), split in
    proportion to the halo masses.  Seeded numpy generator."""
    rng = np.random.default_rng(seed_)
    w = 10.0 ** rng.uniform(0, 2, n_halos)
    counts = np.floor(n * w / w.sum()).astype(np.int64)
    counts[: n - counts.sum()] += 1
    # inverse CDF of the truncated NFW mass profile m(x) = ln(1+x) - x/(1+x)
    xs = np.linspace(0.0, conc, 4097)
    mx = np.log1p(xs) - xs / (1 + xs)
    parts = []
    for c in counts:
        centre = rng.uniform(-5.0, 5.0, 3)
        u = rng.uniform(0, mx[-1], c)
        r = np.interp(u, mx, xs) * (rmax / conc)
        v = rng.normal(size=(c, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        parts.append(centre + v * r[:, None])
    p = np.concatenate(parts)[rng.permutation(n)]
    return tuple(np.ascontiguousarray(p[:, j], dtype=dtype) for j in range(3))
