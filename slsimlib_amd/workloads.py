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


def c4_specs(n_real, seed0=4000, n_lo=1e4, n_hi=1e6, sides=(64, 128, 256, 512, 1024),
             max_planes=12):
    """Batched form: n_real independent realizations, each with 1..max_planes
    planes evenly in D_C over [0.1, min(2, zs - 0.1)], log-uniform particle
    counts in [n_lo, n_hi] per plane (NFW-field generator, float32
    positions), a grid side from `sides` and zs uniform in [1, 4].  Returns
    one dict per realization (planes' z and n, side, zs)."""
    rng = np.random.default_rng(seed0)
    specs = []
    for r in range(n_real):
        k = int(rng.integers(1, max_planes + 1))
        zs = float(rng.uniform(1.0, 4.0))
        side = int(rng.choice(list(sides)))
        zl = planes_evenly_in_Dc(0.1, min(2.0, zs - 0.1), k)
        ns = [int(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi)))) for _ in range(k)]
        specs.append({"realization": r, "z": zl, "n": ns, "side": side, "zs": zs})
    return specs


def build_c4_lens(ctx, spec, theta=0.1, mass_scale=1.0, keep_particles=False):
    """The lens of one batched-form realization (config 4 seeds)."""
    lens = Lens(ctx, spec["zs"], OM, OL, H)
    parts = []
    for j, (z, n) in enumerate(zip(spec["z"], spec["n"])):
        x, y, m, h = field_plane(4, j, z, n, soft=5e-4, realization=spec["realization"],
                                 dtype=np.float32, mass_scale=mass_scale)
        lens.moveinMainHalo(LensHaloParticles(ctx, x, y, m, h, z, theta=theta))
        if keep_particles:
            parts.append((x, y, m, h, np.float32(z)))
    return lens, parts


def c5_stream(n_requests, n_models, seed=5999, sides=(128, 256, 512, 1024, 2048),
              probs=(0.4, 0.3, 0.15, 0.1, 0.05)):
    """Serving form's request mix: per request a lens id in [0, n_models),
    zs uniform in [1, 4], a centre uniform in the 5' field (radians) and a
    grid side from `sides` with probabilities `probs`; the grid resolution is
    fixed at 300"/2048 per pixel.  Returns (lens ids, zs, centres (n, 2),
    sides, resolution in radians)."""
    rng = np.random.default_rng(seed)
    side = rng.choice(list(sides), size=n_requests, p=list(probs))
    lid = rng.integers(0, n_models, n_requests)
    zs = rng.uniform(1.0, 4.0, n_requests)
    cen = (rng.random((n_requests, 2)) - 0.5) * 300 * arcsecTOradians
    return lid, zs, cen, side.astype(np.int64), 300.0 / 2048 * arcsecTOradians
