#!/usr/bin/env python3
"""Benchmark of the multi-plane particle ray shooter on one MI355X.

Workload (the configuration the headline metric is quoted on): 10 lens planes
evenly spaced in comoving distance between z=0.1 and z=2.0, source at z=3,
each plane 2^20 equal-mass particles from 50 projected NFW halos (masses
log-uniform in 1e11-1e14 Msun) over a 5-arcmin field, softening 0.5 kpc,
opening angle 0.1, bucket 20; a GridMap of 2048 x 2048 rays over 300 arcsec.

A step is one GridMap shoot of all 4,194,304 rays through the 10 planes on
the FP32-compensated walk (plane recursion FP64), with the trees already
resident in HBM and the outputs (y, A) written to HBM.  Also reported: the
plain FP64 path on the same GPU, the error of both against each other and
against the FP64 direct sum on a 65,536-ray subsample, the roofline of the
walk kernel, and the CPU oracle (the reference algorithm restated in C++,
std::thread over contiguous ray chunks) timed on this host.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3     # MI355X FP32 vector, AMD published
PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 vector, AMD published
# algorithmic flops per unit of the reference's per-ray criterion
FLOPS = {"visit": 8, "far": 49, "leafp": 25, "bigtest": 7, "bspline": 55, "plane_step": 46}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_err(d, ref):
    d = np.asarray(d).reshape(len(d), -1)
    ref = np.asarray(ref).reshape(len(ref), -1)
    nr = np.linalg.norm(ref, axis=1)
    floor = 1e-3 * np.sqrt(np.mean(nr ** 2))
    e = np.linalg.norm(d, axis=1) / np.maximum(nr, floor if floor > 0 else 1e-300)
    return float(e.max()), float(np.sqrt(np.mean(e ** 2))), float(np.quantile(e, 0.999))


def fields(x, y, A):
    return (x - y, 1 - 0.5 * (A[:, 0] + A[:, 3]),
            np.stack([0.5 * (A[:, 0] - A[:, 3]), 0.5 * (A[:, 1] + A[:, 2])], 1))


def compare(x, y, A, yr, Ar, xr=None):
    """Relative differences per ray (|d| / max(|ref|, 1e-3 rms|ref|)): max, rms
    and 99.9th percentile; xr = the reference rays' image positions if they differ."""
    a, k, g = fields(x, y, A)
    ar, kr, gr = fields(x if xr is None else xr, yr, Ar)
    out = {}
    for name, d, r in (("alpha", a - ar, ar), ("kappa", k - kr, kr), ("gamma", g - gr, gr)):
        mx, rms, p999 = rel_err(d, r)
        out.update({name + "_max": mx, name + "_rms": rms, name + "_p999": p999})
    return out


def c4_requests(S, W, ctx, n_real, seed0=4000):
    """Batched form (one realization per request): W.c4_specs' realizations,
    the particles generated on the host first, then the planes' trees built
    on the device; returns the requests, the particle count and the host
    generation and device build times."""
    specs = W.c4_specs(n_real, seed0=seed0)
    t0 = time.perf_counter()
    planes = [[W.field_plane(4, j, z, n, soft=5e-4, realization=sp["realization"],
                             dtype=np.float32) for j, (z, n) in enumerate(zip(sp["z"], sp["n"]))]
              for sp in specs]
    gen_s = time.perf_counter() - t0
    reqs, nparts, dev_ms = [], 0, 0.0
    t0 = time.perf_counter()
    for sp, pl in zip(specs, planes):
        lens = S.Lens(ctx, sp["zs"])
        for (x, y, m, h), z in zip(pl, sp["z"]):
            halo = S.LensHaloParticles(ctx, x, y, m, h, z, theta=0.1)
            dev_ms += halo.info()["build_ms"]
            lens.moveinMainHalo(halo)
            nparts += x.size
        reqs.append((lens, sp["side"], (0.0, 0.0), W.grid_range(300.0), sp["zs"]))
    build_s = time.perf_counter() - t0
    return reqs, nparts, gen_s, build_s, dev_ms


def run_batched(S, W, ctx, n_real, precision):
    reqs, nparts, gen_s, build_s, dev_ms = c4_requests(S, W, ctx, n_real)
    nrays = sum(r[1] ** 2 for r in reqs)
    b = S.Batch(ctx, reqs, precision)
    b.set_schedule(tile_log2=3, queue=0, order=0, graph=0)
    b.run()
    fixed = min(b.run()["walk_ms"] for _ in range(2))
    best, base_ms, best_ms = b.autotune(repeats=1)
    runs = [b.run() for _ in range(2)]
    tuned = min(r["walk_ms"] for r in runs)
    tuned_total = min(r["total_ms"] for r in runs)
    out = {"realizations": n_real, "rays": nrays, "particles": nparts,
           "host_generation_s": gen_s, "build_wall_s": build_s, "device_build_ms": dev_ms,
           "fixed_schedule_ms": fixed, "fixed_rays_per_s": nrays / (fixed * 1e-3),
           "autotuned_schedule": best, "autotuned_ms": tuned,
           "autotuned_rays_per_s": nrays / (tuned * 1e-3),
           "autotuned_total_ms": tuned_total, "n_guarded": runs[-1]["n_guarded"],
           "autotune_gain": fixed / tuned,
           "rays_per_s_incl_device_build": nrays / (dev_ms * 1e-3 + tuned_total * 1e-3)}
    log("batched:", out)
    for r in reqs:
        r[0].close()
    return out


def run_serving(S, W, ctx, n_models, n_requests, load, precision, seed=5999, rate=None):
    """Serving form: resident 10-plane x 2^20-particle lenses; Poisson arrivals
    of GridMap requests (lens, zs in [1,4], centre in the 5' field, side from
    {128,...,2048} with p = {.4,.3,.15,.1,.05}, 300"/2048 per pixel)."""
    from concurrent.futures import ThreadPoolExecutor
    t0 = time.perf_counter()
    lenses = []
    for r in range(n_models):
        lens, _, _ = W.build_field_lens(ctx, config=5, nplanes=10, n=1 << 20, zs=3.0,
                                        realization=r)
        lenses.append(lens)
    build_s = time.perf_counter() - t0
    lid, zs, cen, sides, res = W.c5_stream(n_requests, n_models, seed=seed)
    rng = np.random.default_rng(seed + 1)
    # capacity from a short closed loop, then Poisson arrivals at `load` of it
    srv = S.Server(ctx, hbm_budget=int(200e9))
    tc = time.perf_counter()
    probe = [srv.submit(lenses[int(lid[i])], int(sides[i]), cen[i], res * (sides[i] - 1),
                        float(zs[i]), precision) for i in range(min(24, n_requests))]
    for rid in probe:
        srv.wait(rid, fetch=False)
    cap_rays = sum(int(sides[i]) ** 2 for i in range(len(probe))) / (time.perf_counter() - tc)
    mean_rays = float(np.mean(sides.astype(np.float64) ** 2))
    if rate is None:
        rate = load * cap_rays / mean_rays
    else:
        load = rate * mean_rays / cap_rays
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    lat, infos = [], []
    lock = threading.Lock()

    kept = {}

    def collect(i, rid, N):
        y, _, info = srv.wait(rid, N)
        with lock:
            infos.append(info)
            if i < 16:
                kept[i] = y

    pool = ThreadPoolExecutor(max_workers=32)
    start = time.perf_counter()
    futs = []
    for i in range(n_requests):
        delay = start + arrivals[i] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        N = int(sides[i])
        rid = srv.submit(lenses[int(lid[i])], N, cen[i], res * (N - 1), float(zs[i]), precision)
        futs.append(pool.submit(collect, i, rid, N))
    for f in futs:
        f.result()
    pool.shutdown()
    lat = np.array([i["t_collect"] - i["t_submit"] for i in infos])
    t_first = min(i["t_submit"] for i in infos)
    t_last = max(i["t_collect"] for i in infos)
    rays = sum(i["rays"] for i in infos)
    st = srv.stats()
    srv.close()
    # the served results of the first 16 requests against lone FP64 GridMaps
    amax = 0.0
    for i, y in kept.items():
        N = int(sides[i])
        g = S.GridMap(lenses[int(lid[i])], N, cen[i], res * (N - 1), S.SLSIM_FP64, zs=float(zs[i]))
        x = g.positions()
        amax = max(amax, rel_err(y - g.y(), x - g.y())[0])
        g.close()
    for lens in lenses:
        lens.close()
    out = {"models": n_models, "requests": n_requests, "build_wall_s": build_s,
           "alpha_max_vs_fp64_first16": amax,
           "capacity_rays_per_s": cap_rays, "load": load, "rate_per_s": rate,
           "throughput_rays_per_s": rays / (t_last - t_first),
           "p50_ms": 1e3 * float(np.percentile(lat, 50)),
           "p99_ms": 1e3 * float(np.percentile(lat, 99)),
           "mean_ms": 1e3 * float(lat.mean()), "admitted": st["admitted"],
           "rejected": st["rejected"], "max_active": st["max_active"],
           "iterations": st["iterations"], "hbm_budget_bytes": st["budget_bytes"]}
    log("serving:", out)
    return out


def run_smoothing(S, W, ctx, n, nsmooth=64, cpu=True):
    """Smoothing sizes (LensHaloParticles::calculate_smoothing) of n 3-D
    particles in 20 NFW spheres, float64 positions: device time including the
    host transfers, and the oracle's restatement of the reference's tree
    search on 16 host threads over a 65,536-particle query subsample."""
    x, y, z = W.particles_3d(9000, n)
    S.LensHaloParticles.calculate_smoothing(ctx, nsmooth, x[:200000], y[:200000], z[:200000])
    t0 = time.perf_counter()
    g = S.LensHaloParticles.calculate_smoothing(ctx, nsmooth, x, y, z)
    tg = time.perf_counter() - t0
    out = {"particles": n, "nsmooth": nsmooth, "gpu_seconds": tg, "gpu_particles_per_s": n / tg}
    if cpu:
        import oracle as O
        which = np.arange(0, n, max(1, n // 65536))
        t0 = time.perf_counter()
        O.smoothing_sizes(x, y, z, nsmooth, which=which[:1], nthreads=16)
        tb = time.perf_counter() - t0                  # tree build (and one query)
        t0 = time.perf_counter()
        o = O.smoothing_sizes(x, y, z, nsmooth, which=which, nthreads=16)
        tc = time.perf_counter() - t0
        q_rate = which.size / max(tc - tb, 1e-9)
        out.update({"cpu_build_seconds": tb, "cpu_seconds_subsample": tc,
                    "cpu_queries": int(which.size), "cpu_queries_per_s": q_rate,
                    "cpu_particles_per_s": n / (tb + n / q_rate), "cpu_threads": 16,
                    "cpu_note": "tree build plus n queries at the subsample's query rate",
                    "equal_to_oracle_on_subsample": bool(np.array_equal(g[which], o))})
        out["gpu_over_cpu"] = out["gpu_particles_per_s"] / out["cpu_particles_per_s"]
    log("smoothing:", out)
    return out


def run_formats(S, W, ctx, n, cpu=True, nsmooth=32):
    """Particle files: n particles in 20 NFW spheres written as a csv5 table
    (sizes given) and a csv4 table (sizes computed), read by
    MakeParticleLenses; the glmb cache reopened; a 2048^2 float64 kappa map
    written as FITS.  The oracle's single-thread restatement of readCSV
    (getline, splitstring, stof, as the reference runs it) is timed on the
    same csv5 file."""
    import tempfile
    x, y, z = W.particles_3d(9100, n, dtype=np.float32)
    d = tempfile.mkdtemp(prefix="slsim_formats_")
    m = np.full(n, np.float32(1e9))
    h = np.full(n, np.float32(1e-3))
    cols = np.stack([x, y, z, m, h], 1).astype(np.float32)
    f5, f4 = os.path.join(d, "p5.csv"), os.path.join(d, "p4.csv")
    np.savetxt(f5, cols, fmt="%.9g", delimiter=",", header="x,y,z,mass,size")
    np.savetxt(f4, cols[:, :4], fmt="%.9g", delimiter=",", header="x,y,z,mass")
    mb = os.path.getsize(f5) / 1e6
    out = {"particles": n, "csv5_MB": mb, "nsmooth": nsmooth}
    t0 = time.perf_counter()
    mk = S.MakeParticleLenses(ctx, f5, S.SLSIM_FMT_CSV5, nsmooth)
    t5 = time.perf_counter() - t0
    out.update({"csv5_read_s": t5, "csv5_particles_per_s": n / t5, "csv5_MB_per_s": mb / t5})
    t0 = time.perf_counter()
    mk4 = S.MakeParticleLenses(ctx, f4, S.SLSIM_FMT_CSV4, nsmooth)
    t4 = time.perf_counter() - t0
    out.update({"csv4_with_device_sizes_s": t4, "csv4_particles_per_s": n / t4})
    t0 = time.perf_counter()
    S.MakeParticleLenses(ctx, f4, S.SLSIM_FMT_CSV4, nsmooth)       # reads the glmb cache
    tg = time.perf_counter() - t0
    out.update({"glmb_cache_read_s": tg, "glmb_particles_per_s": n / tg})
    img = np.random.default_rng(3).normal(size=(2048, 2048))
    t0 = time.perf_counter()
    S.write_fits_pixelmap(os.path.join(d, "k.fits"), img, 1e-7, (0.0, 0.0))
    tf = time.perf_counter() - t0
    out.update({"fits_2048sq_f64_s": tf, "fits_MB_per_s": img.nbytes / 1e6 / tf})
    if cpu:
        from oracle import formats as F
        t0 = time.perf_counter()
        rc, q, _ = F.read_csv(f5, 5)
        tc = time.perf_counter() - t0
        out.update({"cpu_csv5_read_s": tc, "cpu_csv5_particles_per_s": n / tc, "cpu_threads": 1,
                    "csv5_speedup": tc / t5,
                    "csv5_equal_to_oracle": bool(rc == 0 and q.tobytes() == mk.data.tobytes())})
    del mk, mk4
    for fn in os.listdir(d):
        os.remove(os.path.join(d, fn))
    os.rmdir(d)
    log("formats:", out)
    return out


def run_halos(S, W, ctx, n, cpu=True, N=2048, field=False):
    """Halo-of-halos plane: one LensHaloHalos of n NFW halos (masses
    log-uniform 1e10-1e14 Msun, Rsize = r200, c = 10 (M/1e12)^-0.1, centres
    uniform in a 5' field) at z = 0.5, source at z = 2, N^2 FP64 GridMap.
    With field=True the same halos as a screened LensPlaneTree field plane
    (screening length a quarter of the field, periodic buffer: nine walks per
    ray, sigma_background the halos' mean surface density).  The oracle's
    restatement of TreeQuadHalos on 16 host threads over a 65,536-ray
    subsample."""
    rng = np.random.default_rng(4242)
    z = 0.5
    m = (10 ** rng.uniform(10, 14, n)).astype(np.float32)
    rs = np.array([S.nfw_r200(float(v), z) for v in m], np.float32)
    c = (10 * (m / 1e12) ** -0.1).astype(np.float32)
    side = 300 * S.arcsecTOradians * S.coorDist(0.0, z) / (1 + z)
    hx = (rng.random(n) - 0.5) * side
    hy = (rng.random(n) - 0.5) * side
    sigma = float(np.sum(m.astype(np.float64))) / side ** 2
    scr = 4.0 / side
    t0 = time.perf_counter()
    if field:
        halo = S.LensPlaneTree(ctx, z, hx, hy, m, rs, c, sigma, scr)
    else:
        halo = S.LensHaloHalos(ctx, hx, hy, m, rs, c, z)
    build_s = time.perf_counter() - t0
    lens = S.Lens(ctx, 2.0)
    if field:
        lens.addFieldPlane(halo)
    else:
        lens.moveinMainHalo(halo)
    rg = W.grid_range(300.0)
    g = S.GridMap(lens, N, (0.0, 0.0), rg, S.SLSIM_FP64, shoot=False)
    g.shoot()
    t1 = time.perf_counter()
    st = g.shoot()
    ctx.synchronize()
    dt = time.perf_counter() - t1
    out = {"halos": n, "field_plane": bool(field), "grid": N, "host_build_s": build_s,
           "rays_per_s": N * N / dt,
           "walk_launch_ms": st["walk_ms"] / max(st["n_walk_launches"], 1),
           "nodes": halo.info()["n_nodes"]}
    if cpu:
        import oracle as O
        L = O.Lens(2.0)
        if field:
            L.addFieldPlane(O.FieldTree(hx, hy, m, rs, c, sigma, scr), z)
        else:
            L.moveinMainHalo(O.HaloTree(hx, hy, m, rs, c), z)
        x = g.positions()[::max(1, (N * N) // 65536)]
        t2 = time.perf_counter()
        yr, Ar = L.shoot(x, nthreads=16)
        tc = time.perf_counter() - t2
        y = g.y()[::max(1, (N * N) // 65536)]
        out.update({"cpu_rays_per_s": x.shape[0] / tc, "cpu_threads": 16,
                    "cpu_sample_rays": int(x.shape[0]),
                    "max_rel_y_vs_oracle": float(np.abs(y - yr).max() / np.abs(yr).max())})
        out["gpu_over_cpu"] = out["rays_per_s"] / out["cpu_rays_per_s"]
    g.close()
    log("halos:", out)
    return out


def run_images(S, W, ctx, nsrc, cpu=True, N=2048):
    """GridMap::find_images: nsrc source positions around the caustics of the
    65,536-particle NFW halo lens on its N^2 GridMap (30" field).  The oracle's
    serial restatement is timed on 8 sources; the reference runs sources on
    threads, so its rate is quoted per host thread and for 16 threads."""
    lens, _ = W.build_c1_lens(ctx)
    g = S.GridMap(lens, N, (0.0, 0.0), W.grid_range(30.0), S.SLSIM_FP64)
    x, y, A = g.positions(), g.y(), g.A()
    rng = np.random.default_rng(77)
    centre = np.argsort((x ** 2).sum(1))[:100000]
    ys = y[rng.choice(centre, nsrc)] + rng.normal(size=(nsrc, 2)) * 1e-8
    g.find_images(ys[:16])
    t0 = time.perf_counter()
    rays, mult, _ = g.find_images(ys, 64)
    tg = time.perf_counter() - t0
    out = {"sources": nsrc, "grid": N, "images": int(rays.shape[0]),
           "max_multiplicity": int(mult.max()), "gpu_seconds": tg, "sources_per_s": nsrc / tg}
    if cpu:
        import oracle as O
        t1 = time.perf_counter()
        r0, m0, _ = O.find_images(N, x, y, A, ys[:8])
        tc = (time.perf_counter() - t1) / 8
        k = int(m0.sum())
        out.update({"cpu_seconds_per_source": tc, "cpu_sources_per_s_16_threads": 16 / tc,
                    "equal_to_oracle_first8": bool(np.array_equal(m0, mult[:8]) and
                                                   np.array_equal(r0, rays[:k]))})
        out["gpu_over_cpu16"] = out["sources_per_s"] / out["cpu_sources_per_s_16_threads"]
    g.close()
    log("images:", out)
    return out


def run_crit(S, W, ctx, N, cpu=True, Nm=768, Nc=128):
    """GridMap::find_crit and find_magnification_contour on the 65,536-particle
    NFW halo lens's N^2 GridMap (30" field), and find_boundaries_of_caustics
    on an Nc^2 grid of the same lens (6" field; the reference tests every grid
    position against every triangle, N^4 work).  The oracle's serial
    restatements (the reference's routines are serial) are timed on the same
    grid outputs, find_crit on an Nm^2 grid (its radial-curve placement is
    quadratic in the number of curves: 46,300 curves at 768^2)."""
    lens, _ = W.build_c1_lens(ctx)

    def crit_on(n):
        g = S.GridMap(lens, n, (0.0, 0.0), W.grid_range(30.0), S.SLSIM_FP64)
        g.find_crit()
        t0 = time.perf_counter()
        res = g.find_crit()
        return g, res, time.perf_counter() - t0

    g, (curves, hits, types, _), tc = crit_on(N)
    t0 = time.perf_counter()
    cont, _, _, _ = g.find_magnification_contour(0.0)
    tm = time.perf_counter() - t0
    out = {"grid": N, "curves": len(curves), "points": int(sum(len(c) for c in curves)),
           "tangential": int(np.sum(types == S.SLSIM_CRIT_TANGENTIAL)),
           "radial": int(np.sum(types == S.SLSIM_CRIT_RADIAL)),
           "pseudo": int(np.sum(types == S.SLSIM_CRIT_PSEUDO)),
           "find_crit_seconds": tc, "magnification_contour_seconds": tm,
           "contours": len(cont)}
    gm, (cm, hm, tym, _), tcm = crit_on(Nm)
    out.update({"grid_m": Nm, "curves_m": len(cm), "find_crit_seconds_m": tcm})
    gc = S.GridMap(lens, Nc, (0.0, 0.0), W.grid_range(6.0), S.SLSIM_FP64)
    gc.find_boundaries_of_caustics()
    t0 = time.perf_counter()
    bc, _, mult = gc.find_boundaries_of_caustics()
    tb = time.perf_counter() - t0
    out.update({"caustics_grid": Nc, "caustic_boundaries": len(bc),
                "boundaries_of_caustics_seconds": tb})
    if cpu:
        import oracle as O
        x, A = g.positions(), g.A()
        t1 = time.perf_counter()
        O.find_magnification_contour(N, x, A, g.range, 0.0)
        out["cpu_magnification_contour_seconds"] = time.perf_counter() - t1
        x, y, A = gm.positions(), gm.y(), gm.A()
        t1 = time.perf_counter()
        c0, h0, t0_, _ = O.find_crit(Nm, x, y, A, gm.range)
        out["cpu_find_crit_seconds_m"] = time.perf_counter() - t1
        out["find_crit_equal_to_oracle_m"] = bool(
            len(c0) == len(cm) and np.array_equal(t0_, tym) and np.array_equal(h0, hm)
            and all(np.array_equal(a, b) for a, b in zip(c0, cm)))
        xc, yc, Ac = gc.positions(), gc.y(), gc.A()
        t1 = time.perf_counter()
        b0, _, m0 = O.find_boundaries_of_caustics(Nc, xc, yc, Ac)
        out["cpu_boundaries_of_caustics_seconds"] = time.perf_counter() - t1
        out["boundaries_of_caustics_equal_to_oracle"] = bool(
            np.array_equal(m0, mult) and len(b0) == len(bc) and
            all(np.array_equal(a, b) for a, b in zip(b0, bc)))
        out["find_crit_speedup_m"] = out["cpu_find_crit_seconds_m"] / tcm
        out["magnification_contour_speedup"] = out["cpu_magnification_contour_seconds"] / tm
        out["boundaries_of_caustics_speedup"] = out["cpu_boundaries_of_caustics_seconds"] / tb
    for h in (gc, gm, g):
        h.close()
    log("crit:", out)
    return out


def run_small_batches(S, W, ctx, nbatch, cpu=True, seed=606):
    """Callers that shoot many small ray lists (adaptive Grid refinement,
    critical-curve searches): nbatch lists of 64-4096 rays (log-uniform)
    near the 65,536-particle NFW halo's centre, FP64.  One rayshooterInternal
    call per list in sequence (the reference's pattern), against the same
    lists submitted by 8 client threads to the server."""
    lens, _ = W.build_c1_lens(ctx)
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(64), np.log(4096), nbatch)).astype(int)
    lists = [(rng.random((int(k), 2)) - 0.5) * W.grid_range(30.0) for k in sizes]
    total = int(sizes.sum())
    lens.rayshooterInternal(lists[0])
    t0 = time.perf_counter()
    for x in lists:
        lens.rayshooterInternal(x)
    t_seq = time.perf_counter() - t0
    srv = S.Server(ctx)
    srv.wait_rays(srv.submit_rays(lens, lists[0], None, S.SLSIM_FP64), lists[0].shape[0])
    lat = []
    lock = threading.Lock()

    def client(c):
        for i in range(c, nbatch, 8):
            t = time.perf_counter()
            rid = srv.submit_rays(lens, lists[i], None, S.SLSIM_FP64)
            srv.wait_rays(rid, lists[i].shape[0])
            with lock:
                lat.append(time.perf_counter() - t)

    t1 = time.perf_counter()
    th = [threading.Thread(target=client, args=(c,)) for c in range(8)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    t_srv = time.perf_counter() - t1
    st = srv.stats()
    srv.close()
    out = {"batches": nbatch, "rays": total, "sequential_rays_per_s": total / t_seq,
           "served_rays_per_s": total / t_srv, "served_over_sequential": t_seq / t_srv,
           "served_p50_ms": 1e3 * float(np.percentile(lat, 50)),
           "served_p99_ms": 1e3 * float(np.percentile(lat, 99)),
           "max_active": st["max_active"], "iterations": st["iterations"]}
    log("small batches:", out)
    return out


def run_precision_form(S, W, ctx, n=1 << 22, nplanes=20, N=4096):
    """Precision form: 20 planes x 4,194,304 float32-position particles,
    opening angle 0.2, 4096^2 rays; FP32C and FP64 trees against the FP64
    direct sum on a 65,536-ray subsample (every 256th ray in raster order)."""
    t0 = time.perf_counter()
    lens, _, _ = W.build_field_lens(ctx, config=3, nplanes=nplanes, n=n, zs=3.0, theta=0.2,
                                    dtype=np.float32)
    build_s = time.perf_counter() - t0
    rg = W.grid_range(300.0)
    out = {"planes": nplanes, "particles_per_plane": n, "grid": N, "theta": 0.2,
           "build_wall_s": build_s}
    res = {}
    for name, prec in (("fp32c", S.SLSIM_FP32C), ("fp64", S.SLSIM_FP64)):
        g = S.GridMap(lens, N, (0.0, 0.0), rg, prec, shoot=False)
        g.shoot()
        ctx.synchronize()
        t1 = time.perf_counter()
        st = g.shoot()
        ctx.synchronize()
        dt = time.perf_counter() - t1
        out[name + "_rays_per_s"] = N * N / dt
        out[name + "_walk_launch_ms"] = st["walk_ms"] / max(st["n_walk_launches"], 1)
        out[name + "_ms"] = 1e3 * dt
        if prec == S.SLSIM_FP32C:
            out["fp32c_guarded_rays"] = st["n_guarded"]
            out["fp32c_guard_ms"] = 1e3 * dt - st["walk_ms"]
        res[name] = (g.y(), g.A())
        g.close()
    x = S.GridMap(lens, N, (0.0, 0.0), rg, shoot=False).positions()
    out["fp32c_vs_fp64"] = compare(x, res["fp32c"][0], res["fp32c"][1], res["fp64"][0],
                                   res["fp64"][1])
    a32 = x - res["fp32c"][0]
    a64 = x - res["fp64"][0]
    n64 = np.linalg.norm(a64, axis=1)
    ea = np.linalg.norm(a32 - a64, axis=1) / np.maximum(n64, 1e-3 * np.sqrt(np.mean(n64 ** 2)))
    out["alpha_rays_over_1e-5"] = int((ea > 1e-5).sum())
    out["fp32c_over_fp64_speed"] = out["fp32c_rays_per_s"] / out["fp64_rays_per_s"]
    sub = np.arange(0, N * N, 256)
    t1 = time.perf_counter()
    yd, Ad = lens.rayshooterInternal(x[sub], precision=S.SLSIM_DIRECT64)
    out["direct_sum_seconds"] = time.perf_counter() - t1
    out["vs_fp64_direct_sum"] = {
        "fp32c": compare(x[sub], res["fp32c"][0][sub], res["fp32c"][1][sub], yd, Ad),
        "fp64_tree": compare(x[sub], res["fp64"][0][sub], res["fp64"][1][sub], yd, Ad)}
    log("precision form:", out)
    lens.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1 << 20, help="particles per plane")
    ap.add_argument("--grid", type=int, default=2048)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads of the CPU baseline (0: every CPU this process may "
                         "use, std::thread::hardware_concurrency)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip FP64, direct-sum, counts")
    ap.add_argument("--smoothing", type=int, default=0, metavar="N",
                    help="also time the smoothing sizes (Nsmooth 64) of N 3-D particles")
    ap.add_argument("--small-batches", type=int, default=0, metavar="B",
                    help="also time B small ray lists: sequential calls vs the server")
    ap.add_argument("--images", type=int, default=0, metavar="S",
                    help="also time GridMap::find_images for S sources on a 2048^2 grid")
    ap.add_argument("--crit", type=int, default=0, metavar="N",
                    help="also time GridMap::find_crit on an N^2 grid (and the other contours)")
    ap.add_argument("--field-plane", type=int, default=0, metavar="N",
                    help="also time a screened LensPlaneTree field plane of N NFW halos")
    ap.add_argument("--halos", type=int, default=0, metavar="N",
                    help="also time a plane of N NFW halos (LensHaloHalos) on a 2048^2 GridMap")
    ap.add_argument("--formats", type=int, default=0, metavar="N",
                    help="also time the particle-file readers (csv5, csv4 with device "
                         "sizes, glmb) on N particles and a 2048^2 FITS write")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run one plain device-to-device copy of MIB MiB (HIP runtime), the "
                         "known-size kernel that scales HBM counters under rocprofv3 --pmc")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the precision, batched and serving forms")
    ap.add_argument("--no-precision-form", action="store_true",
                    help="skip the 20-plane x 2^22-particle, 4096^2-ray precision form")
    ap.add_argument("--batched", type=int, default=256,
                    help="realizations of the batched form (0: skip)")
    ap.add_argument("--serving-models", type=int, default=64,
                    help="resident lenses of the serving form (0: skip)")
    ap.add_argument("--serving-requests", type=int, default=2000)
    ap.add_argument("--serving-load", type=float, default=0.8)
    ap.add_argument("--serving-rate", type=float, default=0.0,
                    help="Poisson arrival rate (requests/s) instead of a fraction of capacity")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build's ray shooter runs on a single "
              "MI355X (multi-GPU sharding is a design note, not implemented); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import slsimlib_amd as S
    from slsimlib_amd import workloads as W

    ctx = S.Context(0)
    t0 = time.perf_counter()
    lens, parts, zl = W.build_field_lens(ctx, config=2, nplanes=10, n=args.n, zs=3.0,
                                         theta=0.1, soft=5e-4, keep_particles=not args.no_cpu)
    build_s = time.perf_counter() - t0
    build_ms_dev = sum(h.info()["build_ms"] for h in lens.halos)
    nodes = sum(h.info()["n_nodes"] for h in lens.halos)
    log("planes built: %.1f s wall (device build %.1f ms), %d cells" % (build_s, build_ms_dev, nodes))

    N = args.grid
    rg = W.grid_range(300.0)
    nrays = N * N
    g32 = S.GridMap(lens, N, (0.0, 0.0), rg, S.SLSIM_FP32C, shoot=False)
    for _ in range(args.warmup):
        g32.shoot()
    ctx.synchronize()
    walk_ms, launches, stats = 0.0, 0, []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = g32.shoot()
        stats.append(st)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    for st in stats:
        walk_ms += st["walk_ms"]
        launches += st["n_walk_launches"]
    value = args.steps * nrays / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    walk_launch_ms = walk_ms / max(launches, 1)
    log("fp32c: %.3e rays/s, %.1f ms/step, walk launch %.2f ms" % (value, ms_per_step, walk_launch_ms))

    out = {
        "metric": "rays/s multi-plane (alpha,kappa,gamma) at |dalpha|<=1e-5 rel vs FP64",
        "value": value, "unit": "rays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32c",
        "data": "synthetic (seeded NFW particle planes)",
        "config": {"workload": "multi-plane field: 10 planes x %d particles, GridMap %dx%d, "
                               "zs=3, theta=0.1, bucket 20" % (args.n, N, N),
                   "planes": 10, "particles_per_plane": args.n, "grid": N,
                   "field_arcsec": 300, "z_source": 3.0, "theta": 0.1,
                   "walk": "FP32 far field with Kahan-compensated accumulation; FP64 opening "
                           "test, cell offsets, near-field terms and plane recursion",
                   "parallelism": "single GPU"},
        "tree_build": {"device_ms_total": build_ms_dev, "cells_total": nodes},
    }

    if not args.no_extras:
        # plain FP64 path on the same GPU
        g64 = S.GridMap(lens, N, (0.0, 0.0), rg, S.SLSIM_FP64, shoot=False)
        g64.shoot()
        ctx.synchronize()
        t0 = time.perf_counter()
        st64 = g64.shoot()
        ctx.synchronize()
        t64 = time.perf_counter() - t0
        out["fp64_plain"] = {"rays_per_s": nrays / t64, "ms_per_step": 1e3 * t64,
                             "walk_launch_ms": st64["walk_ms"] / max(st64["n_walk_launches"], 1)}
        log("fp64: %.3e rays/s" % (nrays / t64))
        x = g32.positions()
        y32, A32 = g32.y(), g32.A()
        y64, A64 = g64.y(), g64.A()
        out["fp32c_vs_fp64"] = compare(x, y32, A32, y64, A64)
        log("fp32c vs fp64:", out["fp32c_vs_fp64"])

        # work of the reference's per-ray criterion -> algorithmic flops
        cnt = g64.counts().astype(np.float64)
        tot = cnt.sum(0)
        tab = lens.tables()
        flops = (FLOPS["visit"] * tot[0] + FLOPS["far"] * tot[1] + FLOPS["leafp"] * tot[2]
                 + FLOPS["bigtest"] * tot[3] + FLOPS["bspline"] * tot[4]
                 + FLOPS["plane_step"] * nrays * tab["nplanes"])
        gather_bytes = 20 * tot[0] + 16 * tot[1] + 12 * tot[2] + 16 * tot[3]
        per_launch = flops / tab["nplanes"]
        achieved = per_launch / (walk_launch_ms * 1e-3) / 1e12
        traffic = None
        tf = os.path.join(ROOT, "profiles", "walk_fp32c_multiplane_traffic.json")
        if os.path.exists(tf):
            traffic = json.load(open(tf)).get("hbm_bytes_per_launch")
        out["roofline"] = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                           "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                           "traffic": traffic, "kernel": "k_plane_step32 (FP32C walk + recursion, one lens plane)",
                           "flops_per_launch": per_launch,
                           "operand_demand_bytes_per_launch": gather_bytes / tab["nplanes"],
                           "operand_demand_note": "per-ray operand bytes of the reference's "
                           "criterion (20 B a visit, 16 B a far cell, 12 B a leaf particle, "
                           "16 B a big-particle test); the bundle reads each once per 64 rays "
                           "through the scalar cache, so this is not an HBM figure",
                           "launch_ms": walk_launch_ms}
        fp64_tf = per_launch / (out["fp64_plain"]["walk_launch_ms"] * 1e-3) / 1e12
        out["fp64_plain"]["roofline_frac_fp64"] = fp64_tf / PEAK_FP64_TFLOPS
        out["work_per_ray"] = {"visits": tot[0] / nrays, "far_cells": tot[1] / nrays,
                               "leaf_particles": tot[2] / nrays, "big_tests": tot[3] / nrays,
                               "bsplines": tot[4] / nrays, "flops": flops / nrays}
        log("roofline:", out["roofline"])
        out["bundle_utilisation"] = g64.bundle_stats()
        log("bundles:", out["bundle_utilisation"])

        # FP64 direct-sum oracle on a 65,536-ray subsample (every 64th ray)
        sub = np.arange(0, nrays, max(1, nrays // 65536))
        t0 = time.perf_counter()
        yd, Ad = lens.rayshooterInternal(x[sub], precision=S.SLSIM_DIRECT64)
        td = time.perf_counter() - t0
        out["error_vs_fp64_direct_sum"] = {
            "rays": int(sub.size), "seconds": td,
            "fp32c": compare(x[sub], y32[sub], A32[sub], yd, Ad),
            "fp64_tree": compare(x[sub], y64[sub], A64[sub], yd, Ad)}
        log("vs direct sum:", out["error_vs_fp64_direct_sum"])

        # host-buffer path (slsim_shoot, rayshooterInternal on host Points):
        # rays in and y, A out over PCIe, rays Morton-ordered into bundles on
        # the device, timed over the same steps as `value`; a ray's result does
        # not depend on its bundle mates, so it equals the GridMap's bit for bit
        lens.rayshooterInternal(x[:65536], precision=S.SLSIM_FP32C)
        yh = np.zeros((nrays, 2))
        Ah = np.zeros((nrays, 4))
        xh = np.ascontiguousarray(x)
        lens.shoot_into(xh, yh, Ah, precision=S.SLSIM_FP32C)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            lens.shoot_into(xh, yh, Ah, precision=S.SLSIM_FP32C)
        th = (time.perf_counter() - t0) / args.steps
        out["host_buffers_fp32c"] = {
            "rays_per_s": nrays / th, "ms_per_step": 1e3 * th, "steps": args.steps,
            "bytes_per_ray_pcie": 16 + 48,
            "bitwise_equal_to_gridmap": bool(np.array_equal(yh, y32) and np.array_equal(Ah, A32))}
        log("host buffers:", out["host_buffers_fp32c"])

        # conditioning of kappa/gamma: the FP64 tree's own response when the
        # input rays move by 1e-7 of the rms deflection (the size of one FP32
        # rounding of a plane's deflection), to read fp32c_vs_fp64 against
        xs = x[sub]
        amp = 1e-7 * np.sqrt(np.mean(np.sum((xs - y64[sub]) ** 2, 1)))
        phi = np.random.default_rng(7).uniform(0, 2 * np.pi, xs.shape[0])
        xn = xs + amp * np.stack([np.cos(phi), np.sin(phi)], 1)
        yn, An = lens.rayshooterInternal(xn, precision=S.SLSIM_FP64)
        out["fp64_response_to_1e-7_ray_nudge"] = compare(xn, yn, An, y64[sub], A64[sub], xr=xs)
        log("fp64 nudge response:", out["fp64_response_to_1e-7_ray_nudge"])

    if not args.no_cpu:
        import oracle as O
        nproc = os.cpu_count() or 1
        try:
            avail = len(os.sched_getaffinity(0))
        except AttributeError:
            avail = nproc
        # every CPU this process may use (--cpu-threads caps it)
        nthreads = avail if args.cpu_threads <= 0 else max(1, min(avail, args.cpu_threads))
        L = O.Lens(3.0)
        t0 = time.perf_counter()
        for (px, py, pm, ph, z) in parts:
            L.moveinMainHalo(O.Tree(px, py, pm, ph, theta=0.1), z)
        cpu_build = time.perf_counter() - t0
        pos = O.gridmap_positions(N, (0.0, 0.0), rg)
        pilot = pos[np.arange(0, nrays, nrays // 512)]
        t0 = time.perf_counter()
        L.shoot(pilot, nthreads=nthreads)
        tp = (time.perf_counter() - t0) / pilot.shape[0]
        nsample = int(min(65536, max(2048, args.cpu_seconds / tp)))
        stride = max(1, nrays // nsample)
        idx = np.arange(0, nrays, stride)[:nsample]
        times, res = [], None
        for _ in range(3):
            t0 = time.perf_counter()
            res = L.shoot(pos[idx], nthreads=nthreads)
            times.append(time.perf_counter() - t0)
            if sum(times) > 2.5 * args.cpu_seconds:
                break
        tcpu = float(np.median(times))
        cpu_rate = idx.size / tcpu
        model = "unknown CPU"
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        model = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
        out["cpu_baseline"] = {"value": cpu_rate, "unit": "rays/s", "cores": nthreads,
                               "kind": "port", "cpu": model, "nproc": nproc,
                               "affinity_cpus": avail,
                               "rays_per_s_per_thread": cpu_rate / nthreads,
                               "sample": "%d rays (every %dth ray of the %dx%d grid), all 10 "
                                         "planes, median of %d runs; tree build %.1f s "
                                         "excluded; std::thread over contiguous ray chunks "
                                         "on %d threads (the host reports %d CPUs)" % (
                                             idx.size, stride, N, N, len(times), cpu_build,
                                             nthreads, nproc)}
        out["gpu_over_cpu"] = value / cpu_rate
        if not args.no_extras:
            out["fp64_vs_cpu_oracle"] = compare(pos[idx], y64[idx], A64[idx], res[0], res[1])
        log("cpu oracle: %.3e rays/s on %d threads" % (cpu_rate, nthreads))

    if args.copy_calibration:
        nb = args.copy_calibration << 20
        src, dst = ctx.alloc(nb), ctx.alloc(nb)
        ctx.d2d(dst, src, nb)
        out["copy_calibration_bytes"] = nb
        ctx.free(src)
        ctx.free(dst)

    if args.smoothing:
        out["smoothing"] = run_smoothing(S, W, ctx, args.smoothing, cpu=not args.no_cpu)
    if args.formats:
        out["formats"] = run_formats(S, W, ctx, args.formats, cpu=not args.no_cpu)
    if args.halos:
        out["halos"] = run_halos(S, W, ctx, args.halos, cpu=not args.no_cpu)
    if args.field_plane:
        out["field_plane"] = run_halos(S, W, ctx, args.field_plane, cpu=not args.no_cpu,
                                       field=True)
    if args.images:
        out["images"] = run_images(S, W, ctx, args.images, cpu=not args.no_cpu)
    if args.small_batches:
        out["small_batches"] = run_small_batches(S, W, ctx, args.small_batches)
    if args.crit:
        out["crit"] = run_crit(S, W, ctx, args.crit, cpu=not args.no_cpu)

    if not args.no_forms:
        lens.close()   # the forms below need the device memory
    if not args.no_forms and not args.no_precision_form:
        out["precision_form"] = run_precision_form(S, W, ctx)
    if not args.no_forms and args.batched:
        out["batched"] = run_batched(S, W, ctx, args.batched, S.SLSIM_FP32C)
    if not args.no_forms and args.serving_models:
        out["serving"] = run_serving(S, W, ctx, args.serving_models, args.serving_requests,
                                     args.serving_load, S.SLSIM_FP32C,
                                     rate=args.serving_rate or None)

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
