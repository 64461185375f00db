/*
 * slsim_amd.h -- C ABI of the MI355X multi-plane particle ray shooter.
 *
 * Plain C types only: opaque handles, pointers and sizes, int status codes.
 * No exception and no allocation crosses this boundary.  Device memory is
 * owned by the library and counted against the context's HBM budget; host
 * buffers are owned by the caller.  Calls on distinct contexts may run
 * concurrently; calls on one context are serialised on its HIP stream.
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to the GLAMER source tree).
 */
#ifndef SLSIM_AMD_H
#define SLSIM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
enum {
  SLSIM_OK = 0,
  SLSIM_ERR_ARG = 1,          /* bad argument (reference: std::invalid_argument) */
  SLSIM_ERR_HIP = 2,          /* a HIP runtime call failed                        */
  SLSIM_ERR_BUDGET = 3,       /* request exceeds the context's HBM budget         */
  SLSIM_ERR_SAME_PLACE = 4,   /* "particles in same place." (quadTree.h:327-329)  */
  SLSIM_ERR_TOO_DEEP = 5,     /* a cell below level 32 still holds > bucket       */
  SLSIM_ERR_NOT_FLAT = 6,     /* Om + OL != 1 (MultiPlane/lens.cpp:43-47)         */
  SLSIM_ERR_STATE = 7,        /* call out of order (e.g. serving not started)     */
  SLSIM_ERR_IO = 8            /* file missing, unreadable or malformed (reference:
                                 std::runtime_error "file reading error" or exit) */
};

/* Arithmetic of the tree walk.  The plane-to-plane recursion is FP64 in both. */
enum {
  SLSIM_FP64 = 0,             /* plain method: FP64 walk, reference arithmetic    */
  SLSIM_FP32C = 1,            /* FP32 walk with Kahan-compensated accumulation,
                                 guarded: rays whose result could differ from the
                                 FP64 walk's by a flipped opening test (an FP64
                                 test within the FP32 ray's error bound of its
                                 threshold, on a cell whose flip matters) and
                                 rays with |alpha| below 1% of the call's rms
                                 are shot again in FP64                           */
  SLSIM_DIRECT64 = 2          /* FP64 direct sum over every particle of a plane, no
                                 tree: the error-column oracle (O(N) per ray)      */
};

typedef struct slsim_ctx_s* slsim_ctx;
typedef struct slsim_plane_s* slsim_plane;
typedef struct slsim_lens_s* slsim_lens;

/* Message of the last failed call on this thread ("" if none). */
const char* slsim_last_error(void);

/* ---------------------------------------------------------------- context */
/* Opens device `device` with an HBM budget in bytes (0 = 90 % of free HBM). */
int slsim_ctx_create(int device, size_t hbm_budget_bytes, slsim_ctx* out);
int slsim_ctx_destroy(slsim_ctx ctx);
/* Bytes currently allocated by the context and its budget. */
int slsim_ctx_memory(slsim_ctx ctx, size_t* used, size_t* budget);
/* Blocks until all work of the context's stream is done. */
int slsim_ctx_synchronize(slsim_ctx ctx);

/* ------------------------------------------------------------ particle tree */
/*
 * Builds the device quadtree of one particle halo: replaces the
 * LensHaloParticles(std::vector<PType>&, float z, ...) constructor and the
 * TreeQuadParticles build it runs (include/particle_halo.h:58-75, 341-417;
 * include/quadTree.h:247-628).  Positions in physical Mpc, masses in Msun and
 * sizes (softening) in Mpc; the arrays may be freed on return.
 * theta: opening angle (reference default 0.1); bucket: leaf size (20);
 * inv_area: uniform compensating sheet (0 = none).
 */
int slsim_plane_create_f64(slsim_ctx ctx, const double* x, const double* y, const float* m,
                           const float* h, size_t n, double theta, int bucket, double inv_area,
                           slsim_plane* out);
/* Same for ParticleType<float> positions. */
int slsim_plane_create_f32(slsim_ctx ctx, const float* x, const float* y, const float* m,
                           const float* h, size_t n, double theta, int bucket, double inv_area,
                           slsim_plane* out);
/* The same with TreeQuadParticles' maximum_range, as the file-reading
 * LensHaloParticles constructor passes its max_range (include/particle_halo.h:
 * 416; include/quadTree.h:266, 370-380): with max_range > 0 a cell at centre
 * c stays a leaf when theta > 1.732 box / max(|c| - max_range, 1.732 box);
 * max_range <= 0 is the plain build. */
int slsim_plane_create_range_f64(slsim_ctx ctx, const double* x, const double* y, const float* m,
                                 const float* h, size_t n, double theta, int bucket,
                                 double inv_area, double max_range, slsim_plane* out);
int slsim_plane_create_range_f32(slsim_ctx ctx, const float* x, const float* y, const float* m,
                                 const float* h, size_t n, double theta, int bucket,
                                 double inv_area, double max_range, slsim_plane* out);
int slsim_plane_destroy(slsim_plane plane);

/* LensHaloHalos<LensHaloNFW>(halos, redshift, cosmo, inv_area, theta_force)
 * (include/particle_halo.h:1092-1227): a main halo made of n symmetric NFW
 * halos, each LensHaloNFW(mass, Rsize, zlens, concentration, fratio = 1)
 * (MultiPlane/lens_halos.cpp:389-440) at physical position (x, y) (what
 * LensHalo::getX returns), on a TreeQuadHalos quadtree with bucket 5 and
 * opening angle theta_force (include/quadTreeHalos.h).  The plane it returns
 * goes into a lens with slsim_lens_add_halo like a particle plane, alone or
 * merged with particle halos at the same redshift; slsim_plane_force gives
 * TreeQuadHalos::force2D and slsim_plane_direct_sum the sum over every halo.
 * Its walk is FP64 in every precision mode; n < 2^22.  Errors: SLSIM_ERR_ARG (Rsize or
 * concentration <= 0, Rsize/rscale >= 99.98 where LensHaloNFW's tables end,
 * negative mass), SLSIM_ERR_SAME_PLACE (all halos on one line), SLSIM_ERR_TOO_DEEP
 * (more than 5 halos in one place). */
int slsim_halos_create_nfw(slsim_ctx ctx, const double* x, const double* y, const float* mass,
                           const float* rsize, const float* concentration, size_t n,
                           float inv_area, float theta_force, slsim_plane* out);
/* LensPlaneTree(z, halos, Nhalos, sigma_background, inv_screening_scale)
 * (MultiPlane/planes.cpp:13-19, force :65-68): a field plane of n symmetric
 * NFW halos (as slsim_halos_create_nfw) on a TreeQuadHalos with
 * inv_area = sigma_background; with inv_screening_scale != 0 the tree has
 * bucket 5, theta 0.1, the periodic buffer (nine walks at the ray shifted by
 * the halos' extents) and its far cells screened by
 * exp(-r^2 inv_screening_scale^2) (include/quadTreeHalos.h:557-735), else
 * theta 0.5 and neither.  It goes into a lens with slsim_lens_add_field_plane;
 * slsim_plane_force gives its force. */
int slsim_plane_tree_create_nfw(slsim_ctx ctx, const double* x, const double* y,
                                const float* mass, const float* rsize,
                                const float* concentration, size_t n, double sigma_background,
                                double inv_screening_scale, slsim_plane* out);

/* Halo kinds of slsim_halos_create. */
enum {
  SLSIM_HALO_NFW = 0,         /* LensHaloNFW, param = concentration (used as float) */
  SLSIM_HALO_POWERLAW = 1     /* LensHaloPowerLaw, param = beta (kappa ~ r^-beta, < 2) */
};
/* A TreeQuadHalos of n symmetric halos of mixed kinds (fratio = 1; at
 * physical positions, float mass and Rsize = Rmax): LensHaloNFW
 * (MultiPlane/lens_halos.cpp:389-440) or LensHaloPowerLaw (lens_halos.cpp:
 * 805-866, profile include/lens_halos.h:979-997).  For a LensHaloHalos pass
 * its float inv_area and theta_force, no screening and periodic = 0 (add it
 * with slsim_lens_add_halo); for a LensPlaneTree pass sigma_background and
 * either (theta 0.1, inv_screening_scale, periodic 1) or (0.5, 0, 0) (add it
 * with slsim_lens_add_field_plane).  Errors as slsim_halos_create_nfw, and
 * SLSIM_ERR_ARG for beta >= 2 (the reference's "bad gamma") or a bad kind. */
int slsim_halos_create(slsim_ctx ctx, const double* x, const double* y, const float* mass,
                       const float* rsize, const int32_t* kind, const double* param, size_t n,
                       double inv_area, double theta_force, double inv_screening_scale,
                       int periodic, slsim_plane* out);

typedef struct {
  int64_t n_particles;
  int64_t n_nodes;
  int64_t n_leaves;
  int64_t n_big_internal;   /* particles registered in internal-node big lists */
  int32_t depth;            /* deepest level */
  int32_t pad;
  int64_t device_bytes;
  double build_ms;          /* device build time (HIP events)               */
  double root_p1[2], root_p2[2];
} slsim_plane_info;
int slsim_plane_get_info(slsim_plane plane, slsim_plane_info* info);

/*
 * Exports the node set in depth-first pre-order (child0 first, the order of
 * walkTree_recur, include/quadTree.h:1079-1086).  Per node:
 *   dbl[13 i + 0..12] = p1x, p1y, p2x, p2y, cx, cy, mass, q0, q1, q2, r2crit, flip, 0
 * flip: the cell's flip error (float): the largest change of the raw tree
 * deflection when the cell's opening test is decided the other way, at
 * the distance where it flips (its multipole against the exact particle
 * sum); the precision form's guard uses it.
 *   ints[6 i + 0..5]  = level, nparticles, nbig, is_leaf, 0, 0
 * nbig counts the reference's big_particles list of the node.
 */
int slsim_plane_export_nodes(slsim_plane plane, double* dbl, int64_t* ints);

/* ------------------------------------------------------------------- lens */
/* Lens(long* seed, PosType zs, const COSMOLOGY&) with a flat LCDM cosmology
 * (MultiPlane/lens.cpp:59-77).  h is H0/100. */
int slsim_lens_create(slsim_ctx ctx, double zsource, double Om, double OL, double h,
                      slsim_lens* out);
int slsim_lens_destroy(slsim_lens lens);
/* Lens::moveinMainHalo(halo, addplanes=true) (include/lens.h:172-185): the
 * lens takes ownership of `plane`, placed at float redshift z. */
int slsim_lens_add_halo(slsim_lens lens, slsim_plane plane, float z);
/* Lens::moveinMainHalo(halo, addplanes=false) (include/lens.h:172-185 ->
 * addMainHaloToNearestPlane, MultiPlane/lens.cpp:822-840): the halo joins the
 * main plane closest in distance, whatever its own redshift.  On a lens with
 * no main plane the reference makes a plane of the halo and then adds the
 * halo to it again, so it deflects twice: kept. */
int slsim_lens_add_halo_nearest(slsim_lens lens, slsim_plane plane, float z);
/* A field plane (MultiPlane/lens.cpp:349-455 builds them in createFieldPlanes)
 * at redshift z (double, as Lens::setFieldDist leaves it): the lens takes
 * ownership of `plane`, a slsim_plane_tree_create_nfw plane, and combinePlanes
 * (lens.cpp:2397-2515) sorts it with the main planes by distance (ahead of a
 * main plane at the same distance). */
int slsim_lens_add_field_plane(slsim_lens lens, slsim_plane plane, double z);
/* Lens::clearMainHalos (MultiPlane/lens.cpp:1025-1036): every main halo and
 * its plane is released and the planes are combined again (field planes
 * stay). */
int slsim_lens_clear_main_halos(slsim_lens lens);
/* Number of lens planes, and the plane tables plane_redshifts, Dl, dDl (each
 * nplanes + 1 long, the source pseudo-plane last) and charge = 4 pi G. */
int slsim_lens_tables(slsim_lens lens, int* nplanes, double* z, double* Dl, double* dDl,
                      double* charge);
/* Lens::ResetSourcePlane(z, nearest=false) (MultiPlane/lens.cpp:2609-2657);
 * z <= 0 reverts to the lens source. */
int slsim_lens_reset_source(slsim_lens lens, double z);
/* Lens::ResetSourcePlane(z, nearest) (MultiPlane/lens.cpp:2609-2657) in full:
 * *index = the plane index the reference returns (Numerical Recipes' locate
 * on the plane distances, one back when the source distance equals the
 * previous plane's or, with nearest, when z is below the redshift halfway in
 * distance to the previous plane), *zs_implant = the implanted source
 * redshift (with nearest, that plane's redshift when there is one).  As the
 * reference, z <= 0 changes nothing but turns the implanted source on. */
int slsim_lens_reset_source_plane(slsim_lens lens, double z, int nearest, int32_t* index,
                                  double* zs_implant);

/* ------------------------------------------------------------- ray shooting */
typedef struct {
  double total_ms;          /* wall time of the call                         */
  double walk_ms;           /* sum of walk-kernel times (HIP events)         */
  double walk_ms_max_plane; /* slowest single walk launch                    */
  int32_t n_walk_launches;
  int32_t n_guarded;        /* SLSIM_FP32C: rays the precision guard shot
                               again in FP64 (see SLSIM_FP32C)              */
} slsim_shoot_stats;

/*
 * Lens::rayshooterInternal(unsigned long N, Point*) and the per-ray-redshift
 * overloads (LinkedPoint* + source_zs, RAY*)
 * (FullRange/internal_rayshooter_multi.cpp:53-251; include/lens.h:913-1103).
 * Host buffers: x (2n, x0 x1 interleaved, radians), zs (n source redshifts or
 * NULL = lens source), y (2n out), A (4n out, row-major a0 a1 a2 a3).
 * Rays are Morton-ordered on the device so that each wavefront walks a
 * spatially coherent bundle; results come back in input order.
 */
int slsim_shoot(slsim_lens lens, size_t n, const double* x, const double* zs, double* y,
                double* A, int precision, slsim_shoot_stats* stats);
/* Same with device pointers (no host transfer); synchronous. */
int slsim_shoot_device(slsim_lens lens, size_t n, const double* x_dev, const double* zs_dev,
                       double* y_dev, double* A_dev, int precision, slsim_shoot_stats* stats);

/*
 * No counterpart in the reference.  slsim_shoot_device for callers that cannot
 * allocate device memory (serving loops, harnesses that forbid hipMalloc):
 * the same results bit for bit, with every device buffer of the call taken
 * from one fixed 128 MiB array per device (held from the first context on
 * the device until the process ends) instead of the context's allocations.
 * Limits: n <= 131072 rays per call (SLSIM_ERR_ARG beyond); at most 256
 * planes; the lens's planes must already be on the device (one call of an
 * allocating entry point after the last change to the lens), else
 * SLSIM_ERR_ARG.  Every call on a device shares that device's array, which
 * keeps the last call's contents: the library runs such calls on one device
 * one at a time; callers must not touch the array's memory themselves.
 */
int slsim_shoot_device_noalloc(slsim_lens lens, size_t n, const double* x_dev,
                               const double* zs_dev, double* y_dev, double* A_dev,
                               int precision, slsim_shoot_stats* stats);

/*
 * GridMap(LensHndl, N, center, range) (TreeCode_link/gridmap.cpp:93-120): the
 * N x N ray grid x_i = center + range*((i%N, i/N)/(N-1) - 0.5) is generated on
 * the device and shot; y_dev (2 N^2) and A_dev (4 N^2) are device buffers in
 * grid order (x fastest).  zs <= 0 uses the lens source.
 */
int slsim_gridmap_shoot(slsim_lens lens, size_t N, double c0, double c1, double range,
                        double zs, double* y_dev, double* A_dev, int precision,
                        slsim_shoot_stats* stats);

/*
 * Work counters of the reference's per-ray criterion over a grid shoot, per
 * ray summed over planes: cnt[5 i + k] = visits, far cells, leaf particles,
 * big-particle tests, B-spline corrections (host buffer, 5 N^2).
 */
int slsim_gridmap_count(slsim_lens lens, size_t N, double c0, double c1, double range,
                        double zs, uint64_t* cnt_host);

/* Bundle utilisation of a grid shoot (64-ray wavefront bundles, 8x8 tiles):
 * out4 = { lane cell evaluations the rays need, wavefront cell evaluations,
 *          leaf-particle interactions the rays need, wavefront leaf-particle
 *          iterations }; utilisation = need / (64 * wavefront). */
int slsim_gridmap_bundle_stats(slsim_lens lens, size_t N, double c0, double c1, double range,
                               double zs, uint64_t* out4);

/* FP64 direct sum of the smoothed-particle kernel of one plane (no tree), in
 * the tree's sign convention: out[5 i] = alpha0, alpha1, kappa, gamma0, gamma1
 * at physical positions rays[2 i] (host buffers).  Error column oracle. */
int slsim_plane_direct_sum(slsim_plane plane, size_t n, const double* rays, double* out);
/* The tree walk of one plane at physical positions (host buffers), same layout. */
int slsim_plane_force(slsim_plane plane, size_t n, const double* rays, double* out,
                      int precision);

/* GridMap::find_images(std::vector<Point_2d>& ys, std::vector<int>&
 * multiplicity) and find_images(Point_2d y, image_points, triangles)
 * (TreeCode_link/gridmap.cpp:927-1032, 1326-1348) on a grid shot by
 * slsim_gridmap_shoot with the same N, centre and range (y_dev, A_dev its
 * outputs): every grid cell is split into the triangles (k, k+N+1, k+1) and
 * (k, k+N+1, k+N); a triangle whose source-plane image contains ys[s] (edges
 * and a vertex counted by the reference's rules) gives an image at its
 * centroid (edge midpoint, vertex) with A the mean of its corners' A.
 * Outputs, source by source and in the reference's order within a source
 * (by cell, first triangle first): rays[8 i] = (x0, x1, y0, y1, A0..A3),
 * triangles[3 i] = the corners' grid indices, multiplicity[s], *nimages.
 * rays and triangles hold nsrc * max_per_source entries (max_per_source in
 * [1, 4096]); SLSIM_ERR_ARG if a source has more images.  N < 65536. */
int slsim_gridmap_find_images(slsim_ctx ctx, size_t N, double c0, double c1, double range,
                              const double* y_dev, const double* A_dev, size_t nsrc,
                              const double* ys, size_t max_per_source, int32_t* multiplicity,
                              double* rays, int64_t* triangles, size_t* nimages);

/* A GridMap's grid for the slsim_grid_* entry points: SLSIM_GRID_SQUARE is
 * GridMap(lens, N, center, range) (nx = ny = N, TreeCode_link/gridmap.cpp:
 * 93-120, points as slsim_gridmap_shoot generates them); SLSIM_GRID_RECT is
 * GridMap(lens, Nx, center, rangeX, rangeY) (gridmap.cpp:26-91: ny = (int)(nx
 * rangeY / rangeX) rows, point ii + jj nx at the lower-left corner plus
 * (ii, jj) rangeX / (nx - 1)).  The grid's y and A (device) are in point
 * order.  nx, ny < 65536. */
enum { SLSIM_GRID_SQUARE = 0, SLSIM_GRID_RECT = 1 };
typedef struct {
  size_t nx, ny;
  double center[2];
  double range_x, range_y;   /* full widths (range_y unused for a square grid) */
  int32_t layout;
  int32_t pad;
} slsim_grid;
/* slsim_gridmap_find_images on any GridMap grid (GridMap::find_images loops
 * rows j < Ngrid_init2 - 1 and cells i < Ngrid_init - 1, gridmap.cpp:965-1032). */
int slsim_grid_find_images(slsim_ctx ctx, const slsim_grid* grid, const double* y_dev,
                           const double* A_dev, size_t nsrc, const double* ys,
                           size_t max_per_source, int32_t* multiplicity, double* rays,
                           int64_t* triangles, size_t* nimages);

/* ---------------------------------------- critical curves and contours */
/* Curve types, CritType's values (include/point.h:35). */
enum {
  SLSIM_CRIT_ND = 0,
  SLSIM_CRIT_TANGENTIAL = 1,
  SLSIM_CRIT_RADIAL = 2,
  SLSIM_CRIT_PSEUDO = 3
};
/* Curves found on a grid (host memory), read with slsim_curves_read. */
typedef struct slsim_curves_s* slsim_curves;
typedef struct {
  size_t n_curves;
  size_t n_points;      /* over all curves */
  int32_t hits_replaced;/* 0: the reference leaves the caller's hits_boundary as it was */
} slsim_curves_info;

/* GridMap::find_crit(points, hits_boundary, crit_type)
 * (TreeCode_link/gridmap.cpp:525-647) over a grid shot by slsim_gridmap_shoot
 * with the same N, centre and range (y_dev, A_dev its outputs): the
 * boundaries (Utilities::find_boundaries, include/concave_hull.h:200-494) of
 * the pixels with A's smaller eigenvalue negative (tangential) and with both
 * negative (radial), each radial curve placed after the tangential curve
 * holding its first point, and for a tangential curve with no radial curve a
 * pseudo curve (the image positions of the convex hull of its island's
 * source positions).  Positions in the grid's units.  *out is released with
 * slsim_curves_destroy. */
int slsim_gridmap_find_crit(slsim_ctx ctx, size_t N, double c0, double c1, double range,
                            const double* y_dev, const double* A_dev, slsim_curves* out);
/* GridMap::find_magnification_contour(curves, hits_boundary, invmag)
 * (gridmap.cpp:493-522): the boundaries of the pixels with det A < invmag. */
int slsim_gridmap_find_magnification_contour(slsim_ctx ctx, size_t N, double c0, double c1,
                                             double range, const double* A_dev, double invmag,
                                             slsim_curves* out);
/* GridMap::find_boundaries_of_caustics(boundaries, hits_edge)
 * (include/gridmap.h:278-296): the boundaries, in pixel units, of the grid
 * positions that have more than one image when taken as sources
 * (find_images' triangle rules).  multiplicity (N*N host ints) may be NULL. */
int slsim_gridmap_find_boundaries_of_caustics(slsim_ctx ctx, size_t N, double c0, double c1,
                                              double range, const double* y_dev,
                                              int32_t* multiplicity, slsim_curves* out);
/* Utilities::find_boundaries(bitmap, nx, points, hits_edge)
 * (include/concave_hull.h:200-494) on a host bitmap of nx x ny bytes (nonzero
 * inside), on the host: the boundaries in pixel units, edge pixels counted
 * as outside, with hits_edge per curve. */
int slsim_find_boundaries(const uint8_t* bitmap, size_t nx, size_t ny, slsim_curves* out);
/* Utilities::convex_hull(P, hull) (include/concave_hull.h:560-609), on the
 * host: indices into points (2 n doubles), counter-clockwise from the
 * smallest x; hull holds 2 n entries. */
int slsim_convex_hull(const double* points, size_t n, int64_t* hull, size_t* nhull);
/* Utilities::inCurve(x, H) (include/concave_hull.h:1022-1063) for each of the
 * m points against the polygon poly (n points), on the host. */
int slsim_in_curve(const double* poly, size_t n, const double* points, size_t m,
                   int32_t* inside);
/* The three above on any GridMap grid (the bitmaps are nx wide). */
int slsim_grid_find_crit(slsim_ctx ctx, const slsim_grid* grid, const double* y_dev,
                         const double* A_dev, slsim_curves* out);
int slsim_grid_find_magnification_contour(slsim_ctx ctx, const slsim_grid* grid,
                                          const double* A_dev, double invmag,
                                          slsim_curves* out);
int slsim_grid_find_boundaries_of_caustics(slsim_ctx ctx, const slsim_grid* grid,
                                           const double* y_dev, int32_t* multiplicity,
                                           slsim_curves* out);
/* Utilities::TightestHull(v) (TreeCode_link/curve_routines.cpp:3519-3592)
 * of the closed curve points (2 n doubles): the outer boundary of the pixels
 * (1/100 of the curve's narrower extent) within 1.58 pixels of the curve,
 * each boundary point moved to the nearest point of the curve; n <= 3 gives
 * the points, a curve of zero width its box corners.  The distance passes run
 * on the device.  *out holds one curve. */
int slsim_tightest_hull(slsim_ctx ctx, const double* points, size_t n, slsim_curves* out);
/* Utilities::windings(x, points, &area) (TreeCode_link/curve_routines.cpp:
 * 1709-1736), on the host: the winding number of the closed curve (2 n
 * doubles) about x and its absolute area. */
int slsim_windings(const double* x, const double* points, size_t n, int32_t* winding,
                   double* area);
int slsim_curves_get_info(slsim_curves curves, slsim_curves_info* info);
/* offsets: n_curves + 1 (points of curve c are offsets[c] .. offsets[c+1]),
 * points: 2 n_points, types and hits_boundary: n_curves each. */
int slsim_curves_read(slsim_curves curves, int64_t* offsets, double* points, int32_t* types,
                      int32_t* hits_boundary);
int slsim_curves_destroy(slsim_curves curves);

/* ------------------------------------------------ adaptive grid (Grid) */
/* Grid (TreeCode_link/grid_maintenance.cpp:22-134, include/grid_maintenance.h:
 * 24-240): an image-plane grid refined cell by cell.  Its points (image and
 * source positions, magnification matrix, cell size) live in device memory in
 * the reference's order: the initial Nx x Ny grid, then each refinement's new
 * points appended.  Point ids are indices into that order. */
typedef struct slsim_agrid_s* slsim_agrid;
typedef struct slsim_server_s* slsim_server;   /* the request loop, below */
typedef struct {
  int64_t npoints;
  int64_t ngrid_init, ngrid_init2, ngrid_block;
  double image_p1[2], image_p2[2];     /* the image points' root box */
  double source_p1[2], source_p2[2];   /* the source points' root box (with buffer) */
} slsim_agrid_info;
/* Grid(lens, N1d, center, range) (square != 0: grid_maintenance.cpp:22-59,
 * points by Utilities::PositionFromIndex) or Grid(lens, Nx, center, rangeX,
 * rangeY) (square == 0: :67-129, Ny = (int)(Nx rangeY / rangeX)): the points
 * shot to the lens's source at `precision`.  The grid uses the lens until
 * slsim_agrid_destroy. */
int slsim_agrid_create(slsim_lens lens, size_t Nx, const double* center, double range_x,
                       double range_y, int square, int precision, slsim_agrid* out);
int slsim_agrid_destroy(slsim_agrid g);
int slsim_agrid_get_info(slsim_agrid g, slsim_agrid_info* info);
/* Host copies of points first .. first+count-1 (x, y: 2 a point; A: 4;
 * gridsize: 1); any output may be NULL. */
int slsim_agrid_read(slsim_agrid g, int64_t first, int64_t count, double* x, double* y,
                     double* A, double* gridsize);
/* The device arrays (valid until the next refinement): no copy. */
int slsim_agrid_device(slsim_agrid g, double** x_dev, double** y_dev, double** A_dev,
                       double** gridsize_dev);
/* Grid::RefineLeaves(lens, points) (grid_maintenance.cpp:681-982): every
 * listed cell (distinct ids) split into its eight sub-cell points, those
 * outside the initial image box or below 1e-10 of their position dropped,
 * the rest shot in one batch and those whose source position leaves the
 * source box dropped, in the reference's order; the parents' cell sizes / 3
 * (restored when none of their points stayed).  New points are ids
 * *first .. *first + *count - 1. */
int slsim_agrid_refine(slsim_agrid g, size_t n, const int64_t* ids, int64_t* first,
                       int64_t* count);
/* Grid::RefineLeaf(lens, point) (grid_maintenance.cpp:543-679): one cell,
 * the reference's swap order for the points it drops. */
int slsim_agrid_refine_one(slsim_agrid g, int64_t id, int64_t* first, int64_t* count);
/* Grid::EinsteinArea (grid_maintenance.cpp:368-378): the sum of gridsize^2
 * over the points with det A < 0 (a device reduction). */
int slsim_agrid_einstein_area(slsim_agrid g, double* area);
/* No counterpart in the reference, whose refinements call rayshooterInternal
 * under a process-wide mutex (grid_maintenance.cpp:808-811): with a server
 * set (NULL: none), each refinement's rays go as one ray-list request of the
 * server (slsim_server_submit_rays), batched with other clients' requests;
 * the points are the same bit for bit.  The server must outlive the grid's
 * use of it. */
int slsim_agrid_set_server(slsim_agrid g, slsim_server server);

/* --------------------------------------------------- batched GridMaps */
/* Batching and serving have no counterpart in the reference, which shoots
 * one GridMap at a time under a process-wide mutex (gridmap.cpp:15,116-119);
 * per-request results are those of slsim_gridmap_shoot. */
typedef struct slsim_batch_s* slsim_batch;
typedef struct {
  int32_t tile_log2;      /* bundle tile 2^t x 2^(6-t) rays: 3 = 8x8, 2 = 4x16, 4 = 16x4 */
  int32_t queue;          /* 0: one wavefront per work item; 1: persistent grid, work queue */
  int32_t order;          /* 0: request order; 1: requests interleaved bundle by bundle;
                             2: requests on the largest planes first                      */
  int32_t graph;          /* 1: replay the whole launch sequence as a captured HIP graph  */
  int32_t blocks_per_cu;  /* persistent grid: blocks per compute unit (queue = 1)         */
  int32_t fuse;           /* 1: one launch in which each wavefront takes its bundle
                             through all of its plane steps (no launch boundary between
                             steps; order and queue apply to the bundles)              */
} slsim_schedule;

/* n independent GridMap requests (lens_i, N_i, center_i (2 per request),
 * range_i, source z_i (<= 0: the lens source)) shot together: one launch per
 * plane step covers every request that still has a plane in front of its
 * source.  Outputs stay in device buffers owned by the batch. */
int slsim_batch_create(slsim_ctx ctx, int n, const slsim_lens* lenses, const int64_t* N,
                       const double* center, const double* range, const double* zs,
                       int precision, slsim_batch* out);
int slsim_batch_set_schedule(slsim_batch b, const slsim_schedule* s);
int slsim_batch_get_schedule(slsim_batch b, slsim_schedule* s);
int slsim_batch_run(slsim_batch b, slsim_shoot_stats* stats);
/* Device pointers of request i's y (2 N^2) and A (4 N^2), grid order. */
int slsim_batch_output(slsim_batch b, int i, double** y_dev, double** A_dev);
/* Autotuner: times schedules (greedy over fused steps, tile, queue, order,
 * queue grid size; graphs on) and keeps the fastest.  base_ms: the fixed
 * schedule {8x8, one launch per step, one wave per item, request order, no
 * graph}; best_ms: the chosen. */
int slsim_batch_autotune(slsim_batch b, int repeats, slsim_schedule* best, double* base_ms,
                         double* best_ms);
int slsim_batch_destroy(slsim_batch b);

/* ----------------------------------------------------------- serving */
typedef struct {
  double t_submit, t_admit, t_done, t_collect;   /* slsim_now() seconds            */
  int64_t rays;
  int32_t iterations;                            /* plane steps the request took   */
  int32_t status;                                /* SLSIM_OK or the failure code   */
} slsim_request_info;
typedef struct {
  int64_t submitted, admitted, rejected, completed, iterations, max_active;
  int64_t in_use_bytes, budget_bytes;
  double busy_s;                                 /* time with requests in flight   */
} slsim_server_stats;

/* Request loop with continuous batching: a scheduler thread keeps the
 * lenses' trees resident, admits queued GridMap requests while their ray
 * state and outputs fit in hbm_budget_bytes, and runs one plane step of
 * every admitted request per launch, so requests that arrive while others
 * are on the GPU join at the next step.  Each result equals the request's
 * slsim_gridmap_shoot result bit for bit. */
int slsim_server_create(slsim_ctx ctx, size_t hbm_budget_bytes, slsim_server* out);
/* Queues a request; SLSIM_ERR_BUDGET if it can never fit the budget. */
int slsim_server_submit(slsim_server s, slsim_lens lens, int64_t N, double c0, double c1,
                        double range, double zs, int precision, uint64_t* id);
/* Blocks until request `id` is done, copies y (2 N^2) and A (4 N^2) to the
 * host buffers (either may be NULL) and releases its device memory. */
/* Lens::rayshooterInternal(Npoints, Point*) and (Npoints, LinkedPoint*,
 * source_zs) (FullRange/internal_rayshooter_multi.cpp:53-251) as a served
 * request: the callers that shoot many small batches (adaptive Grid
 * refinement, critical curves: TreeCode_link/grid_maintenance.cpp:45-822,
 * find_crit.cpp:1072,1118) join the continuous batching.  n rays at x (2 per
 * ray, radians) with per-ray source redshifts zs (n values, or null for the
 * lens's source), both copied at submit; slsim_server_wait then returns y (2 n)
 * and A (4 n), bitwise those of slsim_shoot. */
int slsim_server_submit_rays(slsim_server s, slsim_lens lens, size_t n, const double* x,
                             const double* zs, int precision, uint64_t* id);
int slsim_server_wait(slsim_server s, uint64_t id, double* y_host, double* A_host,
                      slsim_request_info* info);
/* Non-blocking: *done = 1 once request `id` has finished (collect it with
 * slsim_server_wait, which then returns at once), 0 while it is queued or
 * running. */
int slsim_server_poll(slsim_server s, uint64_t id, int* done);
int slsim_server_get_stats(slsim_server s, slsim_server_stats* st);
int slsim_server_destroy(slsim_server s);
/* Monotonic clock (seconds) of the request timestamps. */
double slsim_now(void);

/* ------------------------------------------------------ smoothing sizes */
/* Replaces LensHaloParticles<PType>::calculate_smoothing(int Nsmooth, PType* pp,
 * size_t Npoints, bool verbose) (include/particle_halo.h:683-730), which sets
 * pp[i].Size = TreeSimple::NNDistance(x_i, Nsmooth + 1) (include/simpleTree.h:
 * 483-620): size_out[i] = distance from particle i to its (nsmooth+1)-th
 * nearest particle in 3-D, itself included, with the differences taken in the
 * coordinate type (double or float) and squared in double, as the reference
 * does.  Positions are host arrays of n; nsmooth in [1, 119].  Errors:
 * SLSIM_ERR_ARG when n <= nsmooth + 1 (the reference exits: "number of
 * neighbors > total number of particles"). */
int slsim_smoothing_sizes_f64(slsim_ctx ctx, const double* x, const double* y, const double* z,
                              size_t n, int nsmooth, float* size_out);
int slsim_smoothing_sizes_f32(slsim_ctx ctx, const float* x, const float* y, const float* z,
                              size_t n, int nsmooth, float* size_out);

/* ------------------------------------------------------- device buffers */
int slsim_device_alloc(slsim_ctx ctx, size_t bytes, void** dev);
int slsim_device_free(slsim_ctx ctx, void* dev);
int slsim_memcpy_d2h(slsim_ctx ctx, void* host, const void* dev, size_t bytes);
int slsim_memcpy_h2d(slsim_ctx ctx, void* dev, const void* host, size_t bytes);
/* Device-to-device copy on the context's stream (the HIP runtime's copy
 * kernel: a known-size reference for HBM traffic counters). */
int slsim_memcpy_d2d(slsim_ctx ctx, void* dst, const void* src, size_t bytes);

/* ---------------------------------------------------------- cosmology */
/* Flat LCDM comoving distance between z1 and z2 (Mpc), the role of
 * COSMOLOGY::coorDist in CosmoLib (MultiPlane/lens.cpp:776, 2489, 2661). */
double slsim_coordist(double Om, double OL, double h, double z1, double z2);
/* Inverse of coorDist(0, z): the redshift at comoving distance D. */
double slsim_inv_coordist(double Om, double OL, double h, double D);

/* ------------------------------------------------- synthetic particles */
/* r200 (physical Mpc) of mass M at redshift z: M = 800/3 pi r^3 rho_crit(z). */
double slsim_nfw_r200(double mass, double z, double Om, double OL, double h);
/* n particle positions of one projected NFW halo (concentration c, truncated
 * at r200) centred on (xc, yc), physical Mpc, from a seeded mt19937_64. */
int slsim_gen_nfw(uint64_t seed, size_t n, double mass, double conc, double z, double Om,
                  double OL, double h, double xc, double yc, double* x, double* y);
/* Superposition of n_halos projected NFW halos with masses log-uniform in
 * [Mmin, Mmax], c = 10 (M/1e12)^-0.1 10^(0.1 N(0,1)), centres uniform in a
 * square of side field_arcsec at angular distance D_A(z); n equal-mass
 * particles split by mass with largest remainders; sizes = softening. */
int slsim_gen_halo_field(uint64_t seed, size_t n, int n_halos, double Mmin, double Mmax,
                         double field_arcsec, double z, double Om, double OL, double h,
                         float softening, double* x, double* y, float* m, float* s);

/* ------------------------------------------------------- particle files */
/* ParticleType<float> (include/particle_types.h:46-58): the record of
 * MakeParticleLenses::data and of glmb files, 24 bytes. */
typedef struct {
  float x[3];
  float Mass;
  float Size;
  int32_t type;
} slsim_particle_f32;

/* SimFileFormat (include/particle_types.h:32-43), same order. */
enum {
  SLSIM_FMT_GLMB = 0,
  SLSIM_FMT_CSV3 = 1,
  SLSIM_FMT_CSV4 = 2,
  SLSIM_FMT_CSV5 = 3,
  SLSIM_FMT_CSV6 = 4,
  SLSIM_FMT_GADGET2 = 5,
  SLSIM_FMT_ASCII = 6
};

/* MakeParticleLenses (include/particle_halo.h:964-1080, MultiPlane/
 * particle_lens.cpp): the particles of one simulation file, grouped by type. */
typedef struct slsim_mpl_s* slsim_mpl;
typedef struct {
  size_t nparticles[6];         /* getnp()                                        */
  float masses[6];              /* getmp(): per-type mass, 0 = individual masses  */
  double z_original;            /* sim_redshift(); -1 for csv files               */
  double cm[3];                 /* getCenterOfMass()                              */
  double bbox_ll[3], bbox_ur[3];/* getBoundingBox()                               */
  int32_t nsmooth;
  int32_t pad;
} slsim_mpl_info;

/* MakeParticleLenses(filename, format, Nsmooth, recenter, ignore_type_in_
 * smoothing) (particle_lens.cpp:22-107): a glmb file is read as is; otherwise
 * the cache "<filename>_S<Nsmooth>.glmb" is read when it exists, else the
 * csv3/csv4/csv5 table or Gadget-2 snapshot is read, smoothing sizes are
 * computed on the device per particle type (csv3, csv4, gadget2) and the cache
 * is written.  Then the centre of mass and bounding box; recenter subtracts the
 * centre of mass.  csv6 and ascii fail with SLSIM_ERR_ARG as in the reference
 * (its csv6 case falls through to the unsupported-format branch).  Errors:
 * SLSIM_ERR_ARG (too few columns, a field stof rejects, bad format),
 * SLSIM_ERR_IO (missing or malformed file). */
int slsim_mpl_create(slsim_ctx ctx, const char* filename, int format, int nsmooth, int recenter,
                     int ignore_type_in_smoothing, slsim_mpl* out);
/* MakeParticleLenses(filename, recenter) for a glmb file (particle_lens.cpp:
 * 109-139; its bounding box starts from the origin, as there). */
int slsim_mpl_create_glmb(slsim_ctx ctx, const char* filename, int recenter, slsim_mpl* out);
int slsim_mpl_destroy(slsim_mpl m);
/* MakeParticleLenses::data: the library's particle array (valid until the next
 * call that changes the particle count, or destroy); writable. */
int slsim_mpl_data(slsim_mpl m, slsim_particle_f32** data, size_t* n);
int slsim_mpl_get_info(slsim_mpl m, slsim_mpl_info* info);
/* Recenter(Point_3d x) (particle_lens.cpp:139-151). */
int slsim_mpl_recenter(slsim_mpl m, const double x[3]);
/* radialCut / cylindricalCut (particle_lens.cpp:447-478, 497-528): drops particles
 * farther than radius, then sorts by type when there is more than one. */
int slsim_mpl_radial_cut(slsim_mpl m, const double center[3], double radius);
int slsim_mpl_cylindrical_cut(slsim_mpl m, const double center[2], double radius);
/* densest_particle() (particle_lens.cpp:480-495): max of Mass/Size^3. */
int slsim_mpl_densest_particle(slsim_mpl m, double x[3]);
/* The unit conversion of CreateHalos(cosmo, redshift, inv_area) (particle_lens.
 * cpp:153-170): positions, sizes and masses times 1/h (h rounded to float, as
 * there).  The caller then builds one plane per type with nparticles > 0 over
 * data[skip, skip + nparticles[i]) (slsim_plane_create_f32, theta 0.1, bucket
 * 20, inv_area), as CreateHalos does with LensHaloParticles. */
int slsim_mpl_physical_units(slsim_mpl m, float hubble);

/* MakeParticleLenses::writeSizesB / readSizesB (particle_halo.h:1037-1073):
 * int32 Nsmooth, uint64 n, uint64 numbytype[6], double redshift, then n
 * slsim_particle_f32 records, native byte order. */
int slsim_glmb_write(const char* filename, const slsim_particle_f32* p, size_t n, int nsmooth,
                     const size_t numbytype[6], double redshift);
int slsim_glmb_read_header(const char* filename, int* nsmooth, size_t* n, size_t numbytype[6],
                           double* redshift);
int slsim_glmb_read(const char* filename, slsim_particle_f32* p, size_t n);

/* GadgetFile<ParticleType<float>> (include/gadget.hh) as readGadget2 uses it:
 * constructor + openFile (gadget.hh:83-268, read_gadget_head :296-330),
 * readBlock("POS") and readBlock("MASS") (:567-610).  SnapFormat 2 (each block
 * behind an 8-byte label record), either byte order, one file (a snapshot
 * split over "<name>.0", "<name>.1", ... is refused).  Particles in file
 * units, type from the block order, Mass from masstab or the MASS block.
 * Errors: SLSIM_ERR_IO (no HEAD/POS/MASS block, sizes that disagree with
 * the header; the reference exits). */
int slsim_gadget2_header(const char* filename, int32_t npart[6], double masstab[6], double* time,
                         double* redshift, double* omega, double* lambda, double* hubble);
int slsim_gadget2_read(const char* filename, slsim_particle_f32* p, size_t n);

/* LensHaloParticles<PType>(simulation_filename, SimFileFormat::ascii,
 * redshift, Nsmooth, cosmo, theta_rotate, recenter, multimass, inv_area,
 * MinPSize, rescale_mass) (particle_halo.h:226-340), all but the tree:
 * readPositionFileASCII (:453-555: "# nparticles N" header, 4 columns
 * x y z mass when multimass, else 3 columns), sizes from "<file>.<Nsmooth>
 * sizes" (readSizesFile :557-632) or computed on the device and written there
 * (writeSizes :734-762, six significant digits), MinPSize, rescale_mass (all
 * particles but the first, as there), comoving -> physical (1/(1+z)), centre
 * of mass, recentring, rotation.  The caller builds the plane from the outputs
 * (slsim_plane_create_*, theta 0.1, bucket 20, inv_area).  Call _count first
 * for n. */
typedef struct {
  double mass;                  /* LensHalo::getMass()                            */
  double mcenter[3];            /* centre of mass (0 after recentring)            */
  double densest_point[3];      /* position of the smallest particle              */
  double rsize;                 /* 1e3, or the largest radius after recentring    */
} slsim_ascii_halo_info;
int slsim_ascii_halo_count(const char* filename, int multimass, size_t* n);
int slsim_ascii_halo_load_f64(slsim_ctx ctx, const char* filename, double redshift, int nsmooth,
                              const double theta_rotate[2], int recenter, int multimass,
                              double min_size, double rescale_mass, size_t n, double* xyz,
                              float* mass, float* size, slsim_ascii_halo_info* info);
int slsim_ascii_halo_load_f32(slsim_ctx ctx, const char* filename, double redshift, int nsmooth,
                              const double theta_rotate[2], int recenter, int multimass,
                              double min_size, double rescale_mass, size_t n, float* xyz,
                              float* mass, float* size, slsim_ascii_halo_info* info);

/* ------------------------------------------------------------ pixel maps */
/* PixelMap<T>::printFITS(filename, flipX, ctype) (ImageProcessing/pixelmap.cpp:
 * 1318-1388), the writer behind GridMap::writeFits (include/gridmap.h:524-531):
 * a FITS primary image (BITPIX -64 for double maps, -32 for float), nx the
 * fast axis, then the keys WCSAXES, CRPIX1/2, CROTA2, CD1_1..CD2_2, Nx, Ny,
 * "range x", RA_global, DEC_global, center_x/y, CRVAL1/2, UNITS (and CTYPE*,
 * RADESYS, CUNIT* when ctype).  Written from the FITS standard (the reference
 * uses cfitsio): keys longer than 8 characters or with spaces use the HIERARCH
 * convention.  A leading '!' in filename is dropped; existing files are
 * replaced.  units: PixelMapUnits (ndef, surfb, count_per_sec, ADU, mass,
 * mass_density). */
int slsim_fits_write_pixelmap(const char* filename, const void* map, int is_double, size_t nx,
                              size_t ny, double resolution, const double center[2], double ra,
                              double dec, int units, int flipx, int ctype);

#ifdef __cplusplus
}
#endif

#endif /* SLSIM_AMD_H */
