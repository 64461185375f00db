"""ctypes declarations of include/slsim_amd.h and loading of libslsim_amd.so."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libslsim_amd.so")

SLSIM_OK = 0
SLSIM_ERR_ARG = 1
SLSIM_ERR_HIP = 2
SLSIM_ERR_BUDGET = 3
SLSIM_ERR_SAME_PLACE = 4
SLSIM_ERR_TOO_DEEP = 5
SLSIM_ERR_NOT_FLAT = 6
SLSIM_ERR_STATE = 7
SLSIM_ERR_IO = 8

SLSIM_FP64 = 0
SLSIM_FP32C = 1
SLSIM_DIRECT64 = 2

ERROR_NAMES = {
    1: "SLSIM_ERR_ARG", 2: "SLSIM_ERR_HIP", 3: "SLSIM_ERR_BUDGET", 4: "SLSIM_ERR_SAME_PLACE",
    5: "SLSIM_ERR_TOO_DEEP", 6: "SLSIM_ERR_NOT_FLAT", 7: "SLSIM_ERR_STATE", 8: "SLSIM_ERR_IO",
}

# SimFileFormat (include/particle_types.h:32-43)
SLSIM_FMT_GLMB, SLSIM_FMT_CSV3, SLSIM_FMT_CSV4, SLSIM_FMT_CSV5, SLSIM_FMT_CSV6, \
    SLSIM_FMT_GADGET2, SLSIM_FMT_ASCII = range(7)


class SlsimError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (ERROR_NAMES.get(code, str(code)), msg))
        self.code = code


if not os.path.exists(LIB_PATH):
    raise ImportError(
        "libslsim_amd.so is not built (%s); run `python slsimlib_amd/build.py` or "
        "__graft_entry__.build()" % LIB_PATH)

lib = ctypes.CDLL(LIB_PATH)

slsim_ctx = ctypes.c_void_p
slsim_plane = ctypes.c_void_p
slsim_lens = ctypes.c_void_p


class slsim_particle_f32(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float * 3), ("Mass", ctypes.c_float), ("Size", ctypes.c_float),
                ("type", ctypes.c_int32)]


class slsim_mpl_info(ctypes.Structure):
    _fields_ = [("nparticles", ctypes.c_size_t * 6), ("masses", ctypes.c_float * 6),
                ("z_original", ctypes.c_double), ("cm", ctypes.c_double * 3),
                ("bbox_ll", ctypes.c_double * 3), ("bbox_ur", ctypes.c_double * 3),
                ("nsmooth", ctypes.c_int32), ("pad", ctypes.c_int32)]


class slsim_ascii_halo_info(ctypes.Structure):
    _fields_ = [("mass", ctypes.c_double), ("mcenter", ctypes.c_double * 3),
                ("densest_point", ctypes.c_double * 3), ("rsize", ctypes.c_double)]


class slsim_plane_info(ctypes.Structure):
    _fields_ = [("n_particles", ctypes.c_int64), ("n_nodes", ctypes.c_int64),
                ("n_leaves", ctypes.c_int64), ("n_big_internal", ctypes.c_int64),
                ("depth", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("device_bytes", ctypes.c_int64), ("build_ms", ctypes.c_double),
                ("root_p1", ctypes.c_double * 2), ("root_p2", ctypes.c_double * 2)]


class slsim_shoot_stats(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("walk_ms", ctypes.c_double),
                ("walk_ms_max_plane", ctypes.c_double), ("n_walk_launches", ctypes.c_int32),
                ("n_guarded", ctypes.c_int32)]


slsim_agrid = ctypes.c_void_p


class slsim_agrid_info(ctypes.Structure):
    _fields_ = [("npoints", ctypes.c_int64), ("ngrid_init", ctypes.c_int64),
                ("ngrid_init2", ctypes.c_int64), ("ngrid_block", ctypes.c_int64),
                ("image_p1", ctypes.c_double * 2), ("image_p2", ctypes.c_double * 2),
                ("source_p1", ctypes.c_double * 2), ("source_p2", ctypes.c_double * 2)]


class slsim_schedule(ctypes.Structure):
    _fields_ = [("tile_log2", ctypes.c_int32), ("queue", ctypes.c_int32),
                ("order", ctypes.c_int32), ("graph", ctypes.c_int32),
                ("blocks_per_cu", ctypes.c_int32), ("fuse", ctypes.c_int32)]


class slsim_request_info(ctypes.Structure):
    _fields_ = [("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_done", ctypes.c_double), ("t_collect", ctypes.c_double),
                ("rays", ctypes.c_int64), ("iterations", ctypes.c_int32),
                ("status", ctypes.c_int32)]


class slsim_server_stats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("admitted", ctypes.c_int64),
                ("rejected", ctypes.c_int64), ("completed", ctypes.c_int64),
                ("iterations", ctypes.c_int64), ("max_active", ctypes.c_int64),
                ("in_use_bytes", ctypes.c_int64), ("budget_bytes", ctypes.c_int64),
                ("busy_s", ctypes.c_double)]


slsim_batch = ctypes.c_void_p
slsim_server = ctypes.c_void_p
slsim_curves = ctypes.c_void_p

# halo kinds of slsim_halos_create
SLSIM_HALO_NFW, SLSIM_HALO_POWERLAW = 0, 1

# CritType (include/point.h:35)
SLSIM_CRIT_ND, SLSIM_CRIT_TANGENTIAL, SLSIM_CRIT_RADIAL, SLSIM_CRIT_PSEUDO = range(4)


SLSIM_GRID_SQUARE, SLSIM_GRID_RECT = 0, 1


class slsim_grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_size_t), ("ny", ctypes.c_size_t),
                ("center", ctypes.c_double * 2), ("range_x", ctypes.c_double),
                ("range_y", ctypes.c_double), ("layout", ctypes.c_int32),
                ("pad", ctypes.c_int32)]


class slsim_curves_info(ctypes.Structure):
    _fields_ = [("n_curves", ctypes.c_size_t), ("n_points", ctypes.c_size_t),
                ("hits_replaced", ctypes.c_int32)]

P = ctypes.POINTER
c_int, c_double, c_float, c_size_t = ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_size_t
c_u64, c_i64, c_vp = ctypes.c_uint64, ctypes.c_int64, ctypes.c_void_p
c_char_p = ctypes.c_char_p

SIGNATURES = {
    "slsim_last_error": (ctypes.c_char_p, []),
    "slsim_ctx_create": (c_int, [c_int, c_size_t, P(slsim_ctx)]),
    "slsim_ctx_destroy": (c_int, [slsim_ctx]),
    "slsim_ctx_memory": (c_int, [slsim_ctx, P(c_size_t), P(c_size_t)]),
    "slsim_ctx_synchronize": (c_int, [slsim_ctx]),
    "slsim_plane_create_f64": (c_int, [slsim_ctx, P(c_double), P(c_double), P(c_float), P(c_float),
                                       c_size_t, c_double, c_int, c_double, P(slsim_plane)]),
    "slsim_plane_create_f32": (c_int, [slsim_ctx, P(c_float), P(c_float), P(c_float), P(c_float),
                                       c_size_t, c_double, c_int, c_double, P(slsim_plane)]),
    "slsim_plane_create_range_f64": (c_int, [slsim_ctx, P(c_double), P(c_double), P(c_float),
                                             P(c_float), c_size_t, c_double, c_int, c_double,
                                             c_double, P(slsim_plane)]),
    "slsim_plane_create_range_f32": (c_int, [slsim_ctx, P(c_float), P(c_float), P(c_float),
                                             P(c_float), c_size_t, c_double, c_int, c_double,
                                             c_double, P(slsim_plane)]),
    "slsim_plane_destroy": (c_int, [slsim_plane]),
    "slsim_plane_get_info": (c_int, [slsim_plane, P(slsim_plane_info)]),
    "slsim_plane_export_nodes": (c_int, [slsim_plane, P(c_double), P(c_i64)]),
    "slsim_lens_create": (c_int, [slsim_ctx, c_double, c_double, c_double, c_double, P(slsim_lens)]),
    "slsim_lens_destroy": (c_int, [slsim_lens]),
    "slsim_lens_add_halo": (c_int, [slsim_lens, slsim_plane, c_float]),
    "slsim_lens_tables": (c_int, [slsim_lens, P(c_int), P(c_double), P(c_double), P(c_double),
                                  P(c_double)]),
    "slsim_lens_reset_source": (c_int, [slsim_lens, c_double]),
    "slsim_shoot": (c_int, [slsim_lens, c_size_t, P(c_double), P(c_double), P(c_double),
                            P(c_double), c_int, P(slsim_shoot_stats)]),
    "slsim_shoot_device": (c_int, [slsim_lens, c_size_t, P(c_double), P(c_double), P(c_double),
                                   P(c_double), c_int, P(slsim_shoot_stats)]),
    "slsim_shoot_device_noalloc": (c_int, [slsim_lens, c_size_t, P(c_double), P(c_double),
                                           P(c_double), P(c_double), c_int,
                                           P(slsim_shoot_stats)]),
    "slsim_gridmap_shoot": (c_int, [slsim_lens, c_size_t, c_double, c_double, c_double, c_double,
                                    P(c_double), P(c_double), c_int, P(slsim_shoot_stats)]),
    "slsim_gridmap_count": (c_int, [slsim_lens, c_size_t, c_double, c_double, c_double, c_double,
                                    P(c_u64)]),
    "slsim_gridmap_bundle_stats": (c_int, [slsim_lens, c_size_t, c_double, c_double, c_double,
                                           c_double, P(c_u64)]),
    "slsim_plane_direct_sum": (c_int, [slsim_plane, c_size_t, P(c_double), P(c_double)]),
    "slsim_plane_force": (c_int, [slsim_plane, c_size_t, P(c_double), P(c_double), c_int]),
    "slsim_batch_create": (c_int, [slsim_ctx, c_int, P(slsim_lens), P(c_i64), P(c_double),
                                   P(c_double), P(c_double), c_int, P(slsim_batch)]),
    "slsim_batch_set_schedule": (c_int, [slsim_batch, P(slsim_schedule)]),
    "slsim_batch_get_schedule": (c_int, [slsim_batch, P(slsim_schedule)]),
    "slsim_batch_run": (c_int, [slsim_batch, P(slsim_shoot_stats)]),
    "slsim_batch_output": (c_int, [slsim_batch, c_int, P(P(c_double)), P(P(c_double))]),
    "slsim_batch_autotune": (c_int, [slsim_batch, c_int, P(slsim_schedule), P(c_double),
                                     P(c_double)]),
    "slsim_batch_destroy": (c_int, [slsim_batch]),
    "slsim_server_create": (c_int, [slsim_ctx, c_size_t, P(slsim_server)]),
    "slsim_server_submit": (c_int, [slsim_server, slsim_lens, c_i64, c_double, c_double,
                                    c_double, c_double, c_int, P(c_u64)]),
    "slsim_server_poll": (c_int, [slsim_server, c_u64, P(c_int)]),
    "slsim_server_submit_rays": (c_int, [slsim_server, slsim_lens, c_size_t, P(c_double),
                                         P(c_double), c_int, P(c_u64)]),
    "slsim_server_wait": (c_int, [slsim_server, c_u64, P(c_double), P(c_double),
                                  P(slsim_request_info)]),
    "slsim_server_get_stats": (c_int, [slsim_server, P(slsim_server_stats)]),
    "slsim_server_destroy": (c_int, [slsim_server]),
    "slsim_now": (c_double, []),
    "slsim_device_alloc": (c_int, [slsim_ctx, c_size_t, P(c_vp)]),
    "slsim_device_free": (c_int, [slsim_ctx, c_vp]),
    "slsim_memcpy_d2h": (c_int, [slsim_ctx, c_vp, c_vp, c_size_t]),
    "slsim_memcpy_h2d": (c_int, [slsim_ctx, c_vp, c_vp, c_size_t]),
    "slsim_memcpy_d2d": (c_int, [slsim_ctx, c_vp, c_vp, c_size_t]),
    "slsim_smoothing_sizes_f64": (c_int, [slsim_ctx, P(c_double), P(c_double), P(c_double),
                                          c_size_t, c_int, P(c_float)]),
    "slsim_smoothing_sizes_f32": (c_int, [slsim_ctx, P(c_float), P(c_float), P(c_float),
                                          c_size_t, c_int, P(c_float)]),
    "slsim_gridmap_find_images": (c_int, [slsim_ctx, c_size_t, c_double, c_double, c_double, c_vp,
                                          c_vp, c_size_t, P(c_double), c_size_t,
                                          P(ctypes.c_int32), P(c_double), P(ctypes.c_int64),
                                          P(c_size_t)]),
    "slsim_gridmap_find_crit": (c_int, [slsim_ctx, c_size_t, c_double, c_double, c_double, c_vp,
                                        c_vp, P(slsim_curves)]),
    "slsim_gridmap_find_magnification_contour": (c_int, [slsim_ctx, c_size_t, c_double, c_double,
                                                         c_double, c_vp, c_double,
                                                         P(slsim_curves)]),
    "slsim_gridmap_find_boundaries_of_caustics": (c_int, [slsim_ctx, c_size_t, c_double, c_double,
                                                          c_double, c_vp, P(ctypes.c_int32),
                                                          P(slsim_curves)]),
    "slsim_grid_find_images": (c_int, [slsim_ctx, P(slsim_grid), c_vp, c_vp, c_size_t,
                                       P(c_double), c_size_t, P(ctypes.c_int32), P(c_double),
                                       P(ctypes.c_int64), P(c_size_t)]),
    "slsim_grid_find_crit": (c_int, [slsim_ctx, P(slsim_grid), c_vp, c_vp, P(slsim_curves)]),
    "slsim_grid_find_magnification_contour": (c_int, [slsim_ctx, P(slsim_grid), c_vp, c_double,
                                                      P(slsim_curves)]),
    "slsim_grid_find_boundaries_of_caustics": (c_int, [slsim_ctx, P(slsim_grid), c_vp,
                                                       P(ctypes.c_int32), P(slsim_curves)]),
    "slsim_find_boundaries": (c_int, [P(ctypes.c_uint8), c_size_t, c_size_t, P(slsim_curves)]),
    "slsim_convex_hull": (c_int, [P(c_double), c_size_t, P(c_i64), P(c_size_t)]),
    "slsim_in_curve": (c_int, [P(c_double), c_size_t, P(c_double), c_size_t, P(ctypes.c_int32)]),
    "slsim_tightest_hull": (c_int, [slsim_ctx, P(c_double), c_size_t, P(slsim_curves)]),
    "slsim_agrid_create": (c_int, [slsim_lens, c_size_t, P(c_double), c_double, c_double, c_int,
                                   c_int, P(slsim_agrid)]),
    "slsim_agrid_destroy": (c_int, [slsim_agrid]),
    "slsim_agrid_get_info": (c_int, [slsim_agrid, P(slsim_agrid_info)]),
    "slsim_agrid_read": (c_int, [slsim_agrid, c_i64, c_i64, P(c_double), P(c_double),
                                 P(c_double), P(c_double)]),
    "slsim_agrid_device": (c_int, [slsim_agrid, P(c_vp), P(c_vp), P(c_vp),
                                   P(c_vp)]),
    "slsim_agrid_refine": (c_int, [slsim_agrid, c_size_t, P(c_i64), P(c_i64), P(c_i64)]),
    "slsim_agrid_refine_one": (c_int, [slsim_agrid, c_i64, P(c_i64), P(c_i64)]),
    "slsim_agrid_einstein_area": (c_int, [slsim_agrid, P(c_double)]),
    "slsim_agrid_set_server": (c_int, [slsim_agrid, c_vp]),
    "slsim_windings": (c_int, [P(c_double), P(c_double), c_size_t, P(ctypes.c_int32),
                               P(c_double)]),
    "slsim_curves_get_info": (c_int, [slsim_curves, P(slsim_curves_info)]),
    "slsim_curves_read": (c_int, [slsim_curves, P(c_i64), P(c_double), P(ctypes.c_int32),
                                  P(ctypes.c_int32)]),
    "slsim_curves_destroy": (c_int, [slsim_curves]),
    "slsim_plane_tree_create_nfw": (c_int, [slsim_ctx, P(c_double), P(c_double), P(c_float),
                                            P(c_float), P(c_float), c_size_t, c_double, c_double,
                                            P(slsim_plane)]),
    "slsim_lens_add_field_plane": (c_int, [slsim_lens, slsim_plane, c_double]),
    "slsim_lens_clear_main_halos": (c_int, [slsim_lens]),
    "slsim_lens_add_halo_nearest": (c_int, [slsim_lens, slsim_plane, c_float]),
    "slsim_lens_reset_source_plane": (c_int, [slsim_lens, c_double, c_int, P(ctypes.c_int32),
                                              P(c_double)]),
    "slsim_halos_create": (c_int, [slsim_ctx, P(c_double), P(c_double), P(c_float), P(c_float),
                                   P(ctypes.c_int32), P(c_double), c_size_t, c_double, c_double,
                                   c_double, c_int, P(slsim_plane)]),
    "slsim_halos_create_nfw": (c_int, [slsim_ctx, P(c_double), P(c_double), P(c_float),
                                       P(c_float), P(c_float), c_size_t, c_float, c_float,
                                       P(slsim_plane)]),
    "slsim_mpl_create": (c_int, [slsim_ctx, c_char_p, c_int, c_int, c_int, c_int, P(c_vp)]),
    "slsim_mpl_create_glmb": (c_int, [slsim_ctx, c_char_p, c_int, P(c_vp)]),
    "slsim_mpl_destroy": (c_int, [c_vp]),
    "slsim_mpl_data": (c_int, [c_vp, P(P(slsim_particle_f32)), P(c_size_t)]),
    "slsim_mpl_get_info": (c_int, [c_vp, P(slsim_mpl_info)]),
    "slsim_mpl_recenter": (c_int, [c_vp, P(c_double)]),
    "slsim_mpl_radial_cut": (c_int, [c_vp, P(c_double), c_double]),
    "slsim_mpl_cylindrical_cut": (c_int, [c_vp, P(c_double), c_double]),
    "slsim_mpl_densest_particle": (c_int, [c_vp, P(c_double)]),
    "slsim_mpl_physical_units": (c_int, [c_vp, c_float]),
    "slsim_glmb_write": (c_int, [c_char_p, c_vp, c_size_t, c_int, P(c_size_t), c_double]),
    "slsim_glmb_read_header": (c_int, [c_char_p, P(c_int), P(c_size_t), P(c_size_t),
                                       P(c_double)]),
    "slsim_glmb_read": (c_int, [c_char_p, c_vp, c_size_t]),
    "slsim_gadget2_header": (c_int, [c_char_p, P(ctypes.c_int32), P(c_double), P(c_double),
                                     P(c_double), P(c_double), P(c_double), P(c_double)]),
    "slsim_gadget2_read": (c_int, [c_char_p, c_vp, c_size_t]),
    "slsim_ascii_halo_count": (c_int, [c_char_p, c_int, P(c_size_t)]),
    "slsim_ascii_halo_load_f64": (c_int, [slsim_ctx, c_char_p, c_double, c_int, P(c_double), c_int,
                                          c_int, c_double, c_double, c_size_t, P(c_double),
                                          P(c_float), P(c_float), P(slsim_ascii_halo_info)]),
    "slsim_ascii_halo_load_f32": (c_int, [slsim_ctx, c_char_p, c_double, c_int, P(c_double), c_int,
                                          c_int, c_double, c_double, c_size_t, P(c_float),
                                          P(c_float), P(c_float), P(slsim_ascii_halo_info)]),
    "slsim_fits_write_pixelmap": (c_int, [c_char_p, c_vp, c_int, c_size_t, c_size_t, c_double,
                                          P(c_double), c_double, c_double, c_int, c_int, c_int]),
    "slsim_coordist": (c_double, [c_double, c_double, c_double, c_double, c_double]),
    "slsim_inv_coordist": (c_double, [c_double, c_double, c_double, c_double]),
    "slsim_nfw_r200": (c_double, [c_double, c_double, c_double, c_double, c_double]),
    "slsim_gen_nfw": (c_int, [c_u64, c_size_t, c_double, c_double, c_double, c_double, c_double,
                              c_double, c_double, c_double, P(c_double), P(c_double)]),
    "slsim_gen_halo_field": (c_int, [c_u64, c_size_t, c_int, c_double, c_double, c_double,
                                     c_double, c_double, c_double, c_double, c_float, P(c_double),
                                     P(c_double), P(c_float), P(c_float)]),
}

for _name, (_res, _args) in SIGNATURES.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def check(rc):
    if rc != SLSIM_OK:
        raise SlsimError(rc, lib.slsim_last_error().decode(errors="replace"))
