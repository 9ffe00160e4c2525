"""ctypes bindings for libsundog.so (the gfx950 build).

The library is built in-tree by sundog_amd/build.py. Loading fails loudly
when it is missing: there is no CPU fallback for any entry point.
"""

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_LIB = os.path.join(HERE, "lib", "libsundog.so")

c_d = ctypes.c_double
c_i = ctypes.c_int
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_u64 = ctypes.c_uint64
c_f = ctypes.c_float
c_p = ctypes.c_void_p
c_s = ctypes.c_char_p
P_d = ctypes.POINTER(ctypes.c_double)
P_f = ctypes.POINTER(ctypes.c_float)
P_i32 = ctypes.POINTER(ctypes.c_int32)
P_u32 = ctypes.POINTER(ctypes.c_uint32)


class XformStep(ctypes.Structure):
    """sundog_xform_step (include/sundog_api.h)."""
    _fields_ = [("kind", c_i32), ("a", c_d), ("b", c_d), ("c", c_d)]


class PhysicsBody(ctypes.Structure):
    _fields_ = [("dynamic", c_i32), ("density", c_d), ("has_velocity", c_i32),
                ("has_angular_velocity", c_i32), ("velocity", c_d * 3),
                ("angular_velocity", c_d * 3), ("friction", c_d),
                ("restitution", c_d), ("thickness", c_d)]


class RenderOptions(ctypes.Structure):
    _fields_ = [("out_path", c_s), ("stats_path", c_s), ("aov_albedo_path", c_s),
                ("aov_normal_path", c_s), ("scene_name", c_s), ("spp", c_i32),
                ("width", c_i32), ("height", c_i32), ("seed", c_i64),
                ("denoise", c_i32), ("transparent_shadows", c_i32), ("clamp", c_d),
                ("physics_time", c_d), ("quiet", c_i32)]


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 256), ("cc_major", c_i32), ("cc_minor", c_i32),
                ("driver_version", c_i32), ("runtime_version", c_i32),
                ("total_mem_mb", c_u64), ("optix_header_version", ctypes.c_uint32),
                ("rtcore_version", ctypes.c_uint32)]


class SessionStats(ctypes.Structure):
    _fields_ = [("rays_traced", c_u64), ("paths", c_u64), ("render_ms", c_d),
                ("build_ms", c_d), ("device_bytes", c_u64), ("stack_depth", c_i32),
                ("tlas_nodes", c_i32), ("blas_nodes", c_i32), ("pad", c_i32),
                ("node_visits", c_u64), ("tri_tests", c_u64), ("quad_tests", c_u64),
                ("inst_transforms", c_u64), ("bvh_builder", c_i32), ("pad2", c_i32), ("bvh_gpu_ms", c_d),
                ("waves", c_i32), ("stack_lds", c_i32), ("sample_groups", c_i32), ("pad3", c_i32)]


class BvhLayout(ctypes.Structure):
    _fields_ = [("num_nodes", c_i32), ("tlas_nodes", c_i32), ("num_tris", c_i32), ("num_instances", c_i32),
                ("tlas_root", ctypes.c_uint32), ("pad", c_i32)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("id", c_i64), ("status", c_i32), ("scene", c_i32), ("width", c_i32), ("height", c_i32),
                ("spp", c_i32), ("pad", c_i32), ("t_submit_ms", c_d), ("t_admit_ms", c_d),
                ("t_done_ms", c_d), ("paths", c_u64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("paths", c_u64), ("rays_traced", c_u64), ("batches", c_u64), ("work_items", c_u64),
                ("kernel_ms", c_d), ("busy_ms", c_d), ("used_bytes", c_u64), ("budget_bytes", c_u64),
                ("active", c_i32), ("waiting", c_i32), ("tile_size", c_i32), ("waves", c_i32),
                ("pass_spp", c_i32), ("blocks_per_cu", c_i32), ("stack", c_i32), ("graph_passes", c_i32),
                ("graph_launches", c_u64)]


class TuneResult(ctypes.Structure):
    _fields_ = [("tile", c_i32), ("waves", c_i32), ("stack", c_i32), ("blocks_per_cu", c_i32),
                ("graph_passes", c_i32), ("trials", c_i32), ("start_paths_per_s", c_d),
                ("best_paths_per_s", c_d)]


def default_options(**kw):
    o = RenderOptions()
    o.out_path = None
    o.stats_path = None
    o.aov_albedo_path = None
    o.aov_normal_path = None
    o.scene_name = b"(scene)"
    o.spp = o.width = o.height = -1
    o.seed = -1
    o.denoise = -1
    o.transparent_shadows = -1
    o.clamp = float("nan")
    o.physics_time = float("nan")
    o.quiet = 1
    for k, v in kw.items():
        if isinstance(v, str):
            v = v.encode()
        setattr(o, k, v)
    return o


# name -> (restype, argtypes); every symbol include/sundog_api.h declares
SIGNATURES = {
    "sundog_last_error": (c_s, []),
    "sundog_scene_create": (c_p, [c_s]),
    "sundog_scene_destroy": (None, [c_p]),
    "sundog_set_render": (c_i, [c_p, c_i, c_i, c_i, c_i, c_d, c_i64, c_d, c_i]),
    "sundog_set_camera": (c_i, [c_p, P_d, P_d, P_d, c_d, c_d, c_d]),
    "sundog_set_background_solid": (c_i, [c_p, P_d]),
    "sundog_set_background_gradient": (c_i, [c_p, P_d, P_d]),
    "sundog_set_background_envmap": (c_i, [c_p, c_s, c_d, c_d, c_i]),
    "sundog_set_physics": (c_i, [c_p, P_d, c_d, c_d, c_d, c_d, c_i, c_i, c_d, c_d]),
    "sundog_add_texture_solid": (c_i, [c_p, P_d]),
    "sundog_add_texture_checker": (c_i, [c_p, P_d, P_d, P_d]),
    "sundog_add_texture_grid": (c_i, [c_p, P_d, P_d, P_d, c_d]),
    "sundog_add_texture_image": (c_i, [c_p, c_s, c_i]),
    "sundog_add_material_lambert": (c_i, [c_p, P_d, c_i]),
    "sundog_add_material_metal": (c_i, [c_p, P_d, c_i, c_d]),
    "sundog_add_material_dielectric": (c_i, [c_p, c_d, P_d, P_d, c_i, c_d]),
    "sundog_add_material_emissive": (c_i, [c_p, P_d, c_i, c_d, c_i]),
    "sundog_add_material_water": (c_i, [c_p, c_d, P_d, c_d, c_d, P_d]),
    "sundog_add_material_plastic": (c_i, [c_p, P_d, c_i, c_d]),
    "sundog_add_mesh": (c_i, [c_p, c_s, c_i, c_s]),
    "sundog_add_object": (c_i, [c_p, c_i, c_i, c_i, c_i, c_i, ctypes.POINTER(XformStep), c_i,
                                c_i, ctypes.POINTER(PhysicsBody)]),
    "sundog_add_flame": (c_i, [c_p, P_d, c_d, c_d, c_d, c_d, c_d, c_i64, c_d]),
    "sundog_add_point_light": (c_i, [c_p, P_d, P_d, c_d]),
    "sundog_add_distant_light": (c_i, [c_p, P_d, P_d]),
    "sundog_scene_validate": (c_i, [c_p]),
    "sundog_probe": (c_i, [ctypes.POINTER(DeviceInfo)]),
    "sundog_render": (c_i, [c_p, ctypes.POINTER(RenderOptions)]),
    "sundog_set_flame_march": (c_i, [c_p, c_i, c_i, c_d, c_i]),
    "sundog_set_water_octaves": (c_i, [c_p, c_i, c_i]),
    "sundog_scene_describe": (c_i, [c_p, ctypes.c_char_p, ctypes.c_size_t,
                                    ctypes.POINTER(ctypes.c_size_t)]),
    "sundog_session_create": (c_p, [c_p, ctypes.POINTER(RenderOptions)]),
    "sundog_session_destroy": (None, [c_p]),
    "sundog_session_reset": (c_i, [c_p]),
    "sundog_session_render": (c_i, [c_p, c_i, c_i]),
    "sundog_session_set_kernel": (c_i, [c_p, c_i]),
    "sundog_session_set_count_rays": (c_i, [c_p, c_i]),
    "sundog_session_set_count_work": (c_i, [c_p, c_i]),
    "sundog_session_set_tuning": (c_i, [c_p, c_s, c_i]),
    "sundog_session_autotune": (c_i, [c_p, c_i, ctypes.POINTER(c_i)]),
    "sundog_session_sync": (c_i, [c_p]),
    "sundog_session_download": (c_i, [c_p, P_f, P_f, P_f]),
    "sundog_session_get_stats": (c_i, [c_p, ctypes.POINTER(SessionStats)]),
    "sundog_session_bvh": (c_i, [c_p, ctypes.POINTER(BvhLayout), c_p, ctypes.c_size_t, P_f, ctypes.c_size_t,
                                 P_i32, c_i]),
    "sundog_session_size": (c_i, [c_p, P_i32, P_i32, P_i32]),
    "sundog_session_stream": (c_p, [c_p]),
    "sundog_session_trace": (c_i, [c_p, P_f, c_i, c_i, P_f, P_i32, P_i32]),
    "sundog_session_shadow": (c_i, [c_p, P_f, c_i, P_f]),
    "sundog_session_flame_field": (c_i, [c_p, c_i, P_f, c_i, P_f, P_f]),
    "sundog_session_env_bins": (c_i, [c_p, P_f, c_i, P_i32, P_i32]),
    "sundog_session_env_bins_mode": (c_i, [c_p, P_f, c_i, c_i, P_i32, P_i32, P_f]),
    "sundog_session_light_samples": (c_i, [c_p, c_i, P_f, c_i, ctypes.c_uint64, ctypes.c_uint64, P_f]),
    "sundog_session_trace_noalloc": (c_i, [c_p, c_p, c_i, c_p, c_p, c_p]),
    "sundog_device_pcg32": (c_i, [c_u64, c_u64, c_i, P_u32]),
    "sundog_server_create": (c_p, [c_u64, c_i, c_i]),
    "sundog_server_destroy": (None, [c_p]),
    "sundog_server_add_scene": (c_i, [c_p, c_p, ctypes.POINTER(RenderOptions)]),
    "sundog_server_set_tuning": (c_i, [c_p, c_s, c_i]),
    "sundog_server_submit": (c_i64, [c_p, c_i, c_i, c_i, c_i, c_i64]),
    "sundog_server_poll": (c_i, [c_p, ctypes.POINTER(RequestInfo), c_i]),
    "sundog_server_fetch": (c_i, [c_p, c_i64, P_f, P_f, P_f]),
    "sundog_server_release": (c_i, [c_p, c_i64]),
    "sundog_server_wait_idle": (c_i, [c_p, c_d]),
    "sundog_server_get_stats": (c_i, [c_p, ctypes.POINTER(ServerStats)]),
    "sundog_server_autotune": (c_i, [c_p, c_i, c_i, c_i, c_i, ctypes.POINTER(TuneResult)]),
    "sundog_server_render_serial": (c_i, [c_p, c_i, c_i, c_i, c_i, c_i64, P_f, ctypes.POINTER(c_d)]),
    "sundog_scene_flatten": (c_p, [c_p, ctypes.POINTER(RenderOptions)]),
    "sundog_flat_free": (None, [c_p]),
    "sundog_pq_oetf": (c_f, [c_f]),
    "sundog_pq_eotf": (c_f, [c_f]),
    "sundog_bt709_to_2020": (None, [P_f, P_f]),
    "sundog_encode_pq12": (c_i, [P_f, c_i, c_i, c_f, ctypes.POINTER(ctypes.c_uint16)]),
    "sundog_denoiser_create": (c_p, [c_u64]),
    "sundog_denoiser_destroy": (None, [c_p]),
    "sundog_denoiser_param_count": (c_i64, [c_p]),
    "sundog_denoiser_get_params": (c_i, [c_p, P_f]),
    "sundog_denoiser_set_params": (c_i, [c_p, P_f]),
    "sundog_denoiser_forward_backward": (c_i, [c_p, P_f, P_f, c_i, c_i, c_i, P_f, P_f, P_f]),
    "sundog_denoiser_train_step": (c_i, [c_p, P_f, P_f, c_i, c_i, c_i, c_f, P_f]),
    "sundog_denoiser_infer": (c_i, [c_p, P_f, c_i, c_i, c_i, P_f]),
    "sundog_denoiser_time_forward": (c_i, [c_p, c_i, c_i, c_i, c_i, ctypes.POINTER(c_d), ctypes.POINTER(c_d)]),
    "sundog_denoiser_save": (c_i, [c_p, c_s]),
    "sundog_denoiser_load": (c_i, [c_p, c_s]),
}

_LIB = None


def load(path=None):
    """Load libsundog.so and bind every declared entry point."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    path = path or os.environ.get("SUNDOG_LIB") or DEFAULT_LIB
    if not os.path.isfile(path):
        raise RuntimeError("libsundog.so not found at %s: build it with "
                           "`python -m sundog_amd.build` (no CPU fallback exists)" % path)
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def last_error(lib=None):
    lib = lib or load()
    msg = lib.sundog_last_error()
    return (msg or b"unknown error").decode("utf-8", "replace")


class SundogError(RuntimeError):
    pass


def check(rc, lib=None):
    """Raise on the ABI's failure convention (set_*: !=0; add_*: <0)."""
    if rc is None or rc != 0:
        raise SundogError(last_error(lib))
    return rc
