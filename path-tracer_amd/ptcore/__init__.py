"""ptcore - Python binding of libptcore.so, the MI355X path-tracing core.

A thin ctypes layer over the C ABI in include/pt_api.h, mirroring the
reference's host objects around its trace kernel:

    Context  <- gpu::request_adapter/request_device   (src/render/gpu.rs:14-76)
    World    <- World::new + take_updates              (src/render/world.rs:589-884)
    Scene    <- WorldResources                         (src/render/world_resources.rs)
    Tracer   <- PathTracer::{new, set_tile, set_max_bounces, trace, display,
                sample_count}                          (src/render/path_tracer.rs:59-358)

The HIP library is required: importing works without it (so CPU-only tools
can load the structs), but every call that needs the library raises
PtcoreUnavailable when it is missing. There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PTCORE_LIB selects another build of the library (e.g. one made with different compile options)
LIB_PATH = os.environ.get("PTCORE_LIB") or os.path.join(_HERE, "libptcore.so")


class PtcoreUnavailable(RuntimeError):
    """libptcore.so is not built or cannot be loaded."""


class PtError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"ptcore error {code}: {message}")
        self.code = code


# status codes (pt_api.h)
PT_OK, PT_ERR_INVALID, PT_ERR_DEVICE, PT_ERR_MEMORY, PT_ERR_STATE, PT_ERR_TIMEOUT, PT_ERR_CANCELLED = range(7)
SCHEDULE_MEGAKERNEL, SCHEDULE_WAVEFRONT = 0, 1
SCENE_CORNELL, SCENE_FIELD, SCENE_LARGE, SCENE_SIZED = 1, 2, 3, 4
EXPOSURE = np.float32(0.00002604166)


# ---------------------------------------------------------------- layouts (pt_layout.h)
class PtCamera(C.Structure):
    _fields_ = [("position", C.c_float * 4), ("lower_left_corner_minus_position", C.c_float * 4),
                ("horizontal", C.c_float * 4), ("vertical", C.c_float * 4)]


class PtSceneUniforms(C.Structure):
    _fields_ = [("total_luminous_flux", C.c_float), ("emissive_triangles_count", C.c_uint32),
                ("light_probabilities", C.c_float * 2)]


class PtSkyboxUniforms(C.Structure):
    _fields_ = [("total_luminous_flux", C.c_float), ("average_luminance", C.c_float),
                ("width", C.c_uint32), ("height", C.c_uint32)]


class PtTraceParameters(C.Structure):
    _fields_ = [("image_size", C.c_uint32 * 2), ("tile_origin", C.c_uint32 * 2), ("max_bounces", C.c_uint32),
                ("sample_count", C.c_uint32), ("pad", C.c_uint32 * 2)]


# numpy structured dtypes of the array layouts
GEOMETRY_DTYPE = np.dtype([("v", "<f4", (9,))])
ATTRIBUTE_DTYPE = np.dtype([("material", "<u4"), ("n", "<f4", (9,)), ("uv", "<f4", (6,)),
                            ("t0", "<f4", (4,)), ("t1", "<f4", (4,)), ("t2", "<f4", (4,))])
MATERIAL_DTYPE = np.dtype([
    ("base_color", "<f4", (4,)), ("base_weight", "<f4"), ("base_metalness", "<f4"),
    ("base_diffuse_roughness", "<f4"), ("emission_weight", "<f4"), ("emission_color_and_luminance", "<f4", (4,)),
    ("specular_color_and_weight", "<f4", (4,)), ("specular_roughness", "<f4"),
    ("specular_roughness_anisotropy", "<f4"), ("specular_haze", "<f4"), ("specular_haze_spread", "<f4"),
    ("specular_ior", "<f4"), ("specular_retroreflectivity", "<f4"), ("normal_scale", "<f4"),
    ("alpha_mode", "<u4"), ("alpha_cutoff", "<f4"), ("pad", "<f4"),
    ("base_color_tex_layer", "<i4"), ("normal_tex_layer", "<i4"), ("metallic_roughness_tex_layer", "<i4"),
    ("emission_tex_layer", "<i4"), ("specular_color_tex_layer", "<i4"), ("specular_weight_tex_layer", "<i4"),
    ("base_color_uv", "<f4", (4,)), ("normal_uv", "<f4", (4,)), ("metallic_roughness_uv", "<f4", (4,)),
    ("emission_uv", "<f4", (4,)), ("specular_color_uv", "<f4", (4,)), ("specular_weight_uv", "<f4", (4,))])
BVH_DTYPE = np.dtype([("aabb_min", "<f4", (3,)), ("left_first", "<u4"), ("aabb_max", "<f4", (3,)),
                      ("tri_count", "<u4")])
SCENE_ALIAS_DTYPE = np.dtype([("threshold", "<f4"), ("alias_index", "<u4"), ("triangle_index", "<u4"),
                              ("selection_probability", "<f4")])
SKY_ALIAS_DTYPE = np.dtype([("threshold", "<f4"), ("alias_index", "<u4"), ("selection_probability", "<f4")])
assert GEOMETRY_DTYPE.itemsize == 36 and ATTRIBUTE_DTYPE.itemsize == 112 and MATERIAL_DTYPE.itemsize == 224
assert BVH_DTYPE.itemsize == 32 and SCENE_ALIAS_DTYPE.itemsize == 16 and SKY_ALIAS_DTYPE.itemsize == 12

_p = C.c_void_p


class ContextInfo(C.Structure):
    _fields_ = [("device", C.c_int32), ("cu_count", C.c_uint32), ("clock_khz", C.c_uint32), ("l2_bytes", C.c_uint32),
                ("hbm_bytes", C.c_uint64), ("arch", C.c_char * 64)]


class SceneDesc(C.Structure):
    """pt_scene_desc: device-layout arrays of one scene (bind group 1)."""
    _fields_ = [
        ("geometries", _p), ("attributes", _p), ("triangle_count", C.c_uint32),
        ("materials", _p), ("material_count", C.c_uint32),
        ("bvh_nodes", _p), ("bvh_node_count", C.c_uint32),
        ("srgb_atlases", _p), ("srgb_atlas_count", C.c_uint32),
        ("linear_atlases", _p), ("linear_atlas_count", C.c_uint32),
        ("atlas_size", C.c_uint32),
        ("camera", PtCamera),
        ("skybox_rgba", _p), ("skybox", PtSkyboxUniforms),
        ("skybox_alias", _p), ("skybox_alias_count", C.c_uint32),
        ("scene_uniforms", PtSceneUniforms),
        ("scene_alias", _p), ("scene_alias_count", C.c_uint32),
    ]


class WorldDesc(C.Structure):
    """pt_world_desc: host scene-preparation input."""
    _fields_ = [
        ("geometries", _p), ("attributes", _p), ("triangle_count", C.c_uint32),
        ("materials", _p), ("material_count", C.c_uint32),
        ("srgb_atlases", _p), ("srgb_atlas_count", C.c_uint32),
        ("linear_atlases", _p), ("linear_atlas_count", C.c_uint32), ("atlas_size", C.c_uint32),
        ("skybox_rgba", _p), ("skybox_width", C.c_uint32), ("skybox_height", C.c_uint32),
        ("skybox_luminance", C.c_float),
        ("camera_position", C.c_float * 3), ("camera_yaw_degrees", C.c_float),
        ("camera_pitch_degrees", C.c_float), ("camera_fov_degrees", C.c_float), ("aspect_ratio", C.c_float),
    ]


class TraceStats(C.Structure):
    _fields_ = [("closest_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("pair_tests", C.c_uint64),
                ("tri_tests", C.c_uint64), ("shadow_pair_tests", C.c_uint64), ("shadow_tri_tests", C.c_uint64),
                ("extend_launches", C.c_uint64), ("extend_ms", C.c_float),
                ("shadow_ms", C.c_float), ("shade_ms", C.c_float), ("total_ms", C.c_float),
                ("extend_busy_ms", C.c_float)]


class WaveConfig(C.Structure):
    _fields_ = [("lds_stack_depth", C.c_uint32), ("chunk_paths", C.c_uint32), ("sort_key", C.c_uint32),
                ("traversal", C.c_uint32), ("streams", C.c_uint32), ("trace_waves_per_cu", C.c_uint32),
                ("shade_blocks_per_cu", C.c_uint32), ("shade_window", C.c_uint32), ("lane_split_paths", C.c_uint32)]


class TuneResult(C.Structure):
    _fields_ = [("best", WaveConfig), ("best_ms", C.c_float), ("initial_ms", C.c_float), ("trials", C.c_uint32)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_active_jobs", C.c_uint32), ("round_paths", C.c_uint32),
                ("job_arena_bytes", C.c_uint64)]


class RenderRequest(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("samples", C.c_uint32),
                ("max_bounces", C.c_uint32)]


class JobInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("rounds", C.c_uint32), ("submit_s", C.c_double), ("admit_s", C.c_double),
                ("done_s", C.c_double), ("footprint_bytes", C.c_uint64), ("result_bytes", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("rounds", C.c_uint64), ("paths", C.c_uint64), ("jobs_done", C.c_uint64),
                ("jobs_rejected", C.c_uint64), ("busy_s", C.c_double), ("bytes_in_use", C.c_uint64),
                ("max_jobs_per_round", C.c_uint32), ("max_rounds_in_flight", C.c_uint32), ("arena_bytes", C.c_uint64)]


class CameraParams(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("yaw", C.c_float), ("pitch", C.c_float), ("fov", C.c_float),
                ("aspect_ratio", C.c_float)]


class WorldUpdates(C.Structure):
    _fields_ = [("camera_changed", C.c_uint32), ("camera", PtCamera),
                ("material_count", C.c_uint32), ("material_indices", _p), ("materials", _p),
                ("scene_uniforms_changed", C.c_uint32), ("scene_uniforms", PtSceneUniforms),
                ("scene_alias_changed", C.c_uint32), ("scene_alias", _p), ("scene_alias_count", C.c_uint32),
                ("skybox_uniforms_changed", C.c_uint32), ("skybox_uniforms", PtSkyboxUniforms),
                ("reset_accumulation", C.c_uint32), ("selected_material", C.c_int32)]


class AssetsSummary(C.Structure):
    _fields_ = [("triangle_count", C.c_uint32), ("material_count", C.c_uint32), ("srgb_atlas_count", C.c_uint32),
                ("linear_atlas_count", C.c_uint32), ("atlas_size", C.c_uint32), ("skybox_width", C.c_uint32),
                ("skybox_height", C.c_uint32), ("warning_count", C.c_uint32)]


SORT_NONE, SORT_OCTANT, SORT_MATERIAL = 0, 1, 2
TRAVERSAL_BATCH, TRAVERSAL_REFILL, TRAVERSAL_FETCH = 0, 1, 2
JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED, JOB_CANCELLED = range(5)

# every symbol include/pt_api.h and include/pt_scenes.h declare: (name, restype, argtypes)
_u32, _i32, _f32 = C.c_uint32, C.c_int32, C.c_float
_SIGNATURES = [
    ("pt_last_error", C.c_char_p, []),
    ("pt_version", C.c_char_p, []),
    ("pt_context_create", _i32, [C.c_int, C.POINTER(_p)]),
    ("pt_context_destroy", _i32, [_p]),
    ("pt_context_synchronize", _i32, [_p]),
    ("pt_context_get_info", _i32, [_p, C.POINTER(ContextInfo)]),
    ("pt_context_measure_gather", _i32, [_p, C.c_uint32, C.POINTER(C.c_double)]),
    ("pt_context_check_fast_math", _i32, [_p, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32,
                                          C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("pt_world_build", _i32, [C.POINTER(WorldDesc), C.POINTER(_p)]),
    ("pt_world_destroy", _i32, [_p]),
    ("pt_world_scene_desc", _i32, [_p, C.POINTER(SceneDesc)]),
    ("pt_world_camera", _i32, [_p, C.c_int, C.POINTER(CameraParams)]),
    ("pt_world_edit_camera", _i32, [_p, C.POINTER(CameraParams)]),
    ("pt_world_material", _i32, [_p, _u32, _p, _p]),
    ("pt_world_edit_material", _i32, [_p, _u32, _p, C.POINTER(C.c_int)]),
    ("pt_world_edit_emission_scale", _i32, [_p, _f32]),
    ("pt_world_emission_scale", _i32, [_p, C.POINTER(_f32)]),
    ("pt_world_edit_skybox_luminance", _i32, [_p, _f32]),
    ("pt_world_pick_material", _i32, [_p, C.POINTER(_f32), _f32, _f32, C.POINTER(_i32)]),
    ("pt_world_take_updates", _i32, [_p, C.POINTER(WorldUpdates)]),
    ("pt_scene_create", _i32, [_p, C.POINTER(SceneDesc), C.POINTER(_p)]),
    ("pt_scene_destroy", _i32, [_p]),
    ("pt_scene_footprint_bytes", _i32, [_p, C.POINTER(C.c_uint64)]),
    ("pt_scene_update_camera", _i32, [_p, C.POINTER(PtCamera)]),
    ("pt_scene_update_materials", _i32, [_p, _p, _p, _u32]),
    ("pt_scene_update_lights", _i32, [_p, C.POINTER(PtSceneUniforms), _p, _u32, C.POINTER(PtSkyboxUniforms)]),
    ("pt_scene_apply_updates", _i32, [_p, C.POINTER(WorldUpdates)]),
    ("pt_scene_intersect", _i32, [_p, _p, _u32, _p, _p, _p, _p, _p, _p]),
    ("pt_scene_occluded", _i32, [_p, _p, _p, _p, _u32, _p]),
    ("pt_tracer_create", _i32, [_p, _p, _u32, _u32, C.POINTER(_p)]),
    ("pt_tracer_destroy", _i32, [_p]),
    ("pt_tracer_set_schedule", _i32, [_p, C.c_int]),
    ("pt_tracer_set_tile", _i32, [_p, _u32, _u32, _u32, _u32, _u32, _u32]),
    ("pt_tracer_set_max_bounces", _i32, [_p, _u32]),
    ("pt_tracer_set_sample_count", _i32, [_p, _u32]),
    ("pt_tracer_sample_count", _u32, [_p]),
    ("pt_tracer_max_bounces", _u32, [_p]),
    ("pt_tracer_trace", _i32, [_p, _u32, C.POINTER(_u32)]),
    ("pt_tracer_trace_all", _i32, [_p, _u32]),
    ("pt_tracer_read_accumulation", _i32, [_p, _p]),
    ("pt_tracer_write_accumulation", _i32, [_p, _p]),
    ("pt_tracer_read_storage", _i32, [_p, _p]),
    ("pt_tracer_display_rgba8", _i32, [_p, _p]),
    ("pt_tracer_primary_hits", _i32, [_p, _u32, _p, _p]),
    ("pt_tracer_set_profiling", _i32, [_p, C.c_int]),
    ("pt_tracer_stats", _i32, [_p, C.POINTER(TraceStats)]),
    ("pt_context_set_wave_config", _i32, [_p, C.POINTER(WaveConfig)]),
    ("pt_context_get_wave_config", _i32, [_p, C.POINTER(WaveConfig)]),
    ("pt_trace_batch", _i32, [C.POINTER(_p), _u32, C.POINTER(_u32)]),
    ("pt_autotune_batch", _i32, [C.POINTER(_p), _u32, C.POINTER(_u32), C.POINTER(TuneResult)]),
    ("pt_tonemap_rgba8", _i32, [_p, _p, _u32, _p]),
    ("pt_render_tile_noalloc", _i32, [_p, C.POINTER(PtTraceParameters), _u32, _u32, _u32, _p, _p]),
    ("pt_server_create", _i32, [_p, C.POINTER(ServerConfig), C.POINTER(_p)]),
    ("pt_server_destroy", _i32, [_p]),
    ("pt_server_submit", _i32, [_p, C.POINTER(RenderRequest), C.POINTER(C.c_uint64)]),
    ("pt_server_wait", _i32, [_p, C.c_uint64, _u32]),
    ("pt_server_info", _i32, [_p, C.c_uint64, C.POINTER(JobInfo)]),
    ("pt_server_result", _i32, [_p, C.c_uint64, _p, C.c_size_t]),
    ("pt_server_cancel", _i32, [_p, C.c_uint64]),
    ("pt_server_stats_get", _i32, [_p, C.POINTER(ServerStats)]),
    ("pt_assets_create", _i32, [C.POINTER(_p)]),
    ("pt_assets_destroy", _i32, [_p]),
    ("pt_assets_load_scene", _i32, [_p, _p, C.c_size_t]),
    ("pt_assets_load_scene_file", _i32, [_p, C.c_char_p]),
    ("pt_assets_load_skybox", _i32, [_p, _p, C.c_size_t]),
    ("pt_assets_load_skybox_file", _i32, [_p, C.c_char_p]),
    ("pt_assets_info", _i32, [_p, C.POINTER(AssetsSummary)]),
    ("pt_assets_warning", C.c_char_p, [_p, _u32]),
    ("pt_assets_world_desc", _i32, [_p, C.POINTER(WorldDesc)]),
    ("pt_image_format_from_path", _i32, [C.c_char_p, C.POINTER(C.c_int)]),
    ("pt_image_encode", _i32, [C.c_int, _p, C.c_size_t, _u32, _u32, _u32, C.POINTER(C.POINTER(C.c_uint8)),
                               C.POINTER(C.c_size_t)]),
    ("pt_image_free", None, [C.POINTER(C.c_uint8)]),
    ("pt_image_save", _i32, [C.c_char_p, _p, _u32, _u32]),
    ("pt_world_export_glb", _i32, [C.POINTER(WorldDesc), C.c_char_p, C.c_char_p, C.POINTER(_u32)]),
    ("pt_scene_save", _i32, [C.POINTER(SceneDesc), C.c_char_p]),
    ("pt_scene_file_load", _i32, [C.c_char_p, C.POINTER(_p)]),
    ("pt_scene_file_desc", _i32, [_p, C.POINTER(SceneDesc)]),
    ("pt_scene_file_destroy", _i32, [_p]),
    ("pt_generate_scene", _i32, [C.c_int, C.c_uint64, _u32, _u32, _u32, C.POINTER(_p)]),
    ("pt_generated_world_desc", _i32, [_p, C.POINTER(WorldDesc)]),
    ("pt_generated_destroy", _i32, [_p]),
]

_lib = None


def lib():
    """Load libptcore.so (raises PtcoreUnavailable if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PtcoreUnavailable(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        try:
            L = C.CDLL(LIB_PATH)
        except OSError as e:
            raise PtcoreUnavailable(f"cannot load {LIB_PATH}: {e}") from e
        for name, res, args in _SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return [name for name, _, _ in _SIGNATURES]


def _check(code):
    if code != PT_OK:
        raise PtError(code, lib().pt_last_error().decode("utf-8", "replace"))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _view(ptr, dtype, count, owner=None):
    """A numpy view of library memory; `owner` (the object that frees it) stays alive with the view."""
    if not ptr or count == 0:
        return np.zeros(0, dtype=dtype)
    buf = (C.c_char * (count * np.dtype(dtype).itemsize)).from_address(ptr)
    buf._owner = owner
    return np.frombuffer(buf, dtype=dtype, count=count)


# ---------------------------------------------------------------- host scene preparation
class Generated:
    """A seeded procedural scene (pt_scenes.h); owns its arrays."""

    def __init__(self, kind, seed, width, height, triangles=0):
        h = C.c_void_p()
        _check(lib().pt_generate_scene(kind, seed, triangles, width, height, C.byref(h)))
        self.handle = h
        self.desc = WorldDesc()
        _check(lib().pt_generated_world_desc(h, C.byref(self.desc)))
        self.desc._owner = self  # the descriptor borrows this object's arrays

    def close(self):
        if self.handle:
            lib().pt_generated_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def arrays(self):
        d = self.desc
        layer = d.atlas_size * d.atlas_size * 4
        return {
            "geometries": _view(d.geometries, GEOMETRY_DTYPE, d.triangle_count, self),
            "attributes": _view(d.attributes, ATTRIBUTE_DTYPE, d.triangle_count, self),
            "materials": _view(d.materials, MATERIAL_DTYPE, d.material_count, self),
            "srgb_atlases": _view(d.srgb_atlases, np.uint8, layer * d.srgb_atlas_count, self),
            "linear_atlases": _view(d.linear_atlases, np.uint8, layer * d.linear_atlas_count, self),
            "skybox": _view(d.skybox_rgba, np.float32, d.skybox_width * d.skybox_height * 4, self).reshape(
                d.skybox_height, d.skybox_width, 4),
        }

    def render_flags(self):
        """The reference's `render` flags that the world's camera and sky luminance become
        (headless.rs:14-64): they are not part of the .glb or the sky file."""
        d = self.desc
        flags = []
        if not any(math.isnan(d.camera_position[k]) for k in range(3)):
            flags += ["--camera-position", ",".join(f"{d.camera_position[k]:.9g}" for k in range(3))]
        for name, v in (("--camera-yaw", d.camera_yaw_degrees), ("--camera-pitch", d.camera_pitch_degrees),
                        ("--camera-fov", d.camera_fov_degrees)):
            if not math.isnan(v):
                flags += [name, f"{v:.9g}"]
        if d.skybox_luminance >= 0:
            flags += ["--skybox-luminance", f"{d.skybox_luminance:.9g}"]
        return flags


def export_glb(world_desc: WorldDesc, glb_path, sky_path=None):
    """pt_world_export_glb: the world input as a .glb, and its sky as a FLOAT OpenEXR (or,
    for a path ending in .hdr, exact RGBE). Returns the count of normal and tangent
    directions that do not load back bit for bit (0 when the round trip is exact)."""
    n = _u32(0)
    _check(lib().pt_world_export_glb(C.byref(world_desc), os.fsencode(str(glb_path)),
                                     None if sky_path is None else os.fsencode(str(sky_path)), C.byref(n)))
    return n.value


def save_scene(scene_desc: SceneDesc, path):
    """pt_scene_save: the device-layout arrays (BVH included) as a .ptscene file."""
    _check(lib().pt_scene_save(C.byref(scene_desc), os.fsencode(str(path))))


class SceneFile:
    """A loaded .ptscene file (pt_scene_file); `desc` borrows its arrays."""

    def __init__(self, path):
        h = C.c_void_p()
        _check(lib().pt_scene_file_load(os.fsencode(str(path)), C.byref(h)))
        self.handle = h
        self.desc = SceneDesc()
        _check(lib().pt_scene_file_desc(h, C.byref(self.desc)))
        self.desc._owner = self  # the descriptor borrows this object's arrays

    def close(self):
        if self.handle:
            lib().pt_scene_file_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class World:
    """Prepared host scene: pt_world_build over a WorldDesc (keeps `keep` alive)."""

    def __init__(self, desc: WorldDesc, keep=None):
        h = C.c_void_p()
        _check(lib().pt_world_build(C.byref(desc), C.byref(h)))
        self.handle = h
        self._keep = keep
        self.desc = SceneDesc()
        _check(lib().pt_world_scene_desc(h, C.byref(self.desc)))
        # the descriptor points into this world's arrays: it keeps the world alive,
        # so World.generate(...).desc stays valid after the World object is dropped
        self.desc._owner = self

    @classmethod
    def generate(cls, kind, seed, width, height, triangles=0):
        g = Generated(kind, seed, width, height, triangles)
        w = cls(g.desc, keep=g)
        return w

    # ---- editing (World, world.rs:589-884)
    def camera(self, initial=False):
        c = CameraParams()
        _check(lib().pt_world_camera(self.handle, 1 if initial else 0, C.byref(c)))
        return dict(position=tuple(c.position), yaw=c.yaw, pitch=c.pitch, fov=c.fov, aspect_ratio=c.aspect_ratio)

    def edit_camera(self, **fields):
        cur = self.camera()
        cur.update(fields)
        c = CameraParams((C.c_float * 3)(*cur["position"]), cur["yaw"], cur["pitch"], cur["fov"], cur["aspect_ratio"])
        _check(lib().pt_world_edit_camera(self.handle, C.byref(c)))

    def material(self, index):
        cur, ini = np.zeros(1, MATERIAL_DTYPE), np.zeros(1, MATERIAL_DTYPE)
        _check(lib().pt_world_material(self.handle, index, _ptr(cur), _ptr(ini)))
        return cur[0], ini[0]

    def edit_material(self, index, material):
        m = np.ascontiguousarray(np.asarray(material, dtype=MATERIAL_DTYPE).reshape(1))
        ok = C.c_int()
        _check(lib().pt_world_edit_material(self.handle, index, _ptr(m), C.byref(ok)))
        return bool(ok.value)

    def edit_emission_scale(self, scale):
        _check(lib().pt_world_edit_emission_scale(self.handle, float(scale)))

    def emission_scale(self):
        v = C.c_float()
        _check(lib().pt_world_emission_scale(self.handle, C.byref(v)))
        return v.value

    def edit_skybox_luminance(self, nits):
        _check(lib().pt_world_edit_skybox_luminance(self.handle, float(nits)))

    def pick_material(self, position, yaw, pitch):
        pos = (C.c_float * 3)(*position)
        m = C.c_int32()
        _check(lib().pt_world_pick_material(self.handle, pos, float(yaw), float(pitch), C.byref(m)))
        return m.value

    def take_updates(self):
        """pt_world_take_updates: the raw struct (valid until the next take) plus numpy copies."""
        u = WorldUpdates()
        _check(lib().pt_world_take_updates(self.handle, C.byref(u)))
        # the edits change the arrays a scene is created from: borrow them again
        _check(lib().pt_world_scene_desc(self.handle, C.byref(self.desc)))
        n = u.material_count
        return {"raw": u, "camera_changed": bool(u.camera_changed), "camera": u.camera,
                "material_indices": _view(u.material_indices, np.uint32, n).copy() if n else np.zeros(0, np.uint32),
                "materials": _view(u.materials, MATERIAL_DTYPE, n).copy() if n else np.zeros(0, MATERIAL_DTYPE),
                "scene_uniforms_changed": bool(u.scene_uniforms_changed), "scene_uniforms": u.scene_uniforms,
                "scene_alias_changed": bool(u.scene_alias_changed),
                "scene_alias": (_view(u.scene_alias, SCENE_ALIAS_DTYPE, u.scene_alias_count).copy()
                                if u.scene_alias_count else np.zeros(0, SCENE_ALIAS_DTYPE)),
                "skybox_uniforms_changed": bool(u.skybox_uniforms_changed), "skybox_uniforms": u.skybox_uniforms,
                "reset_accumulation": bool(u.reset_accumulation), "selected_material": u.selected_material}

    def close(self):
        if self.handle:
            lib().pt_world_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def arrays(self):
        d = self.desc
        layer = d.atlas_size * d.atlas_size * 4
        return {
            "geometries": _view(d.geometries, GEOMETRY_DTYPE, d.triangle_count, self),
            "attributes": _view(d.attributes, ATTRIBUTE_DTYPE, d.triangle_count, self),
            "materials": _view(d.materials, MATERIAL_DTYPE, d.material_count, self),
            "bvh_nodes": _view(d.bvh_nodes, BVH_DTYPE, d.bvh_node_count, self),
            "srgb_atlases": _view(d.srgb_atlases, np.uint8, layer * d.srgb_atlas_count, self),
            "linear_atlases": _view(d.linear_atlases, np.uint8, layer * d.linear_atlas_count, self),
            "skybox": _view(d.skybox_rgba, np.float32, d.skybox.width * d.skybox.height * 4, self),
            "skybox_alias": _view(d.skybox_alias, SKY_ALIAS_DTYPE, d.skybox_alias_count, self),
            "scene_alias": _view(d.scene_alias, SCENE_ALIAS_DTYPE, d.scene_alias_count, self),
        }


def default_material(**fields):
    """A GpuMaterial with the reference's defaults (world.rs:199-309), metalness 0."""
    m = np.zeros(1, MATERIAL_DTYPE)[0]
    m["base_color"] = 1.0
    m["base_weight"] = 1.0
    m["base_diffuse_roughness"] = 0.5
    m["emission_weight"] = 1.0
    m["emission_color_and_luminance"] = (0.0, 0.0, 0.0, 1.0)
    m["specular_color_and_weight"] = 1.0
    m["specular_roughness"] = 0.3
    m["specular_haze_spread"] = 0.3
    m["specular_ior"] = 1.5
    m["normal_scale"] = 1.0
    m["alpha_cutoff"] = 0.5
    for k in ("base_color_tex_layer", "normal_tex_layer", "metallic_roughness_tex_layer", "emission_tex_layer",
              "specular_color_tex_layer", "specular_weight_tex_layer"):
        m[k] = -1
    for k, v in fields.items():
        m[k] = v
    return m


def triangles_from_vertices(verts, material_ids):
    """Geometry + attribute arrays for flat triangles. verts: (N, 3, 3)."""
    verts = np.asarray(verts, np.float32).reshape(-1, 3, 3)
    n = len(verts)
    geo = np.zeros(n, GEOMETRY_DTYPE)
    geo["v"] = verts.reshape(n, 9)
    attr = np.zeros(n, ATTRIBUTE_DTYPE)
    nrm = np.cross(verts[:, 1] - verts[:, 0], verts[:, 2] - verts[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    attr["material"] = np.asarray(material_ids, np.uint32)
    attr["n"] = np.repeat(nrm, 3, axis=0).reshape(n, 9)
    attr["uv"] = np.tile(np.array([0, 0, 1, 0, 0, 1], np.float32), (n, 1))
    t = np.where(np.abs(nrm[:, :1]) < 0.9, np.array([[1.0, 0, 0]]), np.array([[0, 0, 1.0]]))
    t = t - (t * nrm).sum(1, keepdims=True) * nrm
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    t4 = np.concatenate([t, np.ones((n, 1))], axis=1).astype(np.float32)
    attr["t0"], attr["t1"], attr["t2"] = t4, t4, t4
    return geo, attr


IMAGE_FORMATS = {"png": 0, "jpeg": 1, "webp": 2}


def image_format_from_path(path):
    """'png', 'jpeg' or 'webp' by extension (render's parse_output, headless.rs:118-127)."""
    f = C.c_int()
    _check(lib().pt_image_format_from_path(os.fsencode(str(path)), C.byref(f)))
    return {v: k for k, v in IMAGE_FORMATS.items()}[f.value]


def _rgb8(rgb):
    rgb = np.asarray(rgb)
    if rgb.dtype != np.uint8 or rgb.ndim != 3 or rgb.shape[2] != 3:
        raise ValueError("expected an HxWx3 uint8 array")
    return rgb if rgb.strides[1:] == (3, 1) else np.ascontiguousarray(rgb)


def encode_image(rgb, fmt="png", quality=0):
    """Bytes of an HxWx3 uint8 image as PNG, JPEG (quality 1..100, 0 = 75) or lossless WebP."""
    rgb = _rgb8(rgb)
    out, size = C.POINTER(C.c_uint8)(), C.c_size_t()
    _check(lib().pt_image_encode(IMAGE_FORMATS[fmt], _ptr(rgb), rgb.strides[0], rgb.shape[1], rgb.shape[0],
                                 quality, C.byref(out), C.byref(size)))
    try:
        return C.string_at(out, size.value)
    finally:
        lib().pt_image_free(out)


def save_image(path, rgb):
    """image::save_buffer of RGB8 pixels, format by extension (headless.rs:163-170)."""
    rgb = np.ascontiguousarray(_rgb8(rgb))
    _check(lib().pt_image_save(os.fsencode(str(path)), _ptr(rgb), rgb.shape[1], rgb.shape[0]))


class Assets:
    """Scene and sky files (pt_assets.h): Scene::new for .glb and the sky image decode."""

    def __init__(self, scene=None, skybox=None):
        h = C.c_void_p()
        _check(lib().pt_assets_create(C.byref(h)))
        self.handle = h
        if scene is not None:
            self.load_scene(scene)
        if skybox is not None:
            self.load_skybox(skybox)

    def _load(self, fn_bytes, fn_file, src):
        if isinstance(src, (bytes, bytearray, memoryview)):
            buf = (C.c_uint8 * len(src)).from_buffer_copy(bytes(src))
            _check(fn_bytes(self.handle, buf, len(src)))
        else:
            _check(fn_file(self.handle, os.fsencode(str(src))))

    def load_scene(self, src):
        """src: bytes of a .glb / embedded .gltf, or a path."""
        self._load(lib().pt_assets_load_scene, lib().pt_assets_load_scene_file, src)

    def load_skybox(self, src):
        """src: bytes of a Radiance .hdr / OpenEXR / PNG, or a path."""
        self._load(lib().pt_assets_load_skybox, lib().pt_assets_load_skybox_file, src)

    def info(self):
        s = AssetsSummary()
        _check(lib().pt_assets_info(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in AssetsSummary._fields_}

    def warnings(self):
        return [lib().pt_assets_warning(self.handle, i).decode("utf-8", "replace")
                for i in range(self.info()["warning_count"])]

    def world_desc(self, skybox_luminance=-1.0, camera_position=None, yaw=None, pitch=None, fov=None, aspect=1.0):
        d = WorldDesc()
        _check(lib().pt_assets_world_desc(self.handle, C.byref(d)))
        d._owner = self  # the descriptor borrows the loaded arrays
        d.skybox_luminance = skybox_luminance
        if camera_position is not None:
            for k in range(3):
                d.camera_position[k] = camera_position[k]
        if yaw is not None:
            d.camera_yaw_degrees = yaw
        if pitch is not None:
            d.camera_pitch_degrees = pitch
        if fov is not None:
            d.camera_fov_degrees = fov
        d.aspect_ratio = aspect
        return d

    def world(self, **options):
        """pt_world_build over the loaded arrays (World::new)."""
        return World(self.world_desc(**options), keep=self)

    def arrays(self):
        d = self.world_desc()
        layer = d.atlas_size * d.atlas_size * 4
        return {"geometries": _view(d.geometries, GEOMETRY_DTYPE, d.triangle_count, self),
                "attributes": _view(d.attributes, ATTRIBUTE_DTYPE, d.triangle_count, self),
                "materials": _view(d.materials, MATERIAL_DTYPE, d.material_count, self),
                "srgb_atlases": _view(d.srgb_atlases, np.uint8, layer * d.srgb_atlas_count, self),
                "linear_atlases": _view(d.linear_atlases, np.uint8, layer * d.linear_atlas_count, self),
                "skybox": _view(d.skybox_rgba, np.float32, d.skybox_width * d.skybox_height * 4, self).reshape(
                    d.skybox_height, d.skybox_width, 4)}

    def close(self):
        if self.handle:
            lib().pt_assets_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WorldInput:
    """A pt_world_desc over numpy arrays (the input of pt_world_build)."""

    def __init__(self, geometries, attributes, materials, skybox_rgba, skybox_luminance=-1.0,
                 camera_position=None, yaw=None, pitch=None, fov=None, aspect=1.0,
                 srgb=None, linear=None, atlas_size=0):
        self.arrays = dict(geometries=np.ascontiguousarray(geometries, GEOMETRY_DTYPE),
                           attributes=np.ascontiguousarray(attributes, ATTRIBUTE_DTYPE),
                           materials=np.ascontiguousarray(materials, MATERIAL_DTYPE),
                           skybox=np.ascontiguousarray(skybox_rgba, np.float32))
        a = self.arrays
        sky = a["skybox"]
        assert sky.ndim == 3 and sky.shape[2] == 4
        d = WorldDesc()
        d.geometries, d.attributes = _ptr(a["geometries"]), _ptr(a["attributes"])
        d.triangle_count = len(a["geometries"])
        d.materials, d.material_count = _ptr(a["materials"]), len(a["materials"])
        self._srgb = None if srgb is None else np.ascontiguousarray(srgb, np.uint8)
        self._linear = None if linear is None else np.ascontiguousarray(linear, np.uint8)
        layer = max(atlas_size * atlas_size * 4, 1)
        d.srgb_atlases = _ptr(self._srgb)
        d.srgb_atlas_count = 0 if self._srgb is None else self._srgb.size // layer
        d.linear_atlases = _ptr(self._linear)
        d.linear_atlas_count = 0 if self._linear is None else self._linear.size // layer
        d.atlas_size = atlas_size
        d.skybox_rgba = _ptr(sky)
        d.skybox_height, d.skybox_width = sky.shape[0], sky.shape[1]
        d.skybox_luminance = skybox_luminance
        nan = float("nan")
        pos = camera_position if camera_position is not None else (nan, nan, nan)
        for k in range(3):
            d.camera_position[k] = pos[k]
        d.camera_yaw_degrees = nan if yaw is None else yaw
        d.camera_pitch_degrees = nan if pitch is None else pitch
        d.camera_fov_degrees = nan if fov is None else fov
        d.aspect_ratio = aspect
        d._owner = self  # the descriptor points into this object's arrays
        self.desc = d

    def build(self):
        return World(self.desc, keep=self)


class SceneArrays:
    """A pt_scene_desc built from numpy arrays (e.g. a hand-made or loaded scene)."""

    def __init__(self, geometries, attributes, materials, bvh_nodes, camera, skybox, skybox_uniforms,
                 skybox_alias, scene_uniforms, scene_alias, srgb=None, linear=None, atlas_size=0):
        self.arrays = dict(geometries=np.ascontiguousarray(geometries), attributes=np.ascontiguousarray(attributes),
                           materials=np.ascontiguousarray(materials), bvh_nodes=np.ascontiguousarray(bvh_nodes),
                           skybox=np.ascontiguousarray(skybox, dtype=np.float32),
                           skybox_alias=np.ascontiguousarray(skybox_alias),
                           scene_alias=np.ascontiguousarray(scene_alias))
        a = self.arrays
        d = SceneDesc()
        d.geometries, d.attributes = _ptr(a["geometries"]), _ptr(a["attributes"])
        d.triangle_count = len(a["geometries"])
        d.materials, d.material_count = _ptr(a["materials"]), len(a["materials"])
        d.bvh_nodes, d.bvh_node_count = _ptr(a["bvh_nodes"]), len(a["bvh_nodes"])
        self._srgb = None if srgb is None else np.ascontiguousarray(srgb, dtype=np.uint8)
        self._linear = None if linear is None else np.ascontiguousarray(linear, dtype=np.uint8)
        layer = atlas_size * atlas_size * 4
        d.srgb_atlases = _ptr(self._srgb)
        d.srgb_atlas_count = 0 if self._srgb is None else self._srgb.size // max(layer, 1)
        d.linear_atlases = _ptr(self._linear)
        d.linear_atlas_count = 0 if self._linear is None else self._linear.size // max(layer, 1)
        d.atlas_size = atlas_size
        d.camera = camera
        d.skybox_rgba = _ptr(a["skybox"])
        d.skybox = skybox_uniforms
        d.skybox_alias, d.skybox_alias_count = _ptr(a["skybox_alias"]), len(a["skybox_alias"])
        d.scene_uniforms = scene_uniforms
        d.scene_alias, d.scene_alias_count = _ptr(a["scene_alias"]), len(a["scene_alias"])
        d._owner = self  # the descriptor points into this object's arrays
        self.desc = d


# ---------------------------------------------------------------- device objects
class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().pt_context_create(device, C.byref(h)))
        self.handle = h

    def synchronize(self):
        _check(lib().pt_context_synchronize(self.handle))

    def info(self):
        """pt_context_get_info: CU count, peak clock, L2 and device memory of the context's GPU."""
        i = ContextInfo()
        _check(lib().pt_context_get_info(self.handle, C.byref(i)))
        return {"device": i.device, "cu_count": i.cu_count, "clock_khz": i.clock_khz, "l2_bytes": i.l2_bytes,
                "hbm_bytes": i.hbm_bytes, "arch": i.arch.decode("utf-8", "replace")}

    def measure_gather(self, lanes_per_segment=1):
        """pt_context_measure_gather: the L1's rate of 16-byte lane accesses per CU per cycle, each lane its
        own 64-byte segment (1) or four lanes per segment (4)."""
        r = C.c_double()
        _check(lib().pt_context_measure_gather(self.handle, lanes_per_segment, C.byref(r)))
        return r.value

    def check_fast_math(self, sqrt_first, sqrt_last, div_pairs=0, div_seed=1):
        """pt_context_check_fast_math: mismatches of the restricted-domain square root against sqrtf over the
        float32 bit patterns sqrt_first..sqrt_last, and of the restricted-domain division against a / b over
        div_pairs seeded operand pairs. Returns (sqrt_mismatches, div_mismatches)."""
        ms, md = C.c_uint64(), C.c_uint64()
        _check(lib().pt_context_check_fast_math(self.handle, sqrt_first, sqrt_last, div_pairs, div_seed,
                                                C.byref(ms), C.byref(md)))
        return ms.value, md.value

    def close(self):
        if self.handle:
            lib().pt_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """A scene resident in device memory (pt_scene)."""

    def __init__(self, ctx: Context, desc: SceneDesc):
        h = C.c_void_p()
        _check(lib().pt_scene_create(ctx.handle, C.byref(desc), C.byref(h)))
        self.handle = h
        self.ctx = ctx

    def footprint(self):
        b = C.c_uint64()
        _check(lib().pt_scene_footprint_bytes(self.handle, C.byref(b)))
        return b.value

    def update_camera(self, camera: PtCamera):
        _check(lib().pt_scene_update_camera(self.handle, C.byref(camera)))

    def apply_updates(self, updates):
        """pt_scene_apply_updates with what World.take_updates returned (call before the next take)."""
        _check(lib().pt_scene_apply_updates(self.handle, C.byref(updates["raw"])))

    def update_materials(self, indices, materials):
        idx = np.ascontiguousarray(indices, dtype=np.uint32)
        m = np.ascontiguousarray(materials, dtype=MATERIAL_DTYPE)
        _check(lib().pt_scene_update_materials(self.handle, _ptr(idx), _ptr(m), len(idx)))

    def intersect(self, rays):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        n = len(rays)
        hit = np.empty(n, np.int32)
        t, v, w = (np.empty(n, np.float32) for _ in range(3))
        pairs, tris = np.empty(n, np.uint32), np.empty(n, np.uint32)
        _check(lib().pt_scene_intersect(self.handle, _ptr(rays), n, _ptr(hit), _ptr(t), _ptr(v), _ptr(w),
                                        _ptr(pairs), _ptr(tris)))
        return hit, t, v, w, pairs, tris

    def occluded(self, rays, tmax, skip=None):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        n = len(rays)
        tmax = np.ascontiguousarray(tmax, dtype=np.float32)
        skip = np.full(n, -1, np.int32) if skip is None else np.ascontiguousarray(skip, dtype=np.int32)
        out = np.empty(n, np.uint8)
        _check(lib().pt_scene_occluded(self.handle, _ptr(rays), _ptr(tmax), _ptr(skip), n, _ptr(out)))
        return out

    def close(self):
        """Free the device scene; PtError (PT_ERR_STATE, scene kept) while server jobs use it."""
        if self.handle:
            _check(lib().pt_scene_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Tracer:
    """PathTracer equivalent over a tile (path_tracer.rs:59-358)."""

    MAX_SAMPLES_PER_TRACE = 16

    def __init__(self, ctx: Context, scene: Scene, tile_w, tile_h, schedule=SCHEDULE_WAVEFRONT):
        h = C.c_void_p()
        _check(lib().pt_tracer_create(ctx.handle, scene.handle, tile_w, tile_h, C.byref(h)))
        self.handle = h
        self.ctx, self.scene = ctx, scene
        self.tile_w, self.tile_h = tile_w, tile_h
        self.set_schedule(schedule)

    def set_schedule(self, schedule):
        _check(lib().pt_tracer_set_schedule(self.handle, schedule))

    def set_tile(self, image_size, origin, size):
        _check(lib().pt_tracer_set_tile(self.handle, image_size[0], image_size[1], origin[0], origin[1],
                                        size[0], size[1]))
        self.tile_w, self.tile_h = size

    def set_max_bounces(self, n):
        _check(lib().pt_tracer_set_max_bounces(self.handle, n))

    def set_sample_count(self, n):
        _check(lib().pt_tracer_set_sample_count(self.handle, n))

    @property
    def sample_count(self):
        return lib().pt_tracer_sample_count(self.handle)

    @property
    def max_bounces(self):
        return lib().pt_tracer_max_bounces(self.handle)

    def trace(self, samples):
        n = C.c_uint32()
        _check(lib().pt_tracer_trace(self.handle, samples, C.byref(n)))
        return n.value

    def trace_all(self, samples):
        _check(lib().pt_tracer_trace_all(self.handle, samples))

    def accumulation(self):
        out = np.empty((self.tile_h, self.tile_w, 4), np.float32)
        _check(lib().pt_tracer_read_accumulation(self.handle, _ptr(out)))
        return out

    def write_accumulation(self, rgba):
        a = np.ascontiguousarray(rgba, dtype=np.float32)
        assert a.size == self.tile_w * self.tile_h * 4
        _check(lib().pt_tracer_write_accumulation(self.handle, _ptr(a)))

    def storage(self):
        out = np.empty((self.tile_h, self.tile_w, 4), np.float32)
        _check(lib().pt_tracer_read_storage(self.handle, _ptr(out)))
        return out

    def display_rgba8(self, out=None):
        """Tone-mapped RGBA8 tile; `out` (tile_h x tile_w x 4 uint8, contiguous) is reused when given."""
        if out is None:
            out = np.empty((self.tile_h, self.tile_w, 4), np.uint8)
        if out.dtype != np.uint8 or out.shape != (self.tile_h, self.tile_w, 4) or not out.flags.c_contiguous:
            raise ValueError("display buffer must be a contiguous tile_h x tile_w x 4 uint8 array")
        _check(lib().pt_tracer_display_rgba8(self.handle, _ptr(out)))
        return out

    def primary_hits(self, sample_count=1):
        tri = np.empty((self.tile_h, self.tile_w), np.int32)
        t = np.empty((self.tile_h, self.tile_w), np.float32)
        _check(lib().pt_tracer_primary_hits(self.handle, sample_count, _ptr(tri), _ptr(t)))
        return tri, t

    def set_profiling(self, level=1):
        """0 off, 1 per-kernel HIP-event timing, 2 timing + traversal counters."""
        _check(lib().pt_tracer_set_profiling(self.handle, int(level)))

    def stats(self):
        s = TraceStats()
        _check(lib().pt_tracer_stats(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in TraceStats._fields_}

    def close(self):
        if self.handle:
            lib().pt_tracer_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


DEFAULT_WAVE_CONFIG = dict(lds_stack_depth=16, chunk_paths=0, sort_key=SORT_NONE, traversal=0, streams=0,
                           trace_waves_per_cu=512, shade_blocks_per_cu=32, shade_window=0, lane_split_paths=1 << 21)


def set_wave_config(ctx: Context, **overrides):
    """Set the context's wavefront tuning; unspecified knobs take the defaults."""
    unknown = set(overrides) - set(DEFAULT_WAVE_CONFIG)
    if unknown:
        raise ValueError(f"unknown wave config fields: {sorted(unknown)}")
    v = {**DEFAULT_WAVE_CONFIG, **overrides}
    cfg = WaveConfig(*(v[k] for k, _ in WaveConfig._fields_))
    _check(lib().pt_context_set_wave_config(ctx.handle, C.byref(cfg)))


def get_wave_config(ctx: Context):
    cfg = WaveConfig()
    _check(lib().pt_context_get_wave_config(ctx.handle, C.byref(cfg)))
    return {k: getattr(cfg, k) for k, _ in WaveConfig._fields_}


def tonemap_rgba8(ctx: Context, rgba):
    """The display pass (ACES + sRGB encode, RGBA8) of float texels (..., 4) on the device."""
    a = np.ascontiguousarray(rgba, dtype=np.float32)
    if a.shape[-1] != 4:
        raise ValueError("texels must have 4 channels")
    out = np.empty(a.shape, np.uint8)
    _check(lib().pt_tonemap_rgba8(ctx.handle, _ptr(a), a.size // 4, _ptr(out)))
    return out


def _tracer_array(tracers, samples):
    n = len(tracers)
    handles = (C.c_void_p * n)(*[t.handle for t in tracers])
    s = (C.c_uint32 * n)(*([samples] * n if isinstance(samples, int) else samples))
    return n, handles, s


def trace_batch(tracers, samples):
    """Trace samples (int or per-tracer list) on every tracer in one mixed wavefront (pt_trace_batch)."""
    n, h, s = _tracer_array(tracers, samples)
    _check(lib().pt_trace_batch(h, n, s))


def autotune_batch(tracers, samples):
    n, h, s = _tracer_array(tracers, samples)
    r = TuneResult()
    _check(lib().pt_autotune_batch(h, n, s, C.byref(r)))
    return {"best": {k: getattr(r.best, k) for k, _ in WaveConfig._fields_}, "best_ms": r.best_ms,
            "initial_ms": r.initial_ms, "trials": r.trials}


class Server:
    """Render server with continuous batching (pt_server_*).

    The server keeps each submitted job's Scene referenced until the job is
    collected (result) or the server is closed, so a scene cannot be garbage
    collected while a job over it is queued or running."""

    def __init__(self, ctx: Context, hbm_budget_bytes=200 << 30, max_active_jobs=64, round_paths=0,
                 job_arena_bytes=0):
        h = C.c_void_p()
        cfg = ServerConfig(hbm_budget_bytes, max_active_jobs, round_paths, job_arena_bytes)
        _check(lib().pt_server_create(ctx.handle, C.byref(cfg), C.byref(h)))
        self.handle = h
        self.ctx = ctx
        self._jobs = {}  # job id -> (scene, width, height)
        self._lock = __import__("threading").Lock()

    def submit(self, scene: Scene, width, height, samples, max_bounces):
        req = RenderRequest(scene.handle, width, height, samples, max_bounces)
        jid = C.c_uint64()
        with self._lock:
            _check(lib().pt_server_submit(self.handle, C.byref(req), C.byref(jid)))
            self._jobs[jid.value] = (scene, width, height)
        return jid.value

    def wait(self, job_id, timeout_ms=60_000):
        _check(lib().pt_server_wait(self.handle, job_id, timeout_ms))

    def info(self, job_id):
        i = JobInfo()
        _check(lib().pt_server_info(self.handle, job_id, C.byref(i)))
        return {k: getattr(i, k) for k, _ in JobInfo._fields_}

    def result(self, job_id):
        """The done job's accumulation (height x width x 4 float32); releases the job."""
        with self._lock:
            entry = self._jobs.get(job_id)
        if entry is None:
            _check(lib().pt_server_result(self.handle, job_id, None, 0))  # raises: unknown or collected
        _, width, height = entry
        out = np.empty((height, width, 4), np.float32)
        try:
            _check(lib().pt_server_result(self.handle, job_id, _ptr(out), out.nbytes))
        except PtError as e:
            if e.code != PT_ERR_STATE:  # the record is gone: failed, cancelled (or collected elsewhere)
                with self._lock:
                    self._jobs.pop(job_id, None)
            raise
        with self._lock:
            self._jobs.pop(job_id, None)
        return out

    def cancel(self, job_id):
        _check(lib().pt_server_cancel(self.handle, job_id))

    def stats(self):
        s = ServerStats()
        _check(lib().pt_server_stats_get(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self.handle:
            lib().pt_server_destroy(self.handle)
            self.handle = None
            self._jobs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gpu_available():
    """True if a HIP device is visible (counting devices does not initialise HIP contexts in torch)."""
    try:
        import torch
        return torch.cuda.is_available() and torch.cuda.device_count() > 0
    except Exception:
        return False


def make_camera(position, yaw_deg=0.0, pitch_deg=0.0, fov_deg=90.0, aspect=1.0):
    """Camera::gpu_camera (world.rs:2159-2173) evaluated in float32, for tests and tools."""
    f = np.float32
    rad = f(math.pi) / f(180.0)
    yaw, pitch, fov = f(yaw_deg) * rad, f(pitch_deg) * rad, f(fov_deg) * rad
    sy, cy, sp, cp = f(math.sin(yaw)), f(math.cos(yaw)), f(math.sin(pitch)), f(math.cos(pitch))
    ry = np.array([[cy, 0, -sy], [0, 1, 0], [sy, 0, cy]], np.float32)  # columns
    rx = np.array([[1, 0, 0], [0, cp, sp], [0, -sp, cp]], np.float32)

    def mulv(m, v):
        return (m[0] * v[0] + m[1] * v[1]) + m[2] * v[2]

    rot = np.array([mulv(ry, rx[c]) for c in range(3)], np.float32)
    height = f(2.0) * f(math.tan(fov * f(0.5)))
    h = mulv(rot, np.array([1, 0, 0], np.float32)) * f(aspect) * height
    v = mulv(rot, np.array([0, -1, 0], np.float32)) * height
    llc = mulv(rot, np.array([0, 0, -1], np.float32)) - (h + v) * f(0.5)
    cam = PtCamera()
    for k in range(3):
        cam.position[k] = position[k]
        cam.lower_left_corner_minus_position[k] = llc[k]
        cam.horizontal[k] = h[k]
        cam.vertical[k] = v[k]
    return cam
