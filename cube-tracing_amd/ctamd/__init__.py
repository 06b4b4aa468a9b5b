"""ctamd -- Python binding of the MI355X-native cube-tracing ray-tracing path.

A thin ctypes layer over ``libctamd.so`` (built from ``cube-tracing_amd/csrc``)
whose C ABI is declared in ``include/ct_abi.h``.  The library holds the whole
hot path: the acceleration-structure manager (ACCEL_STRUCT_MNGR,
src/accel_struct_mngr.hpp:18-664 in the reference), the GPU LBVH build, the
traversal kernels and the ReSTIR DI frame passes.  This module only converts
arguments and raises :class:`CtError` on a negative status; there is no
fallback path -- without the built library every call raises.
"""

from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libctamd.so")

CT_OK = 0
CT_E_INVALID = -1
CT_E_CAPACITY = -2
CT_E_HIP = -3
CT_E_UNSUPPORTED = -4
CT_E_STATE = -5
CT_E_NODEVICE = -6

CT_MAX_INSTANCES = 983040
CT_MAX_PRIMITIVES = 134217728
CT_VOXEL_EXTENT = 0.03125
CT_INVALID_INDEX = 0xFFFFFFFF
CT_MAX_NEIGHBORS = 8

CT_TAA_BIT = 1 << 0
CT_RIS_POINT_LIGHT_BIT = 1 << 1
CT_RIS_ENV_LIGHT_BIT = 1 << 2
CT_RIS_CUBE_LIGHT_BIT = 1 << 3
CT_RIS_BRDF_BIT = 1 << 4
CT_REMAP_BIT = 1 << 30

CT_AS_IDLE, CT_AS_UPDATING, CT_AS_SWITCHING, CT_AS_SETTLING = 0, 1, 2, 4

CT_LIGHT_NONE, CT_LIGHT_POINT, CT_LIGHT_CUBE, CT_LIGHT_ENV_MAP = 0, 1, 2, 5

CT_MATERIAL_LAMBERTIAN, CT_MATERIAL_METAL, CT_MATERIAL_DIELECTRIC, CT_MATERIAL_CONSTANT_MEDIUM = 0, 1, 2, 3

CT_TASK_BUILD_BLAS_FROM_CPU = 0
CT_TASK_DELETE_PRIMITIVE_BLAS_FROM_CPU = 1
CT_TASK_DELETE_BLAS_FROM_CPU = 2
CT_TASK_UPDATE_BLAS_FROM_CPU = 3
CT_TASK_DELETE_PRIMITIVE_BLAS_FROM_GPU = 4
CT_TASK_UNDO_OP_CPU = 5

CT_BUF_IMAGE, CT_BUF_PREV_IMAGE, CT_BUF_TAA_FRAME, CT_BUF_DI = 0, 1, 2, 3
CT_BUF_PREV_DI, CT_BUF_INT_RESERVOIR, CT_BUF_PREV_RESERVOIR, CT_BUF_VELOCITY = 4, 5, 6, 7
CT_BUF_PATH_RESERVOIR, CT_BUF_TEMPORAL_PATH_RESERVOIR, CT_BUF_INDIRECT = 8, 9, 10

CT_SHIFT_RECONNECTION, CT_SHIFT_HYBRID = 0, 2
CT_MAX_PT_NEIGHBORS = 8
CT_MAX_PT_DEPTH = 14
CT_ADJOINT_MAX_DEPTH = 7

(CT_AXIS_X_BOTTOM_TOP, CT_AXIS_Y_BOTTOM_TOP, CT_AXIS_Z_BOTTOM_TOP,
 CT_AXIS_X_TOP_BOTTOM, CT_AXIS_Y_TOP_BOTTOM, CT_AXIS_Z_TOP_BOTTOM) = range(6)

f32, u32, i32 = C.c_float, C.c_uint32, C.c_int32


# --------------------------------------------------------------- structs
class Aabb(C.Structure):
    _fields_ = [("minimum", f32 * 3), ("maximum", f32 * 3)]


class ObjectInfo(C.Structure):
    _fields_ = [("instance_id", u32), ("primitive_id", u32)]


class Instance(C.Structure):
    _fields_ = [("transform", f32 * 16), ("first_primitive_index", u32), ("primitive_count", u32)]


class Primitive(C.Structure):
    _fields_ = [("material_index", u32), ("light_index", u32)]


class Material(C.Structure):
    _fields_ = [("type", u32), ("ambient", f32 * 3), ("diffuse", f32 * 3), ("specular", f32 * 3),
                ("transmittance", f32 * 3), ("emission", f32 * 3), ("shininess", f32),
                ("roughness", f32), ("ior", f32), ("dissolve", f32), ("illum", i32)]


class Light(C.Structure):
    _fields_ = [("position", f32 * 3), ("emissive", f32 * 3), ("instance_info", ObjectInfo),
                ("size", f32), ("type", u32)]


class LightConfig(C.Structure):
    _fields_ = [("point_light_count", u32), ("point_light_pdf", f32), ("cube_light_count", u32),
                ("cube_light_pdf", f32), ("sphere_light_count", u32), ("sphere_light_pdf", f32),
                ("analytic_light_count", u32), ("analytic_light_pdf", f32), ("env_map_count", u32),
                ("env_map_pdf", f32), ("brdf_count", u32), ("brdf_pdf", f32), ("light_count", u32)]


class Reservoir(C.Structure):
    _fields_ = [("Y", u32), ("l_type", u32), ("seed", u32), ("W_y", f32), ("W_sum", f32), ("M", f32),
                ("F", f32 * 3)]


class DiInfo(C.Structure):
    _fields_ = [("position", f32 * 3), ("distance", f32), ("normal", f32 * 3), ("ray_origin", f32 * 3),
                ("seed", u32), ("instance_hit", ObjectInfo), ("mat_index", u32), ("confidence", f32)]


class AccelDesc(C.Structure):
    _fields_ = [("max_instance_count", u32), ("max_primitive_count", u32),
                ("max_cube_light_count", u32), ("device", i32)]


class Staging(C.Structure):
    _fields_ = [("instances", C.POINTER(Instance)), ("primitives", C.POINTER(Primitive)),
                ("aabbs", C.POINTER(Aabb)), ("cube_lights", C.POINTER(Light)),
                ("host_instance_count", u32), ("host_primitive_count", u32), ("cube_light_count", u32),
                ("max_instance_count", u32), ("max_primitive_count", u32), ("max_cube_light_count", u32)]


class _TaskBuild(C.Structure):
    _fields_ = [("instance_count", u32), ("primitive_count", u32), ("transform", f32 * 16)]


class _TaskDelPrim(C.Structure):
    _fields_ = [("instance_index", u32), ("del_primitive_index", u32), ("remap_primitive_index", u32),
                ("del_light_index", u32), ("remap_light_index", u32), ("remap_primitive_light_index", u32)]


class _TaskDelete(C.Structure):
    _fields_ = [("instance_index", u32), ("first_primitive_index", u32), ("deleted_primitive_count", u32)]


class _TaskUpdate(C.Structure):
    _fields_ = [("instance_index", u32), ("transform", f32 * 16), ("primitive_count", u32),
                ("primitive_index_buf_offset", u32), ("aabb_buf_offset", u32)]


class _TaskDelGpu(C.Structure):
    _fields_ = [("instance_index", u32), ("del_prim_count", u32)]


class _TaskUnion(C.Union):
    _fields_ = [("blas_build_from_cpu", _TaskBuild), ("blas_delete_primitive_from_cpu", _TaskDelPrim),
                ("blas_delete_from_cpu", _TaskDelete), ("blas_update", _TaskUpdate),
                ("blas_del_prim_gpu", _TaskDelGpu)]


class Task(C.Structure):
    _fields_ = [("type", u32), ("u", _TaskUnion)]


class Tlas(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("leaves", C.c_void_p), ("root_ref", u32), ("instance_count", u32),
                ("depth", u32), ("generation", u32)]


class World(C.Structure):
    _fields_ = [("instances", C.c_void_p), ("primitives", C.c_void_p), ("aabbs", C.c_void_p),
                ("cube_lights", C.c_void_p), ("remapped_primitives", C.c_void_p),
                ("remapped_cube_lights", C.c_void_p), ("instance_count", u32), ("primitive_capacity", u32),
                ("cube_light_count", u32), ("cube_light_capacity", u32)]


class BvhArrays(C.Structure):
    _fields_ = [("keys", C.POINTER(u32)), ("ids", C.POINTER(u32)), ("child", C.POINTER(u32)),
                ("boxes", C.POINTER(f32)), ("parent", C.POINTER(u32)), ("n", u32), ("depth", u32),
                ("root_ref", u32), ("pad", u32)]


class TraceStats(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("nodes", C.c_uint64), ("leaves", C.c_uint64), ("instances", C.c_uint64)]


class CtxDesc(C.Structure):
    _fields_ = [("max_width", u32), ("max_height", u32), ("device", i32), ("pad", u32)]


class RestirParams(C.Structure):
    _fields_ = [("min_ris_point_samples", u32), ("max_ris_point_samples", u32),
                ("min_ris_env_samples", u32), ("max_ris_env_samples", u32),
                ("min_ris_cube_samples", u32), ("max_ris_cube_samples", u32),
                ("max_influence_from_the_past", f32), ("min_neighbors", u32), ("max_neighbors", u32),
                ("min_neighbor_radius", f32), ("max_neighbor_radius", f32), ("max_distance_to_hit", f32),
                ("taa_alpha", f32), ("taa_truncate_dispatch", u32), ("direct_emittance", u32),
                ("indirect", u32), ("max_depth", u32), ("pt_temporal", u32), ("pt_spatial", u32),
                ("pt_shift_mapping", u32), ("pt_neighbor_count", u32), ("pt_neighbor_radius", i32),
                ("pt_m_cap", f32), ("near_field_distance", f32), ("roughness_threshold", f32),
                ("pt_russian_roulette", u32), ("pad", u32)]


class ObjectHit(C.Structure):
    _fields_ = [("object", ObjectInfo), ("hit", f32 * 3)]


class PathReservoir(C.Structure):
    _fields_ = [("M", f32), ("weight", f32), ("path_flags", u32), ("rc_random_seed", u32), ("F", f32 * 3),
                ("light_pdf", f32), ("cached_jacobian", f32 * 3), ("init_random_seed", u32),
                ("rc_vertex_hit", ObjectHit), ("rc_vertex_wi", f32 * 3), ("rc_vertex_irradiance", f32 * 3)]


class FrameParams(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("frame_number", u32), ("is_active", u32), ("time", f32),
                ("is_afternoon", u32), ("has_previous", u32), ("pad", u32), ("inv_view", f32 * 16),
                ("inv_proj", f32 * 16), ("prev_view_proj", f32 * 16), ("restir", RestirParams)]


class FrameStats(C.Structure):
    _fields_ = [("primary_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("shadow_rays_ris", C.c_uint64),
                ("shadow_rays_temporal", C.c_uint64), ("shadow_rays_spatial", C.c_uint64),
                ("nodes", C.c_uint64), ("leaves", C.c_uint64), ("instances", C.c_uint64),
                ("closest_nodes", C.c_uint64), ("closest_leaves", C.c_uint64), ("closest_instances", C.c_uint64),
                ("ms_total", f32), ("ms_primary", f32), ("ms_ris", f32), ("ms_temporal", f32),
                ("ms_spatial", f32), ("ms_taa", f32), ("ms_trace_any", f32), ("ms_trace_closest", f32),
                ("bounce_rays", C.c_uint64), ("shadow_rays_pt", C.c_uint64), ("pt_nodes", C.c_uint64),
                ("pt_leaves", C.c_uint64), ("pt_instances", C.c_uint64), ("ms_pt", f32), ("ms_pt_spatial", f32),
                ("ms_pt_segments", f32), ("pad", f32)]


class AdjointView(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("inv_view", f32 * 16), ("inv_proj", f32 * 16)]


class AdjointDesc(C.Structure):
    _fields_ = [("n_views", u32), ("spp", u32), ("max_depth", u32), ("seed", u32)]


class BatchTuning(C.Structure):
    _fields_ = [("blocks_per_cu", u32), ("ms", f32), ("candidates", u32), ("groups", u32), ("far_ray", f32)]


class SceneDesc(C.Structure):
    _fields_ = [("grid", u32 * 3), ("chunk", u32 * 3), ("occupied_count", u32), ("emissive_count", u32),
                ("le_min", f32), ("le_max", f32), ("albedo_min", f32), ("albedo_max", f32), ("seed", u32),
                ("palette_size", u32), ("voxel_extent", f32), ("palette_emissive", u32)]


class VoxDesc(C.Structure):
    _fields_ = [("axis", u32), ("chunk", u32 * 3), ("model", C.c_int32), ("voxel_extent", f32), ("flux", f32),
                ("pad", u32)]


class SceneInfo(C.Structure):
    _fields_ = [("instance_count", u32), ("primitive_count", u32), ("material_count", u32),
                ("cube_light_count", u32), ("instances", C.POINTER(Instance)), ("aabbs", C.POINTER(Aabb)),
                ("primitives", C.POINTER(Primitive)), ("materials", C.POINTER(Material)),
                ("cube_lights", C.POINTER(Light)), ("grid_min", f32 * 3), ("grid_max", f32 * 3)]


STRUCT_SIZES = {"ct_aabb": (Aabb, 24), "ct_instance": (Instance, 72), "ct_primitive": (Primitive, 8),
                "ct_material": (Material, 84), "ct_light": (Light, 40), "ct_light_config": (LightConfig, 52),
                "ct_reservoir": (Reservoir, 36), "ct_di_info": (DiInfo, 60), "ct_path_reservoir": (PathReservoir, 92),
                "ct_restir_params": (RestirParams, 108), "ct_adjoint_view": (AdjointView, 136),
                "ct_adjoint_desc": (AdjointDesc, 16), "ct_batch_tuning": (BatchTuning, 20),
                "ct_task": (Task, 84), "ct_world": (World, 64), "ct_vox_desc": (VoxDesc, 32)}


# ----------------------------------------------------------------- loader
class CtError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"ct error {code}: {message}")
        self.code = code


_lib = None

_SIGS = {
    "ct_last_error_message": ([], C.c_char_p),
    "ct_device_available": ([], C.c_int),
    "ct_accel_create": ([C.POINTER(AccelDesc), C.POINTER(C.c_void_p)], C.c_int),
    "ct_accel_destroy": ([C.c_void_p], C.c_int),
    "ct_accel_get_staging": ([C.c_void_p, C.POINTER(Staging)], C.c_int),
    "ct_accel_set_cube_light_count": ([C.c_void_p, u32], C.c_int),
    "ct_accel_request_aabb_host_buffer_count": ([C.c_void_p, u32, C.POINTER(C.POINTER(Aabb)), C.POINTER(u32)], C.c_int),
    "ct_accel_request_primitive_index_host_buffer_count": ([C.c_void_p, u32, C.POINTER(C.POINTER(u32)),
                                                            C.POINTER(u32)], C.c_int),
    "ct_accel_read_world": ([C.c_void_p, C.POINTER(Aabb), C.POINTER(Primitive), u32, C.POINTER(Light), u32,
                             C.POINTER(u32), C.POINTER(u32)], C.c_int),
    "ct_accel_task_queue_add": ([C.c_void_p, C.POINTER(Task)], C.c_int),
    "ct_accel_update_scene": ([C.c_void_p, C.c_int, C.c_void_p, C.POINTER(C.c_int)], C.c_int),
    "ct_accel_get_status": ([C.c_void_p, C.POINTER(u32)], C.c_int),
    "ct_accel_is_remapping_active": ([C.c_void_p, C.POINTER(C.c_int)], C.c_int),
    "ct_accel_get_tlas": ([C.c_void_p, C.c_int, C.POINTER(Tlas)], C.c_int),
    "ct_accel_get_world": ([C.c_void_p, C.c_int, C.POINTER(World)], C.c_int),
    "ct_accel_read_bvh": ([C.c_void_p, u32, C.POINTER(BvhArrays)], C.c_int),
    "ct_accel_get_world_to_object": ([C.c_void_p, u32, C.POINTER(f32)], C.c_int),
    "ct_trace_closest": ([C.c_void_p, u32, C.c_void_p, C.c_void_p, f32, C.c_void_p, C.c_void_p,
                          C.POINTER(TraceStats)], C.c_int),
    "ct_trace_any": ([C.c_void_p, u32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                      C.POINTER(TraceStats)], C.c_int),
    "ct_trace_any_noalloc": ([C.c_void_p, u32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                              C.c_void_p], C.c_int),
    "ct_accel_read_instances": ([C.c_void_p, C.POINTER(Instance), u32, C.POINTER(u32)], C.c_int),
    "ct_adjoint_render": ([C.c_void_p, C.POINTER(AdjointDesc), C.POINTER(AdjointView), C.c_void_p, C.c_void_p,
                           C.c_void_p, C.c_void_p], C.c_int),
    "ct_adjoint_backward": ([C.c_void_p, C.POINTER(AdjointDesc), C.POINTER(AdjointView), C.c_void_p, C.c_void_p,
                             C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "ct_trace_bounce": ([C.c_void_p, u32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                         C.c_void_p, C.c_void_p, C.POINTER(TraceStats)], C.c_int),
    "ct_ctx_create": ([C.POINTER(CtxDesc), C.POINTER(C.c_void_p)], C.c_int),
    "ct_ctx_destroy": ([C.c_void_p], C.c_int),
    "ct_ctx_set_materials": ([C.c_void_p, C.POINTER(Material), u32], C.c_int),
    "ct_ctx_append_materials": ([C.c_void_p, C.POINTER(Material), u32, C.POINTER(u32)], C.c_int),
    "ct_ctx_set_point_lights": ([C.c_void_p, C.POINTER(Light), u32], C.c_int),
    "ct_ctx_set_env_lights": ([C.c_void_p, C.POINTER(Light), u32], C.c_int),
    "ct_ctx_set_light_config": ([C.c_void_p, C.POINTER(LightConfig)], C.c_int),
    "ct_ctx_get_light_config": ([C.c_void_p, C.POINTER(LightConfig)], C.c_int),
    "ct_ctx_reset": ([C.c_void_p, C.c_void_p], C.c_int),
    "ct_restir_params_reference": ([C.POINTER(RestirParams)], None),
    "ct_frame_render": ([C.c_void_p, C.c_void_p, C.POINTER(FrameParams), C.c_void_p], C.c_int),
    "ct_frame_read": ([C.c_void_p, u32, C.c_void_p, C.c_size_t], C.c_int),
    "ct_ctx_set_instrumentation": ([C.c_void_p, u32], C.c_int),
    "ct_frame_get_stats": ([C.c_void_p, C.POINTER(FrameStats)], C.c_int),
    "ct_frame_adjoint": ([C.c_void_p, C.c_void_p, C.POINTER(FrameParams), C.c_void_p, C.c_void_p, C.c_void_p,
                          C.c_void_p], C.c_int),
    "ct_batch_create": ([u32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)], C.c_int),
    "ct_batch_destroy": ([C.c_void_p], C.c_int),
    "ct_batch_render": ([C.c_void_p, C.POINTER(FrameParams), C.c_int, C.c_void_p], C.c_int),
    "ct_batch_get_stats": ([C.c_void_p, C.POINTER(FrameStats)], C.c_int),
    "ct_batch_autotune": ([C.c_void_p, C.POINTER(FrameParams), u32, C.POINTER(BatchTuning)], C.c_int),
    "ct_camera_matrices": ([C.POINTER(f32), C.POINTER(f32), C.POINTER(f32), f32, f32, f32, f32,
                            C.POINTER(f32), C.POINTER(f32)], C.c_int),
    "ct_camera_prev_view_proj": ([C.POINTER(f32), C.POINTER(f32), C.POINTER(f32)], C.c_int),
    "ct_mat4_inverse": ([C.POINTER(f32), C.POINTER(f32)], C.c_int),
    "ct_scene_generate": ([C.POINTER(SceneDesc), C.POINTER(C.c_void_p)], C.c_int),
    "ct_scene_get_info": ([C.c_void_p, C.POINTER(SceneInfo)], C.c_int),
    "ct_scene_destroy": ([C.c_void_p], C.c_int),
    "ct_scene_load": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "ct_scene_bind": ([C.c_void_p, C.c_void_p], C.c_int),
    "ct_brush_apply": ([C.c_void_p, C.c_void_p, u32, u32, u32, C.c_void_p], C.c_int),
    "ct_brush_mark": ([C.c_void_p, C.POINTER(ObjectInfo), u32, C.c_void_p], C.c_int),
    "ct_brush_get_counters": ([C.c_void_p, C.POINTER(u32), C.POINTER(u32)], C.c_int),
    "ct_accel_process_voxel_modifications": ([C.c_void_p, C.c_void_p, C.POINTER(u32), C.POINTER(u32)], C.c_int),
    "ct_path_trace": ([C.c_void_p, C.POINTER(AdjointDesc), C.POINTER(AdjointView), C.POINTER(Material), u32,
                       C.c_void_p, C.c_void_p], C.c_int),
    "ct_pt_random_replay": ([C.c_void_p, C.c_void_p, C.POINTER(FrameParams), C.c_int32, C.c_int32, C.c_void_p,
                             C.c_void_p], C.c_int),
    "ct_scene_load_transformed": ([C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(f32), C.c_void_p], C.c_int),
    "ct_scene_from_vox": ([C.c_char_p, C.c_size_t, C.POINTER(VoxDesc), C.POINTER(C.c_void_p)], C.c_int),
}


def lib():
    """Load libctamd.so; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CtError(CT_E_NODEVICE, f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        # PyTorch-ROCm ships its own HIP runtime.  Two HIP runtimes in one
        # process do not share devices, so when torch is installed it is
        # loaded first and the library binds to the same runtime (device
        # pointers then pass freely between torch tensors and the ABI).
        try:
            import torch  # noqa: F401
        except Exception:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(rc: int):
    if rc != CT_OK:
        msg = lib().ct_last_error_message()
        raise CtError(rc, msg.decode() if msg else "")


def device_available() -> bool:
    return bool(lib().ct_device_available())


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(f32))


# ------------------------------------------------------------------- math
def restir_params_reference() -> RestirParams:
    p = RestirParams()
    lib().ct_restir_params_reference(C.byref(p))
    return p


def camera_matrices(position, forward, up, fov_radians, aspect, z_near=0.001, z_far=1000.0):
    """inv_view, inv_proj (y.y negated) as camera.h + main.cpp:1452-1460 build them."""
    pos = np.asarray(position, np.float32)
    fwd = np.asarray(forward, np.float32)
    upv = np.asarray(up, np.float32)
    iv = np.zeros(16, np.float32)
    ip = np.zeros(16, np.float32)
    check(lib().ct_camera_matrices(_fp(pos), _fp(fwd), _fp(upv), fov_radians, aspect, z_near, z_far, _fp(iv), _fp(ip)))
    return iv, ip


def prev_view_proj(prev_inv_view, prev_inv_proj) -> np.ndarray:
    a = np.ascontiguousarray(prev_inv_view, np.float32)
    b = np.ascontiguousarray(prev_inv_proj, np.float32)
    out = np.zeros(16, np.float32)
    check(lib().ct_camera_prev_view_proj(_fp(a), _fp(b), _fp(out)))
    return out


def mat4_inverse(m) -> np.ndarray:
    a = np.ascontiguousarray(m, np.float32)
    out = np.zeros(16, np.float32)
    check(lib().ct_mat4_inverse(_fp(a), _fp(out)))
    return out


# ------------------------------------------------------------------ scene
def _arr(ptr, ctype, n):
    if n == 0:
        return (ctype * 0)()
    return C.cast(ptr, C.POINTER(ctype * n)).contents


class Scene:
    """A synthetic voxel scene generated by the library (host side).  The
    instance, AABB, primitive, material and light arrays are copied out of
    the library on first use."""

    def __init__(self, desc: SceneDesc = None, handle=None):
        self.desc = desc
        h = C.c_void_p()
        if handle is None:
            check(lib().ct_scene_generate(C.byref(desc), C.byref(h)))
        else:
            h = handle
        self.handle = h
        info = SceneInfo()
        check(lib().ct_scene_get_info(h, C.byref(info)))
        self.info = info
        self.grid_min = np.array(info.grid_min[:], np.float32)
        self.grid_max = np.array(info.grid_max[:], np.float32)
        self._cache = {}

    @classmethod
    def from_vox(cls, data, axis=0, chunk=(0, 0, 0), model=-1, voxel_extent=0.0, flux=0.0):
        """A MagicaVoxel .vox file (bytes, or a path) as a scene, with the
        map loader's rules (map_loader.cpp:97-325); axis 0 = X_BOTTOM_TOP
        as main.cpp:1009 loads it."""
        if not isinstance(data, (bytes, bytearray)):
            with open(data, "rb") as fh:
                data = fh.read()
        d = VoxDesc()
        d.axis = axis
        d.chunk[:] = list(chunk)
        d.model = model
        d.voxel_extent = voxel_extent
        d.flux = flux
        h = C.c_void_p()
        check(lib().ct_scene_from_vox(bytes(data), len(data), C.byref(d), C.byref(h)))
        return cls(None, handle=h)

    def _get(self, name):
        if name not in self._cache:
            info = self.info
            ni, np_, nm, nl = info.instance_count, info.primitive_count, info.material_count, info.cube_light_count
            if name == "instances":
                v = (Instance * ni).from_buffer_copy(_arr(info.instances, Instance, ni))
            elif name == "aabbs":
                v = np.ctypeslib.as_array(C.cast(info.aabbs, C.POINTER(f32)), shape=(np_ * 6,)).reshape(np_, 6).copy()
            elif name == "primitives":
                v = np.ctypeslib.as_array(C.cast(info.primitives, C.POINTER(u32)), shape=(np_ * 2,)).reshape(np_, 2).copy()
            elif name == "materials":
                v = (Material * nm).from_buffer_copy(_arr(info.materials, Material, nm))
            else:
                v = (Light * nl).from_buffer_copy(_arr(info.cube_lights, Light, nl)) if nl else (Light * 0)()
            self._cache[name] = v
        return self._cache[name]

    instances = property(lambda self: self._get("instances"))
    aabbs = property(lambda self: self._get("aabbs"))
    primitives = property(lambda self: self._get("primitives"))
    materials = property(lambda self: self._get("materials"))
    lights = property(lambda self: self._get("lights"))

    def close(self):
        if self.handle:
            lib().ct_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Accel:
    """ACCEL_STRUCT_MNGR equivalent: owns BLAS/TLAS and the scene buffers on the GPU."""

    def __init__(self, max_instances=4096, max_primitives=1 << 22, max_cube_lights=1 << 18, device=0):
        d = AccelDesc(max_instances, max_primitives, max_cube_lights, device)
        h = C.c_void_p()
        check(lib().ct_accel_create(C.byref(d), C.byref(h)))
        self.handle = h

    def staging(self) -> Staging:
        s = Staging()
        check(lib().ct_accel_get_staging(self.handle, C.byref(s)))
        return s

    def task_queue_add(self, task: Task):
        check(lib().ct_accel_task_queue_add(self.handle, C.byref(task)))

    def update_scene(self, synchronize=True) -> bool:
        did = C.c_int(0)
        check(lib().ct_accel_update_scene(self.handle, 1 if synchronize else 0, None, C.byref(did)))
        return bool(did.value)

    def tlas(self, previous=False) -> Tlas:
        """get_current_tlas, or get_previous_tlas with previous=True (the
        TLAS from before the last update while SETTLING)."""
        t = Tlas()
        check(lib().ct_accel_get_tlas(self.handle, 1 if previous else 0, C.byref(t)))
        return t

    def world_to_object(self, instance: int) -> np.ndarray:
        out = np.zeros(16, np.float32)
        check(lib().ct_accel_get_world_to_object(self.handle, instance, _fp(out)))
        return out

    def read_instances(self):
        cnt = u32(0)
        check(lib().ct_accel_read_instances(self.handle, None, 0, C.byref(cnt)))
        out = (Instance * max(cnt.value, 1))()
        check(lib().ct_accel_read_instances(self.handle, out, cnt.value, C.byref(cnt)))
        return out[: cnt.value]

    def status(self) -> int:
        s = u32(0)
        check(lib().ct_accel_get_status(self.handle, C.byref(s)))
        return s.value

    def remapping_active(self) -> bool:
        a = C.c_int(0)
        check(lib().ct_accel_is_remapping_active(self.handle, C.byref(a)))
        return bool(a.value)

    def world(self, previous=False) -> World:
        w = World()
        check(lib().ct_accel_get_world(self.handle, 1 if previous else 0, C.byref(w)))
        return w

    def read_world(self) -> dict:
        """Device contents of the WORLD buffers: aabbs [cap, 6], primitives
        [cap, 2] by slot, the cube lights held (stale records past the
        count included), and the two remapping tables."""
        w = self.world()
        cap, lcap = w.primitive_capacity, max(w.cube_light_capacity, 1)
        aabbs = np.zeros((max(cap, 1), 6), np.float32)
        prims = np.zeros((max(cap, 1), 2), np.uint32)
        lights = (Light * lcap)()
        rp = np.zeros(max(cap, 1), np.uint32)
        rl = np.zeros(lcap, np.uint32)
        check(lib().ct_accel_read_world(self.handle, aabbs.ctypes.data_as(C.POINTER(Aabb)),
                                        prims.ctypes.data_as(C.POINTER(Primitive)), cap, lights, lcap,
                                        rp.ctypes.data_as(C.POINTER(u32)), rl.ctypes.data_as(C.POINTER(u32))))
        return {"aabbs": aabbs[:cap], "prims": prims[:cap], "lights": list(lights), "light_count": w.cube_light_count,
                "remap_prims": rp[:cap], "remap_lights": rl, "capacity": cap}

    # -- tasks (TASK kinds, accel_struct_mngr.hpp:40-107)
    def delete_blas(self, instance: int):
        t = Task()
        t.type = CT_TASK_DELETE_BLAS_FROM_CPU
        t.u.blas_delete_from_cpu.instance_index = instance
        self.task_queue_add(t)

    def delete_primitive(self, instance: int, primitive: int):
        t = Task()
        t.type = CT_TASK_DELETE_PRIMITIVE_BLAS_FROM_CPU
        t.u.blas_delete_primitive_from_cpu.instance_index = instance
        t.u.blas_delete_primitive_from_cpu.del_primitive_index = primitive
        self.task_queue_add(t)

    def delete_primitives_from_gpu(self, instance: int, count: int):
        t = Task()
        t.type = CT_TASK_DELETE_PRIMITIVE_BLAS_FROM_GPU
        t.u.blas_del_prim_gpu.instance_index = instance
        t.u.blas_del_prim_gpu.del_prim_count = count
        self.task_queue_add(t)

    def update_blas(self, instance: int, transform=None, edits=()):
        """transform (column-major 16, multiplied onto the instance's) and
        edits [(primitive index, aabb min+max)] staged for the task."""
        t = Task()
        t.type = CT_TASK_UPDATE_BLAS_FROM_CPU
        t.u.blas_update.instance_index = instance
        tr = np.eye(4, dtype=np.float32).reshape(16) if transform is None else np.asarray(transform, np.float32)
        t.u.blas_update.transform[:] = tr.reshape(16).tolist()
        edits = list(edits)
        if edits:
            pbase, poff = C.POINTER(u32)(), u32(0)
            check(lib().ct_accel_request_primitive_index_host_buffer_count(self.handle, len(edits), C.byref(pbase),
                                                                            C.byref(poff)))
            abase, aoff = C.POINTER(Aabb)(), u32(0)
            check(lib().ct_accel_request_aabb_host_buffer_count(self.handle, len(edits), C.byref(abase),
                                                                 C.byref(aoff)))
            for k, (p, box) in enumerate(edits):
                pbase[poff.value + k] = int(p)
                b = abase[aoff.value + k]
                b.minimum[:] = [float(v) for v in box[:3]]
                b.maximum[:] = [float(v) for v in box[3:6]]
            t.u.blas_update.primitive_count = len(edits)
            t.u.blas_update.primitive_index_buf_offset = poff.value
            t.u.blas_update.aabb_buf_offset = aoff.value
        self.task_queue_add(t)

    def undo(self):
        t = Task()
        t.type = CT_TASK_UNDO_OP_CPU
        self.task_queue_add(t)

    def brush_mark(self, voxels):
        """Marks (instance id, primitive id) pairs for deletion (delete_primitive_from_instance)."""
        v = np.ascontiguousarray(voxels, np.uint32).reshape(-1, 2)
        arr = (ObjectInfo * max(len(v), 1))()
        C.memmove(arr, v.ctypes.data, 8 * len(v))
        check(lib().ct_brush_mark(self.handle, arr, len(v), None))

    def brush_counters(self):
        """(newly marked instances, newly marked primitives), BRUSH_COUNTER."""
        i, p = u32(0), u32(0)
        check(lib().ct_brush_get_counters(self.handle, C.byref(i), C.byref(p)))
        return i.value, p.value

    def process_voxel_modifications(self):
        """Compacts the marked instances on the device and queues their
        DELETE_PRIMITIVE_BLAS_FROM_GPU tasks; returns (tasks, voxels)."""
        i, p = u32(0), u32(0)
        check(lib().ct_accel_process_voxel_modifications(self.handle, None, C.byref(i), C.byref(p)))
        return i.value, p.value

    def read_bvh(self, instance: int, n: int) -> dict:
        keys = np.zeros(n, np.uint32)
        ids = np.zeros(n, np.uint32)
        child = np.zeros(max(2 * (n - 1), 1), np.uint32)
        boxes = np.zeros(max(12 * (n - 1), 1), np.float32)
        parent = np.zeros(2 * n - 1, np.uint32)
        a = BvhArrays(keys.ctypes.data_as(C.POINTER(u32)), ids.ctypes.data_as(C.POINTER(u32)),
                      child.ctypes.data_as(C.POINTER(u32)), boxes.ctypes.data_as(C.POINTER(f32)),
                      parent.ctypes.data_as(C.POINTER(u32)), 0, 0, 0, 0)
        check(lib().ct_accel_read_bvh(self.handle, instance, C.byref(a)))
        return {"keys": keys, "ids": ids, "child": child[: 2 * (n - 1)], "boxes": boxes[: 12 * (n - 1)],
                "parent": parent, "depth": a.depth, "root_ref": a.root_ref, "n": a.n}

    def trace_closest(self, origins, directions, tmax=1e9, stats=False):
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(directions, np.float32)
        n = o.shape[0]
        t = np.zeros(n, np.float32)
        ids = np.zeros((n, 2), np.uint32)
        st = TraceStats()
        check(lib().ct_trace_closest(self.handle, n, o.ctypes.data, d.ctypes.data, tmax, t.ctypes.data,
                                     ids.ctypes.data, C.byref(st) if stats else None))
        return (t, ids, st) if stats else (t, ids)

    def trace_any(self, origins, directions, distances, targets=None, stats=False):
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(directions, np.float32)
        dist = np.ascontiguousarray(distances, np.float32)
        n = o.shape[0]
        tg = None if targets is None else np.ascontiguousarray(targets, np.uint32)
        occ = np.zeros(n, np.uint8)
        st = TraceStats()
        check(lib().ct_trace_any(self.handle, n, o.ctypes.data, d.ctypes.data, dist.ctypes.data,
                                 None if tg is None else tg.ctypes.data, occ.ctypes.data,
                                 C.byref(st) if stats else None))
        return (occ, st) if stats else occ

    def trace_bounce(self, origins, directions, stats=False):
        """intersect() of bounce.glsl: ids, position, normal, distance, wo per ray."""
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(directions, np.float32)
        n = o.shape[0]
        ids = np.zeros((n, 2), np.uint32)
        pos = np.zeros((n, 3), np.float32)
        nrm = np.zeros((n, 3), np.float32)
        dist = np.zeros(n, np.float32)
        wo = np.zeros((n, 3), np.float32)
        st = TraceStats()
        check(lib().ct_trace_bounce(self.handle, n, o.ctypes.data, d.ctypes.data, ids.ctypes.data, pos.ctypes.data,
                                    nrm.ctypes.data, dist.ctypes.data, wo.ctypes.data,
                                    C.byref(st) if stats else None))
        out = {"ids": ids, "pos": pos, "nrm": nrm, "dist": dist, "wo": wo}
        if stats:
            out["stats"] = st
        return out

    def close(self):
        if self.handle:
            lib().ct_accel_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ctx:
    """Frame context: DI/reservoir/image buffers and the per-frame passes."""

    def __init__(self, max_width, max_height, device=0):
        d = CtxDesc(max_width, max_height, device, 0)
        h = C.c_void_p()
        check(lib().ct_ctx_create(C.byref(d), C.byref(h)))
        self.handle = h

    def light_config(self) -> LightConfig:
        c = LightConfig()
        check(lib().ct_ctx_get_light_config(self.handle, C.byref(c)))
        return c

    def reset(self):
        check(lib().ct_ctx_reset(self.handle, None))

    def set_instrumentation(self, flags: int):
        check(lib().ct_ctx_set_instrumentation(self.handle, flags))

    def render(self, accel: Accel, params: FrameParams, stream=None):
        """Renders one frame on `stream` (a hipStream_t as an int; None: the
        context's own stream)."""
        check(lib().ct_frame_render(self.handle, accel.handle, C.byref(params), C.c_void_p(stream) if stream else None))

    def stats(self) -> FrameStats:
        s = FrameStats()
        check(lib().ct_frame_get_stats(self.handle, C.byref(s)))
        return s

    def read(self, which: int, w: int, h: int):
        n = w * h
        if which in (CT_BUF_IMAGE, CT_BUF_PREV_IMAGE, CT_BUF_TAA_FRAME):
            out = np.zeros((n, 4), np.float32)
        elif which == CT_BUF_VELOCITY:
            out = np.zeros((n, 2), np.float32)
        elif which in (CT_BUF_DI, CT_BUF_PREV_DI):
            out = (DiInfo * n)()
        elif which in (CT_BUF_PATH_RESERVOIR, CT_BUF_TEMPORAL_PATH_RESERVOIR):
            out = (PathReservoir * n)()
        elif which == CT_BUF_INDIRECT:
            out = np.zeros((n, 3), np.float32)
        else:
            out = (Reservoir * n)()
        nbytes = out.nbytes if isinstance(out, np.ndarray) else C.sizeof(out)
        ptr = out.ctypes.data if isinstance(out, np.ndarray) else C.addressof(out)
        check(lib().ct_frame_read(self.handle, which, ptr, nbytes))
        return out

    def close(self):
        if self.handle:
            lib().ct_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Several views (context + accel each) rendered with one launch per pass."""

    def __init__(self, ctxs, accels):
        n = len(ctxs)
        self.ctxs, self.accels = list(ctxs), list(accels)
        c = (C.c_void_p * n)(*[x.handle for x in ctxs])
        a = (C.c_void_p * n)(*[x.handle for x in accels])
        h = C.c_void_p()
        check(lib().ct_batch_create(n, c, a, C.byref(h)))
        self.handle = h

    def render(self, params, use_graph=True):
        arr = (FrameParams * len(params))(*params)
        check(lib().ct_batch_render(self.handle, arr, 1 if use_graph else 0, None))

    def stats(self) -> FrameStats:
        s = FrameStats()
        check(lib().ct_batch_get_stats(self.handle, C.byref(s)))
        return s

    def autotune(self, params, repeats=3) -> BatchTuning:
        arr = (FrameParams * len(params))(*params)
        t = BatchTuning()
        check(lib().ct_batch_autotune(self.handle, arr, repeats, C.byref(t)))
        return t

    def close(self):
        if getattr(self, "handle", None):
            lib().ct_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_scene(scene: Scene, accel: Accel, ctx: Ctx, transform=None):
    """load_scene (main.cpp:1045-1126): staging, BUILD_BLAS_FROM_CPU, update_scene(true);
    with `transform` (4x4, column-major as INSTANCE.transform) it is load_model's
    placed build (main.cpp:1006-1043)."""
    if transform is None:
        check(lib().ct_scene_load(scene.handle, accel.handle, ctx.handle, None))
    else:
        t = (f32 * 16)(*np.asarray(transform, np.float32).reshape(16).tolist())
        check(lib().ct_scene_load_transformed(scene.handle, accel.handle, ctx.handle, t, None))


def brush_apply(ctx: Ctx, accel: Accel, x: int, y: int, window: int):
    """The squared brush (squared_brush.slang) at pixel (x, y) of ctx's last frame."""
    check(lib().ct_brush_apply(ctx.handle, accel.handle, x, y, window, None))


def pt_random_replay(ctx: Ctx, accel: Accel, params, dx=0, dy=0):
    """Random replay of the last frame's path reservoirs (ct_pt_random_replay):
    (H*W, 12) u32 per pixel -- ok, length reached, reservoir path length,
    reconnection length, reconnection vertex (instance, primitive, hit bits),
    previous vertex (instance, primitive), 0."""
    import torch
    n = params.width * params.height
    out = torch.zeros((n, 12), dtype=torch.int32, device="cuda")
    check(lib().ct_pt_random_replay(ctx.handle, accel.handle, C.byref(params), dx, dy, C.c_void_p(out.data_ptr()),
                                    None))
    torch.cuda.synchronize()
    return out.cpu().numpy().view(np.uint32)


def bind_scene(scene: Scene, ctx: Ctx):
    """Materials and light configuration of a loaded scene for another (fresh) context."""
    check(lib().ct_scene_bind(scene.handle, ctx.handle))


from .configs import *  # noqa: E402,F401,F403
