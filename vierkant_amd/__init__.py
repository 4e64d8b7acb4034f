"""vierkant_amd - MI355X (gfx950) ray-intersection and path-tracing core behind a C ABI.

This module is the Python binding of ``libvkrt.so`` (declared in ``include/vkrt.h``). It loads the
in-tree shared library and raises if it is missing: there is no CPU fallback. Arrays are numpy arrays
(host memory, staged by the library) or DeviceArray buffers in HBM (passed as device pointers).
"""
import atexit
import ctypes
import pathlib
import weakref

import numpy as np

PKG = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG / "libvkrt.so"

OK, E_INVALID_ARG, E_OOM, E_HIP, E_UNSUPPORTED, E_NO_DEVICE, E_INTERNAL = 0, -1, -2, -3, -4, -5, -6
HOST_PTRS, DEVICE_PTRS, ASYNC = 0, 1, 2
CULL_BACK = 8
REJECT, INTERSECT, INSIDE = 0, 1, 2
INVALID_PRIM = 0xFFFFFFFF
BUILDER_LBVH, BUILDER_PLOC = 0, 1

RAY_DTYPE = np.dtype([("ox", "<f4"), ("oy", "<f4"), ("oz", "<f4"), ("tmin", "<f4"),
                      ("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4"), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("prim", "<u4")])
TRI_RESULT_DTYPE = np.dtype([("type", "<u4"), ("distance", "<f4"), ("u", "<f4"), ("v", "<f4")])
RESULT_DTYPE = np.dtype([("type", "<u4"), ("distance", "<f4")])


class VkrtError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"vkrt error {code}: {message}")
        self.code = code


class BvhParams(ctypes.Structure):
    _fields_ = [("builder", ctypes.c_uint32), ("max_leaf_size", ctypes.c_uint32), ("width", ctypes.c_uint32),
                ("ploc_radius", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 4)]


class TraceConfig(ctypes.Structure):
    _fields_ = [("block_size", ctypes.c_uint32), ("lds_stack_depth", ctypes.c_uint32),
                ("traversal", ctypes.c_uint32), ("xcd_chunk", ctypes.c_uint32)]


TRAVERSAL_STEP, TRAVERSAL_POSTPONED = 0, 1


class PtSettings(ctypes.Structure):
    """PBRPathTracer::settings_t subset (vkrt_pt_settings)"""
    _fields_ = [("resolution", ctypes.c_uint32 * 2), ("max_num_batches", ctypes.c_uint32),
                ("num_samples", ctypes.c_uint32), ("max_trace_depth", ctypes.c_uint32),
                ("max_path_beta", ctypes.c_float), ("suspend_trace_when_done", ctypes.c_uint32),
                ("disable_material", ctypes.c_uint32), ("suppress_reset", ctypes.c_uint32),
                ("denoising", ctypes.c_uint32), ("tonemap", ctypes.c_uint32), ("gamma", ctypes.c_float),
                ("exposure", ctypes.c_float), ("mis_mode", ctypes.c_uint32),
                ("suppress_refractive_caustics", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 1)]


class Texture(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("rgba", ctypes.c_void_p),
                ("address_mode_u", ctypes.c_uint32), ("address_mode_v", ctypes.c_uint32),
                ("min_filter", ctypes.c_uint32), ("mag_filter", ctypes.c_uint32)]


ADDRESS_REPEAT, ADDRESS_MIRRORED_REPEAT, ADDRESS_CLAMP_TO_EDGE = 0, 1, 2
FILTER_LINEAR, FILTER_NEAREST = 0, 1


class ModelView(ctypes.Structure):
    _fields_ = [("positions", ctypes.c_void_p), ("normals", ctypes.c_void_p), ("texcoords", ctypes.c_void_p),
                ("tangents", ctypes.c_void_p), ("n_vert", ctypes.c_uint32), ("indices", ctypes.c_void_p),
                ("n_tri", ctypes.c_uint32),
                ("tri_object", ctypes.c_void_p), ("object_material", ctypes.c_void_p), ("n_objects", ctypes.c_uint32),
                ("materials", ctypes.c_void_p), ("n_materials", ctypes.c_uint32), ("textures", ctypes.c_void_p),
                ("n_textures", ctypes.c_uint32), ("n_skipped_textures", ctypes.c_uint32)]


class PtScene(ctypes.Structure):
    _fields_ = [("positions", ctypes.c_void_p), ("normals", ctypes.c_void_p), ("n_vert", ctypes.c_uint32),
                ("indices", ctypes.c_void_p), ("n_tri", ctypes.c_uint32), ("tri_object", ctypes.c_void_p),
                ("object_material", ctypes.c_void_p), ("n_objects", ctypes.c_uint32),
                ("materials", ctypes.c_void_p), ("n_materials", ctypes.c_uint32), ("environment", ctypes.c_float),
                ("texcoords", ctypes.c_void_p), ("textures", ctypes.c_void_p), ("n_textures", ctypes.c_uint32),
                ("tangents", ctypes.c_void_p)]


#: vkrt_pt_instance: object to world = translation + rotate(rotation (x, y, z, w), scale * p)
PT_INSTANCE_DTYPE = np.dtype([("translation", "<f4", 3), ("rotation", "<f4", 4), ("scale", "<f4", 3),
                              ("mesh", "<u4"), ("reserved", "<u4", 3)])


def pt_instances(meshes, translations=None, rotations=None, scales=None):
    """PT_INSTANCE_DTYPE records: one per mesh index, identity parts where an argument is None"""
    meshes = np.asarray(meshes, np.uint32).reshape(-1)
    out = np.zeros(len(meshes), PT_INSTANCE_DTYPE)
    out["mesh"] = meshes
    out["rotation"] = (0.0, 0.0, 0.0, 1.0) if rotations is None else np.asarray(rotations, np.float32)
    out["scale"] = 1.0 if scales is None else np.asarray(scales, np.float32)
    if translations is not None:
        out["translation"] = np.asarray(translations, np.float32)
    return out


class Camera(ctypes.Structure):
    _fields_ = [("eye", ctypes.c_float * 3), ("fov_y", ctypes.c_float), ("forward", ctypes.c_float * 3),
                ("aperture", ctypes.c_float), ("up", ctypes.c_float * 3), ("focal_distance", ctypes.c_float)]


class PtStats(ctypes.Structure):
    _fields_ = [("raytrace_ms", ctypes.c_float), ("traced", ctypes.c_uint32), ("segments", ctypes.c_uint64),
                ("samples", ctypes.c_uint64)]


MATERIAL_DTYPE = np.dtype([("color", "<f4", 4), ("emission", "<f4", 4), ("metalness", "<f4"), ("roughness", "<f4"),
                           ("transmission", "<f4"), ("ior", "<f4"), ("clearcoat", "<f4"),
                           ("clearcoat_roughness", "<f4"), ("sheen_roughness", "<f4"),
                           ("diffuse_transmission", "<f4"), ("sheen_color", "<f4", 4),
                           ("diffuse_transmission_color", "<f4", 4), ("two_sided", "<u4"), ("color_texture", "<u4"),
                           ("alpha_mode", "<u4"), ("alpha_cutoff", "<f4"), ("normal_texture", "<u4"),
                           ("emission_texture", "<u4"), ("ao_rough_metal_texture", "<u4"),
                           ("transmission_texture", "<u4"), ("diffuse_transmission_texture", "<u4"),
                           ("diffuse_transmission_color_texture", "<u4"), ("attenuation_color", "<f4", 3),
                           ("attenuation_distance", "<f4"), ("reserved", "<u4", 2)])
assert MATERIAL_DTYPE.itemsize == 160
ALPHA_OPAQUE, ALPHA_MASK, ALPHA_BLEND = 0, 1, 2
PIXEL_DTYPE = np.dtype([("radiance", "<f4", 3), ("alpha", "<f4"), ("depth", "<f4"), ("normal", "<u4"),
                        ("object_id", "<u4"), ("albedo", "<u4")])


def material(color=(0.8, 0.8, 0.8), emission=(0.0, 0.0, 0.0), emission_strength=0.0, metalness=0.0, roughness=1.0,
             transmission=0.0, ior=1.5, two_sided=False, alpha=1.0, color_texture=0, alpha_mode=0,
             alpha_cutoff=0.5, attenuation_color=(1.0, 1.0, 1.0), attenuation_distance=0.0, **texture_slots):
    """one vkrt_material record (material_t defaults: alpha 1, no clearcoat/sheen/diffuse transmission);
    color_texture is 1 + the index of the scene's base color texture (0: none); attenuation_distance > 0
    gives the material an interior medium absorbing as attenuation_color ** (d / attenuation_distance)"""
    m = np.zeros((), dtype=MATERIAL_DTYPE)
    m["attenuation_color"], m["attenuation_distance"] = attenuation_color, attenuation_distance
    m["color"] = (*color, alpha)
    m["color_texture"], m["alpha_mode"], m["alpha_cutoff"] = color_texture, alpha_mode, alpha_cutoff
    for k, v in texture_slots.items():  # normal_texture, emission_texture, ao_rough_metal_texture, ...
        m[k] = v
    m["emission"] = (*emission, emission_strength)
    m["metalness"], m["roughness"], m["transmission"], m["ior"] = metalness, roughness, transmission, ior
    m["two_sided"] = 1 if two_sided else 0
    return m


# light_t (vkrt_light): 80 bytes, the layout the shaders read
LIGHT_DTYPE = np.dtype([("position", "<f4", 3), ("type", "<u4"), ("color", "<f4", 3), ("intensity", "<f4"),
                        ("direction", "<f4", 3), ("range", "<f4"), ("spot_angle_scale", "<f4"),
                        ("spot_angle_offset", "<f4"), ("angular_size", "<f4"), ("size_y", "<f4"),
                        ("tangent", "<f4", 3), ("size_x", "<f4")])
assert LIGHT_DTYPE.itemsize == 80
LIGHT_OMNI, LIGHT_SPOT, LIGHT_DIRECTIONAL, LIGHT_AREA, LIGHT_RECT, LIGHT_SPHERE, LIGHT_TUBE, LIGHT_DISK = range(8)
MIS_FULL, MIS_NEE_ONLY, MIS_BSDF_ONLY = 0, 1, 2
PT_LAUNCH_PERSISTENT, PT_LAUNCH_FULL_GRID = 0, 1
PT_TAIL_DEPTH_DEFAULT, PT_NO_TAIL = 4, 0xFFFFFFFF


def light(type, position=(0.0, 0.0, 0.0), direction=(0.0, 0.0, -1.0), color=(1.0, 1.0, 1.0), intensity=1.0,
          size=(1.0, 1.0), tangent=(1.0, 0.0, 0.0), angular_size=0.0, range=np.inf, inner_cone_angle=0.0,
          outer_cone_angle=np.pi / 4):
    """one light_t record as convert_light fills it (include/vierkant/punctual_light.hpp:91-111): spot angle
    scale/offset from the cone angles, size = (size_x, size_y)"""
    l = np.zeros((), dtype=LIGHT_DTYPE)
    l["type"], l["position"], l["direction"], l["color"] = type, position, direction, color
    l["intensity"], l["tangent"], l["angular_size"] = intensity, tangent, angular_size
    l["range"] = range if range > 0 else np.inf
    scale = np.float32(1.0) / max(np.float32(0.001), np.float32(np.cos(np.float32(inner_cone_angle))) -
                                  np.float32(np.cos(np.float32(outer_cone_angle))))
    l["spot_angle_scale"] = scale
    l["spot_angle_offset"] = -np.float32(np.cos(np.float32(outer_cone_angle))) * np.float32(scale)
    l["size_x"], l["size_y"] = size
    return l


class BatchItem(ctypes.Structure):
    _fields_ = [("bvh", ctypes.c_void_p), ("rays", ctypes.c_void_p), ("hits", ctypes.c_void_p),
                ("n_rays", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("block_size", ctypes.c_uint32), ("lds_stack_depth", ctypes.c_uint32), ("tile", ctypes.c_uint32),
                ("blocks_per_cu", ctypes.c_uint32), ("xcd_chunk", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 3)]

    def shape(self):
        return (self.block_size, self.lds_stack_depth, self.tile, self.blocks_per_cu, self.xcd_chunk)


class BvhInfo(ctypes.Structure):
    _fields_ = [("n_tri", ctypes.c_uint32), ("n_nodes", ctypes.c_uint32), ("n_leaves", ctypes.c_uint32),
                ("depth", ctypes.c_uint32), ("sah_cost", ctypes.c_float), ("build_ms", ctypes.c_float),
                ("device_bytes", ctypes.c_uint64), ("bounds", ctypes.c_float * 6),
                ("n_wide_nodes", ctypes.c_uint32), ("wide_stack", ctypes.c_uint32)]


class BvhCheck(ctypes.Structure):
    _fields_ = [("ok", ctypes.c_uint32), ("wide", ctypes.c_uint32), ("tris_reached", ctypes.c_uint32),
                ("tris_not_once", ctypes.c_uint32), ("box_violations", ctypes.c_uint32),
                ("bad_refs", ctypes.c_uint32), ("stack_bound", ctypes.c_uint32), ("stack_ok", ctypes.c_uint32),
                ("bad_kind", ctypes.c_uint32), ("bad_node", ctypes.c_uint32), ("bad_slot", ctypes.c_uint32),
                ("bad_tri", ctypes.c_uint32), ("bad_box", ctypes.c_float * 6), ("bad_outer", ctypes.c_float * 6)]

    def as_dict(self):
        d = {name: int(getattr(self, name)) for name, _ in self._fields_[:12]}
        if self.box_violations:
            d["bad_box"] = [float(x) for x in self.bad_box]
            d["bad_outer"] = [float(x) for x in self.bad_outer]
        return d


class RcParams(ctypes.Structure):
    _fields_ = [("log2_table", ctypes.c_uint32), ("base_resolution", ctypes.c_uint32),
                ("max_resolution", ctypes.c_uint32), ("learning_rate", ctypes.c_float), ("beta1", ctypes.c_float),
                ("beta2", ctypes.c_float), ("epsilon", ctypes.c_float), ("reserved", ctypes.c_uint32 * 9)]


RC_GRAD_ONLY = 4


class PtCacheConfig(ctypes.Structure):
    _fields_ = [("cache_depth", ctypes.c_uint32), ("train_stride", ctypes.c_uint32),
                ("max_train_records", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 5)]

_lib = None
_vp, _u32, _i32, _f32, _u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_float, ctypes.c_uint64

_SIGNATURES = {
    "vkrt_version": (_i32, []),
    "vkrt_device_count": (_i32, []),
    "vkrt_create": (_i32, [_i32, _vp, ctypes.POINTER(_vp)]),
    "vkrt_destroy": (None, [_vp]),
    "vkrt_last_error": (ctypes.c_char_p, [_vp]),
    "vkrt_sync": (_i32, [_vp]),
    "vkrt_stream": (_vp, [_vp]),
    "vkrt_device_alloc": (_i32, [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "vkrt_device_free": (_i32, [_vp, _vp]),
    "vkrt_copy": (_i32, [_vp, _vp, _vp, ctypes.c_size_t, _i32, _i32]),
    "vkrt_intersect_triangles": (_i32, [_vp, _vp, _vp, _u32, _vp, _i32]),
    "vkrt_intersect_aabbs": (_i32, [_vp, _vp, _vp, _u32, _vp, _i32]),
    "vkrt_intersect_obbs": (_i32, [_vp, _vp, _vp, _u32, _vp, _i32]),
    "vkrt_intersect_spheres": (_i32, [_vp, _vp, _vp, _u32, _vp, _i32]),
    "vkrt_intersect_planes": (_i32, [_vp, _vp, _vp, _u32, _vp, _i32]),
    "vkrt_tri_box_overlap": (_i32, [_vp, _vp, _vp, _u32, _vp, _i32]),
    "vkrt_tri_tri_overlap": (_i32, [_vp, _vp, _vp, _u32, _vp, _i32]),
    "vkrt_compute_aabb": (_i32, [_vp, _vp, _u32, _vp, _i32]),
    "vkrt_compute_bounding_sphere": (_i32, [_vp, _vp, _u32, _vp, _i32]),
    "vkrt_intersect_brute_closest": (_i32, [_vp, _vp, _u32, _vp, _u32, _vp, _i32]),
    "vkrt_intersect_brute_closest_noalloc": (_i32, [_vp, _vp, _u32, _vp, _u32, _vp, _i32]),
    "vkrt_intersect_brute_any": (_i32, [_vp, _vp, _u32, _vp, _u32, _vp, _i32]),
    "vkrt_bvh_build": (_i32, [_vp, _vp, _u32, _u32, _vp, _u32, _vp, _i32, ctypes.POINTER(_vp)]),
    "vkrt_bvh_refit": (_i32, [_vp, _vp, _vp, _u32, _i32]),
    "vkrt_bvh_get_info": (_i32, [_vp, ctypes.POINTER(BvhInfo)]),
    "vkrt_bvh_destroy": (None, [_vp]),
    "vkrt_bvh_validate": (_i32, [_vp, _vp, _vp, _u32, _i32, ctypes.POINTER(BvhCheck)]),
    "vkrt_set_trace_config": (_i32, [_vp, ctypes.POINTER(TraceConfig)]),
    "vkrt_get_trace_config": (_i32, [_vp, ctypes.POINTER(TraceConfig)]),
    "vkrt_trace_closest": (_i32, [_vp, _vp, _vp, _u32, _vp, _i32]),
    "vkrt_trace_any": (_i32, [_vp, _vp, _vp, _u32, _vp, _i32]),
    "vkrt_batch_create": (_i32, [_vp, _vp, _u32, ctypes.POINTER(_vp)]),
    "vkrt_batch_destroy": (None, [_vp]),
    "vkrt_batch_set_config": (_i32, [_vp, ctypes.POINTER(BatchConfig)]),
    "vkrt_batch_get_config": (_i32, [_vp, ctypes.POINTER(BatchConfig)]),
    "vkrt_batch_launch": (_i32, [_vp, _i32]),
    "vkrt_batch_autotune": (_i32, [_vp, _u32, ctypes.POINTER(BatchConfig), ctypes.POINTER(ctypes.c_float)]),
    "vkrt_batch_total_rays": (_u64, [_vp]),
    "vkrt_batch_candidates": (_u32, [_vp, _u32]),
    "vkrt_batch_closest": (_i32, [_vp, _vp, _u32, _i32]),
    "vkrt_pt_default_settings": (None, [ctypes.POINTER(PtSettings)]),
    "vkrt_pt_create": (_i32, [_vp, ctypes.POINTER(PtSettings), _u32, ctypes.POINTER(_vp)]),
    "vkrt_pt_destroy": (None, [_vp]),
    "vkrt_pt_set_settings": (_i32, [_vp, ctypes.POINTER(PtSettings)]),
    "vkrt_pt_get_settings": (_i32, [_vp, ctypes.POINTER(PtSettings)]),
    "vkrt_pt_set_scene": (_i32, [_vp, ctypes.POINTER(PtScene)]),
    "vkrt_pt_set_lights": (_i32, [_vp, _vp, _u32]),
    "vkrt_pt_set_launch_mode": (_i32, [_vp, _i32, _u32]),
    "vkrt_pt_set_instanced_scene": (_i32, [_vp, ctypes.POINTER(PtScene), _vp, _u32, _vp, _u32]),
    "vkrt_model_load": (_i32, [ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "vkrt_model_error": (ctypes.c_char_p, []),
    "vkrt_model_get": (_i32, [_vp, ctypes.POINTER(ModelView)]),
    "vkrt_model_pt_scene": (_i32, [_vp, ctypes.c_float, ctypes.POINTER(PtScene)]),
    "vkrt_model_destroy": (None, [_vp]),
    "vkrt_pt_render": (_i32, [_vp, ctypes.POINTER(Camera)]),
    "vkrt_pt_reset_accumulator": (_i32, [_vp]),
    "vkrt_pt_current_batch": (_u32, [_vp]),
    "vkrt_pt_read_pixels": (_i32, [_vp, _vp, _i32]),
    "vkrt_pt_get_stats": (_i32, [_vp, ctypes.POINTER(PtStats)]),
    "vkrt_pt_read_image": (_i32, [_vp, _vp, _i32]),
    "vkrt_denoise": (_i32, [_vp, _vp, _u32, _u32, _u32, _vp, _i32]),
    "vkrt_tonemap": (_i32, [_vp, _vp, _u32, ctypes.c_float, ctypes.c_float, _vp, _i32]),
    "vkrt_rc_default_params": (None, [ctypes.POINTER(RcParams)]),
    "vkrt_rc_create": (_i32, [_vp, ctypes.POINTER(RcParams), _u32, ctypes.POINTER(_vp)]),
    "vkrt_rc_destroy": (None, [_vp]),
    "vkrt_rc_get_config": (_i32, [_vp, ctypes.POINTER(RcParams), ctypes.POINTER(_u32), ctypes.POINTER(_u64)]),
    "vkrt_rc_set_bounds": (_i32, [_vp, _vp, _vp]),
    "vkrt_rc_query": (_i32, [_vp, _vp, _u32, _vp, _i32]),
    "vkrt_rc_train": (_i32, [_vp, _vp, _vp, _u32, _i32, ctypes.POINTER(ctypes.c_float)]),
    "vkrt_rc_get_params": (_i32, [_vp, _vp, _vp]),
    "vkrt_rc_set_params": (_i32, [_vp, _vp, _vp]),
    "vkrt_rc_get_gradients": (_i32, [_vp, _vp, _vp]),
    "vkrt_rc_step": (_u32, [_vp]),
    "vkrt_pt_default_cache_config": (None, [ctypes.POINTER(PtCacheConfig)]),
    "vkrt_pt_trace_transmittance": (_i32, [_vp, _vp, _u32, _vp, _i32]),
    "vkrt_pt_set_radiance_cache": (_i32, [_vp, _vp, ctypes.POINTER(PtCacheConfig)]),
    "vkrt_instance_set_transform": (None, [_vp, _vp]),
    "vkrt_scene_create": (_i32, [_vp, _vp, _u32, _vp, _u32, ctypes.POINTER(_vp)]),
    "vkrt_scene_set_instances": (_i32, [_vp, _vp, _u32]),
    "vkrt_scene_destroy": (None, [_vp]),
    "vkrt_scene_trace_closest": (_i32, [_vp, _vp, _vp, _u32, _vp, _i32]),
    "vkrt_scene_trace_any": (_i32, [_vp, _vp, _vp, _u32, _vp, _i32]),
    "vkrt_scene_get_info": (_i32, [_vp, ctypes.POINTER(_u32), ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    "vkrt_icosphere_size": (_i32, [_u32, ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    "vkrt_icosphere": (_i32, [_f32, _u32, _vp, _vp]),
    "vkrt_plane_size": (_i32, [_u32, _u32, ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    "vkrt_plane": (_i32, [_f32, _f32, _u32, _u32, _vp, _vp]),
}


def lib():
    """Load libvkrt.so (built by ``python -m vierkant_amd.build``); raise if it is missing."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m vierkant_amd.build` "
                               "(there is no CPU fallback)")
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def declared_symbols():
    return list(_SIGNATURES)


def _ptr(a):
    """(pointer, is_device) of a numpy array, a DeviceArray or a torch tensor."""
    if a is None:
        return None, False
    if isinstance(a, DeviceArray):
        return a.ptr, True
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("arrays must be C-contiguous")
        return a.ctypes.data, False
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensors must be contiguous")
        return a.data_ptr(), a.is_cuda
    raise TypeError(f"unsupported array type {type(a)}")


def _flags(arrays, extra=0):
    kinds = {_ptr(a)[1] for a in arrays if a is not None}
    if len(kinds) > 1:
        raise ValueError("mixing host and device arrays in one call")
    return (DEVICE_PTRS if kinds == {True} else HOST_PTRS) | extra


def _n(a, per=1):
    if isinstance(a, (np.ndarray, DeviceArray)):
        return a.nbytes // (4 * per)
    return a.numel() * a.element_size() // (4 * per)


class DeviceArray:
    """A device (HBM) buffer owned through the context (vkrt_device_alloc / vkrt_device_free), typed
    like a numpy array; pass it wherever the binding takes an array to use the device-pointer path."""

    def __init__(self, ctx, shape, dtype):
        self.ctx, self.dtype = ctx, np.dtype(dtype)
        self.shape = (shape,) if isinstance(shape, int) else tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        h = _vp()
        ctx.check(lib().vkrt_device_alloc(ctx.handle, self.nbytes, ctypes.byref(h)))
        self.ptr = h.value
        ctx._adopt(self)

    @classmethod
    def from_host(cls, ctx, a):
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.shape, a.dtype)
        ctx.check(lib().vkrt_copy(ctx.handle, d.ptr, a.ctypes.data, a.nbytes, 0, 0))
        return d

    def host(self):
        out = np.empty(self.shape, self.dtype)
        self.ctx.check(lib().vkrt_copy(self.ctx.handle, out.ctypes.data, self.ptr, self.nbytes, 1, 0))
        return out

    def __len__(self):
        return self.shape[0]

    def free(self):
        if self.ptr:
            lib().vkrt_device_free(self.ctx.handle, self.ptr)
            self.ptr = None

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HipEvents:
    """hipEvent pairs recorded on a stream, through the HIP runtime libvkrt.so itself uses."""

    def __init__(self, n):
        lib()
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(_vp)]
        self.hip.hipEventRecord.argtypes = [_vp, _vp]
        self.hip.hipEventSynchronize.argtypes = [_vp]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), _vp, _vp]
        self.hip.hipEventDestroy.argtypes = [_vp]
        self.ev = []
        for _ in range(n):
            e = _vp()
            if self.hip.hipEventCreate(ctypes.byref(e)) != 0:
                raise RuntimeError("hipEventCreate failed")
            self.ev.append(e)

    def record(self, i, stream):
        if self.hip.hipEventRecord(self.ev[i], stream) != 0:
            raise RuntimeError("hipEventRecord failed")

    def sync(self, i):
        if self.hip.hipEventSynchronize(self.ev[i]) != 0:
            raise RuntimeError("hipEventSynchronize failed")

    def ms(self, a, b):
        out = ctypes.c_float()
        if self.hip.hipEventElapsedTime(ctypes.byref(out), self.ev[a], self.ev[b]) != 0:
            raise RuntimeError("hipEventElapsedTime failed")
        return out.value

    def close(self):
        for e in self.ev:
            self.hip.hipEventDestroy(e)
        self.ev = []


class Context:
    """A vkrt context on one HIP device (vkrt_create / vkrt_destroy)."""

    def __init__(self, device=0, stream=None):
        self._h = _vp()
        self._children = []
        rc = lib().vkrt_create(device, stream, ctypes.byref(self._h))
        if rc != OK:
            raise VkrtError(rc, "vkrt_create failed (no gfx950 device?)")
        _live_contexts.add(self)

    def _adopt(self, obj):
        """Track an object that holds native state of this context, so that close() can release it first."""
        self._children = [r for r in self._children if r() is not None]
        self._children.append(weakref.ref(obj))

    def close(self):
        """Release every batch, path tracer, BVH and device array made on this context, newest first (a
        batch refers to its BVHs and arrays, every object to the context), then the context itself."""
        if self._h:
            for r in reversed(self._children):
                obj = r()
                if obj is not None:
                    obj.close()
            self._children = []
            lib().vkrt_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    @property
    def stream(self):
        return lib().vkrt_stream(self._h)

    def check(self, rc):
        if rc != OK:
            raise VkrtError(rc, lib().vkrt_last_error(self._h).decode())

    def sync(self):
        self.check(lib().vkrt_sync(self._h))

    # -- pairwise tests -------------------------------------------------------------------------
    def _pairwise(self, fn, prims, per, rays, out_dtype, out=None):
        n = _n(rays, 8)
        if out is None:
            out = np.zeros(n, dtype=out_dtype)
        self.check(fn(self._h, _ptr(prims)[0], _ptr(rays)[0], n, _ptr(out)[0], _flags([prims, rays, out])))
        return out

    def intersect_triangles(self, tris, rays, out=None):
        return self._pairwise(lib().vkrt_intersect_triangles, tris, 9, rays, TRI_RESULT_DTYPE, out)

    def intersect_aabbs(self, boxes, rays, out=None):
        return self._pairwise(lib().vkrt_intersect_aabbs, boxes, 6, rays, RESULT_DTYPE, out)

    def intersect_obbs(self, obbs, rays, out=None):
        return self._pairwise(lib().vkrt_intersect_obbs, obbs, 15, rays, RESULT_DTYPE, out)

    def intersect_spheres(self, spheres, rays, out=None):
        return self._pairwise(lib().vkrt_intersect_spheres, spheres, 4, rays, RESULT_DTYPE, out)

    def intersect_planes(self, planes, rays, out=None):
        return self._pairwise(lib().vkrt_intersect_planes, planes, 4, rays, RESULT_DTYPE, out)

    def tri_box_overlap(self, tris, boxes, out=None):
        n = _n(tris, 9)
        if out is None:
            out = np.zeros(n, dtype=np.uint32)
        self.check(lib().vkrt_tri_box_overlap(self._h, _ptr(tris)[0], _ptr(boxes)[0], n, _ptr(out)[0],
                                              _flags([tris, boxes, out])))
        return out

    def tri_tri_overlap(self, tris_a, tris_b, out=None):
        n = _n(tris_a, 9)
        if out is None:
            out = np.zeros(n, dtype=np.uint32)
        self.check(lib().vkrt_tri_tri_overlap(self._h, _ptr(tris_a)[0], _ptr(tris_b)[0], n, _ptr(out)[0],
                                              _flags([tris_a, tris_b, out])))
        return out

    def compute_aabb(self, points, out=None):
        n = _n(points, 3)
        if out is None:
            out = np.zeros(6, dtype=np.float32)
        self.check(lib().vkrt_compute_aabb(self._h, _ptr(points)[0], n, _ptr(out)[0], _flags([points, out])))
        return out

    def denoise(self, pixels, denoise=True):
        """the a-trous denoise pass over a (H, W) PIXEL_DTYPE buffer: (H, W, 4) float32"""
        px = np.ascontiguousarray(pixels, PIXEL_DTYPE)
        h, w = px.shape
        out = np.zeros((h, w, 4), np.float32)
        self.check(lib().vkrt_denoise(self._h, px.ctypes.data, w, h, 1 if denoise else 0, out.ctypes.data, 0))
        return out

    def tonemap(self, rgba, exposure=2.0, gamma=1.0):
        """the exposure/gamma composite over (..., 4) float32 texels, each channel rounded to half"""
        a = np.ascontiguousarray(rgba, np.float32)
        assert a.shape[-1] == 4
        out = np.empty_like(a)
        self.check(lib().vkrt_tonemap(self._h, a.ctypes.data, a.size // 4, exposure, gamma, out.ctypes.data, 0))
        return out

    def compute_bounding_sphere(self, points, out=None):
        """(center.x, center.y, center.z, radius): compute_bounding_sphere of the reference, bit-identical"""
        n = _n(points, 3)
        if out is None:
            out = np.zeros(4, dtype=np.float32)
        self.check(lib().vkrt_compute_bounding_sphere(self._h, _ptr(points)[0], n, _ptr(out)[0],
                                                      _flags([points, out])))
        return out

    # -- brute force ----------------------------------------------------------------------------
    def brute_closest(self, tris, rays, out=None, flags=0):
        n_tri, n_ray = _n(tris, 9), _n(rays, 8)
        if out is None:
            out = np.zeros(n_ray, dtype=HIT_DTYPE)
        self.check(lib().vkrt_intersect_brute_closest(self._h, _ptr(tris)[0], n_tri, _ptr(rays)[0], n_ray,
                                                      _ptr(out)[0], _flags([tris, rays, out], flags)))
        return out

    def brute_any(self, tris, rays, out=None, flags=0):
        n_tri, n_ray = _n(tris, 9), _n(rays, 8)
        if out is None:
            out = np.zeros(n_ray, dtype=np.uint8)
        self.check(lib().vkrt_intersect_brute_any(self._h, _ptr(tris)[0], n_tri, _ptr(rays)[0], n_ray,
                                                  _ptr(out)[0], _flags([tris, rays, out], flags)))
        return out

    # -- BVH ------------------------------------------------------------------------------------
    def bvh_build(self, positions, indices, builder=BUILDER_LBVH, max_leaf_size=0, stride=12, width=0,
                  ploc_radius=0):
        """width: 0 = traverse the 4-wide collapse when its stack bound fits, 2 = the binary tree only;
        ploc_radius: PLOC's neighbour search radius (0 = 16)"""
        params = BvhParams(builder, max_leaf_size, width, ploc_radius)
        n_vert = _n(positions) * 4 // stride
        n_tri = _n(indices, 3)
        h = _vp()
        self.check(lib().vkrt_bvh_build(self._h, _ptr(positions)[0], n_vert, stride, _ptr(indices)[0], n_tri,
                                        ctypes.byref(params), _flags([positions, indices]), ctypes.byref(h)))
        return Bvh(self, h)

    def set_trace_config(self, block_size, lds_stack_depth, traversal=None, xcd_chunk=None):
        """launch shape of the trace kernels; traversal TRAVERSAL_*, xcd_chunk 0 or a power of two up to 256
        (None keeps the current value)"""
        cur = self.get_trace_config()
        traversal = cur[2] if traversal is None else traversal
        xcd_chunk = cur[3] if xcd_chunk is None else xcd_chunk
        self.check(lib().vkrt_set_trace_config(self._h, ctypes.byref(TraceConfig(block_size, lds_stack_depth,
                                                                                 traversal, xcd_chunk))))

    def get_trace_config(self):
        c = TraceConfig()
        self.check(lib().vkrt_get_trace_config(self._h, ctypes.byref(c)))
        return c.block_size, c.lds_stack_depth, c.traversal, c.xcd_chunk

    def trace_closest(self, bvh, rays, out=None, flags=0):
        n = _n(rays, 8)
        if out is None:
            out = np.zeros(n, dtype=HIT_DTYPE)
        self.check(lib().vkrt_trace_closest(self._h, bvh.handle, _ptr(rays)[0], n, _ptr(out)[0],
                                            _flags([rays, out], flags)))
        return out

    def trace_any(self, bvh, rays, out=None, flags=0):
        n = _n(rays, 8)
        if out is None:
            out = np.zeros(n, dtype=np.uint8)
        self.check(lib().vkrt_trace_any(self._h, bvh.handle, _ptr(rays)[0], n, _ptr(out)[0],
                                        _flags([rays, out], flags)))
        return out


_live_contexts = weakref.WeakSet()


@atexit.register
def _close_live_contexts():
    """At interpreter exit the finalizer can release a context before the objects that use it; close every
    live context (and with it its dependents, newest first) while the module is still intact."""
    for c in list(_live_contexts):
        c.close()


class Bvh:
    """A BVH built by vkrt_bvh_build; destroyed with the object."""

    def __init__(self, ctx, handle):
        self.ctx, self._h = ctx, handle
        ctx._adopt(self)

    @property
    def handle(self):
        return self._h

    def info(self):
        info = BvhInfo()
        self.ctx.check(lib().vkrt_bvh_get_info(self._h, ctypes.byref(info)))
        return info

    def refit(self, positions, stride=12):
        self.ctx.check(lib().vkrt_bvh_refit(self.ctx.handle, self._h, _ptr(positions)[0], stride,
                                            _flags([positions])))

    def validate(self, positions, stride=12):
        """host-side invariant check (vkrt_bvh_validate) against the mesh positions; returns a dict"""
        out = BvhCheck()
        self.ctx.check(lib().vkrt_bvh_validate(self.ctx.handle, self._h, _ptr(positions)[0], stride,
                                               _flags([positions]), ctypes.byref(out)))
        return out.as_dict()

    def close(self):
        if self._h:
            lib().vkrt_bvh_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """many (bvh, device rays, device hits) items traced by one persistent launch, replayed as a hipGraph"""

    def __init__(self, ctx, items):
        """items: list of (Bvh, DeviceArray rays, DeviceArray hits)"""
        self.ctx, self._items = ctx, items
        arr = (BatchItem * len(items))(*[BatchItem(b.handle, r.ptr, h.ptr, len(r), 0) for b, r, h in items])
        self._h = _vp()
        ctx.check(lib().vkrt_batch_create(ctx.handle, arr, len(items), ctypes.byref(self._h)))
        ctx._adopt(self)

    def launch(self, flags=0):
        self.ctx.check(lib().vkrt_batch_launch(self._h, flags))

    @property
    def config(self):
        """(block_size, lds_stack_depth, tile, blocks_per_cu, xcd_chunk)"""
        c = BatchConfig()
        self.ctx.check(lib().vkrt_batch_get_config(self._h, ctypes.byref(c)))
        return c.shape()

    @config.setter
    def config(self, cfg):
        self.ctx.check(lib().vkrt_batch_set_config(self._h, ctypes.byref(BatchConfig(*cfg))))

    def autotune(self, repeats=2):
        best, ms = BatchConfig(), ctypes.c_float()
        self.ctx.check(lib().vkrt_batch_autotune(self._h, repeats, ctypes.byref(best), ctypes.byref(ms)))
        return best.shape(), ms.value

    @staticmethod
    def candidates():
        """every launch shape the autotuner can pick"""
        n = lib().vkrt_batch_candidates(None, 0)
        arr = (BatchConfig * n)()
        lib().vkrt_batch_candidates(arr, n)
        return [c.shape() for c in arr]

    def total_rays(self):
        return lib().vkrt_batch_total_rays(self._h)

    def close(self):
        if self._h:
            lib().vkrt_batch_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_settings():
    s = PtSettings()
    lib().vkrt_pt_default_settings(ctypes.byref(s))
    return s


def make_camera(eye, forward, up=(0.0, 1.0, 0.0), fov_y=0.7853981633974483, aperture=0.0, focal_distance=1.0):
    return Camera((ctypes.c_float * 3)(*eye), fov_y, (ctypes.c_float * 3)(*forward), aperture,
                  (ctypes.c_float * 3)(*up), focal_distance)


def _pt_scene_view(scene):
    """a vkrt_pt_scene over contiguous copies of a scene dict, with the arrays it points into"""
    arr = {k: np.ascontiguousarray(scene[k]) for k in
           ("positions", "normals", "indices", "tri_object", "object_material", "materials")}
    uv = scene.get("texcoords")
    uv = np.ascontiguousarray(uv, np.float32) if uv is not None and len(uv) else None
    tg = scene.get("tangents")
    tg = np.ascontiguousarray(tg, np.float32) if tg is not None and len(tg) else None
    imgs = [np.ascontiguousarray(t, np.uint8) for t in scene.get("textures", [])]
    smp = list(scene.get("texture_samplers") or [(0, 0, 0, 0)] * len(imgs))
    tex = (Texture * max(1, len(imgs)))(*[Texture(t.shape[1], t.shape[0], t.ctypes.data, *s)
                                          for t, s in zip(imgs, smp)])
    sc = PtScene(arr["positions"].ctypes.data, arr["normals"].ctypes.data, len(arr["positions"]),
                 arr["indices"].ctypes.data, len(arr["indices"]) // 3, arr["tri_object"].ctypes.data,
                 arr["object_material"].ctypes.data, len(arr["object_material"]),
                 arr["materials"].ctypes.data, len(arr["materials"]), float(scene.get("environment", 1.0)),
                 uv.ctypes.data if uv is not None else None, ctypes.addressof(tex) if imgs else None, len(imgs),
                 tg.ctypes.data if tg is not None else None)
    return sc, (arr, uv, tg, imgs, tex)


class PathTracer:
    """PBRPathTracer on the GPU: `settings` (settings_t subset), render(camera) = render_scene,
    current_batch(), reset_accumulator(), pixels() = the pixel_data_t buffer."""

    def __init__(self, ctx, settings=None, seed=0):
        self.ctx = ctx
        self._h = _vp()
        s = settings if settings is not None else default_settings()
        ctx.check(lib().vkrt_pt_create(ctx.handle, ctypes.byref(s), seed, ctypes.byref(self._h)))
        self._keep = None
        ctx._adopt(self)

    @property
    def settings(self):
        s = PtSettings()
        self.ctx.check(lib().vkrt_pt_get_settings(self._h, ctypes.byref(s)))
        return s

    @settings.setter
    def settings(self, s):
        self.ctx.check(lib().vkrt_pt_set_settings(self._h, ctypes.byref(s)))

    def set_scene(self, scene):
        """scene: dict with positions [n,3] f32, normals [n,3] f32, indices u32, tri_object u32,
        object_material u32, materials (MATERIAL_DTYPE array), environment float; optional texcoords [n,2] f32,
        tangents [n,3] f32 (for normal maps) and textures (a list of [h,w,4] uint8 RGBA images, referenced by the
        materials' texture slots as index + 1) with texture_samplers (one (address_mode_u, address_mode_v,
        min_filter, mag_filter) per texture; default repeat and linear)"""
        sc, keep = _pt_scene_view(scene)
        self.ctx.check(lib().vkrt_pt_set_scene(self._h, ctypes.byref(sc)))
        del keep

    def set_instanced_scene(self, scene, mesh_first, instances):
        """a two-level scene: `scene` as for set_scene, its triangles split into meshes at mesh_first
        (n_meshes + 1 ascending triangle offsets, 0 first and the triangle count last), each mesh a bottom
        level BVH, placed by `instances` (PT_INSTANCE_DTYPE records, see pt_instances)"""
        sc, keep = _pt_scene_view(scene)
        mf = np.ascontiguousarray(mesh_first, np.uint32).reshape(-1)
        ins = np.ascontiguousarray(instances, PT_INSTANCE_DTYPE).reshape(-1)
        if len(mf) < 2:
            raise ValueError("mesh_first needs n_meshes + 1 offsets")
        self.ctx.check(lib().vkrt_pt_set_instanced_scene(self._h, ctypes.byref(sc), mf.ctypes.data, len(mf) - 1,
                                                         ins.ctypes.data if len(ins) else None, len(ins)))
        del keep

    def set_launch_mode(self, mode, tail_depth=PT_TAIL_DEPTH_DEFAULT):
        """PT_LAUNCH_FULL_GRID (default) or PT_LAUNCH_PERSISTENT per-depth launches; paths alive at tail_depth
        finish in one lane each (PT_NO_TAIL: per-depth launches to the end). Every choice gives the same image"""
        self.ctx.check(lib().vkrt_pt_set_launch_mode(self._h, mode, tail_depth))

    def set_lights(self, lights):
        """the frame's light list (LIGHT_DTYPE records; empty clears it)"""
        arr = np.ascontiguousarray(np.asarray(lights, dtype=LIGHT_DTYPE).reshape(-1))
        self.ctx.check(lib().vkrt_pt_set_lights(self._h, arr.ctypes.data if len(arr) else None, len(arr)))

    def render(self, camera):
        self.ctx.check(lib().vkrt_pt_render(self._h, ctypes.byref(camera)))

    def reset_accumulator(self):
        self.ctx.check(lib().vkrt_pt_reset_accumulator(self._h))

    def current_batch(self):
        return lib().vkrt_pt_current_batch(self._h)

    def pixels(self):
        s = self.settings
        out = np.zeros(s.resolution[0] * s.resolution[1], dtype=PIXEL_DTYPE)
        self.ctx.check(lib().vkrt_pt_read_pixels(self._h, out.ctypes.data, 0))
        return out.reshape(s.resolution[1], s.resolution[0])

    def image(self):
        """the output image (H, W, 4): denoised rgb when settings.denoising, else radiance and alpha; then the
        exposure/gamma composite (half-rounded) when settings.tonemap"""
        s = self.settings
        out = np.zeros((s.resolution[1], s.resolution[0], 4), np.float32)
        self.ctx.check(lib().vkrt_pt_read_image(self._h, out.ctypes.data, 0))
        return out

    def trace_transmittance(self, rays, cull_back=False):
        """transmittance_test through the scene: (n, 3) rgb transmittance per ray"""
        r = np.ascontiguousarray(rays)
        out = np.zeros((len(r), 4), np.float32)
        self.ctx.check(lib().vkrt_pt_trace_transmittance(self._h, r.ctypes.data, len(r), out.ctypes.data,
                                                         CULL_BACK if cull_back else 0))
        return out[:, :3]

    def set_radiance_cache(self, cache, **cfg):
        """attach a RadianceCache (None detaches); cfg: cache_depth, train_stride, max_train_records"""
        c = PtCacheConfig()
        lib().vkrt_pt_default_cache_config(ctypes.byref(c))
        for k, v in cfg.items():
            setattr(c, k, v)
        self.ctx.check(lib().vkrt_pt_set_radiance_cache(self._h, None if cache is None else cache.handle,
                                                        ctypes.byref(c)))
        self._cache = cache

    def stats(self):
        st = PtStats()
        self.ctx.check(lib().vkrt_pt_get_stats(self._h, ctypes.byref(st)))
        return st

    def close(self):
        if self._h:
            lib().vkrt_pt_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RadianceCache:
    """Neural radiance cache (hash grid + MLP on the matrix cores): query, train, parameters."""

    RECORD_FLOATS = 16

    def __init__(self, ctx, params=None, seed=0):
        self.ctx = ctx
        if params is None:
            params = default_rc_params()
        self._h = _vp()
        ctx.check(lib().vkrt_rc_create(ctx.handle, ctypes.byref(params), seed, ctypes.byref(self._h)))
        ctx._adopt(self)
        p, nm, ng = RcParams(), _u32(), _u64()
        ctx.check(lib().vkrt_rc_get_config(self._h, ctypes.byref(p), ctypes.byref(nm), ctypes.byref(ng)))
        self.params, self.n_mlp, self.n_grid = p, nm.value, ng.value

    @property
    def handle(self):
        return self._h

    def set_bounds(self, bmin, bmax):
        lo = np.ascontiguousarray(bmin, np.float32)
        hi = np.ascontiguousarray(bmax, np.float32)
        self.ctx.check(lib().vkrt_rc_set_bounds(self._h, lo.ctypes.data, hi.ctypes.data))

    def query(self, records):
        rec = np.ascontiguousarray(records, np.float32).reshape(-1, 16)
        out = np.zeros((len(rec), 4), np.float32)
        self.ctx.check(lib().vkrt_rc_query(self._h, rec.ctypes.data, len(rec), out.ctypes.data, 0))
        return out[:, :3]

    def train(self, records, targets, grad_only=False):
        rec = np.ascontiguousarray(records, np.float32).reshape(-1, 16)
        tg = np.zeros((len(rec), 4), np.float32)
        if len(rec):
            tg[:, :3] = np.asarray(targets, np.float32).reshape(len(rec), -1)[:, :3]
        loss = ctypes.c_float()
        self.ctx.check(lib().vkrt_rc_train(self._h, rec.ctypes.data, tg.ctypes.data, len(rec),
                                           RC_GRAD_ONLY if grad_only else 0, ctypes.byref(loss)))
        return loss.value

    def get_params(self):
        mlp = np.zeros(self.n_mlp, np.float32)
        grid = np.zeros(self.n_grid, np.float32)
        self.ctx.check(lib().vkrt_rc_get_params(self._h, mlp.ctypes.data, grid.ctypes.data))
        return mlp, grid.reshape(16, -1, 2)

    def set_params(self, mlp=None, grid=None):
        m = None if mlp is None else np.ascontiguousarray(mlp, np.float32)
        g = None if grid is None else np.ascontiguousarray(grid, np.float32)
        self.ctx.check(lib().vkrt_rc_set_params(self._h, None if m is None else m.ctypes.data,
                                                None if g is None else g.ctypes.data))

    def get_gradients(self):
        mlp = np.zeros(self.n_mlp, np.float32)
        grid = np.zeros(self.n_grid, np.float32)
        self.ctx.check(lib().vkrt_rc_get_gradients(self._h, mlp.ctypes.data, grid.ctypes.data))
        return mlp, grid.reshape(16, -1, 2)

    def step(self):
        return lib().vkrt_rc_step(self._h)

    def close(self):
        if self._h:
            lib().vkrt_rc_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_rc_params(**kw):
    p = RcParams()
    lib().vkrt_rc_default_params(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


INSTANCE_DTYPE = np.dtype([("object_to_world", np.float32, 12), ("world_to_object", np.float32, 12),
                           ("blas", np.uint32), ("custom_index", np.uint32), ("reserved", np.uint32, 2)])
INSTANCE_HIT_DTYPE = np.dtype([("t", np.float32), ("u", np.float32), ("v", np.float32), ("prim", np.uint32),
                               ("instance", np.uint32), ("custom_index", np.uint32), ("reserved", np.uint32, 2)])


def instance(object_to_world, blas, custom_index=0):
    """one vkrt_instance record: the 3x4 object-to-world matrix and its inverse (vkrt_instance_set_transform)"""
    rec = np.zeros(1, INSTANCE_DTYPE)
    m = np.ascontiguousarray(np.asarray(object_to_world, np.float32).reshape(12))
    lib().vkrt_instance_set_transform(rec.ctypes.data, m.ctypes.data)
    rec["blas"], rec["custom_index"] = blas, custom_index
    return rec[0]


class Scene:
    """two-level scene: instances (INSTANCE_DTYPE records) of mesh BVHs"""

    def __init__(self, ctx, meshes, instances):
        self.ctx, self._meshes = ctx, list(meshes)
        arr = (_vp * max(1, len(self._meshes)))(*[m.handle for m in self._meshes])
        inst = np.ascontiguousarray(instances, INSTANCE_DTYPE)
        self._h = _vp()
        ctx.check(lib().vkrt_scene_create(ctx.handle, arr, len(self._meshes), inst.ctypes.data, len(inst),
                                          ctypes.byref(self._h)))
        ctx._adopt(self)

    def set_instances(self, instances):
        inst = np.ascontiguousarray(instances, INSTANCE_DTYPE)
        self.ctx.check(lib().vkrt_scene_set_instances(self._h, inst.ctypes.data, len(inst)))

    def info(self):
        a, b, c = _u32(), _u32(), _u32()
        self.ctx.check(lib().vkrt_scene_get_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def trace_closest(self, rays, out=None, flags=0):
        n = _n(rays, 8)
        if out is None:
            out = np.zeros(n, INSTANCE_HIT_DTYPE)
        self.ctx.check(lib().vkrt_scene_trace_closest(self.ctx.handle, self._h, _ptr(rays)[0], n, _ptr(out)[0],
                                                      _flags([rays, out], flags)))
        return out

    def trace_any(self, rays, out=None, flags=0):
        n = _n(rays, 8)
        if out is None:
            out = np.zeros(n, np.uint8)
        self.ctx.check(lib().vkrt_scene_trace_any(self.ctx.handle, self._h, _ptr(rays)[0], n, _ptr(out)[0],
                                                  _flags([rays, out], flags)))
        return out

    def close(self):
        if self._h:
            lib().vkrt_scene_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# -- scene files (host) -------------------------------------------------------------------------
def load_model(path, environment=1.0):
    """an OBJ (+ MTL) or glTF 2.0 (.gltf / .glb) file as a scene dict for PathTracer.set_scene (copies of the
    library's arrays: positions, normals, texcoords, indices, tri_object, object_material, materials,
    environment, n_textures); raises VkrtError with the loader's message"""
    h = _vp()
    rc = lib().vkrt_model_load(str(path).encode(), ctypes.byref(h))
    if rc != OK:
        raise VkrtError(rc, lib().vkrt_model_error().decode())
    try:
        v = ModelView()
        if lib().vkrt_model_get(h, ctypes.byref(v)) != OK:
            raise VkrtError(E_INVALID_ARG, "vkrt_model_get failed")

        def arr(ptr, n, dtype):
            if n == 0 or not ptr:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(np.ctypeslib.as_ctypes_type(dtype))),
                                         (n,)).copy()

        mats = np.zeros(v.n_materials, MATERIAL_DTYPE)
        if v.n_materials:
            ctypes.memmove(mats.ctypes.data, v.materials, v.n_materials * MATERIAL_DTYPE.itemsize)
        textures, samplers = [], []
        if v.n_textures:
            tv = ctypes.cast(v.textures, ctypes.POINTER(Texture))
            for k in range(v.n_textures):
                t = tv[k]
                textures.append(arr(t.rgba, 4 * t.width * t.height, np.uint8).reshape(t.height, t.width, 4))
                samplers.append((t.address_mode_u, t.address_mode_v, t.min_filter, t.mag_filter))
        return {"positions": arr(v.positions, 3 * v.n_vert, np.float32).reshape(-1, 3),
                "normals": arr(v.normals, 3 * v.n_vert, np.float32).reshape(-1, 3),
                "texcoords": arr(v.texcoords, 2 * v.n_vert, np.float32).reshape(-1, 2),
                "tangents": arr(v.tangents, 3 * v.n_vert, np.float32).reshape(-1, 3),
                "indices": arr(v.indices, 3 * v.n_tri, np.uint32),
                "tri_object": arr(v.tri_object, v.n_tri, np.uint32),
                "object_material": arr(v.object_material, v.n_objects, np.uint32),
                "materials": mats, "environment": float(environment), "textures": textures,
                "texture_samplers": samplers,
                "n_skipped_textures": int(v.n_skipped_textures)}
    finally:
        lib().vkrt_model_destroy(h)


# -- procedural geometry (host) ----------------------------------------------------------------
def icosphere(radius, subdiv):
    """Geometry::IcoSphere(radius, subdiv): (positions float32 [n,3], indices uint32 [m])."""
    nv, ni = _u32(), _u32()
    if lib().vkrt_icosphere_size(subdiv, ctypes.byref(nv), ctypes.byref(ni)) != OK:
        raise ValueError("bad subdivision count")
    pos = np.zeros((nv.value, 3), dtype=np.float32)
    idx = np.zeros(ni.value, dtype=np.uint32)
    if lib().vkrt_icosphere(ctypes.c_float(radius), subdiv, pos.ctypes.data, idx.ctypes.data) != OK:
        raise RuntimeError("vkrt_icosphere failed")
    return pos, idx


def plane(width, height, segments_w, segments_h):
    """Geometry::Plane(width, height, segments_w, segments_h) in the XZ plane."""
    nv, ni = _u32(), _u32()
    if lib().vkrt_plane_size(segments_w, segments_h, ctypes.byref(nv), ctypes.byref(ni)) != OK:
        raise ValueError("bad segment counts")
    pos = np.zeros((nv.value, 3), dtype=np.float32)
    idx = np.zeros(ni.value, dtype=np.uint32)
    if lib().vkrt_plane(ctypes.c_float(width), ctypes.c_float(height), segments_w, segments_h,
                        pos.ctypes.data, idx.ctypes.data) != OK:
        raise RuntimeError("vkrt_plane failed")
    return pos, idx
