"""ctypes binding of the C ABI (include/rm.h, include/rm_host.h).

Loads raymond_amd/libraymond_amd.so built in-tree.  There is no fallback:
if the library is missing the import of this module raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# RM_LIBRARY selects another build of the library (kernel variants under test)
LIB_PATH = os.environ.get("RM_LIBRARY") or os.path.join(HERE, "libraymond_amd.so")

RAY_CAMERA, RAY_REFLECTION, RAY_TRANSMISSION, RAY_SHADOW = 1, 2, 4, 8
RAY_VOLUME, RAY_DIFFUSE, RAY_GLOSSY, RAY_SINGULAR, RAY_ALL = 16, 32, 64, 128, 255
SAMPLING_BSDF, SAMPLING_NEE, SAMPLING_MIS = 0, 1, 2
TONEMAP_LINEAR, TONEMAP_HABLE, TONEMAP_ACES = 0, 1, 2


class RmError(RuntimeError):
    pass


class DeviceInfo(C.Structure):
    _fields_ = [("compute_units", C.c_int), ("wavefront_size", C.c_int), ("global_memory_bytes", C.c_uint64),
                ("name", C.c_char * 64), ("arch", C.c_char * 32)]


class InstanceDesc(C.Structure):
    _fields_ = [("vertex_offset", C.c_uint32), ("vertex_count", C.c_uint32), ("face_offset", C.c_uint32),
                ("face_count", C.c_uint32), ("transform", C.c_float * 16), ("visibility", C.c_uint32),
                ("emissive", C.c_uint32), ("has_bounds", C.c_uint32), ("bounds_min", C.c_float * 3),
                ("bounds_max", C.c_float * 3)]


class LightInfo(C.Structure):
    _fields_ = [("shader_index", C.c_uint32), ("casts_shadows", C.c_uint32), ("uses_mis", C.c_uint32)]


class AreaLight(C.Structure):
    _fields_ = [("info", LightInfo), ("transform", C.c_float * 12), ("color", C.c_float * 3),
                ("is_circular", C.c_uint32)]


class PointLight(C.Structure):
    _fields_ = [("info", LightInfo), ("location", C.c_float * 3), ("radius", C.c_float), ("color", C.c_float * 3)]


class SunLight(C.Structure):
    _fields_ = [("info", LightInfo), ("direction", C.c_float * 3), ("cos_angle", C.c_float),
                ("color", C.c_float * 3)]


class SpotLight(C.Structure):
    _fields_ = [("info", LightInfo), ("location", C.c_float * 3), ("direction", C.c_float * 3),
                ("radius", C.c_float), ("color", C.c_float * 3), ("spot_size", C.c_float),
                ("spot_blend", C.c_float)]


class CameraDesc(C.Structure):
    _fields_ = [("transform", C.c_float * 16), ("near_clip", C.c_float), ("far_clip", C.c_float),
                ("focal_length", C.c_float), ("shift", C.c_float * 2)]


class TextureDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("channels", C.c_uint32),
                ("pixels", C.POINTER(C.c_float))]


class LensSurface(C.Structure):
    _fields_ = [("radius", C.c_float), ("thickness", C.c_float), ("aperture", C.c_float), ("ior_d", C.c_float),
                ("abbe", C.c_float)]


class LensDesc(C.Structure):
    _fields_ = [("num_surfaces", C.c_uint32), ("surfaces", C.POINTER(LensSurface)), ("stop_index", C.c_int32),
                ("relative_stop", C.c_float), ("aperture_blades", C.c_int32), ("sensor_scale", C.c_float),
                ("camera_scale", C.c_float), ("focus", C.c_float), ("spectral", C.c_uint32)]


def lens_desc(surfaces, stop_index=0, relative_stop=1.0, aperture_blades=7, sensor_scale=1.0, camera_scale=0.001,
              focus=0.0, spectral=True):
    """rm_lens_desc from surfaces (n, 5): radius, thickness, aperture, n_d,
    V_d, object side first, the image plane last (defaults as the
    reference's uniforms, Renderer.swift:333-343)."""
    s = np.ascontiguousarray(surfaces, np.float32).reshape(-1, 5)
    arr = (LensSurface * s.shape[0])(*[LensSurface(*map(float, row)) for row in s])
    d = LensDesc(s.shape[0], arr, stop_index, relative_stop, aperture_blades, sensor_scale, camera_scale, focus,
                 int(bool(spectral)))
    d._keep = arr
    return d


class SceneDesc(C.Structure):
    _fields_ = [
        ("num_vertices", C.c_uint32), ("vertices", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
        ("texcoords", C.POINTER(C.c_float)),
        ("num_faces", C.c_uint32), ("indices", C.POINTER(C.c_uint32)), ("face_materials", C.POINTER(C.c_uint16)),
        ("num_instances", C.c_uint32), ("instances", C.POINTER(InstanceDesc)),
        ("num_area_lights", C.c_uint32), ("area_lights", C.POINTER(AreaLight)),
        ("num_point_lights", C.c_uint32), ("point_lights", C.POINTER(PointLight)),
        ("num_sun_lights", C.c_uint32), ("sun_lights", C.POINTER(SunLight)),
        ("num_spot_lights", C.c_uint32), ("spot_lights", C.POINTER(SpotLight)),
        ("world_shader", C.c_uint32), ("envmap_resolution", C.c_uint32),
        ("camera", CameraDesc),
        ("num_textures", C.c_uint32), ("textures", C.POINTER(TextureDesc)),
    ]


class SceneStats(C.Structure):
    _fields_ = [("triangles", C.c_uint64), ("bvh_nodes", C.c_uint64), ("bvh_depth", C.c_uint32),
                ("bvh_build_seconds", C.c_double), ("jit_seconds", C.c_double), ("envmap_seconds", C.c_double),
                ("device_bytes", C.c_uint64), ("stored_triangles", C.c_uint64), ("two_level", C.c_uint32)]


class Uniforms(C.Structure):
    _fields_ = [("sampling_mode", C.c_uint32), ("accumulate", C.c_uint32), ("exposure", C.c_float),
                ("tonemapping", C.c_uint32), ("rr_depth", C.c_uint32), ("accumulate_f64", C.c_uint32),
                ("timing", C.c_uint32), ("traversal_stats", C.c_uint32), ("material_sort", C.c_uint32),
                ("use_graph", C.c_uint32), ("trace_blocks_per_cu", C.c_uint32), ("sort_key_shift", C.c_uint32)]


class RenderStats(C.Structure):
    _fields_ = [("rays", C.c_uint64 * 33), ("shadow_rays", C.c_uint64 * 32), ("total_rays", C.c_uint64),
                ("closest_total", C.c_uint64), ("shadow_total", C.c_uint64), ("samples", C.c_uint64), ("ms_total", C.c_double), ("ms_trace_closest", C.c_double),
                ("ms_shade", C.c_double), ("ms_trace_any", C.c_double),
                ("trace_closest_launches", C.c_uint32), ("trace_any_launches", C.c_uint32),
                ("iterations", C.c_uint64), ("nodes_closest", C.c_uint64), ("tris_closest", C.c_uint64), ("nodes_any", C.c_uint64),
                ("tris_any", C.c_uint64), ("ms_sort", C.c_double)]


class CacheDesc(C.Structure):
    _fields_ = [("log2_table_size", C.c_uint32), ("seed", C.c_uint32), ("learning_rate", C.c_float),
                ("train_shift", C.c_uint32), ("batch_size", C.c_uint32), ("steps_per_launch", C.c_uint32)]


class CacheStats(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("samples", C.c_uint64), ("queries", C.c_uint64),
                ("train_records", C.c_uint64), ("last_loss", C.c_float), ("ms_infer", C.c_double),
                ("ms_train", C.c_double)]


class JobDesc(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("samples", C.c_uint32),
                ("max_depth", C.c_uint32), ("seed", C.c_uint32)]


BATCH_SHORTEST_FIRST, BATCH_SAMPLE_MAJOR = 0, 1


class BatchConfig(C.Structure):
    _fields_ = [("queue_paths", C.c_uint32), ("material_sort", C.c_uint32), ("rr_depth", C.c_uint32),
                ("timing", C.c_uint32), ("use_graph", C.c_uint32), ("order", C.c_uint32),
                ("trace_blocks_per_cu", C.c_uint32), ("sort_key_shift", C.c_uint32)]


class BatchStats(C.Structure):
    _fields_ = [("total_rays", C.c_uint64), ("closest_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("samples", C.c_uint64), ("iterations", C.c_uint64), ("loops", C.c_uint32),
                ("queue_paths", C.c_uint32), ("ms_total", C.c_double), ("makespan_ms", C.c_double),
                ("ms_trace_closest", C.c_double), ("ms_shade", C.c_double), ("ms_trace_any", C.c_double),
                ("ms_sort", C.c_double)]


class BatchTuneEntry(C.Structure):
    _fields_ = [("queue_paths", C.c_uint32), ("material_sort", C.c_uint32), ("trace_blocks_per_cu", C.c_uint32),
                ("sort_key_shift", C.c_uint32), ("mrays_per_s", C.c_double)]


CACHE_FEATURES = 12
CACHE_MLP_PARAMS = 9216


EXPORTS = {
    # name: (restype, argtypes)
    "rm_last_error": (C.c_char_p, []),
    "rm_version": (C.c_char_p, []),
    "rm_device_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "rm_device_destroy": (None, [C.c_void_p]),
    "rm_device_get_info": (C.c_int, [C.c_void_p, C.POINTER(DeviceInfo)]),
    "rm_scene_create": (C.c_int, [C.c_void_p, C.POINTER(SceneDesc), C.POINTER(C.c_void_p)]),
    "rm_scene_destroy": (None, [C.c_void_p]),
    "rm_scene_build_accel": (C.c_int, [C.c_void_p, C.c_int]),
    "rm_scene_build_accel_two_level": (C.c_int, [C.c_void_p, C.c_int]),
    "rm_scene_compile_materials": (C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t, C.c_char_p, C.c_char_p,
                                             C.c_size_t]),
    "rm_scene_build_envmap": (C.c_int, [C.c_void_p]),
    "rm_scene_set_camera": (C.c_int, [C.c_void_p, C.POINTER(CameraDesc)]),
    "rm_scene_get_stats": (C.c_int, [C.c_void_p, C.POINTER(SceneStats)]),
    "rm_scene_read_envmap": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]),
    "rm_scene_read_light_tables": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]),
    "rm_trace_closest": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p]),
    "rm_trace_any": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "rm_trace_closest_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_void_p]),
    "rm_renderer_create": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                     C.POINTER(C.c_void_p)]),
    "rm_renderer_destroy": (None, [C.c_void_p]),
    "rm_renderer_set_uniforms": (C.c_int, [C.c_void_p, C.POINTER(Uniforms)]),
    "rm_renderer_reset": (C.c_int, [C.c_void_p]),
    "rm_renderer_set_lens": (C.c_int, [C.c_void_p, C.c_void_p]),
    "rm_render": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(RenderStats)]),
    "rm_read_image": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "rm_read_image_f64": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "rm_resolve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "rm_renderer_frame_index": (C.c_uint32, [C.c_void_p]),
    "rm_device_copy_probe": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_double)]),
    "rm_cache_create": (C.c_int, [C.c_void_p, C.POINTER(CacheDesc), C.POINTER(C.c_void_p)]),
    "rm_cache_destroy": (None, [C.c_void_p]),
    "rm_cache_get_stats": (C.c_int, [C.c_void_p, C.POINTER(CacheStats)]),
    "rm_cache_table_floats": (C.c_uint64, [C.c_void_p]),
    "rm_cache_read_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]),
    "rm_cache_write_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]),
    "rm_cache_infer": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "rm_cache_gradients": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                     C.c_void_p]),
    "rm_cache_train": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]),
    "rm_renderer_set_cache": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "rm_batch_create": (C.c_int, [C.c_void_p, C.POINTER(JobDesc), C.c_uint32, C.POINTER(C.c_void_p)]),
    "rm_batch_destroy": (None, [C.c_void_p]),
    "rm_batch_render": (C.c_int, [C.c_void_p, C.POINTER(BatchConfig), C.POINTER(BatchStats)]),
    "rm_batch_job_times": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "rm_batch_read_image": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_size_t]),
    "rm_batch_autotune": (C.c_int, [C.c_void_p, C.POINTER(BatchConfig), C.POINTER(BatchTuneEntry), C.c_uint32,
                                    C.POINTER(C.c_uint32), C.POINTER(C.c_int)]),
    "rmh_codegen": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "rmh_identifier": (C.c_int, [C.c_char_p, C.c_char_p, C.c_size_t]),
    "rmh_codegen_textures": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "rmh_sky_nishita_texture": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                          C.c_float]),
    "rmh_sky_nishita_sun": (C.c_int, [C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_void_p, C.c_void_p]),
    "rmh_free": (None, [C.c_void_p]),
    "rmh_ply_read": (C.c_int, [C.c_char_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "rmh_ply_free": (None, [C.c_void_p]),
    "rmh_save_exr": (C.c_int, [C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_float, C.POINTER(C.c_void_p)]),
    "rmh_jit_precompile": (C.c_int, [C.c_char_p, C.c_size_t, C.c_char_p, C.POINTER(C.c_int), C.c_char_p, C.c_size_t]),
    "rmh_bvh8_check": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                 C.c_char_p, C.c_size_t]),
    "rmh_two_level_trace": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                      C.c_void_p, C.c_uint32, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.POINTER(C.c_uint64), C.c_char_p, C.c_size_t]),
}


def load_library(path=LIB_PATH):
    if not os.path.exists(path):
        raise ImportError("raymond_amd native library not built: %s (run raymond_amd/build.py)" % path)
    lib = C.CDLL(path)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


LIB = load_library()


def _check(status):
    if status != 0:
        raise RmError("%d: %s" % (status, LIB.rm_last_error().decode(errors="replace")))


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


def codegen(materials_json, noinline=False):
    """Node-graph JSON -> material source (MaterialBuilder.swift:32-904)."""
    src = C.c_void_p()
    err = C.c_void_p()
    status = LIB.rmh_codegen(materials_json.encode(), int(noinline), C.byref(src), C.byref(err))
    if status != 0:
        msg = C.cast(err, C.c_char_p).value.decode() if err.value else "codegen failed"
        LIB.rmh_free(err)
        raise RmError(msg)
    out = C.cast(src, C.c_char_p).value.decode()
    LIB.rmh_free(src)
    return out


def codegen_textures(materials_json):
    """The texture slots of a material library, in slot order
    (rmh_codegen_textures): dicts with "kind" "image" and "path", or "nishita"
    and the sky's atmosphere."""
    import json
    out = C.c_void_p()
    err = C.c_void_p()
    if LIB.rmh_codegen_textures(materials_json.encode(), C.byref(out), C.byref(err)) != 0:
        msg = C.cast(err, C.c_char_p).value.decode() if err.value else "codegen failed"
        LIB.rmh_free(err)
        raise RmError(msg)
    text = C.cast(out, C.c_char_p).value.decode()
    LIB.rmh_free(out)
    return json.loads(text)


def sky_nishita_texture(width=512, height=512, sun_elevation=0.2617994, altitude=0.0, air_density=1.0,
                        dust_density=1.0, ozone_density=1.0):
    """The Nishita sky texture (rmh_sky_nishita_texture): (height, width, 4)
    float32, CIE XYZ in RGB, alpha 1."""
    out = np.zeros((height, width, 4), np.float32)
    out[..., 3] = 1.0
    _check(LIB.rmh_sky_nishita_texture(out.ctypes.data, width, height, sun_elevation, altitude, air_density,
                                       dust_density, ozone_density))
    return out


def sky_nishita_sun(sun_elevation, angular_diameter, altitude=0.0, air_density=1.0, dust_density=1.0):
    """XYZ radiance of the sun disc's lower and upper edge (rmh_sky_nishita_sun)."""
    bottom = np.zeros(3, np.float32)
    top = np.zeros(3, np.float32)
    _check(LIB.rmh_sky_nishita_sun(sun_elevation, angular_diameter, altitude, air_density, dust_density,
                                   bottom.ctypes.data, top.ctypes.data))
    return bottom, top


class PlyMesh(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("num_faces", C.c_uint32), ("vertices", C.POINTER(C.c_float)),
                ("normals", C.POINTER(C.c_float)), ("texcoords", C.POINTER(C.c_float)),
                ("indices", C.POINTER(C.c_uint32)), ("face_palette", C.POINTER(C.c_uint8)),
                ("bounds_min", C.c_float * 3), ("bounds_max", C.c_float * 3)]


def read_ply(path):
    """The exporter's ASCII PLY (rmh_ply_read; PLYReader.m:1-169): a dict of
    vertices (n, 3), normals (n, 3), texcoords (n, 2), faces (m, 3),
    face_palette (m,) and the reader's bounds_min / bounds_max."""
    mesh = PlyMesh()
    err = C.c_void_p()
    if LIB.rmh_ply_read(os.fsencode(path), C.byref(mesh), C.byref(err)) != 0:
        msg = C.cast(err, C.c_char_p).value.decode() if err.value else "PLY read failed"
        LIB.rmh_free(err)
        raise RmError(msg)
    try:
        nv, nf = mesh.num_vertices, mesh.num_faces

        def take(ptr, n, k, dtype):
            if n == 0:
                return np.zeros((0, k), dtype)
            return np.ctypeslib.as_array(ptr, shape=(n * k,)).reshape(n, k).astype(dtype, copy=True)

        return dict(vertices=take(mesh.vertices, nv, 3, np.float32), normals=take(mesh.normals, nv, 3, np.float32),
                    texcoords=take(mesh.texcoords, nv, 2, np.float32), faces=take(mesh.indices, nf, 3, np.uint32),
                    face_palette=take(mesh.face_palette, nf, 1, np.uint8).reshape(-1),
                    bounds_min=np.array(mesh.bounds_min[:], np.float32),
                    bounds_max=np.array(mesh.bounds_max[:], np.float32))
    finally:
        LIB.rmh_ply_free(C.byref(mesh))


def save_exr(path, rgba, norm=1.0):
    """Writes an (h, w, 4) float image as OpenEXR (rmh_save_exr): RGB * norm,
    alpha 1."""
    img = np.ascontiguousarray(rgba, np.float32)
    if img.ndim != 3 or img.shape[2] != 4:
        raise ValueError("expected an (h, w, 4) image")
    err = C.c_void_p()
    if LIB.rmh_save_exr(os.fsencode(path), img.ctypes.data, img.shape[1], img.shape[0], float(norm),
                        C.byref(err)) != 0:
        msg = C.cast(err, C.c_char_p).value.decode() if err.value else "EXR write failed"
        LIB.rmh_free(err)
        raise RmError(msg)


def identifier(name):
    buf = C.create_string_buffer(256)
    _check(LIB.rmh_identifier(name.encode(), buf, 256))
    return buf.value.decode()


def jit_precompile(source, options=""):
    """Compiles material source into the on-disk code-object cache without a
    device (rmh_jit_precompile); returns True if the entry already existed."""
    b = source.encode()
    hit = C.c_int(0)
    log = C.create_string_buffer(1 << 16)
    status = LIB.rmh_jit_precompile(b, len(b), options.encode(), C.byref(hit), log, len(log))
    if status != 0:
        raise RmError("material compile failed:\n%s" % log.value.decode())
    return bool(hit.value)


def bvh8_check(tris, threads=0):
    """Host build + structural check of the BVH8 (rmh_bvh8_check)."""
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    nodes, depth = C.c_uint32(), C.c_uint32()
    err = C.create_string_buffer(256)
    status = LIB.rmh_bvh8_check(tris.ctypes.data, tris.shape[0], threads, C.byref(nodes), C.byref(depth), err, 256)
    if status != 0:
        raise RmError(err.value.decode())
    return nodes.value, depth.value


def two_level_trace(vertices, faces, instances, rays, any_hit=False, threads=0):
    """Host build of the two-level form and a CPU trace through it
    (rmh_two_level_trace).  instances: InstanceDesc array (vertex_offset,
    face_offset, face_count and transform used).  Returns (t, inst, prim,
    bary, stored_triangles)."""
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    f = np.ascontiguousarray(faces, np.uint32).reshape(-1, 3)
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = r.shape[0]
    t = np.zeros(n, np.float32)
    inst = np.zeros(n, np.uint32)
    prim = np.zeros(n, np.uint32)
    bary = np.zeros((n, 2), np.float32)
    stored = C.c_uint64()
    err = C.create_string_buffer(256)
    status = LIB.rmh_two_level_trace(v.ctypes.data, v.shape[0], f.ctypes.data, f.shape[0], C.addressof(instances),
                                     len(instances), r.ctypes.data, n, int(bool(any_hit)), threads, t.ctypes.data,
                                     inst.ctypes.data, prim.ctypes.data, bary.ctypes.data, C.byref(stored), err, 256)
    if status != 0:
        raise RmError(err.value.decode())
    return t, inst, prim, bary, stored.value


class Device:
    def __init__(self, ordinal=0):
        self.handle = C.c_void_p()
        _check(LIB.rm_device_create(ordinal, C.byref(self.handle)))

    def copy_probe(self, nbytes):
        """Streaming device copy of nbytes (reads and writes nbytes); returns ms."""
        ms = C.c_double(0.0)
        _check(LIB.rm_device_copy_probe(self.handle, nbytes, C.byref(ms)))
        return ms.value

    def info(self):
        i = DeviceInfo()
        _check(LIB.rm_device_get_info(self.handle, C.byref(i)))
        return i

    def close(self):
        if self.handle:
            LIB.rm_device_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        self.close()


class Scene:
    """A device-resident scene (the reference's Scene, SceneLoader.swift:6-27)."""

    def __init__(self, device, desc):
        self.device = device
        self.desc = desc
        self.handle = C.c_void_p()
        sd = desc.to_c()
        _check(LIB.rm_scene_create(device.handle, C.byref(sd), C.byref(self.handle)))

    def build_accel(self, threads=0, two_level=False):
        if two_level:
            _check(LIB.rm_scene_build_accel_two_level(self.handle, threads))
        else:
            _check(LIB.rm_scene_build_accel(self.handle, threads))

    def compile_materials(self, source, options=""):
        log = C.create_string_buffer(1 << 16)
        b = source.encode()
        status = LIB.rm_scene_compile_materials(self.handle, b, len(b), options.encode(), log, len(log))
        if status != 0:
            raise RmError("material compile failed: %s\n%s" % (LIB.rm_last_error().decode(), log.value.decode()))
        return log.value.decode()

    def build_envmap(self):
        _check(LIB.rm_scene_build_envmap(self.handle))

    def stats(self):
        s = SceneStats()
        _check(LIB.rm_scene_get_stats(self.handle, C.byref(s)))
        return s

    def read_envmap(self):
        res = self.desc.envmap_resolution or 2048
        levels = int(np.log2(res))
        mip = np.zeros(sum(4 ** k for k in range(levels + 1)), np.float32)
        pdfs = np.zeros(res * res, np.float32)
        _check(LIB.rm_scene_read_envmap(self.handle, pdfs.ctypes.data, pdfs.size, mip.ctypes.data, mip.size))
        return pdfs, mip

    def read_light_tables(self, n_faces, n_lights):
        faces = np.zeros(max(n_faces, 1), np.float32)
        areas = np.zeros(max(n_lights, 1), np.float32)
        _check(LIB.rm_scene_read_light_tables(self.handle, faces.ctypes.data, faces.size, areas.ctypes.data,
                                              areas.size))
        return faces[:n_faces], areas[:n_lights]

    def trace_closest(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        t = np.zeros(n, np.float32)
        inst = np.zeros(n, np.uint32)
        prim = np.zeros(n, np.uint32)
        bary = np.zeros((n, 2), np.float32)
        _check(LIB.rm_trace_closest(self.handle, rays.ctypes.data, n, t.ctypes.data, inst.ctypes.data,
                                    prim.ctypes.data, bary.ctypes.data))
        return t, inst, prim, bary

    def trace_any(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        occ = np.zeros(n, np.uint32)
        _check(LIB.rm_trace_any(self.handle, rays.ctypes.data, n, occ.ctypes.data))
        return occ

    def trace_closest_noalloc(self, d_rays, n, d_t, d_inst, d_prim, d_bary, stream=None):
        """rm_trace_closest_noalloc on device arrays (addresses as ints): only
        enqueues the work on `stream` (None: the device's stream)."""
        _check(LIB.rm_trace_closest_noalloc(self.handle, d_rays, n, d_t, d_inst, d_prim, d_bary, stream))

    def close(self):
        if self.handle:
            LIB.rm_scene_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        self.close()


class Renderer:
    """Wavefront renderer (host/Renderer.swift): execute/setSize/reset/saveFrame."""

    def __init__(self, scene, width, height, max_depth=8, samples_per_launch=1):
        self.scene = scene
        self.width, self.height = width, height
        self.handle = C.c_void_p()
        _check(LIB.rm_renderer_create(scene.handle, width, height, max_depth, samples_per_launch,
                                      C.byref(self.handle)))
        self.uniforms = Uniforms(sampling_mode=SAMPLING_MIS, accumulate=1, exposure=1.0,
                                 tonemapping=TONEMAP_LINEAR, rr_depth=0)
        self.set_uniforms()

    def set_uniforms(self, **kw):
        for k, v in kw.items():
            setattr(self.uniforms, k, v)
        _check(LIB.rm_renderer_set_uniforms(self.handle, C.byref(self.uniforms)))

    def reset(self):
        _check(LIB.rm_renderer_reset(self.handle))

    def set_lens(self, lens):
        """Renderer.setLens: an api.lens_desc(...) or None for the pinhole camera."""
        self._lens = lens
        _check(LIB.rm_renderer_set_lens(self.handle, C.byref(lens) if lens is not None else None))

    def execute(self, samples=1):
        st = RenderStats()
        _check(LIB.rm_render(self.handle, samples, C.byref(st)))
        return st

    def image(self):
        out = np.zeros((self.height, self.width, 4), np.float32)
        _check(LIB.rm_read_image(self.handle, out.ctypes.data, out.size))
        return out

    def image_f64(self):
        out = np.zeros((self.height, self.width, 4), np.float64)
        _check(LIB.rm_read_image_f64(self.handle, out.ctypes.data, out.size))
        return out

    def resolve(self):
        out = np.zeros((self.height, self.width, 4), np.float32)
        _check(LIB.rm_resolve(self.handle, out.ctypes.data, out.size))
        return out

    @property
    def frame_index(self):
        return LIB.rm_renderer_frame_index(self.handle)

    def save_frame(self, path):
        """Renderer.saveFrame (Renderer.swift:804-807): the accumulation buffer
        as OpenEXR, scaled by 1 / (frame index - 1) when accumulating, as the
        reference scales it (one frame fewer than normalizeImage divides by)."""
        norm = 1.0
        if self.uniforms.accumulate:
            if self.frame_index < 2:
                raise RmError("save_frame needs at least two accumulated frames (the scale is 1 / (frames - 1))")
            norm = float(np.float32(1.0) / np.float32(self.frame_index - 1))
        save_exr(path, self.image(), norm)

    def set_cache(self, cache, mode=1):
        """Attach a RadianceCache (mode 1: query and train, 2: query only) or detach (None)."""
        self._cache = cache
        _check(LIB.rm_renderer_set_cache(self.handle, cache.handle if cache else None, mode if cache else 0))

    def close(self):
        if self.handle:
            LIB.rm_renderer_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        self.close()


def _rows(a, k):
    return np.ascontiguousarray(a, np.float32).reshape(-1, k)


class RadianceCache:
    """Learned radiance cache (rm_cache_*): hash-grid encoding + 48-64-64-3 MLP."""

    def __init__(self, device, log2_table_size=19, seed=0, learning_rate=1e-2, train_shift=4, batch_size=8192,
                 steps_per_launch=2):
        self.device = device
        self.desc = CacheDesc(log2_table_size, seed, learning_rate, train_shift, batch_size, steps_per_launch)
        self.handle = C.c_void_p()
        _check(LIB.rm_cache_create(device.handle, C.byref(self.desc), C.byref(self.handle)))
        self.table_floats = int(LIB.rm_cache_table_floats(self.handle))

    def stats(self):
        s = CacheStats()
        _check(LIB.rm_cache_get_stats(self.handle, C.byref(s)))
        return s

    def read_params(self):
        mlp = np.zeros(CACHE_MLP_PARAMS, np.float32)
        table = np.zeros(self.table_floats, np.float32)
        _check(LIB.rm_cache_read_params(self.handle, mlp.ctypes.data, mlp.size, table.ctypes.data, table.size))
        return mlp, table

    def write_params(self, mlp=None, table=None):
        m = None if mlp is None else np.ascontiguousarray(mlp, np.float32)
        t = None if table is None else np.ascontiguousarray(table, np.float32)
        _check(LIB.rm_cache_write_params(self.handle, None if m is None else m.ctypes.data, 0 if m is None else m.size,
                                         None if t is None else t.ctypes.data, 0 if t is None else t.size))

    def infer(self, features):
        f = _rows(features, CACHE_FEATURES)
        out = np.zeros((f.shape[0], 3), np.float32)
        _check(LIB.rm_cache_infer(self.handle, f.ctypes.data, f.shape[0], out.ctypes.data))
        return out

    def gradients(self, features, targets, table=True):
        f = _rows(features, CACHE_FEATURES)
        t = _rows(targets, 3)
        gm = np.zeros(CACHE_MLP_PARAMS, np.float32)
        gt = np.zeros(self.table_floats, np.float32) if table else None
        loss = np.zeros(1, np.float32)
        _check(LIB.rm_cache_gradients(self.handle, f.ctypes.data, t.ctypes.data, f.shape[0], gm.ctypes.data,
                                      None if gt is None else gt.ctypes.data, loss.ctypes.data))
        return gm, gt, float(loss[0])

    def train(self, features, targets, steps=1):
        f = _rows(features, CACHE_FEATURES)
        t = _rows(targets, 3)
        loss = np.zeros(1, np.float32)
        _check(LIB.rm_cache_train(self.handle, f.ctypes.data, t.ctypes.data, f.shape[0], steps, loss.ctypes.data))
        return float(loss[0])

    def close(self):
        if self.handle:
            LIB.rm_cache_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        self.close()


class Batch:
    """Many jobs in one bounce loop (rm_batch_*).  jobs: [(Scene, width, height, samples, max_depth, seed)]."""

    def __init__(self, device, jobs):
        self.device = device
        self.scenes = [j[0] for j in jobs]  # kept alive: the batch points at their envmaps
        self.sizes = [(j[1], j[2], j[3]) for j in jobs]
        arr = (JobDesc * len(jobs))()
        for k, (scene, w, h, spp, depth, seed) in enumerate(jobs):
            arr[k] = JobDesc(scene.handle.value, w, h, spp, depth, seed)
        self.handle = C.c_void_p()
        _check(LIB.rm_batch_create(device.handle, arr, len(jobs), C.byref(self.handle)))

    def render(self, queue_paths=0, material_sort=0, rr_depth=0, timing=0, use_graph=0, order=BATCH_SHORTEST_FIRST,
               trace_blocks_per_cu=0, sort_key_shift=0):
        cfg = BatchConfig(queue_paths, material_sort, rr_depth, timing, use_graph, order, trace_blocks_per_cu,
                          sort_key_shift)
        st = BatchStats()
        _check(LIB.rm_batch_render(self.handle, C.byref(cfg), C.byref(st)))
        return st

    def job_times(self):
        out = np.zeros(len(self.sizes), np.float64)
        _check(LIB.rm_batch_job_times(self.handle, out.ctypes.data, out.size))
        return out

    def image(self, job):
        w, h, _ = self.sizes[job]
        out = np.zeros((h, w, 4), np.float32)
        _check(LIB.rm_batch_read_image(self.handle, job, out.ctypes.data, out.size))
        return out

    def autotune(self, rr_depth=0):
        cfg = BatchConfig(0, 0, rr_depth, 0, 0)
        entries = (BatchTuneEntry * 16)()
        count = C.c_uint32(0)
        hit = C.c_int(0)
        _check(LIB.rm_batch_autotune(self.handle, C.byref(cfg), entries, 16, C.byref(count), C.byref(hit)))
        tried = [(e.queue_paths, e.material_sort, e.trace_blocks_per_cu, e.sort_key_shift, e.mrays_per_s)
                 for e in entries[:count.value]]
        return cfg, tried, bool(hit.value)

    def close(self):
        if self.handle:
            LIB.rm_batch_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        self.close()
