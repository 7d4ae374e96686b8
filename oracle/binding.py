"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Imported by tests/ and by bench.py's cpu_baseline leg, never by the product.
compile_materials() builds the generated material source (the same text the
GPU compiles with hiprtc) with g++ into a shared object the oracle loads.
"""
import ctypes as C
import hashlib
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
LIB_PATH = os.path.join(BUILD, "liboracle.so")
DEVICE_DIR = os.path.join(REPO, "raymond_amd", "csrc", "device")

CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fvisibility=hidden", "-fvisibility-inlines-hidden",
            "-Wno-unused-function"]


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


class RenderArgs(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("max_depth", C.c_uint32),
                ("sampling_mode", C.c_uint32), ("rr_depth", C.c_uint32), ("frame_begin", C.c_uint32),
                ("frames", C.c_uint32), ("pixels", C.POINTER(C.c_uint32)), ("num_pixels", C.c_uint32),
                ("threads", C.c_int), ("brute_force", C.c_int), ("lens", C.c_void_p)]


class RenderResult(C.Structure):
    _fields_ = [("rays", C.c_uint64 * 33), ("shadow_rays", C.c_uint64 * 32), ("total_rays", C.c_uint64),
                ("paths", C.c_uint64), ("seconds", C.c_double), ("threads_used", C.c_int)]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_scene_create.restype = C.c_void_p
        L.or_scene_create.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_size_t]
        L.or_scene_destroy.argtypes = [C.c_void_p]
        L.or_build_envmap.argtypes = [C.c_void_p, C.c_int]
        L.or_set_envmap.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_read_envmap.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_read_light_tables.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_trace_closest.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_int, C.c_int]
        L.or_trace_any.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_int, C.c_int]
        L.or_scene_set_two_level.argtypes = [C.c_void_p, C.c_int]
        L.or_render.argtypes = [C.c_void_p, C.POINTER(RenderArgs), C.c_void_p, C.POINTER(RenderResult)]
        L.or_render_f64.argtypes = [C.c_void_p, C.POINTER(RenderArgs), C.c_void_p, C.POINTER(RenderResult)]
        _LIB = L
    return _LIB


def compile_materials(source):
    """g++ build of the generated material text plus exported entry points."""
    wrapper = (
        '#include "material_prelude.h"\n' + source +
        '\nextern "C" __attribute__((visibility("default"))) void or_material_surface('
        'uint32_t i, const rm::SceneView* c, rm::ShadingContext* s) { rm::shade_surface(i, *c, *s); }\n'
        'extern "C" __attribute__((visibility("default"))) void or_material_light('
        'uint32_t i, const rm::SceneView* c, rm::ShadingContext* s) { rm::shade_light(i, *c, *s); }\n')
    h = hashlib.sha1()
    h.update(wrapper.encode())
    for name in sorted(os.listdir(DEVICE_DIR)):
        if name.endswith(".h"):
            with open(os.path.join(DEVICE_DIR, name), "rb") as f:
                h.update(f.read())
    h.update(" ".join(CXXFLAGS).encode())
    key = h.hexdigest()[:16]
    os.makedirs(BUILD, exist_ok=True)
    so = os.path.join(BUILD, "materials_%s.so" % key)
    if not os.path.exists(so):
        src = os.path.join(BUILD, "materials_%s.cpp" % key)
        with open(src, "w") as f:
            f.write(wrapper)
        tmp = so + ".tmp%d" % os.getpid()
        subprocess.run(["g++"] + CXXFLAGS + ["-I", DEVICE_DIR, "-shared", "-o", tmp, src], check=True)
        os.replace(tmp, so)
    return so


class OracleScene:
    def __init__(self, desc, material_source=None):
        L = lib()
        self.desc = desc
        so = compile_materials(material_source if material_source is not None else desc.material_source())
        sd = desc.to_c()
        err = C.create_string_buffer(512)
        self.handle = L.or_scene_create(C.addressof(sd), so.encode(), err, 512)
        if not self.handle:
            raise RuntimeError("oracle scene: " + err.value.decode())

    def set_two_level(self, on=True):
        """Hits through the two-level form (every instance in its object space)."""
        if lib().or_scene_set_two_level(self.handle, int(bool(on))) != 0:
            raise RuntimeError("oracle scene: singular instance transform")

    def build_envmap(self, threads=0):
        lib().or_build_envmap(self.handle, threads)

    def set_envmap(self, pdfs, mip):
        pdfs = np.ascontiguousarray(pdfs, np.float32)
        mip = np.ascontiguousarray(mip, np.float32)
        lib().or_set_envmap(self.handle, pdfs.ctypes.data, mip.ctypes.data)

    def read_envmap(self):
        res = self.desc.envmap_resolution or 2048
        levels = int(np.log2(res))
        mip = np.zeros(sum(4 ** k for k in range(levels + 1)), np.float32)
        pdfs = np.zeros(res * res, np.float32)
        if lib().or_read_envmap(self.handle, pdfs.ctypes.data, mip.ctypes.data) != 0:
            raise RuntimeError("oracle envmap not built")
        return pdfs, mip

    def read_light_tables(self, n_faces, n_lights):
        faces = np.zeros(max(n_faces, 1), np.float32)
        areas = np.zeros(max(n_lights, 1), np.float32)
        lib().or_read_light_tables(self.handle, faces.ctypes.data, areas.ctypes.data)
        return faces[:n_faces], areas[:n_lights]

    def trace_closest(self, rays, brute_force=False, threads=0):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        t = np.zeros(n, np.float32)
        inst = np.zeros(n, np.uint32)
        prim = np.zeros(n, np.uint32)
        bary = np.zeros((n, 2), np.float32)
        lib().or_trace_closest(self.handle, rays.ctypes.data, n, t.ctypes.data, inst.ctypes.data,
                               prim.ctypes.data, bary.ctypes.data, int(brute_force), threads)
        return t, inst, prim, bary

    def trace_any(self, rays, brute_force=False, threads=0):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        occ = np.zeros(n, np.uint32)
        lib().or_trace_any(self.handle, rays.ctypes.data, n, occ.ctypes.data, int(brute_force), threads)
        return occ

    def render(self, width, height, max_depth, frames=1, frame_begin=0, sampling_mode=2, rr_depth=0,
               pixels=None, threads=0, brute_force=False, image=None, lens=None):
        """lens: an rm_lens_desc structure (raymond_amd.api.lens_desc) or None."""
        if image is None:
            image = np.zeros((height, width, 4), np.float32)
        args = RenderArgs(width, height, max_depth, sampling_mode, rr_depth, frame_begin, frames, None, 0,
                          threads, int(brute_force), C.addressof(lens) if lens is not None else None)
        if pixels is not None:
            pixels = np.ascontiguousarray(pixels, np.uint32)
            args.pixels = pixels.ctypes.data_as(C.POINTER(C.c_uint32))
            args.num_pixels = pixels.size
        res = RenderResult()
        st = lib().or_render(self.handle, C.byref(args), image.ctypes.data, C.byref(res))
        if st != 0:
            raise RuntimeError("oracle render failed (%d)" % st)
        return image, res

    def render_f64(self, width, height, max_depth, frames=1, frame_begin=0, sampling_mode=2, rr_depth=0,
                   threads=0):
        """The same paths as render(), each sample summed in FP32 and the
        samples accumulated in FP64 (the GPU's accumulate_f64 image)."""
        image = np.zeros((height, width, 4), np.float64)
        args = RenderArgs(width, height, max_depth, sampling_mode, rr_depth, frame_begin, frames, None, 0,
                          threads, 0)
        res = RenderResult()
        st = lib().or_render_f64(self.handle, C.byref(args), image.ctypes.data, C.byref(res))
        if st != 0:
            raise RuntimeError("oracle render failed (%d)" % st)
        return image, res

    def close(self):
        if self.handle:
            lib().or_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


# ---------------------------------------------------------------------------
# Radiance cache (rcache_ref.cpp): double-precision forward and gradients.

def _rc_lib():
    L = lib()
    if not getattr(L, "_rc_ready", False):
        L.or_rc_round_half.restype = C.c_float
        L.or_rc_round_half.argtypes = [C.c_float]
        L.or_rc_atan2.restype = C.c_float
        L.or_rc_atan2.argtypes = [C.c_float, C.c_float]
        L.or_rc_acos.restype = C.c_float
        L.or_rc_acos.argtypes = [C.c_float]
        L.or_rc_grid.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p]
        L.or_rc_index.restype = C.c_uint32
        L.or_rc_index.argtypes = [C.c_uint32, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32]
        L.or_rc_forward.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
        L.or_rc_gradients.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                      C.c_void_p, C.c_void_p, C.c_void_p]
        L._rc_ready = True
    return L


def rc_round_half(x):
    return np.array([_rc_lib().or_rc_round_half(float(v)) for v in np.ravel(x)], np.float32)


def rc_atan2(y, x):
    return _rc_lib().or_rc_atan2(y, x)


def rc_acos(z):
    return _rc_lib().or_rc_acos(z)


def rc_grid(log2_table):
    res = np.zeros(16, np.uint32)
    dense = np.zeros(16, np.uint32)
    _rc_lib().or_rc_grid(log2_table, res.ctypes.data, dense.ctypes.data)
    return res, dense


def rc_index(log2_table, level, x, y, z):
    return _rc_lib().or_rc_index(log2_table, level, x, y, z)


def rc_forward(log2_table, mlp, table, features):
    mlp = np.ascontiguousarray(mlp, np.float32)
    table = np.ascontiguousarray(table, np.float32)
    f = np.ascontiguousarray(features, np.float32).reshape(-1, 12)
    out = np.zeros((f.shape[0], 3), np.float64)
    _rc_lib().or_rc_forward(log2_table, mlp.ctypes.data, table.ctypes.data, f.ctypes.data, f.shape[0],
                            out.ctypes.data)
    return out


def rc_gradients(log2_table, mlp, table, features, targets):
    mlp = np.ascontiguousarray(mlp, np.float32)
    table = np.ascontiguousarray(table, np.float32)
    f = np.ascontiguousarray(features, np.float32).reshape(-1, 12)
    t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
    gm = np.zeros(mlp.size, np.float64)
    gt = np.zeros(table.size, np.float64)
    loss = np.zeros(1, np.float64)
    _rc_lib().or_rc_gradients(log2_table, mlp.ctypes.data, table.ctypes.data, f.ctypes.data, t.ctypes.data,
                              f.shape[0], gm.ctypes.data, gt.ctypes.data, loss.ctypes.data)
    return gm, gt, float(loss[0])
