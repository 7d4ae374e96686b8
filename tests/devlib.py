"""Host build of the shared device library (tests/native/devlib_shim.cpp) for
CPU property tests.  Compiled with the same strict-IEEE flags as the oracle."""
import ctypes as C
import hashlib
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
DEVICE = os.path.join(ROOT, "raymond_amd", "csrc", "device")
SRC = os.path.join(HERE, "native", "devlib_shim.cpp")
BUILD = os.path.join(HERE, "native", "_build")
FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-shared", "-Wno-unused-function"]

_LIBS = {}


def lib(plant=None):
    """The shim over the device headers.  plant = (header, old, new): a build
    over a copy of the headers with `old` replaced by `new` in `header` (a
    mistake planted to show that a check catches it); `old` must occur."""
    key = plant
    if key in _LIBS:
        return _LIBS[key]
    texts = {}
    for f in sorted(os.listdir(DEVICE)):
        if f.endswith(".h"):
            with open(os.path.join(DEVICE, f)) as fh:
                texts[f] = fh.read()
    if plant is not None:
        header, old, new = plant
        assert old in texts[header], "planted mistake does not apply: %r" % (plant,)
        texts[header] = texts[header].replace(old, new)
    h = hashlib.sha1()
    with open(SRC, "rb") as fh:
        h.update(fh.read())
    for f, t in texts.items():
        h.update(f.encode())
        h.update(t.encode())
    os.makedirs(BUILD, exist_ok=True)
    tag = h.hexdigest()[:12]
    so = os.path.join(BUILD, "devlib_%s.so" % tag)
    if not os.path.exists(so):
        inc = DEVICE
        if plant is not None:
            inc = os.path.join(BUILD, "planted_%s" % tag)
            os.makedirs(inc, exist_ok=True)
            for f, t in texts.items():
                with open(os.path.join(inc, f), "w") as fh:
                    fh.write(t)
        tmp = so + ".tmp%d" % os.getpid()
        subprocess.run(["g++"] + FLAGS + ["-I", inc, "-o", tmp, SRC], check=True)
        os.replace(tmp, so)
    L = C.CDLL(so)
    for name, args in (("shim_hash_uint", [C.c_uint32]), ("shim_hash_uint2", [C.c_uint32] * 2),
                       ("shim_hash_uint3", [C.c_uint32] * 3), ("shim_hash_uint4", [C.c_uint32] * 4)):
        getattr(L, name).restype = C.c_uint32
        getattr(L, name).argtypes = args
    L.shim_shading_context.restype = C.c_uint32
    L.shim_fractal4.restype = C.c_float
    L.shim_fractal4.argtypes = [C.c_void_p, C.c_float, C.c_float]
    _LIBS[key] = L
    return L


def f32(a):
    return np.ascontiguousarray(a, np.float32)


def u32(a):
    return np.ascontiguousarray(a, np.uint32)


def tea32(a, b):
    a, b = u32(a), u32(b)
    out = np.zeros(a.size, np.uint32)
    lib().shim_tea32(C.c_void_p(a.ctypes.data), C.c_void_p(b.ctypes.data), C.c_void_p(out.ctypes.data), a.size)
    return out


def tea_float(a, b):
    a, b = u32(a), u32(b)
    out = np.zeros(a.size, np.float32)
    lib().shim_tea_float(C.c_void_p(a.ctypes.data), C.c_void_p(b.ctypes.data), C.c_void_p(out.ctypes.data), a.size)
    return out


def prng_sequence(a, b, count):
    out = np.zeros(count, np.float32)
    lib().shim_prng_sequence(C.c_uint32(a), C.c_uint32(b), C.c_void_p(out.ctypes.data), count)
    return out


ELEM = {"sin": 0, "cos": 1, "atan2": 2, "acos": 3, "log": 4, "exp": 5, "pow": 6, "tan": 7, "asin": 8, "atan": 9}


def elem(name, x, y=None):
    x = f32(x)
    y = f32(np.zeros_like(x) if y is None else y)
    out = np.zeros(x.size, np.float32)
    lib().shim_elem(ELEM[name], C.c_void_p(x.ctypes.data), C.c_void_p(y.ctypes.data), C.c_void_p(out.ctypes.data),
                    x.size)
    return out


def warp(fn, x, out_dim):
    x = f32(x)
    n = x.shape[0]
    out = np.zeros((n, out_dim), np.float32)
    lib().shim_warp(fn, C.c_void_p(x.ctypes.data), C.c_void_p(out.ctypes.data), n)
    return out


def bsdf_eval(params, lobe, wo, wi):
    p, wo, wi = f32(params), f32(wo), f32(wi)
    n = wo.shape[0]
    val = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    lib().shim_bsdf_eval(C.c_void_p(p.ctypes.data), lobe, C.c_void_p(wo.ctypes.data), C.c_void_p(wi.ctypes.data),
                         C.c_void_p(val.ctypes.data), C.c_void_p(pdf.ctypes.data), n)
    return val, pdf


def bsdf_sample(params, lobe, wo, rnd):
    p, wo, rnd = f32(params), f32(wo), f32(rnd)
    n = wo.shape[0]
    wi = np.zeros((n, 3), np.float32)
    w = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    flags = np.zeros(n, np.uint32)
    lib().shim_bsdf_sample(C.c_void_p(p.ctypes.data), lobe, C.c_void_p(wo.ctypes.data), C.c_void_p(rnd.ctypes.data),
                           C.c_void_p(wi.ctypes.data), C.c_void_p(w.ctypes.data), C.c_void_p(pdf.ctypes.data),
                           C.c_void_p(flags.ctypes.data), n)
    return wi, w, pdf, flags


def intersect(rays, tris):
    rays, tris = f32(rays), f32(tris)
    n = rays.shape[0]
    hit = np.zeros(n, np.uint32)
    t = np.zeros(n, np.float32)
    b0 = np.zeros(n, np.float32)
    b1 = np.zeros(n, np.float32)
    lib().shim_intersect(C.c_void_p(rays.ctypes.data), C.c_void_p(tris.ctypes.data), n, C.c_void_p(hit.ctypes.data),
                         C.c_void_p(t.ctypes.data), C.c_void_p(b0.ctypes.data), C.c_void_p(b1.ctypes.data))
    return hit.astype(bool), t, b0, b1


def world_sample(pdfs, mip, res, uv):
    pdfs, mip, uv = f32(pdfs), f32(mip), f32(uv)
    n = uv.shape[0]
    d = np.zeros((n, 3), np.float32)
    p = np.zeros(n, np.float32)
    lib().shim_world_sample(C.c_void_p(pdfs.ctypes.data), C.c_void_p(mip.ctypes.data), res,
                            C.c_void_p(uv.ctypes.data), C.c_void_p(d.ctypes.data), C.c_void_p(p.ctypes.data), n)
    return d, p


def world_pdf(pdfs, mip, res, dirs):
    pdfs, mip, dirs = f32(pdfs), f32(mip), f32(dirs)
    n = dirs.shape[0]
    p = np.zeros(n, np.float32)
    lib().shim_world_pdf(C.c_void_p(pdfs.ctypes.data), C.c_void_p(mip.ctypes.data), res,
                         C.c_void_p(dirs.ctypes.data), C.c_void_p(p.ctypes.data), n)
    return p


def noise(p, octaves=2.0, roughness=0.5):
    p = f32(p)
    n = p.shape[0]
    per = np.zeros(n, np.float32)
    fbm = np.zeros(n, np.float32)
    lib().shim_noise(C.c_void_p(p.ctypes.data), n, C.c_float(octaves), C.c_float(roughness),
                     C.c_void_p(per.ctypes.data), C.c_void_p(fbm.ctypes.data))
    return per, fbm


def hsv_roundtrip(rgb):
    rgb = f32(rgb)
    out = np.zeros_like(rgb)
    lib().shim_hsv(C.c_void_p(rgb.ctypes.data), C.c_void_p(out.ctypes.data), rgb.shape[0])
    return out


def _p(a):
    return C.c_void_p(a.ctypes.data)


def lobe_eval(kind, params, wo, wi, L=None):
    """One lobe (1 Specular, 2 Transmission, 3 Clearcoat) from explicit
    parameters: value (n, 3) and pdf (n,) of wo -> wi in the shading frame."""
    L = L or lib()
    p, wo, wi = f32(params), f32(wo), f32(wi)
    n = wo.shape[0]
    val = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    L.shim_lobe_eval(kind, _p(p), _p(wo), _p(wi), _p(val), _p(pdf), n)
    return val, pdf


def lobe_sample(kind, params, wo, rnd, L=None):
    L = L or lib()
    p, wo, rnd = f32(params), f32(wo), f32(rnd)
    n = wo.shape[0]
    wi = np.zeros((n, 3), np.float32)
    w = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    L.shim_lobe_sample(kind, _p(p), _p(wo), _p(rnd), _p(wi), _p(w), _p(pdf), n)
    return wi, w, pdf


PRINCIPLED_IN = ["base_r", "base_g", "base_b", "metallic", "roughness", "specular", "specular_tint", "sheen",
                 "sheen_tint", "clearcoat", "clearcoat_roughness", "transmission", "ior", "anisotropic",
                 "transmission_roughness", "alpha", "emission_r", "emission_g", "emission_b", "emission_strength"]


def principled(**kw):
    """BsdfPrincipled::compute; returns a dict of the UberShader fields."""
    L = kw.pop("L", None) or lib()
    d = dict(base_r=0.8, base_g=0.8, base_b=0.8, metallic=0.0, roughness=0.5, specular=0.5, specular_tint=0.0,
             sheen=0.0, sheen_tint=0.5, clearcoat=0.0, clearcoat_roughness=0.03, transmission=0.0, ior=1.45,
             anisotropic=0.0, transmission_roughness=0.0, alpha=1.0, emission_r=0.0, emission_g=0.0, emission_b=0.0,
             emission_strength=0.0)
    for k, v in kw.items():
        assert k in d, k
        d[k] = v
    x = f32([d[k] for k in PRINCIPLED_IN])
    o = np.zeros(34, np.float32)
    L.shim_principled(_p(x), _p(o))
    o = o.astype(np.float64)
    return dict(lobe_prob=o[0:4], diffuse_weight=o[4:7], sheen_weight=o[7:10], diffuse_roughness=o[10],
                alpha_x=o[11], alpha_y=o[12], cspec0=o[13:16], specular_ior=o[16], specular_weight=o[17],
                reflection_alpha=o[18], transmission_alpha=o[19], transmission_base=o[20:23],
                transmission_cspec0=o[23:26], transmission_ior=o[26], transmission_weight=o[27],
                clearcoat_alpha=o[28], clearcoat_weight=o[29], alpha=o[30], emission=o[31:34])


def shading_context(verts, normals, uvs, faces, mats, m, nm, bounds, prim, b0, b1, L=None):
    """ShadingContext::build; returns (dict of vectors, shader index)."""
    L = L or lib()
    verts, normals, uvs = f32(verts), f32(normals), f32(uvs)
    faces = u32(faces)
    mats = np.ascontiguousarray(mats, np.uint16)
    m = f32(np.asarray(m, np.float64).T.reshape(-1))      # column-major
    nm = f32(np.asarray(nm, np.float64).T.reshape(-1))
    bounds = f32(bounds)
    out = np.zeros(24, np.float32)
    idx = L.shim_shading_context(_p(verts), _p(normals), _p(uvs), _p(faces), _p(mats), _p(m), _p(nm), _p(bounds),
                                 C.c_uint32(prim), C.c_float(b0), C.c_float(b1), _p(out))
    names = ["uv", "position", "generated", "object", "normal", "true_normal", "tu", "tv"]
    return {k: out[3 * i:3 * i + 3].astype(np.float64) for i, k in enumerate(names)}, int(idx)


def shape_light(verts, faces, cdf, area, m, nm, origin, seeds_a, seeds_b, L=None):
    """ShapeLight::sample from `origin`: (n, 7) = direction (3), distance,
    pdf, weight, pdf recomputed from the sample's shading context."""
    L = L or lib()
    verts = f32(verts)
    normals = f32(np.zeros_like(verts))
    uvs = f32(np.zeros((verts.shape[0], 2)))
    faces = u32(faces)
    cdf = f32(cdf)
    m = f32(np.asarray(m, np.float64).T.reshape(-1))
    nm = f32(np.asarray(nm, np.float64).T.reshape(-1))
    origin = f32(origin)
    sa, sb = u32(seeds_a), u32(seeds_b)
    n = sa.size
    out = np.zeros((n, 7), np.float32)
    L.shim_shape_light(_p(verts), _p(normals), _p(uvs), _p(faces), faces.shape[0], _p(cdf), C.c_float(area), _p(m),
                       _p(nm), _p(origin), _p(sa), _p(sb), n, _p(out))
    return out


def perlin(dim, p, L=None):
    """Signed, unscaled Perlin noise of dimension 1-4 at points p (n, dim)."""
    L = L or lib()
    p = np.asarray(p, np.float32).reshape(-1, dim)
    q = np.zeros((p.shape[0], 4), np.float32)
    q[:, :dim] = p
    out = np.zeros(p.shape[0], np.float32)
    L.shim_perlin(dim, _p(q), _p(out), p.shape[0])
    return out


def fractal4(p, octaves, roughness, L=None):
    L = L or lib()
    q = f32(p)
    return float(L.shim_fractal4(_p(q), octaves, roughness))


def ramp4(pos, colors, fac, L=None):
    L = L or lib()
    pos, colors, fac = f32(pos), f32(colors), f32(fac)
    out = np.zeros((fac.size, 4), np.float32)
    L.shim_ramp4(_p(pos), _p(colors), _p(fac), _p(out), fac.size)
    return out


MATH_OPS = ["ADD", "SUBTRACT", "MULTIPLY", "DIVIDE", "MULTIPLY_ADD", "POWER", "MINIMUM", "MAXIMUM", "TANGENT",
            "LESS_THAN", "GREATER_THAN", "MODULO"]


def math_node(op, a, b, c=0.0, clamp=False, L=None):
    L = L or lib()
    a = f32(np.atleast_1d(a))
    b = f32(np.broadcast_to(b, a.shape))
    c = f32(np.broadcast_to(c, a.shape))
    out = np.zeros(a.size, np.float32)
    L.shim_math(MATH_OPS.index(op), int(clamp), _p(a), _p(b), _p(c), _p(out), a.size)
    return out


BLENDS = ["MIX", "ADD", "MULTIPLY", "SCREEN", "OVERLAY", "SUB", "COLOR", "LIGHTEN", "DARKEN", "VALUE"]


def mix_node(blend, c1, c2, fac, clamp=False, L=None):
    L = L or lib()
    c1, c2 = f32(c1).reshape(-1, 4), f32(c2).reshape(-1, 4)
    fac = f32(np.broadcast_to(fac, (c1.shape[0],)))
    out = np.zeros_like(c1)
    L.shim_mix(BLENDS.index(blend), int(clamp), _p(c1), _p(c2), _p(fac), _p(out), c1.shape[0])
    return out


def mix_shader(fac, e0, e1, rnd, L=None):
    L = L or lib()
    e0, e1, rnd = f32(e0), f32(e1), f32(rnd)
    out = np.zeros((rnd.size, 5), np.float32)
    L.shim_mix_shader(C.c_float(fac), _p(e0), _p(e1), _p(rnd), _p(out), rnd.size)
    return out


def texture_sample(pixels, uv, L=None):
    """Bilinear repeat lookups of an (h, w, c) texture at uv (n, 2)."""
    L = L or lib()
    px = f32(pixels)
    uv = f32(uv).reshape(-1, 2)
    out = np.zeros((uv.shape[0], 4), np.float32)
    L.shim_texture_sample(_p(px), px.shape[1], px.shape[0], px.shape[2], _p(uv), _p(out), uv.shape[0])
    return out


def sky_nishita(pixels, data, dirs, L=None):
    L = L or lib()
    px, data, dirs = f32(pixels), f32(data), f32(dirs).reshape(-1, 3)
    out = np.zeros((dirs.shape[0], 3), np.float32)
    L.shim_sky_nishita(_p(px), px.shape[1], px.shape[0], _p(data), _p(dirs), _p(out), dirs.shape[0])
    return out


# lens camera (camera.h): settings = (stop_index, relative_stop, blades,
# sensor_scale, camera_scale, focus, spectral)
def lens_settings(stop_index=0, relative_stop=1.0, blades=7, sensor_scale=1.0, camera_scale=1.0, focus=0.0,
                  spectral=0):
    return f32([stop_index, relative_stop, blades, sensor_scale, camera_scale, focus, spectral])


def lens_aperture(pos, relative_stop, blades, L=None):
    L = L or lib()
    pos = f32(pos).reshape(-1, 2)
    out = np.zeros(pos.shape[0], np.uint32)
    L.shim_lens_aperture(_p(pos), pos.shape[0], C.c_float(relative_stop), blades, _p(out))
    return out.astype(bool)


def lens_refract(d, n, eta, L=None):
    L = L or lib()
    d, n = f32(d).reshape(-1, 3), f32(n).reshape(-1, 3)
    eta = f32(np.broadcast_to(eta, (d.shape[0],)))
    out = np.zeros_like(d)
    ok = np.zeros(d.shape[0], np.uint32)
    L.shim_lens_refract(_p(d), _p(n), _p(eta), d.shape[0], _p(out), _p(ok))
    return out, ok.astype(bool)


def lens_ior(ior_d, abbe, lam, L=None):
    L = L or lib()
    lam = f32(lam).ravel()
    out = np.zeros_like(lam)
    L.shim_lens_ior(C.c_float(ior_d), C.c_float(abbe), _p(lam), _p(out), lam.size)
    return out


def lens_trace(surfaces, settings, o, d, lam, L=None):
    L = L or lib()
    s, st = f32(surfaces).reshape(-1, 5), f32(settings)
    o, d = f32(o).reshape(-1, 3), f32(d).reshape(-1, 3)
    lam = f32(np.broadcast_to(lam, (o.shape[0],)))
    oo, dd = np.zeros_like(o), np.zeros_like(d)
    ok = np.zeros(o.shape[0], np.uint32)
    L.shim_lens_trace(_p(s), s.shape[0], _p(st), _p(o), _p(d), _p(lam), _p(oo), _p(dd), _p(ok), o.shape[0])
    return oo, dd, ok.astype(bool)


def lens_camera_ray(surfaces, settings, W, H, x, y, seed, n, L=None):
    L = L or lib()
    s, st = f32(surfaces).reshape(-1, 5), f32(settings)
    o, d, w = (np.zeros((n, 3), np.float32) for _ in range(3))
    ok = np.zeros(n, np.uint32)
    L.shim_lens_camera_ray(_p(s), s.shape[0], _p(st), W, H, x, y, seed, _p(o), _p(d), _p(w), _p(ok), n)
    return o, d, w, ok.astype(bool)


def spectral(nm, L=None):
    L = L or lib()
    nm = f32(nm).ravel()
    xyz, rgb = np.zeros((nm.size, 3), np.float32), np.zeros((nm.size, 3), np.float32)
    norm = np.zeros(3, np.float32)
    L.shim_spectral(_p(nm), nm.size, _p(xyz), _p(rgb), _p(norm))
    return xyz, rgb, norm


def xyz_to_rgb(xyz, L=None):
    L = L or lib()
    xyz = f32(xyz).reshape(-1, 3)
    out = np.zeros_like(xyz)
    L.shim_xyz_to_rgb(_p(xyz), _p(out), xyz.shape[0])
    return out


def two_level_brute(vertices, faces, ranges, transforms, rays, any_hit=False, L=None):
    """Closest (or any) hits by exhaustion over instances (vertex offset,
    face offset, face count) with column-major object-to-world transforms."""
    L = L or lib()
    v, r, xf = f32(vertices).reshape(-1, 3), f32(rays).reshape(-1, 8), f32(transforms).reshape(-1, 16)
    fa = np.ascontiguousarray(faces, np.uint32).reshape(-1, 3)
    rg = np.ascontiguousarray(ranges, np.uint32).reshape(-1, 3)
    n = r.shape[0]
    t = np.zeros(n, np.float32)
    inst = np.zeros(n, np.uint32)
    prim = np.zeros(n, np.uint32)
    bary = np.zeros((n, 2), np.float32)
    L.shim_two_level_brute(_p(v), _p(fa), _p(rg), _p(xf), rg.shape[0], _p(r), n, int(bool(any_hit)), _p(t), _p(inst),
                           _p(prim), _p(bary))
    return t, inst, prim, bary
