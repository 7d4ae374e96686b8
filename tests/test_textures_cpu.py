"""Texture slots: image and environment textures and the Nishita sky
(device/nodes/nodes.hpp:443-564, 646-870; host/sky/sky_nishita.cpp).

* the bilinear repeat sampler against a float64 numpy restatement;
* the sky precompute against the reference's own sky_nishita.cpp, compiled
  from the reference checkout into oracle/_ref/ (oracle/ref.mk);
* sky radiance lookups against a numpy restatement of the texture layout,
  sun disc and horizon fade;
* image nodes through the codegen and the oracle in closed form, and the
  image decoders against encoders written here.
"""
import ctypes as C
import math
import os
import struct
import zlib

import numpy as np
import pytest

import raymond_amd.api as api
import raymond_amd.images as images
import raymond_amd.scenes as scenes
from tests import devlib
from tests.test_nodes_cpu import _in, check, emitter_scene, node
from oracle import binding

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SKY = os.path.join(ROOT, "oracle", "_ref", "libsky_ref.so")


# ---------------------------------------------------------------------------
# Sampler

def bilinear_repeat(tex, uv):
    """Texel centres at (i + 1/2) / size, repeat addressing, float64."""
    h, w, c = tex.shape
    t = tex.astype(np.float64)
    if c == 1:
        t = np.concatenate([t, t, t, np.ones_like(t)], 2)
    x = uv[:, 0].astype(np.float64) * w - 0.5
    y = uv[:, 1].astype(np.float64) * h - 0.5
    x0, y0 = np.floor(x), np.floor(y)
    fx, fy = (x - x0)[:, None], (y - y0)[:, None]
    x0, y0 = x0.astype(np.int64) % w, y0.astype(np.int64) % h
    x1, y1 = (x0 + 1) % w, (y0 + 1) % h
    return ((t[y0, x0] * (1 - fx) + t[y0, x1] * fx) * (1 - fy) + (t[y1, x0] * (1 - fx) + t[y1, x1] * fx) * fy)


def check_sampler(L=None):
    rng = np.random.default_rng(3)
    for shape in ((5, 7, 4), (3, 8, 1), (1, 1, 4), (16, 2, 4)):
        tex = rng.uniform(0, 1, shape).astype(np.float32)
        uv = np.concatenate([rng.uniform(-3, 3, (2000, 2)), rng.integers(-4, 5, (64, 2)) / np.array(shape[1::-1])])
        got = devlib.texture_sample(tex, uv, L)
        want = bilinear_repeat(tex, uv.astype(np.float32))
        assert np.abs(got - want).max() < 2e-6, (shape, np.abs(got - want).max())


def test_sampler_bilinear_repeat():
    check_sampler()


# ---------------------------------------------------------------------------
# Nishita sky: our precompute against the reference's own code

SKY_CASES = [(0.25, 100.0, 1.0, 1.0, 1.0), (0.02, 2000.0, 2.0, 0.5, 1.0), (1.2, 0.0, 1.0, 3.0, 0.3),
             (-0.05, 0.0, 1.0, 1.0, 1.0)]


def ref_sky():
    if not os.path.exists(REF_SKY):
        pytest.skip("oracle/_ref/libsky_ref.so is built from the reference checkout (oracle/ref.mk), absent here")
    L = C.CDLL(REF_SKY)
    L.SKY_nishita_skymodel_precompute_texture.argtypes = [C.c_void_p] + [C.c_int] * 5 + [C.c_float] * 5
    L.SKY_nishita_skymodel_precompute_sun.argtypes = [C.c_float] * 5 + [C.c_void_p, C.c_void_p]
    return L


@pytest.mark.parametrize("params", SKY_CASES)
def test_nishita_texture_matches_reference_code(params):
    ref = ref_sky()
    w, h = 96, 64
    want = np.zeros((h, w, 4), np.float32)
    ref.SKY_nishita_skymodel_precompute_texture(want.ctypes.data, 4, 0, h, w, h, *params)
    got = api.sky_nishita_texture(w, h, *params)
    d = np.abs(got[..., :3] - want[..., :3])
    assert np.all(d <= 1e-4 * np.abs(want[..., :3]) + 1e-9), d.max()
    assert want[..., :3].max() > 0.1


@pytest.mark.parametrize("params", SKY_CASES)
def test_nishita_sun_matches_reference_code(params):
    ref = ref_sky()
    elev, alt, air, dust, _ = params
    for size in (0.00951, 0.05):
        b0, t0 = np.zeros(3, np.float32), np.zeros(3, np.float32)
        ref.SKY_nishita_skymodel_precompute_sun(elev, size, alt, air, dust, b0.ctypes.data, t0.ctypes.data)
        b1, t1 = api.sky_nishita_sun(elev, size, alt, air, dust)
        assert np.allclose(b1, b0, rtol=1e-5, atol=0) and np.allclose(t1, t0, rtol=1e-5, atol=0), (b0, b1, t0, t1)


def test_nishita_sky_is_physical():
    """Independent of the reference: the zenith sky is bluer than the
    horizon's (Rayleigh scattering), the radiance rises toward the sun's
    azimuth at the horizon, and a thicker atmosphere dims the sun."""
    tex = api.sky_nishita_texture(64, 32, 0.3, 0.0, 1.0, 1.0, 1.0)
    rgb = np.einsum("ij,hwj->hwi", np.array([[3.240479, -1.537150, -0.498535], [-0.969256, 1.875991, 0.041556],
                                             [0.055648, -0.204043, 1.057311]]), tex[..., :3])
    zen, hor = rgb[-1].mean(0), rgb[4].mean(0)
    assert zen[2] / zen[0] > hor[2] / hor[0]
    row = tex[4, :, 1]  # longitude -pi .. pi; the sun sits at longitude 0 (the middle)
    assert row[32] > row[0] and row[31] > row[1]
    b1, _ = api.sky_nishita_sun(0.1, 0.01, 0.0, 1.0, 1.0)
    b2, _ = api.sky_nishita_sun(0.1, 0.01, 0.0, 4.0, 1.0)
    assert b2[1] < b1[1]


def sky_radiance_np(dirs, data, tex):
    """The sky node's lookup in numpy: above the horizon the sun disc (edge
    colours blended by elevation, limb darkening 1 - 0.6 (1 - sqrt(1 - f^2)))
    or the texture at (azimuth, sqrt(elevation / (pi/2))); below it the
    horizon row faded by (1 + 2.5 z)^3 down to z = -0.4; XYZ to linear sRGB."""
    out = np.zeros((dirs.shape[0], 3))
    m = np.array([[3.240479, -1.537150, -0.498535], [-0.969256, 1.875991, 0.041556], [0.055648, -0.204043, 1.057311]])
    elev_s, rot, diam, inten = data[6], data[7], data[8], data[9]
    half = diam / 2
    sun = np.array([math.cos(elev_s) * math.cos(rot + math.pi / 2), math.cos(elev_s) * math.sin(rot + math.pi / 2),
                    math.sin(elev_s)])
    for k, d in enumerate(dirs.astype(np.float64)):
        az = math.atan2(d[0], d[1])
        elev = math.pi / 2 - math.acos(d[2])
        u = (az + math.pi + rot) / (2 * math.pi)
        u = u - 1 if u > 1 else u
        if d[2] >= 0:
            ang = 2 * math.atan2(np.linalg.norm(d - sun), np.linalg.norm(d + sun))
            if ang < half and diam > 0:
                bottom, top = np.array(data[0:3]), np.array(data[3:6])
                xyz = np.zeros(3)
                if elev_s - half > 0:
                    y = (elev - elev_s) / diam + 0.5
                    xyz = (bottom * (1 - y) + top * y) * inten
                elif elev_s + half > 0:
                    y = elev / (elev_s + half)
                    xyz = (bottom * (1 - y) + top * y) * inten
                f = ang / half
                xyz = xyz * (1 - 0.6 * (1 - math.sqrt(1 - f * f)))
            else:
                xyz = bilinear_repeat(tex, np.array([[u, math.sqrt(elev / (math.pi / 2))]]))[0, :3]
        elif d[2] >= -0.4:
            xyz = bilinear_repeat(tex, np.array([[u, 1e-3]]))[0, :3] * (1 + 2.5 * d[2]) ** 3
        else:
            xyz = np.zeros(3)
        out[k] = m @ xyz
    return out


def check_sky_lookup(L=None):
    rng = np.random.default_rng(5)
    tex = api.sky_nishita_texture(64, 32, 0.2, 0.0, 1.0, 1.0, 1.0)
    bottom, top = api.sky_nishita_sun(0.2, 0.05, 0.0, 1.0, 1.0)
    data = np.r_[bottom, top, 0.2, 0.7, 0.05, 1.0].astype(np.float32)
    sun = np.array([math.cos(0.2) * math.cos(0.7 + math.pi / 2), math.cos(0.2) * math.sin(0.7 + math.pi / 2),
                    math.sin(0.2)])
    dirs = np.concatenate([rng.normal(size=(1500, 3)), sun + rng.normal(size=(300, 3)) * 0.01])
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    dirs = dirs.astype(np.float32)
    got = devlib.sky_nishita(tex, data, dirs, L)
    want = sky_radiance_np(dirs, data, tex)
    assert np.all(np.abs(got - want) <= 2e-4 * np.abs(want).max(axis=1, keepdims=True) + 1e-6), \
        np.abs(got - want).max()
    assert np.sum(np.abs(want).sum(1) > 1e3) > 10  # some directions fall in the sun disc


def test_sky_lookup_closed_form():
    check_sky_lookup()


# ---------------------------------------------------------------------------
# Image decoders

def png_bytes(a, ctype, depth=8, filters=(0, 1, 2, 3, 4)):
    """A PNG of integer array a (h, w, c) with a different filter per row."""
    h, w = a.shape[:2]
    raw = a.astype(">u2" if depth == 16 else np.uint8).reshape(h, -1).view(np.uint8)
    bpp = raw.shape[1] // w
    out = b""
    prev = np.zeros(raw.shape[1], np.int32)
    for y in range(h):
        cur = raw[y].astype(np.int32)
        f = filters[y % len(filters)]
        left = np.concatenate([np.zeros(bpp, np.int32), cur[:-bpp]])
        upleft = np.concatenate([np.zeros(bpp, np.int32), prev[:-bpp]])
        if f == 0:
            enc = cur
        elif f == 1:
            enc = cur - left
        elif f == 2:
            enc = cur - prev
        elif f == 3:
            enc = cur - (left + prev) // 2
        else:
            pa, pb, pc = np.abs(prev - upleft), np.abs(left - upleft), np.abs(left + prev - 2 * upleft)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, prev, upleft))
            enc = cur - pred
        out += bytes([f]) + (enc & 0xff).astype(np.uint8).tobytes()
        prev = cur

    def chunk(kind, data):
        return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xffffffff)

    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(out)) + chunk(b"IEND", b""))


def exr_bytes(rgba, half=False, compression=0):
    """A scanline OpenEXR of float RGBA, uncompressed (0) or ZIP (3)."""
    h, w, _ = rgba.shape

    def attr(name, typ, data):
        return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(data)) + data

    ptype = 1 if half else 2
    chl = b"".join(c.encode() + b"\0" + struct.pack("<iBBBBii", ptype, 0, 0, 0, 0, 1, 1) for c in "ABGR") + b"\0"
    hdr = (attr("channels", "chlist", chl) + attr("compression", "compression", bytes([compression])) +
           attr("dataWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1)) +
           attr("displayWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1)) +
           attr("lineOrder", "lineOrder", b"\0") + attr("pixelAspectRatio", "float", struct.pack("<f", 1)) +
           attr("screenWindowCenter", "v2f", struct.pack("<ff", 0, 0)) +
           attr("screenWindowWidth", "float", struct.pack("<f", 1)) + b"\0")
    lines = 16 if compression == 3 else 1
    nb = (h + lines - 1) // lines
    head = struct.pack("<II", 20000630, 2) + hdr
    blocks = []
    for b in range(nb):
        rows = rgba[b * lines:(b + 1) * lines]
        data = b"".join(rows[r, :, i].astype("<f2" if half else "<f4").tobytes()
                        for r in range(rows.shape[0]) for i in (3, 2, 1, 0))
        if compression == 3:
            z = np.frombuffer(data, np.uint8)
            t = np.concatenate([z[0::2], z[1::2]]).astype(np.int32)
            d = np.concatenate([t[:1], (t[1:] - t[:-1] + 128) & 0xff]).astype(np.uint8)
            packed = zlib.compress(d.tobytes())
            if len(packed) < len(data):  # a block that does not shrink is stored as it is
                data = packed
        blocks.append(struct.pack("<ii", b * lines, len(data)) + data)
    off = len(head) + 8 * nb
    offsets = []
    for blk in blocks:
        offsets.append(off)
        off += len(blk)
    return head + struct.pack("<%dQ" % nb, *offsets) + b"".join(blocks)


def test_png_decoder(tmp_path):
    rng = np.random.default_rng(7)
    for ctype, c, depth in ((2, 3, 8), (6, 4, 8), (0, 1, 8), (4, 2, 8), (2, 3, 16), (0, 1, 16)):
        a = rng.integers(0, 256 if depth == 8 else 65536, (9, 11, c))
        p = tmp_path / ("i%d_%d.png" % (ctype, depth))
        p.write_bytes(png_bytes(a, ctype, depth))
        got = images.load(str(p))
        v = a / (255.0 if depth == 8 else 65535.0)
        if c == 1:
            want = v
        else:
            want = np.ones((9, 11, 4))
            if c == 2:
                want[..., :3] = v[..., :1]
                want[..., 3] = v[..., 1]
            else:
                want[..., :c] = v
        assert got.shape == want.shape and np.abs(got - want).max() < 1e-6, (ctype, depth)


@pytest.mark.parametrize("half,compression", [(False, 0), (True, 0), (False, 3), (True, 3)])
def test_exr_decoder(tmp_path, half, compression):
    rng = np.random.default_rng(8)
    a = rng.uniform(-2, 40, (37, 13, 4)).astype(np.float16 if half else np.float32).astype(np.float32)
    p = tmp_path / "i.exr"
    p.write_bytes(exr_bytes(a, half, compression))
    assert np.array_equal(images.load(str(p)), a)


# ---------------------------------------------------------------------------
# Image nodes through the codegen and the oracle

def texture_value(tmp_path, node_type, params, vector, img, name="t.png"):
    """Emission = the image node's colour at a constant vector."""
    p = tmp_path / name
    if name.endswith(".png"):
        p.write_bytes(png_bytes(img, 6 if img.shape[2] == 4 else 0))
    else:
        p.write_bytes(exr_bytes(img))
    const = node("COMBXYZ", {"X": _in("VALUE", vector[0]), "Y": _in("VALUE", vector[1]),
                             "Z": _in("VALUE", vector[2])})
    tex = node(node_type, {"Vector": _in("VECTOR", [0, 0, 0], link=("V", "Vector"))},
               dict(params, filepath=str(p)))
    d = emitter_scene({"V": const, "N": tex}, ("N", "Color"))
    orc = binding.OracleScene(d)
    orc.build_envmap()
    img_out, _ = orc.render(4, 4, 1, frames=1)
    return img_out[..., :3].astype(np.float64)


def srgb_to_linear(c):
    return np.where(c < 0.04045, np.maximum(c, 0) / 12.92, ((c + 0.055) / 1.055) ** 2.4)


@pytest.mark.parametrize("colorspace", ["sRGB", "Linear", "Non-Color"])
def test_image_texture_flat(tmp_path, colorspace):
    rng = np.random.default_rng(9)
    img = rng.integers(0, 256, (6, 5, 4))
    uv = (0.37, 0.81, 0.0)
    got = texture_value(tmp_path, "TEX_IMAGE", {"interpolation": "Linear", "projection": "FLAT",
                                                "extension": "REPEAT", "colorspace": colorspace,
                                                "alpha": "STRAIGHT"}, uv, img)
    # flat projection flips v (nodes.hpp:505-507)
    c = bilinear_repeat((img / 255.0).astype(np.float32), np.array([[uv[0], 1 - uv[1]]], np.float32))[0, :3]
    check(got, srgb_to_linear(c) if colorspace == "sRGB" else c, rel=1e-5)


def test_image_texture_grey_png(tmp_path):
    img = np.arange(12).reshape(3, 4, 1) * 20
    got = texture_value(tmp_path, "TEX_IMAGE", {"interpolation": "Linear", "projection": "FLAT",
                                                "colorspace": "Linear"}, (0.6, 0.2, 0.0), img)
    c = bilinear_repeat((img / 255.0).astype(np.float32), np.array([[0.6, 0.8]], np.float32))[0, :3]
    check(got, c, rel=1e-5)


def test_environment_texture_equirect(tmp_path):
    rng = np.random.default_rng(10)
    img = rng.uniform(0, 3, (8, 16, 4)).astype(np.float32)
    d = np.array([0.3, -0.5, 0.6])
    got = texture_value(tmp_path, "TEX_ENVIRONMENT", {"interpolation": "Linear", "projection": "EQUIRECTANGULAR",
                                                      "colorspace": "Linear"}, d, img, name="e.exr")
    u = (math.atan2(d[0], d[1]) + math.pi / 2) / (2 * math.pi)
    v = math.acos(d[2] / np.linalg.norm(d)) / math.pi
    check(got, bilinear_repeat(img, np.array([[u, v]], np.float32))[0, :3], rel=1e-5)


def test_texture_slots_and_missing_files(tmp_path):
    """One slot per distinct file in first-use order, a sky slot per sky
    node; a file that cannot be read becomes one black texel."""
    tex = lambda f: node("TEX_IMAGE", {}, {"filepath": f, "interpolation": "Linear", "projection": "FLAT",
                                            "colorspace": "Linear"})
    graph = {"A": tex("a.png"), "B": tex("b.png"), "C": tex("a.png"),
             "M": node("MIX_RGB", {"Color1": _in("RGBA", [0, 0, 0, 1], link=("A", "Color")),
                                   "Color2": _in("RGBA", [0, 0, 0, 1], link=("B", "Color")),
                                   "Fac": _in("VALUE", 0.5, link=("C", "Alpha"))}, {"blend_type": "MIX"})}
    d = emitter_scene(graph, ("M", "Color"))
    d.light_materials = [{"name": "sky", "nodes": {
        "S": node("TEX_SKY", {}, {"sky_type": "NISHITA", "sun_elevation": 0.3, "sun_rotation": 0.5,
                                  "sun_disc": True, "sun_size": 0.01, "sun_intensity": 1.0, "altitude": 10.0,
                                  "air_density": 1.0, "dust_density": 2.0, "ozone_density": 1.0}),
        "Background": node("BACKGROUND", {"Color": _in("RGBA", [0, 0, 0, 1], link=("S", "Color")),
                                          "Strength": _in("VALUE", 1.0)}),
        "World Output": scenes.output_material(("Background", "Background")) | {"type": "OUTPUT_WORLD"}}}]
    slots = api.codegen_textures(d.materials_json())
    assert [s["kind"] for s in slots] == ["image", "image", "nishita"]
    assert [os.path.basename(s["path"]) for s in slots[:2]] == ["a.png", "b.png"]
    assert slots[2]["dust_density"] == 2.0 and abs(slots[2]["sun_elevation"] - 0.3) < 1e-7
    src = d.material_source()
    b, t = api.sky_nishita_sun(0.3, 0.01, 10.0, 1.0, 2.0)
    for k, v in enumerate(list(b) + list(t) + [0.3, 0.5, 0.01, 1.0]):
        line = [l for l in src.splitlines() if ".data[%d] = " % k in l][0]
        assert abs(float(line.split("=")[1].strip(" f;")) - v) <= 1e-6 * max(abs(v), 1e-30), (k, line, v)
    d.texture_dir = str(tmp_path)
    with pytest.warns(UserWarning, match="black texel"):
        texs = d.resolve_textures()
    assert [t.shape for t in texs[:2]] == [(1, 1, 1), (1, 1, 1)] and texs[2].shape == (512, 512, 4)


def test_hosek_wilkie_maps_to_nishita():
    """The reference renders a Hosek-Wilkie sky as Nishita with a fixed
    atmosphere, no sun disc and radiance x 0.08 (MaterialBuilder.swift:335-352)."""
    g = {"S": node("TEX_SKY", {}, {"sky_type": "HOSEK_WILKIE", "sun_direction": [0.0, 0.6, 0.8]}),
         "Background": node("BACKGROUND", {"Color": _in("RGBA", [0, 0, 0, 1], link=("S", "Color")),
                                           "Strength": _in("VALUE", 1.0)}),
         "World Output": {"type": "OUTPUT_WORLD", "parameters": {},
                          "inputs": {"Surface": _in("SHADER", link=("Background", "Background"))}}}
    d = emitter_scene({"N": node("VALUE", {}, {"value": 0.5})}, ("N", "Value"))
    d.light_materials = [{"name": "sky", "nodes": g}]
    slot = api.codegen_textures(d.materials_json())[0]
    assert slot["kind"] == "nishita" and slot["air_density"] == 2.0 and slot["dust_density"] == 0.5
    assert abs(slot["sun_elevation"] - math.asin(0.8)) < 1e-6
    src = d.material_source()
    assert ".scale = 0.0799999982f;" in src and ".data[8] = -1.0f;" in src
    with pytest.raises(api.RmError, match="unsupported sky type"):
        g["S"]["parameters"]["sky_type"] = "PREETHAM"
        d.material_source()


# ---------------------------------------------------------------------------
# Planted mistakes

PLANTS = [
    ("texel centre without the half offset", "texture.h", "const float fx = uv.x * (float)t.width - 0.5f;",
     "const float fx = uv.x * (float)t.width;", check_sampler),
    ("clamped instead of repeated", "texture.h", "const int x1 = wrap_index(x0 + 1, t.width)",
     "const int x1 = min_(x0 + 1, t.width - 1)", check_sampler),
    ("sky row from elevation without the square root", "nodes.h",
     "const float y = sqrt_(dir_elevation / (0.5f * kPi));", "const float y = dir_elevation / (0.5f * kPi);",
     check_sky_lookup),
    ("sun azimuth without the quarter turn", "nodes.h",
     "geographical_to_direction(sun_elevation, sun_rotation + 0.5f * kPi)",
     "geographical_to_direction(sun_elevation, sun_rotation)", check_sky_lookup),
]


@pytest.mark.parametrize("name,header,old,new,check_fn", PLANTS, ids=[p[0] for p in PLANTS])
def test_planted_mistakes_fail(name, header, old, new, check_fn):
    L = devlib.lib(plant=(header, old, new))
    with pytest.raises(AssertionError):
        check_fn(L)
