"""Synthetic scenes for the benchmark configurations, as scene descriptions for
the C ABI (include/rm.h) plus node-graph materials in the Blender exporter's
JSON schema (scripts/blender_exporter/raymond_blender/materials.py:300-560).

Entity, material and light order is fixed by construction (the reference's
order follows Swift dictionary iteration, EntityBuilder.swift:50-52,85, and
is not reproducible).
"""
import ctypes as C
import json
import os
import warnings
import math

import numpy as np

from . import api

# ---------------------------------------------------------------------------
# Node-graph helpers (exporter schema)


def _in(t, value=None, link=None):
    d = {"type": t}
    if value is not None:
        d["value"] = value
    if link is not None:
        d["links"] = [[link[0], link[1]]]
    return d


PRINCIPLED_DEFAULTS = {
    "Base Color": ("RGBA", [0.8, 0.8, 0.8, 1.0]), "Subsurface": ("VALUE", 0.0),
    "Subsurface Radius": ("VECTOR", [1.0, 0.2, 0.1]), "Subsurface Color": ("RGBA", [0.8, 0.8, 0.8, 1.0]),
    "Subsurface IOR": ("VALUE", 1.4), "Subsurface Anisotropy": ("VALUE", 0.0), "Metallic": ("VALUE", 0.0),
    "Specular": ("VALUE", 0.5), "Specular Tint": ("VALUE", 0.0), "Roughness": ("VALUE", 0.5),
    "Anisotropic": ("VALUE", 0.0), "Anisotropic Rotation": ("VALUE", 0.0), "Sheen": ("VALUE", 0.0),
    "Sheen Tint": ("VALUE", 0.5), "Clearcoat": ("VALUE", 0.0), "Clearcoat Roughness": ("VALUE", 0.03),
    "IOR": ("VALUE", 1.45), "Transmission": ("VALUE", 0.0), "Transmission Roughness": ("VALUE", 0.0),
    "Emission": ("RGBA", [0.0, 0.0, 0.0, 1.0]), "Emission Strength": ("VALUE", 1.0), "Alpha": ("VALUE", 1.0),
    "Normal": ("VECTOR", [0.0, 0.0, 0.0]), "Clearcoat Normal": ("VECTOR", [0.0, 0.0, 0.0]),
    "Tangent": ("VECTOR", [0.0, 0.0, 0.0]), "Weight": ("VALUE", 0.0),
}


def principled_node(**overrides):
    inputs = {}
    for k, (t, v) in PRINCIPLED_DEFAULTS.items():
        inputs[k] = _in(t, v)
    for k, v in overrides.items():
        name = k.replace("_", " ")
        t = PRINCIPLED_DEFAULTS[name][0]
        if isinstance(v, tuple) and v and v[0] == "link":
            inputs[name] = _in(t, PRINCIPLED_DEFAULTS[name][1], link=(v[1], v[2]))
        else:
            inputs[name] = _in(t, v)
    return {"type": "BSDF_PRINCIPLED", "parameters": {"distribution": "GGX", "subsurface_method": "RANDOM_WALK"},
            "inputs": inputs}


def output_material(link):
    return {"type": "OUTPUT_MATERIAL", "parameters": {},
            "inputs": {"Surface": _in("SHADER", link=link), "Volume": _in("SHADER"),
                       "Displacement": _in("VECTOR", [0.0, 0.0, 0.0]), "Thickness": _in("VALUE", 0.0)}}


def diffuse_material(name, rgb, roughness=0.0):
    return {"name": name, "nodes": {
        "Diffuse BSDF": {"type": "BSDF_DIFFUSE", "parameters": {},
                         "inputs": {"Color": _in("RGBA", list(rgb) + [1.0]), "Roughness": _in("VALUE", roughness),
                                    "Normal": _in("VECTOR", [0.0, 0.0, 0.0]), "Weight": _in("VALUE", 0.0)}},
        "Material Output": output_material(("Diffuse BSDF", "BSDF")),
    }}


def principled_material(name, **kw):
    return {"name": name, "nodes": {
        "Principled BSDF": principled_node(**kw),
        "Material Output": output_material(("Principled BSDF", "BSDF")),
    }}


def emission_material(name, rgb, strength):
    return {"name": name, "nodes": {
        "Emission": {"type": "EMISSION", "parameters": {},
                     "inputs": {"Color": _in("RGBA", list(rgb) + [1.0]), "Strength": _in("VALUE", strength),
                                "Weight": _in("VALUE", 0.0)}},
        "Material Output": output_material(("Emission", "Emission")),
    }}


def background_world(name, rgb, strength=1.0):
    return {"name": name, "nodes": {
        "Background": {"type": "BACKGROUND", "parameters": {},
                       "inputs": {"Color": _in("RGBA", list(rgb) + [1.0]), "Strength": _in("VALUE", strength),
                                  "Weight": _in("VALUE", 0.0)}},
        "World Output": {"type": "OUTPUT_WORLD", "parameters": {},
                         "inputs": {"Surface": _in("SHADER", link=("Background", "Background")),
                                    "Volume": _in("SHADER"), "Thickness": _in("VALUE", 0.0)}},
    }}


SKY_PARAMS = dict(sun_direction=[0.35, -0.45, 0.82], sun_strength=1000.0, sun_sharpness=2000.0,
                  zenith=[0.15, 0.3, 0.8, 1.0], horizon=[0.7, 0.8, 0.95, 1.0], ground=[0.25, 0.22, 0.2, 1.0])


def sky_world(name="sky", p=SKY_PARAMS):
    return {"name": name, "nodes": {
        "Sky": {"type": "TEX_SKY_ANALYTIC", "parameters": {},
                "inputs": {"Vector": _in("VECTOR", [0.0, 0.0, 0.0]),
                           "Sun Direction": _in("VECTOR", p["sun_direction"]),
                           "Sun Strength": _in("VALUE", p["sun_strength"]),
                           "Sun Sharpness": _in("VALUE", p["sun_sharpness"]),
                           "Zenith Color": _in("RGBA", p["zenith"]), "Horizon Color": _in("RGBA", p["horizon"]),
                           "Ground Color": _in("RGBA", p["ground"])}},
        "Background": {"type": "BACKGROUND", "parameters": {},
                       "inputs": {"Color": _in("RGBA", [0.0, 0.0, 0.0, 1.0], link=("Sky", "Color")),
                                  "Strength": _in("VALUE", 1.0), "Weight": _in("VALUE", 0.0)}},
        "World Output": {"type": "OUTPUT_WORLD", "parameters": {},
                         "inputs": {"Surface": _in("SHADER", link=("Background", "Background")),
                                    "Volume": _in("SHADER"), "Thickness": _in("VALUE", 0.0)}},
    }}


# ---------------------------------------------------------------------------
# Scene description


def translate(t):
    m = np.eye(4, dtype=np.float64)
    m[:3, 3] = t
    return m


def rotate_z(deg):
    a = math.radians(deg)
    m = np.eye(4)
    m[0, 0], m[0, 1], m[1, 0], m[1, 1] = math.cos(a), -math.sin(a), math.sin(a), math.cos(a)
    return m


def scale(s):
    m = np.eye(4)
    m[0, 0], m[1, 1], m[2, 2] = s
    return m


def look_at(eye, target, up=(0.0, 0.0, 1.0)):
    """Camera-to-world matrix; the camera looks along its -z axis with +y up."""
    eye, target, up = (np.asarray(v, np.float64) for v in (eye, target, up))
    back = eye - target
    back /= np.linalg.norm(back)
    right = np.cross(up, back)
    right /= np.linalg.norm(right)
    upv = np.cross(back, right)
    m = np.eye(4)
    m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = right, upv, back, eye
    return m


def col_major(m):
    return [float(x) for x in np.asarray(m, np.float32).T.reshape(-1)]


class Shape:
    """Triangle mesh in object space; faces index its own vertices."""

    def __init__(self, vertices, normals, faces, materials, texcoords=None):
        self.vertices = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        self.normals = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        self.faces = np.ascontiguousarray(faces, np.uint32).reshape(-1, 3)
        self.materials = np.ascontiguousarray(np.broadcast_to(np.asarray(materials, np.uint16),
                                                              (self.faces.shape[0],)))
        if texcoords is None:
            texcoords = np.zeros((self.vertices.shape[0], 2), np.float32)
        self.texcoords = np.ascontiguousarray(texcoords, np.float32).reshape(-1, 2)
        # object-space bounds of the "Generated" coordinates; None: the exact
        # bounds of the vertices (a loaded PLY brings its reader's bounds)
        self.bounds = None


class SceneDescription:
    """Host-side scene: shapes, entities, lights, camera and materials."""

    def __init__(self, name):
        self.name = name
        self.shapes = []        # Shape
        self.entities = []      # (shape index, 4x4 transform, visibility, emissive)
        self.area_lights, self.point_lights, self.sun_lights, self.spot_lights = [], [], [], []
        self.surface_materials = []
        self.light_materials = []
        self.world_shader = 0
        self.camera = None      # dict(transform, near, far, focal, shift)
        self.envmap_resolution = 2048
        # texture slots of the material library: None resolves them from the
        # generated code (image files relative to texture_dir, Nishita skies
        # generated); or a list of (h, w, 1 | 4) float arrays in slot order
        self.textures = None
        self.texture_dir = None
        self._keep = []

    def add_shape(self, shape):
        self.shapes.append(shape)
        return len(self.shapes) - 1

    def add_entity(self, shape, transform=None, visibility=api.RAY_ALL, emissive=False):
        self.entities.append((shape, np.eye(4) if transform is None else transform, visibility, emissive))

    def materials_json(self):
        return json.dumps({"surface": self.surface_materials, "light": self.light_materials})

    def material_source(self, noinline=None):
        """Generated material code; large graph sets are emitted as out-of-line
        functions to bound the shading kernel's registers and compile time."""
        if noinline is None:
            noinline = sum(len(m["nodes"]) for m in self.surface_materials) > 400
        return api.codegen(self.materials_json(), noinline)

    def resolve_textures(self):
        """The texture slots the generated material code reads, in slot order
        (api.codegen_textures): image files loaded (a file that cannot be read
        becomes one black texel, as in the reference), Nishita sky textures
        generated at 512 x 512."""
        from raymond_amd import images
        out = []
        for s in api.codegen_textures(self.materials_json()):
            if s["kind"] == "nishita":
                out.append(api.sky_nishita_texture(512, 512, s["sun_elevation"], s["altitude"], s["air_density"],
                                                   s["dust_density"], s["ozone_density"]))
                continue
            path = s["path"]
            if not os.path.isabs(path):
                path = os.path.join(self.texture_dir or os.getcwd(), path)
            try:
                out.append(images.load(path))
            except (OSError, ValueError) as e:
                warnings.warn("texture %s: %s; using a black texel" % (path, e))
                out.append(images.placeholder())
        return out

    def arrays(self):
        """Concatenated geometry and per-instance ranges."""
        vo, fo = [], []
        nv = nf = 0
        for s in self.shapes:
            vo.append(nv)
            fo.append(nf)
            nv += s.vertices.shape[0]
            nf += s.faces.shape[0]
        verts = np.concatenate([s.vertices for s in self.shapes]) if self.shapes else np.zeros((0, 3), np.float32)
        norms = np.concatenate([s.normals for s in self.shapes]) if self.shapes else np.zeros((0, 3), np.float32)
        uvs = np.concatenate([s.texcoords for s in self.shapes]) if self.shapes else np.zeros((0, 2), np.float32)
        faces = np.concatenate([s.faces for s in self.shapes]) if self.shapes else np.zeros((0, 3), np.uint32)
        mats = np.concatenate([s.materials for s in self.shapes]) if self.shapes else np.zeros(0, np.uint16)
        inst = []
        for (si, m, vis, em) in self.entities:
            s = self.shapes[si]
            inst.append(dict(vertex_offset=vo[si], vertex_count=s.vertices.shape[0], face_offset=fo[si],
                             face_count=s.faces.shape[0], transform=col_major(m), visibility=vis,
                             emissive=int(em), bounds=s.bounds))
        return (np.ascontiguousarray(verts, np.float32), np.ascontiguousarray(norms, np.float32),
                np.ascontiguousarray(uvs, np.float32), np.ascontiguousarray(faces, np.uint32),
                np.ascontiguousarray(mats, np.uint16), inst)

    def camera_c(self):
        c = self.camera
        return api.CameraDesc((C.c_float * 16)(*col_major(c["transform"])), c.get("near", 1e-3),
                              c.get("far", 1e4), c["focal"], (C.c_float * 2)(*c.get("shift", (0.0, 0.0))))

    def to_c(self):
        verts, norms, uvs, faces, mats, inst = self.arrays()
        self._keep = [verts, norms, uvs, faces, mats]
        sd = api.SceneDesc()
        sd.num_vertices = verts.shape[0]
        sd.vertices = verts.ctypes.data_as(C.POINTER(C.c_float))
        sd.normals = norms.ctypes.data_as(C.POINTER(C.c_float))
        sd.texcoords = uvs.ctypes.data_as(C.POINTER(C.c_float))
        sd.num_faces = faces.shape[0]
        sd.indices = faces.ctypes.data_as(C.POINTER(C.c_uint32))
        sd.face_materials = mats.ctypes.data_as(C.POINTER(C.c_uint16))
        arr = (api.InstanceDesc * max(1, len(inst)))()
        for k, d in enumerate(inst):
            b = d["bounds"]
            arr[k] = api.InstanceDesc(d["vertex_offset"], d["vertex_count"], d["face_offset"], d["face_count"],
                                      (C.c_float * 16)(*d["transform"]), d["visibility"], d["emissive"],
                                      int(b is not None), (C.c_float * 3)(*(b[0] if b is not None else (0, 0, 0))),
                                      (C.c_float * 3)(*(b[1] if b is not None else (0, 0, 0))))
        self._keep.append(arr)
        sd.num_instances = len(inst)
        sd.instances = arr

        def light_array(ctype, items):
            a = (ctype * max(1, len(items)))(*items)
            self._keep.append(a)
            return len(items), a

        sd.num_area_lights, sd.area_lights = light_array(api.AreaLight, self.area_lights)
        sd.num_point_lights, sd.point_lights = light_array(api.PointLight, self.point_lights)
        sd.num_sun_lights, sd.sun_lights = light_array(api.SunLight, self.sun_lights)
        sd.num_spot_lights, sd.spot_lights = light_array(api.SpotLight, self.spot_lights)
        sd.world_shader = self.world_shader
        sd.envmap_resolution = self.envmap_resolution
        sd.camera = self.camera_c()
        if self.textures is None:
            self.textures = self.resolve_textures()
        tex = (api.TextureDesc * max(1, len(self.textures)))()
        for k, t in enumerate(self.textures):
            t = np.ascontiguousarray(t, np.float32)
            if t.ndim != 3 or t.shape[2] not in (1, 4):
                raise ValueError("texture %d must be (h, w, 1) or (h, w, 4)" % k)
            self._keep.append(t)
            tex[k] = api.TextureDesc(t.shape[1], t.shape[0], t.shape[2], t.ctypes.data_as(C.POINTER(C.c_float)))
        self._keep.append(tex)
        sd.num_textures = len(self.textures)
        sd.textures = tex
        return sd

    @property
    def num_triangles(self):
        return sum(self.shapes[e[0]].faces.shape[0] for e in self.entities)


# ---------------------------------------------------------------------------
# Geometry builders


def quad(p0, p1, p2, p3, material):
    """Two triangles (p0 p1 p2), (p0 p2 p3) with a flat normal."""
    v = np.array([p0, p1, p2, p3], np.float64)
    n = np.cross(v[1] - v[0], v[2] - v[0])
    n /= np.linalg.norm(n)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    return v, np.repeat(n[None], 4, 0), np.array([[0, 1, 2], [0, 2, 3]]), [material, material], uv


def merge(parts):
    vs, ns, fs, ms, uvs = [], [], [], [], []
    off = 0
    for (v, n, f, m, uv) in parts:
        vs.append(v)
        ns.append(n)
        fs.append(f + off)
        ms.extend(m)
        uvs.append(uv)
        off += v.shape[0]
    return Shape(np.concatenate(vs), np.concatenate(ns), np.concatenate(fs), ms, np.concatenate(uvs))


def grid_quad(p0, p1, p2, p3, material, k):
    """The quad p0 p1 p2 p3 as a k x k grid of quads (2 k^2 triangles), flat normal."""
    if k <= 1:
        return quad(p0, p1, p2, p3, material)
    P = np.array([p0, p1, p2, p3], np.float64)
    n = np.cross(P[1] - P[0], P[2] - P[0])
    n /= np.linalg.norm(n)
    t = np.linspace(0.0, 1.0, k + 1)
    u, v = np.meshgrid(t, t, indexing="xy")
    u, v = u.ravel(), v.ravel()
    w = np.stack([(1 - u) * (1 - v), u * (1 - v), u * v, (1 - u) * v], 1)
    verts = w @ P
    idx = np.arange((k + 1) * (k + 1)).reshape(k + 1, k + 1)
    a, b = idx[:-1, :-1].ravel(), idx[:-1, 1:].ravel()
    c, d = idx[1:, 1:].ravel(), idx[1:, :-1].ravel()
    faces = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
    uv = np.stack([u, v], 1).astype(np.float32)
    return verts, np.repeat(n[None], verts.shape[0], 0), faces, [material] * faces.shape[0], uv


def unit_cube(material):
    """Cube [-0.5, 0.5]^3 with outward normals, 12 triangles."""
    h = 0.5
    P = lambda x, y, z: (x * h, y * h, z * h)
    faces = [
        quad(P(-1, -1, -1), P(-1, 1, -1), P(1, 1, -1), P(1, -1, -1), material),   # bottom (-z)
        quad(P(-1, -1, 1), P(1, -1, 1), P(1, 1, 1), P(-1, 1, 1), material),       # top (+z)
        quad(P(-1, -1, -1), P(1, -1, -1), P(1, -1, 1), P(-1, -1, 1), material),   # front (-y)
        quad(P(1, 1, -1), P(-1, 1, -1), P(-1, 1, 1), P(1, 1, 1), material),       # back (+y)
        quad(P(-1, 1, -1), P(-1, -1, -1), P(-1, -1, 1), P(-1, 1, 1), material),   # left (-x)
        quad(P(1, -1, -1), P(1, 1, -1), P(1, 1, 1), P(1, -1, 1), material),       # right (+x)
    ]
    return merge(faces)


def cornell_box(width=512, height=512):
    """Unit Cornell box, 34 triangles + a 0.5x0.5 emitter (2 triangles), z up."""
    d = SceneDescription("cornell")
    WHITE, RED, GREEN, TALL, LIGHT = 0, 1, 2, 3, 4
    d.surface_materials = [
        diffuse_material("white", (0.73, 0.73, 0.73)),
        diffuse_material("red", (0.65, 0.05, 0.05)),
        diffuse_material("green", (0.12, 0.45, 0.15)),
        principled_material("tall box", Base_Color=[0.73, 0.73, 0.73, 1.0], Metallic=0.0, Roughness=0.4,
                            Specular=0.5),
        emission_material("light", (1.0, 1.0, 1.0), 15.0),
    ]
    d.light_materials = [background_world("world", (0.0, 0.0, 0.0), 1.0)]
    d.world_shader = 0
    walls = merge([
        quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), WHITE),   # floor, normal +z
        quad((0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1), WHITE),   # ceiling, normal -z
        quad((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1), WHITE),   # back wall, normal -y
        quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1), RED),     # left wall, normal +x
        quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0), GREEN),   # right wall, normal -x
    ])
    s_walls = d.add_shape(walls)
    s_short = d.add_shape(unit_cube(WHITE))
    s_tall = d.add_shape(unit_cube(TALL))
    light = merge([quad((-0.5, -0.5, 0), (-0.5, 0.5, 0), (0.5, 0.5, 0), (0.5, -0.5, 0), LIGHT)])  # normal -z
    s_light = d.add_shape(light)
    d.add_entity(s_walls)
    d.add_entity(s_short, translate((0.66, 0.32, 0.15)) @ rotate_z(-17.0) @ scale((0.3, 0.3, 0.3)))
    d.add_entity(s_tall, translate((0.34, 0.64, 0.3)) @ rotate_z(17.0) @ scale((0.3, 0.3, 0.6)))
    d.add_entity(s_light, translate((0.5, 0.5, 0.999)) @ scale((0.5, 0.5, 1.0)), emissive=True)
    d.camera = dict(transform=look_at((0.5, -1.35, 0.5), (0.5, 0.5, 0.5)), near=1e-3, far=1e4, focal=2.4)
    d.width, d.height = width, height
    return d


def tessellated_cornell_box(width, height, triangles):
    """The Cornell box of cornell_box() with walls and boxes tessellated so that the
    scene has about `triangles` triangles (the emitter stays one quad)."""
    d = cornell_box(width, height)
    k = max(1, int(round((triangles / 34.0) ** 0.5)))
    WHITE, RED, GREEN, TALL = 0, 1, 2, 3
    d.shapes[0] = merge([
        grid_quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), WHITE, k),
        grid_quad((0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1), WHITE, k),
        grid_quad((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1), WHITE, k),
        grid_quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1), RED, k),
        grid_quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0), GREEN, k),
    ])
    h = 0.5

    def cube(material):
        P = lambda x, y, z: (x * h, y * h, z * h)
        return merge([
            grid_quad(P(-1, -1, -1), P(-1, 1, -1), P(1, 1, -1), P(1, -1, -1), material, k),
            grid_quad(P(-1, -1, 1), P(1, -1, 1), P(1, 1, 1), P(-1, 1, 1), material, k),
            grid_quad(P(-1, -1, -1), P(1, -1, -1), P(1, -1, 1), P(-1, -1, 1), material, k),
            grid_quad(P(1, 1, -1), P(-1, 1, -1), P(-1, 1, 1), P(1, 1, 1), material, k),
            grid_quad(P(-1, 1, -1), P(-1, -1, -1), P(-1, -1, 1), P(-1, 1, 1), material, k),
            grid_quad(P(1, -1, -1), P(1, 1, -1), P(1, 1, 1), P(1, -1, 1), material, k),
        ])

    d.shapes[1] = cube(WHITE)
    d.shapes[2] = cube(TALL)
    d.name = "cornell-%d" % d.num_triangles
    return d


def light_info(shader_index, casts_shadows=True, uses_mis=False):
    return api.LightInfo(shader_index, int(casts_shadows), int(uses_mis))


def area_light(info, center, size, color, circular=False, rotation=None):
    """Lamp of `size` (x, y) around `center`; uv in [-0.5, 0.5]^2 maps through
    the 3x4 affine rows.  It emits toward -normal, the -z axis of `rotation`."""
    R = np.eye(3) if rotation is None else np.asarray(rotation, np.float64)[:3, :3]
    M = np.zeros((3, 4))
    M[:, 0], M[:, 1], M[:, 2] = R[:, 0] * size[0], R[:, 1] * size[1], R[:, 2]
    M[:, 3] = center
    return api.AreaLight(info, (C.c_float * 12)(*M.reshape(-1)), (C.c_float * 3)(*color), int(circular))


def point_light(info, location, radius, color):
    return api.PointLight(info, (C.c_float * 3)(*location), radius, (C.c_float * 3)(*color))


def sun_light(info, direction, cos_angle, color):
    """`direction` points from the scene toward the sun."""
    d = np.asarray(direction, np.float64)
    d /= np.linalg.norm(d)
    return api.SunLight(info, (C.c_float * 3)(*d), cos_angle, (C.c_float * 3)(*color))


def spot_light(info, location, direction, radius, color, spot_size, spot_blend):
    """`direction` points from the lit region back toward the lamp (the
    attenuation is its dot product with the receiver-to-lamp direction);
    `spot_size` is the cosine of the cone's half angle, `spot_blend` the
    smoothing width in cosine."""
    d = np.asarray(direction, np.float64)
    d /= np.linalg.norm(d)
    return api.SpotLight(info, (C.c_float * 3)(*location), (C.c_float * 3)(*d), radius, (C.c_float * 3)(*color),
                         spot_size, spot_blend)


def lamp_box(width=64, height=64):
    """The Cornell box lit by one lamp of every analytic type instead of its
    emissive quad: a rectangular and a circular area lamp under the ceiling,
    a point lamp with a radius, a spot lamp and a sun shining in through the
    open front.  Lamps use two emission shaders; some take part in MIS."""
    d = cornell_box(width, height)
    d.name = "lamp box"
    d.entities = d.entities[:3]
    d.shapes = d.shapes[:3]
    d.light_materials = [background_world("world", (0.0, 0.0, 0.0), 1.0),
                         emission_material("warm", (1.0, 0.85, 0.6), 4.0),
                         emission_material("cool", (0.6, 0.8, 1.0), 2.0)]
    d.area_lights = [
        area_light(light_info(1, uses_mis=True), (0.3, 0.5, 0.98), (0.25, 0.4), (1.0, 1.0, 1.0)),
        area_light(light_info(2), (0.72, 0.35, 0.97), (0.2, 0.2), (1.5, 1.5, 1.5), circular=True),
    ]
    d.point_lights = [point_light(light_info(1, uses_mis=True), (0.5, 0.3, 0.8), 0.05, (3.0, 3.0, 3.0))]
    d.sun_lights = [sun_light(light_info(2), (0.2, -1.0, 0.6), 0.995, (0.8, 0.8, 0.8))]
    d.spot_lights = [spot_light(light_info(1, casts_shadows=True), (0.25, 0.25, 0.9), (0.3, 0.2, 1.0), 0.02,
                                (6.0, 6.0, 6.0), 0.85, 0.05)]
    return d


def batch_jobs(n=64, seed=4):
    """n Cornell-box jobs with triangle counts log-uniform in
    [1k, 2M], resolutions from {256, 512, 768, 1024}^2, 16-256 spp and depth
    6-12.  Returns [(scene description, width, height, samples, depth)]."""
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(n):
        tris = int(np.exp(rng.uniform(np.log(1e3), np.log(2e6))))
        res = int(rng.choice([256, 512, 768, 1024]))
        spp = int(rng.integers(16, 257))
        depth = int(rng.integers(6, 13))
        d = tessellated_cornell_box(res, res, tris)
        d.envmap_resolution = 64  # black world
        jobs.append((d, res, res, spp, depth))
    return jobs


def icosphere(level):
    """Unit icosphere by midpoint subdivision: 10*4^level + 2 vertices, 20*4^level faces."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
                  [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mid])
        nf = f.shape[0]
        a, b, c = inv[:nf] + base, inv[nf:2 * nf] + base, inv[2 * nf:] + base
        f = np.concatenate([np.stack([f[:, 0], a, c], 1), np.stack([f[:, 1], b, a], 1),
                            np.stack([f[:, 2], c, b], 1), np.stack([a, b, c], 1)])
    return v, f


def value_noise(p, seed, octaves=3):
    """Seeded 3-D value noise in [-1, 1]: hashed lattice values, smoothstep-trilinear, fBm octaves."""
    out = np.zeros(p.shape[0])
    amp, freq, norm = 1.0, 4.0, 0.0
    for o in range(octaves):
        q = p * freq
        i = np.floor(q).astype(np.int64)
        fr = q - i
        w = fr * fr * (3.0 - 2.0 * fr)

        def lattice(dx, dy, dz):
            h = ((i[:, 0] + dx) * 73856093) ^ ((i[:, 1] + dy) * 19349663) ^ ((i[:, 2] + dz) * 83492791)
            h = (h ^ (seed * 2654435761 + o * 97531)) & 0xFFFFFFFF
            h = (h * 2246822519) & 0xFFFFFFFF
            h ^= h >> 13
            h = (h * 3266489917) & 0xFFFFFFFF
            h ^= h >> 16
            return h / 4294967295.0 * 2.0 - 1.0

        acc = 0.0
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    wx = w[:, 0] if dx else 1 - w[:, 0]
                    wy = w[:, 1] if dy else 1 - w[:, 1]
                    wz = w[:, 2] if dz else 1 - w[:, 2]
                    acc = acc + wx * wy * wz * lattice(dx, dy, dz)
        out += amp * acc
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return out / norm


def vertex_normals(v, f):
    fn = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    n = np.zeros_like(v)
    for k in range(3):
        np.add.at(n, f[:, k], fn)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    return n


def sphere_field(width=1920, height=1080, grid=8, level=6, radius=0.4, amplitude=0.05):
    """grid x grid displaced icospheres (seed 1000+i), jittered grid (seed 7),
    Principled per instance (seed 42), ground plane, analytic sky."""
    d = SceneDescription("spheres")
    rng_mat = np.random.default_rng(42)
    mats = []
    n = grid * grid
    for i in range(n):
        base = rng_mat.uniform(0, 1, 3)
        metallic = 1.0 if rng_mat.uniform() < 0.3 else 0.0
        rough = rng_mat.uniform(0.05, 1.0)
        coat = rng_mat.uniform(0, 1)
        mats.append(principled_material("sphere %02d" % i, Base_Color=[float(x) for x in base] + [1.0],
                                        Metallic=metallic, Roughness=float(rough), Specular=0.5,
                                        Clearcoat=float(coat)))
    mats.append(diffuse_material("ground", (0.5, 0.5, 0.5)))
    d.surface_materials = mats
    d.light_materials = [sky_world()]
    d.world_shader = 0
    v0, f0 = icosphere(level)
    rng_grid = np.random.default_rng(7)
    jitter = rng_grid.uniform(-0.25, 0.25, (n, 2))
    for i in range(n):
        disp = radius * (1.0 + amplitude * value_noise(v0, 1000 + i) / radius)
        v = v0 * disp[:, None]
        s = Shape(v, vertex_normals(v, f0), f0, i)
        si = d.add_shape(s)
        gx, gy = i % grid, i // grid
        cx = (gx - (grid - 1) / 2.0) + jitter[i, 0]
        cy = (gy - (grid - 1) / 2.0) + jitter[i, 1]
        d.add_entity(si, translate((cx, cy, radius)))
    g = 50.0
    ground = merge([quad((-g, -g, 0), (g, -g, 0), (g, g, 0), (-g, g, 0), n)])
    d.add_entity(d.add_shape(ground))
    d.camera = dict(transform=look_at((0.0, -8.5, 5.5), (0.0, 0.3, 0.0)), near=1e-3, far=1e4, focal=2.0)
    d.width, d.height = width, height
    return d


def instanced_field(width=1920, height=1080, grid=16, shapes=4, level=5, radius=0.4, amplitude=0.05):
    """grid x grid instances of `shapes` displaced icospheres (seeds 1000+i),
    each instance a turn about z (seed 9), a scale of 0.6-1.4 and a jittered
    grid position (seed 7); Principled per shape (seed 42), ground plane,
    analytic sky.  Built two-level, each shape's triangles are stored once."""
    d = SceneDescription("instances")
    rng_mat = np.random.default_rng(42)
    mats = []
    for i in range(shapes):
        base = rng_mat.uniform(0, 1, 3)
        mats.append(principled_material("shape %02d" % i, Base_Color=[float(x) for x in base] + [1.0],
                                        Metallic=1.0 if i % 3 == 1 else 0.0,
                                        Roughness=float(rng_mat.uniform(0.05, 1.0)), Specular=0.5,
                                        Clearcoat=float(rng_mat.uniform(0, 1))))
    mats.append(diffuse_material("ground", (0.5, 0.5, 0.5)))
    d.surface_materials = mats
    d.light_materials = [sky_world()]
    d.world_shader = 0
    v0, f0 = icosphere(level)
    ids = []
    for i in range(shapes):
        v = v0 * (radius * (1.0 + amplitude * value_noise(v0, 1000 + i) / radius))[:, None]
        ids.append(d.add_shape(Shape(v, vertex_normals(v, f0), f0, i)))
    rng = np.random.default_rng(9)
    jitter = np.random.default_rng(7).uniform(-0.25, 0.25, (grid * grid, 2))
    spacing = 8.0 / grid
    for i in range(grid * grid):
        gx, gy = i % grid, i // grid
        c = np.array([(gx - (grid - 1) / 2.0) * spacing + jitter[i, 0] * spacing,
                      (gy - (grid - 1) / 2.0) * spacing + jitter[i, 1] * spacing, 0.0])
        sc = rng.uniform(0.6, 1.4) * spacing
        c[2] = radius * sc
        d.add_entity(ids[i % shapes], translate(c) @ rotate_z(rng.uniform(0, 360)) @ scale((sc, sc, sc)))
    g = 50.0
    d.add_entity(d.add_shape(merge([quad((-g, -g, 0), (g, -g, 0), (g, g, 0), (-g, g, 0), shapes)])))
    d.camera = dict(transform=look_at((0.0, -8.5, 5.5), (0.0, 0.3, 0.0)), near=1e-3, far=1e4, focal=2.0)
    d.width, d.height = width, height
    return d


def singlet_lens(focal_mm=50.0, f_number=4.0):
    """A biconvex N-BK7 singlet (n_d 1.5168, V_d 64.17) with a stop behind it,
    as lens surfaces for api.lens_desc: object medium, front and back faces,
    stop, image plane.  Equal radii R = 2 f (n - 1) give the thin-lens focal
    length f; the stop sits 5 mm behind the lens, the image plane at f."""
    n = 1.5168
    R = 2.0 * focal_mm * (n - 1.0)
    ap = focal_mm / (2.0 * f_number)
    return np.array([[0.0, 0.0, 100.0, 1.0, 0.0],
                     [R, 4.0, ap * 1.2, n, 64.17],
                     [-R, 5.0, ap * 1.2, 1.0, 0.0],
                     [0.0, focal_mm - 5.0, ap, 1.0, 0.0],
                     [0.0, 0.0, 100.0, 1.0, 0.0]], np.float32)


def build_device_scene(device, desc, threads=0, envmap=True, noinline=None, two_level=False):
    """SceneLoader.loadScene: materials (JIT), shapes/entities (upload + BVH,
    flat or two-level), lights (envmap)."""
    scene = api.Scene(device, desc)
    scene.compile_materials(desc.material_source(noinline))
    scene.build_accel(threads, two_level=two_level)
    if envmap:
        scene.build_envmap()
    return scene


# ---------------------------------------------------------------------------
# Seeded random node graphs


def _node(t, inputs, params=None):
    return {"type": t, "parameters": params or {}, "inputs": inputs}


def random_graph_material(name, rng, min_nodes=20, max_nodes=60):
    """A random Cycles-style graph of 20-60 nodes from {noise, voronoi, 4-stop ramp,
    math add/mul/pow, mix RGB, bump, mix shader, principled}; the output is a
    principled BSDF or a mix of two.  Every node is reachable from the output.

    The tail (principled, clamps, mix shader, folds of unused outputs) varies
    in size, so a draw whose node count leaves [min_nodes, max_nodes] is
    redrawn from the same generator; the result stays a function of the seed."""
    for _ in range(64):
        graph = _random_graph_attempt(name, rng, int(rng.integers(min_nodes, max_nodes + 1)))
        if min_nodes <= len(graph["nodes"]) <= max_nodes:
            return graph
    raise RuntimeError("random_graph_material: no graph within [%d, %d] nodes" % (min_nodes, max_nodes))


def _random_graph_attempt(name, rng, target):
    nodes = {}
    producers = {"value": [], "color": [], "vector": []}
    consumed = set()
    counter = [0]

    def new(prefix):
        counter[0] += 1
        return "%s %d" % (prefix, counter[0])

    def pick(kind, allow_const=True):
        pool = producers[kind]
        fresh = [p for p in pool if p not in consumed]
        if fresh and rng.uniform() < 0.8:
            p = fresh[int(rng.integers(len(fresh)))]
        elif pool and (not allow_const or rng.uniform() < 0.7):
            p = pool[int(rng.integers(len(pool)))]
        else:
            return None
        consumed.add(p)
        return p

    def val_input(kind, const):
        p = pick(kind)
        t = {"value": "VALUE", "color": "RGBA", "vector": "VECTOR"}[kind]
        if p is None:
            return _in(t, const)
        return _in(t, const, link=p)

    tc = new("Texture Coordinate")
    nodes[tc] = _node("TEX_COORD", {})
    producers["vector"].append((tc, "Generated"))
    kinds = ["noise", "voronoi", "ramp", "math", "mix", "bump"]
    weights = np.array([0.25, 0.08, 0.17, 0.25, 0.2, 0.05])
    def unconsumed():
        return sum(1 for kind in ("color", "value") for p in producers[kind] if p not in consumed)

    # leave room for the principled/mix/output tail (<= 9 nodes) and one fold per unused output
    while len(nodes) + unconsumed() + 9 < target:
        k = kinds[int(rng.choice(len(kinds), p=weights / weights.sum()))]
        if k == "noise":
            n = new("Noise Texture")
            nodes[n] = _node("TEX_NOISE", {
                "Vector": val_input("vector", [0.0, 0.0, 0.0]), "W": _in("VALUE", 0.0),
                "Scale": _in("VALUE", float(rng.uniform(1, 12))), "Detail": _in("VALUE", float(rng.uniform(0, 4))),
                "Roughness": _in("VALUE", float(rng.uniform(0.3, 0.7))),
                "Distortion": _in("VALUE", float(rng.choice([0.0, rng.uniform(0, 1)])))},
                {"noise_dimensions": "3D"})
            producers["value"].append((n, "Fac"))
            producers["color"].append((n, "Color"))
        elif k == "voronoi":
            n = new("Voronoi Texture")
            nodes[n] = _node("TEX_VORONOI", {
                "Vector": val_input("vector", [0.0, 0.0, 0.0]), "Scale": _in("VALUE", float(rng.uniform(1, 8))),
                "Smoothness": _in("VALUE", 1.0), "Exponent": _in("VALUE", 0.5), "Randomness": _in("VALUE", 1.0),
                "W": _in("VALUE", 0.0)}, {"distance": "EUCLIDEAN", "feature": "F1", "dimensions": "3D"})
            producers["value"].append((n, "Distance"))
            producers["color"].append((n, "Color"))
        elif k == "ramp":
            n = new("Color Ramp")
            pos = np.sort(rng.uniform(0, 1, 4))
            els = [{"position": float(p), "color": [float(x) for x in rng.uniform(0, 1, 3)] + [1.0]} for p in pos]
            nodes[n] = _node("VALTORGB", {"Fac": val_input("value", 0.5)},
                             {"color_mode": "RGB", "interpolation": "LINEAR", "elements": els})
            producers["color"].append((n, "Color"))
        elif k == "math":
            n = new("Math")
            op = ["ADD", "MULTIPLY", "POWER"][int(rng.integers(3))]
            nodes[n] = _node("MATH", {"Value": val_input("value", float(rng.uniform(0, 1))),
                                      "Value_001": val_input("value", float(rng.uniform(0.2, 2.0))),
                                      "Value_002": _in("VALUE", 0.5)},
                             {"operation": op, "use_clamp": bool(rng.uniform() < 0.5)})
            producers["value"].append((n, "Value"))
        elif k == "mix":
            n = new("Mix RGB")
            blend = ["MIX", "ADD", "MULTIPLY", "SCREEN", "OVERLAY", "LIGHTEN", "DARKEN"][int(rng.integers(7))]
            nodes[n] = _node("MIX_RGB", {"Fac": val_input("value", float(rng.uniform(0, 1))),
                                         "Color1": val_input("color", [float(x) for x in rng.uniform(0, 1, 3)] + [1.0]),
                                         "Color2": val_input("color", [float(x) for x in rng.uniform(0, 1, 3)] + [1.0])},
                             {"blend_type": blend, "use_clamp": bool(rng.uniform() < 0.5)})
            producers["color"].append((n, "Color"))
        else:
            n = new("Bump")
            nodes[n] = _node("BUMP", {"Strength": _in("VALUE", 1.0), "Distance": _in("VALUE", 1.0),
                                      "Height": val_input("value", 1.0), "Normal": _in("VECTOR", [0.0, 0.0, 0.0])})
            producers["vector"].append((n, "Normal"))

    def clamp01(src):
        m = new("Math")
        nodes[m] = _node("MATH", {"Value": _in("VALUE", 0.5, link=src), "Value_001": _in("VALUE", 0.0),
                                  "Value_002": _in("VALUE", 0.5)}, {"operation": "ADD", "use_clamp": True})
        return (m, "Value")

    def principled():
        p = new("Principled BSDF")
        base = pick("color", allow_const=False)
        rough = pick("value", allow_const=False)
        metal = pick("value") if rng.uniform() < 0.4 else None
        normal = pick("vector") if producers["vector"] else None
        kw = dict(Specular=0.5, Roughness=float(rng.uniform(0.1, 0.9)), Metallic=0.0)
        node = principled_node(**kw)
        if base:
            node["inputs"]["Base Color"] = _in("RGBA", [0.8, 0.8, 0.8, 1.0], link=base)
        if rough:
            node["inputs"]["Roughness"] = _in("VALUE", 0.5, link=clamp01(rough))
        if metal:
            node["inputs"]["Metallic"] = _in("VALUE", 0.0, link=clamp01(metal))
        if normal and normal[1] == "Normal":
            node["inputs"]["Normal"] = _in("VECTOR", [0.0, 0.0, 0.0], link=normal)
        node["inputs"]["Clearcoat"] = _in("VALUE", float(rng.choice([0.0, rng.uniform(0, 1)])))
        nodes[p] = node
        return (p, "BSDF")

    shader = principled()
    if rng.uniform() < 0.5:
        s2 = principled()
        mx = new("Mix Shader")
        nodes[mx] = _node("MIX_SHADER", {"Fac": val_input("value", float(rng.uniform(0.2, 0.8))),
                                         "Shader": _in("SHADER", link=shader), "Shader_001": _in("SHADER", link=s2)})
        shader = (mx, "Shader")
    # fold producers that nothing consumes into the base-colour chain so every node is reachable
    leftovers = [p for kind in ("color", "value") for p in producers[kind] if p not in consumed]
    pnode = [k for k in nodes if k.startswith("Principled")][0]
    for lp in leftovers:
        mx = new("Mix RGB")
        cur = nodes[pnode]["inputs"]["Base Color"]
        inputs = {"Fac": _in("VALUE", float(rng.uniform(0.1, 0.5))), "Color2": _in("RGBA", [0.5, 0.5, 0.5, 1.0], link=lp)}
        inputs["Color1"] = cur if "links" in cur else _in("RGBA", cur.get("value", [0.8, 0.8, 0.8, 1.0]))
        inputs["Color1"]["type"] = "RGBA"
        nodes[mx] = _node("MIX_RGB", inputs, {"blend_type": "MIX", "use_clamp": True})
        nodes[pnode]["inputs"]["Base Color"] = _in("RGBA", [0.8, 0.8, 0.8, 1.0], link=(mx, "Color"))
        consumed.add(lp)
    nodes[new("Material Output")] = output_material(shader)
    return {"name": name, "nodes": nodes}


def shader_graph_scene(width=1920, height=1080, materials=32, seed=3):
    """The Cornell-box geometry, `materials` random node graphs assigned per triangle."""
    d = cornell_box(width, height)
    rng = np.random.default_rng(seed)
    graphs = [random_graph_material("graph %02d" % i, rng) for i in range(materials)]
    light = d.surface_materials[4]
    d.surface_materials = graphs + [light]
    light_index = materials
    for si, s in enumerate(d.shapes):
        if si == 3:  # emitter quad keeps its emission material
            s.materials = np.full(s.faces.shape[0], light_index, np.uint16)
        else:
            s.materials = rng.integers(0, materials, s.faces.shape[0]).astype(np.uint16)
    d.name = "graphs"
    return d


def count_emitted_nodes(source):
    """Nodes per material function in generated source (for reporting)."""
    counts = []
    for body in source.split("(const SceneView& ctx, ShadingContext& shading) {")[1:]:
        body = body.split("\n}\n")[0]
        if "switch (index)" not in body:
            counts.append(body.count(".compute(ctx, shading);"))
    return counts


def bundled_material_sources(big=False):
    """(name, material source) of the scenes the tests, the smoke check and the
    benchmark workloads build (the 32-graph scene as the benchmark's inlined
    module; `big` adds its out-of-line form), for warming the code-object
    cache ahead of time."""
    out = {}
    for name, desc in [("cornell", cornell_box(64, 64)),
                       ("small spheres", sphere_field(96, 54, grid=2, level=3)),
                       ("graphs, 8 materials", shader_graph_scene(48, 48, materials=8, seed=3)),
                       ("lamp box", lamp_box(48, 48)),
                       ("sphere field", sphere_field(1920, 1080))]:
        out[name] = desc.material_source()
    g = shader_graph_scene(1920, 1080)
    out["graphs, inlined"] = g.material_source(False)
    if big:
        out["graphs, out of line"] = g.material_source(True)
    # scenes with equal materials share one module
    seen, jobs = set(), []
    for name, src in out.items():
        if src not in seen:
            seen.add(src)
            jobs.append((name, src))
    return jobs
