"""Scene ingestion: the scene files of the reference's Blender exporter, built
into a SceneDescription the way the reference's host builds its scene.

Input (scripts/blender_exporter/raymond_blender/exporter.py:50-69): one JSON
file with the objects "materials", "shapes", "entities", "lights", "camera"
and "render", each keyed by a unique name, plus the meshes it names as ASCII
PLY files (shapes.py:19-123) at paths relative to the JSON file.  The
reference parses the JSON with Rayjay, an external package that is not part
of its sources; the meaning of each field below follows the exporter that
writes it and the builders that read it:

* materials: node graphs (materials.py:273-560), indexed per kind in the order
  they are first requested (MaterialBuilder.swift:876-879): light materials
  first, for every light (LightBuilder.swift:27-34), then surface materials,
  shape by shape in the order the entities name their shapes
  (EntityBuilder.swift:44-52, ShapeBuilder.swift:212-235);
* shapes: type "ply" only (ShapeBuilder.swift:213-215); a face's material is
  the shape's material list at the face's palette slot (PLYReader.m:159-160);
  a shape with any emissive material becomes a shape light for every entity
  that uses it (EntityBuilder.swift:100-114);
* entities: shape, 4x4 matrix (row-major rows, object to world), visibility
  flags (EntityBuilder.swift:54-64);
* lights: AREA, POINT, SPOT, SUN and one WORLD, converted as
  LightBuilder.swift:183-223 does (power folded into the colour, cone and
  sun angles halved and turned into cosines, circular area lamps scaled by
  4/pi);
* camera: transform, clip range, shift, and the focal length divided by half
  the film width (default 50 mm; no camera: the default camera,
  SceneLoader.swift:31-44, 128-140).

Orders: the reference iterates Swift dictionaries, whose order varies from
run to run, for instance, material and light indices
(EntityBuilder.swift:50,85; LightBuilder.swift:53).  This loader uses the
order of the JSON objects instead, so the same file always gives the same
indices.

Material emission: the reference asks Rayjay whether a material has surface
emission (MaterialBuilder.swift:885-887).  Here a surface material emits when
the node graph behind its output's Surface input holds an Emission node or a
Principled BSDF whose emission colour and strength are linked or non-zero.
"""
import json
import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import api
from . import scenes

# EntityBuilder.swift:54-64: exported visibility flags -> ray flags
_VISIBILITY = (("camera", api.RAY_CAMERA), ("diffuse", api.RAY_DIFFUSE), ("glossy", api.RAY_GLOSSY),
               ("transmission", api.RAY_TRANSMISSION), ("volume", api.RAY_VOLUME), ("shadow", api.RAY_SHADOW))


class SceneFormatError(ValueError):
    pass


def _f32(x):
    return np.float32(x)


def _matrix(flat, what):
    m = np.asarray(flat, np.float64)
    if m.shape != (16,):
        raise SceneFormatError("%s: expected 16 numbers" % what)
    return m.reshape(4, 4)


def _visibility(v):
    mask = 0
    for key, flag in _VISIBILITY:
        if v.get(key, False):
            mask |= flag
    return mask


def _normalize(v):
    """simd normalize in float arithmetic: v * (1 / sqrt(dot(v, v)))."""
    v = np.asarray(v, np.float32)
    d = np.float32(v[0] * v[0] + v[1] * v[1]) + np.float32(v[2] * v[2])
    return v * np.float32(np.float32(1.0) / np.sqrt(np.float32(d)))


def _cos_f32(x):
    return np.float32(math.cos(float(np.float32(x))))


def _linked_or_nonzero(inp):
    if inp is None:
        return False
    if inp.get("links"):
        return True
    v = inp.get("value", 0.0)
    if isinstance(v, list):
        return any(float(c) != 0.0 for c in v[:3])
    return float(v) != 0.0


def has_surface_emission(graph):
    """True when the graph behind the material output's Surface input emits."""
    outputs = [n for n in graph.values() if n.get("type") == "OUTPUT_MATERIAL"]
    if not outputs:
        return False
    seen, stack = set(), []
    for out in outputs:
        for link in (out.get("inputs", {}).get("Surface", {}).get("links") or []):
            stack.append(link[0])
    while stack:
        name = stack.pop()
        if name in seen or name not in graph:
            continue
        seen.add(name)
        node = graph[name]
        t = node.get("type")
        inputs = node.get("inputs", {})
        if t == "EMISSION":
            return True
        if t == "BSDF_PRINCIPLED" and _linked_or_nonzero(inputs.get("Emission")) and \
                _linked_or_nonzero(inputs.get("Emission Strength")):
            return True
        for inp in inputs.values():
            for link in (inp.get("links") or []):
                stack.append(link[0])
    return False


class _MaterialIndex:
    """Per-kind indices in first-request order (MaterialBuilder.swift:876-879)."""

    def __init__(self, library):
        self.library = library
        self.names = {"surface": [], "light": []}
        self.index = {"surface": {}, "light": {}}

    def get(self, kind, name):
        if name not in self.library:
            raise SceneFormatError("unknown material '%s'" % name)
        idx = self.index[kind]
        if name not in idx:
            idx[name] = len(self.names[kind])
            self.names[kind].append(name)
        return idx[name]

    def materials(self, kind):
        return [{"name": n, "nodes": self.library[n]} for n in self.names[kind]]


def load_scene(path, envmap_resolution=2048):
    """Builds a SceneDescription from the exporter's scene JSON at `path`.
    `resolution` (width, height) from the file's "render" object is kept on
    the description."""
    with open(path) as f:
        data = json.load(f)
    base = os.path.dirname(os.path.abspath(path))
    materials = data.get("materials") or {}
    shapes = data.get("shapes") or {}
    entities = data.get("entities") or {}
    lights = data.get("lights") or {}
    mats = _MaterialIndex(materials)
    desc = scenes.SceneDescription(os.path.splitext(os.path.basename(path))[0])
    # image paths of the exporter are relative to the scene file
    desc.texture_dir = os.path.dirname(os.path.abspath(path))
    desc.envmap_resolution = envmap_resolution

    # light materials first (LightBuilder.swift:27-34)
    for name, light in lights.items():
        if "material" not in light:
            raise SceneFormatError("light '%s' has no material" % name)
        mats.get("light", light["material"])

    # shapes, in the order the entities name them (EntityBuilder.swift:44-52)
    order, named = [], set()
    for ename, ent in entities.items():
        sname = ent.get("shape")
        if sname not in shapes:
            raise SceneFormatError("entity '%s' names unknown shape '%s'" % (ename, sname))
        if sname not in named:
            named.add(sname)
            if shapes[sname].get("type") != "ply":
                raise SceneFormatError("shape '%s': unsupported type '%s'" % (sname, shapes[sname].get("type")))
            order.append(sname)
    # the meshes parse in parallel (the reader releases the GIL)
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1, max(1, len(order)))) as pool:
        meshes = list(pool.map(lambda n: api.read_ply(os.path.join(base, shapes[n]["filepath"])), order))
    shape_index, shape_emissive = {}, []
    for sname, mesh in zip(order, meshes):
        sh = shapes[sname]
        palette = np.array([mats.get("surface", m) for m in sh.get("materials", [])], np.uint16)
        emissive = any(has_surface_emission(materials[m]) for m in sh.get("materials", []))
        if mesh["face_palette"].size and int(mesh["face_palette"].max()) >= palette.size:
            raise SceneFormatError("shape '%s': face material slot beyond its %d materials" % (sname, palette.size))
        shape = scenes.Shape(mesh["vertices"], mesh["normals"], mesh["faces"],
                             palette[mesh["face_palette"]] if palette.size else np.zeros(0, np.uint16),
                             mesh["texcoords"])
        shape.bounds = (mesh["bounds_min"], mesh["bounds_max"])
        shape_index[sname] = desc.add_shape(shape)
        shape_emissive.append(emissive)

    # entities (EntityBuilder.swift:66-76, 97-141)
    for ename, ent in entities.items():
        si = shape_index[ent["shape"]]
        desc.add_entity(si, _matrix(ent.get("matrix", np.eye(4).reshape(-1)), "entity '%s' matrix" % ename),
                        _visibility(ent.get("visibility", {})), emissive=shape_emissive[si])

    # lights (LightBuilder.swift:36-47, 183-223, 240-249)
    world = None
    for name, light in lights.items():
        t = light.get("type")
        p = light.get("parameters", {})
        info = scenes.light_info(mats.get("light", light["material"]), bool(light.get("cast_shadows", True)), False)
        if t == "WORLD":
            if world is None:
                world = mats.get("light", light["material"])
            continue
        color = np.asarray(p.get("color", [1.0, 1.0, 1.0]), np.float32)[:3]
        power = _f32(p.get("power", 1.0))
        if t == "AREA":
            circular = bool(p.get("is_circular", False))
            norm = np.float32(np.float32(4.0) / np.float32(math.pi)) if circular else np.float32(1.0)
            col = (norm * color) * power
            m = _matrix(p["transform"], "light '%s' transform" % name)
            desc.area_lights.append(api.AreaLight(info, (api.C.c_float * 12)(*m[:3].reshape(-1)),
                                                  (api.C.c_float * 3)(*col), int(circular)))
        elif t == "POINT":
            desc.point_lights.append(api.PointLight(info, (api.C.c_float * 3)(*p["location"][:3]),
                                                    float(p.get("radius", 0.0)),
                                                    (api.C.c_float * 3)(*(color * power))))
        elif t == "SPOT":
            angle = _cos_f32(_f32(p["spot_size"]) / np.float32(2.0))
            blend = (np.float32(1.0) - angle) * _f32(p.get("spot_blend", 0.0))
            desc.spot_lights.append(api.SpotLight(info, (api.C.c_float * 3)(*p["location"][:3]),
                                                  (api.C.c_float * 3)(*_normalize(p["direction"][:3])),
                                                  float(p.get("radius", 0.0)),
                                                  (api.C.c_float * 3)(*(color * power)), float(angle), float(blend)))
        elif t == "SUN":
            desc.sun_lights.append(api.SunLight(info, (api.C.c_float * 3)(*_normalize(p["direction"][:3])),
                                                float(_cos_f32(_f32(p.get("angle", 0.0)) / np.float32(2.0))),
                                                (api.C.c_float * 3)(*(color * power))))
        else:
            raise SceneFormatError("light '%s': unsupported type '%s'" % (name, t))
    if world is None:
        # LightBuilder.swift:240 requires a world light
        raise SceneFormatError("the scene has no WORLD light")
    desc.world_shader = world
    desc.surface_materials = mats.materials("surface")
    desc.light_materials = mats.materials("light")

    # camera (SceneLoader.swift:31-44, 128-140)
    cam = data.get("camera")
    if cam:
        film_w = _f32(cam.get("film", {}).get("width", 36.0))
        focal = _f32(cam.get("focal_length", 50.0)) / (film_w / np.float32(2.0))
        desc.camera = dict(transform=_matrix(cam["transform"], "camera transform"), near=float(cam["near_clip"]),
                           far=float(cam["far_clip"]), focal=float(focal),
                           shift=tuple(float(x) for x in cam.get("shift", (0.0, 0.0))))
    else:
        desc.camera = dict(transform=np.eye(4), near=0.1, far=100.0, focal=2.5, shift=(0.0, 0.0))
    res = (data.get("render") or {}).get("resolution")
    desc.resolution = (int(res["width"]), int(res["height"])) if res else None
    return desc


def _ply_text(shape, palette_slots):
    """ASCII PLY in the exporter's layout (shapes.py:19-33, 98-123)."""
    lines = ["ply", "format ascii 1.0", "comment Copyright raymond exporter research team",
             "element vertex %d" % shape.vertices.shape[0], "property float x", "property float y",
             "property float z", "property float nx", "property float ny", "property float nz", "property float s",
             "property float t", "element face %d" % shape.faces.shape[0],
             "property list uchar uint vertex_indices", "property uchar material_index", "end_header"]
    out = ["\n".join(lines) + "\n"]
    for v, n, t in zip(shape.vertices, shape.normals, shape.texcoords):
        out.append("%.6f %.6f %.6f %.6f %.6f %.6f %.6f %.6f\n" % (v[0], v[1], v[2], n[0], n[1], n[2], t[0], t[1]))
    for f, m in zip(shape.faces, palette_slots):
        out.append("3 %d %d %d %d\n" % (f[0], f[1], f[2], m))
    return "".join(out)


def _flat(m):
    return [float(x) for x in np.asarray(m, np.float64).reshape(-1)]


def write_scene(desc, path, lights=None):
    """Writes `desc` as exporter scene files: `path` (JSON) and one PLY per
    shape under meshes/ next to it.  Shapes keep their material slots in
    first-use order.  `lights` gives the exported light objects (type,
    material, parameters, ...) to write; by default the description's world
    shader becomes the one WORLD light and lamps are not written."""
    base = os.path.dirname(os.path.abspath(path))
    os.makedirs(os.path.join(base, "meshes"), exist_ok=True)
    materials = {}
    for m in list(desc.surface_materials) + list(desc.light_materials):
        if m["name"] in materials and materials[m["name"]] != m["nodes"]:
            raise SceneFormatError("two materials named '%s'" % m["name"])
        materials[m["name"]] = m["nodes"]
    shapes, entities = {}, {}
    for si, s in enumerate(desc.shapes):
        used = []
        for mi in s.materials:
            if int(mi) not in used:
                used.append(int(mi))
        slot = {mi: k for k, mi in enumerate(used)}
        name = "shape%d" % si
        rel = os.path.join("meshes", name + ".ply")
        with open(os.path.join(base, rel), "w") as f:
            f.write(_ply_text(s, [slot[int(mi)] for mi in s.materials]))
        shapes[name] = {"type": "ply", "filepath": rel,
                        "materials": [desc.surface_materials[mi]["name"] for mi in used]}
    for k, (si, m, vis, em) in enumerate(desc.entities):
        entities["entity%d" % k] = {
            "shape": "shape%d" % si,
            "visibility": {key: bool(vis & flag) for key, flag in _VISIBILITY},
            "matrix": _flat(m)}
    if lights is None:
        lights = {"World": {"type": "WORLD", "material": desc.light_materials[desc.world_shader]["name"],
                            "visibility": {key: True for key, _ in _VISIBILITY}, "cast_shadows": True,
                            "use_mis": True, "parameters": {}}}
    cam = desc.camera
    camera = None
    if cam is not None:
        # focal = focal_length / (film width / 2) with a 36 mm film
        camera = {"near_clip": cam.get("near", 1e-3), "far_clip": cam.get("far", 1e4),
                  "film": {"width": 36.0, "height": 24.0}, "transform": _flat(cam["transform"]),
                  "shift": list(cam.get("shift", (0.0, 0.0))), "type": "perspective",
                  "focal_length": float(cam["focal"]) * 18.0}
    res = getattr(desc, "resolution", None)
    data = {"materials": materials, "shapes": shapes, "entities": entities, "lights": lights, "camera": camera,
            "render": {"resolution": {"width": res[0], "height": res[1]}} if res else None}
    with open(path, "w") as f:
        json.dump(data, f, indent=2)
