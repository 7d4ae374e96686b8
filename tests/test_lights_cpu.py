"""Closed-form checks of the oracle's analytic lamps (lights/AreaLight.metal:7-38,
PointLight.metal:9-40, SunLight.metal:9-37, SpotLight.metal:9-54).

A camera looks straight down on a diffuse floor in a black world lit by one
lamp; the mean of many samples per pixel is compared with the lamp's
contribution integrated by quadrature over the pixel footprint and the lamp,
which does not use the oracle's samplers, warps or random numbers.  What the
quadrature integrates is the reference's estimator, quirks included:
  * area lamps weigh a sample by color/4 * cos_lamp / distance (the geometry
    term divides the unnormalised direction's dot product by distance^2,
    AreaLight.metal:24-33), uniformly over the lamp's uv square or disk;
  * point and spot lamps weigh by color / (4 pi distance^2); a spot lamp's
    attenuation is the cosine to its axis, zero at or inside spot_size,
    smoothed over spot_blend (SpotLight.metal:9-20);
  * a sun samples its cone uniformly in solid angle with weight color.
The camera path hits the floor, picks the world or the lamp (probability 1/2
each, weight x2) and shades the lamp sample with the Lambertian floor:
pixel = albedo/pi * E(x), with E the lamp term above times cos at the floor.
"""
import math

import numpy as np
import pytest

import raymond_amd.scenes as scenes
from oracle import binding

W = H = 16
SPP = 512
EYE_Z, FOCAL = 2.0, 2.0
ALBEDO = 0.5


def floor_scene(kind, light):
    d = scenes.SceneDescription("floor")
    d.surface_materials = [scenes.diffuse_material("floor", (ALBEDO, ALBEDO, ALBEDO))]
    d.light_materials = [scenes.background_world("world", (0.0, 0.0, 0.0), 1.0),
                         scenes.emission_material("lamp", (1.0, 1.0, 1.0), 1.0)]
    d.world_shader = 0
    d.add_shape(scenes.merge([scenes.quad((-3, -3, 0), (3, -3, 0), (3, 3, 0), (-3, 3, 0), 0)]))
    d.add_entity(0)
    setattr(d, kind + "_lights", [light])
    d.camera = dict(transform=scenes.look_at((0.0, 0.0, EYE_Z), (0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)),
                    near=1e-3, far=1e4, focal=FOCAL)
    d.envmap_resolution = 16
    return d


def render(desc):
    orc = binding.OracleScene(desc)
    orc.build_envmap()
    img, _ = orc.render(W, H, 2, frames=SPP)
    orc.close()
    return img[..., 0] / SPP


def floor_points(sub=6):
    """Floor hit points of sub x sub jittered camera rays per pixel: the camera
    at height EYE_Z looks down -z with +y up, so pixel (x, y) at sub-pixel
    offset (rx, ry) sees (u, v * H/W) * EYE_Z / FOCAL with
    u = 2 (x + rx)/W - 1 and v = 1 - 2 (y + ry)/H (generate.metal:108-121)."""
    r = (np.arange(sub) + 0.5) / sub
    x = np.arange(W)[None, :, None, None] + r[None, None, None, :]
    y = np.arange(H)[:, None, None, None] + r[None, None, :, None]
    u = 2.0 * x / W - 1.0
    v = 1.0 - 2.0 * y / H
    px = np.broadcast_to(u * EYE_Z / FOCAL, (H, W, sub, sub))
    py = np.broadcast_to(v * (H / W) * EYE_Z / FOCAL, (H, W, sub, sub))
    return np.stack([px, py, np.zeros_like(px)], -1)  # (H, W, sub, sub, 3)


def expected(lamp_term):
    """Pixel mean of albedo/pi * lamp_term(x) over the footprint; lamp_term
    returns E at floor points of shape (..., 3)."""
    pts = floor_points()
    E = lamp_term(pts.reshape(-1, 3)).reshape(pts.shape[:-1])
    return (ALBEDO / math.pi) * E.mean(axis=(2, 3))


def disk_nodes(n_r=24, n_t=48):
    """Midpoint polar quadrature of the unit disk: points and weights summing to 1."""
    r_edges = np.linspace(0.0, 1.0, n_r + 1)
    rm = 0.5 * (r_edges[1:] + r_edges[:-1])
    t = (np.arange(n_t) + 0.5) * 2.0 * math.pi / n_t
    R, T = np.meshgrid(rm, t, indexing="ij")
    w = np.repeat(((r_edges[1:] ** 2 - r_edges[:-1] ** 2))[:, None], n_t, 1) / n_t
    return np.stack([R * np.cos(T), R * np.sin(T)], -1).reshape(-1, 2), w.reshape(-1)


def check(img, ref):
    """Image mean within 1.5 %, 4x4 blocks within 6 % of the peak block
    (about 4.5 standard deviations of the light pick's 0-or-2x noise)."""
    assert img.sum() == pytest.approx(ref.sum(), rel=0.015)
    b_img = img.reshape(H // 4, 4, W // 4, 4).mean(axis=(1, 3))
    b_ref = ref.reshape(H // 4, 4, W // 4, 4).mean(axis=(1, 3))
    assert np.abs(b_img - b_ref).max() <= 0.06 * b_ref.max()


def cos_floor(p, x):
    d = p - x
    return np.maximum(d[..., 2], 0.0) / np.linalg.norm(d, axis=-1)


@pytest.mark.parametrize("radius", [0.0, 0.3])
def test_point_lamp(radius):
    S, L = 20.0, np.array([0.3, -0.2, 1.0])
    img = render(floor_scene("point", scenes.point_light(scenes.light_info(1), L, radius, (S, S, S))))

    def term(x):
        if radius == 0.0:
            d2 = ((L - x) ** 2).sum(-1)
            return S / (4.0 * math.pi * d2) * cos_floor(L, x)
        # uniform disk of `radius` facing the receiver (rotation-invariant, so
        # any basis perpendicular to the lamp-to-receiver axis)
        q, w = disk_nodes()
        n = (x - L) / np.linalg.norm(x - L, axis=-1, keepdims=True)
        a = np.cross(n, [1.0, 0.0, 0.0])
        a /= np.linalg.norm(a, axis=-1, keepdims=True)
        b = np.cross(n, a)
        p = L + radius * (q[:, 0, None, None] * a + q[:, 1, None, None] * b)  # (nodes, pts, 3)
        d2 = ((p - x) ** 2).sum(-1)
        return (w[:, None] * S / (4.0 * math.pi * d2) * cos_floor(p, x)).sum(0)

    ref = expected(term)
    assert ref.max() / ref.min() > 1.5  # the test sees the falloff
    check(img, ref)


@pytest.mark.parametrize("tilt,cos_angle", [(0.0, 1.0), (0.0, 0.8), (35.0, 0.95)])
def test_sun_lamp(tilt, cos_angle):
    S = 2.0
    b = math.radians(tilt)
    img = render(floor_scene("sun", scenes.sun_light(scenes.light_info(1), (math.sin(b), 0.0, math.cos(b)),
                                                     cos_angle, (S, S, S))))
    # uniform cone about the sun axis: E[dir] = axis * (1 + cos_angle)/2
    ref = expected(lambda x: np.full(x.shape[0], S * math.cos(b) * (1.0 + cos_angle) / 2.0))
    check(img, ref)


@pytest.mark.parametrize("circular", [False, True])
def test_area_lamp(circular):
    S, center, size = 10.0, np.array([0.2, 0.1, 0.8]), (0.8, 0.5)
    img = render(floor_scene("area", scenes.area_light(scenes.light_info(1), center, size, (S, S, S),
                                                       circular=circular)))
    if circular:
        q, w = disk_nodes()
        uv = 0.5 * q  # disk of radius 1/2 about the uv square's centre
    else:
        g = (np.arange(48) + 0.5) / 48 - 0.5
        U, V = np.meshgrid(g, g, indexing="ij")
        uv = np.stack([U.ravel(), V.ravel()], -1)
        w = np.full(uv.shape[0], 1.0 / uv.shape[0])
    p = center + np.stack([uv[:, 0] * size[0], uv[:, 1] * size[1], np.zeros(len(uv))], -1)  # normal +z

    def term(x):
        d = p[:, None, :] - x[None]
        dist = np.linalg.norm(d, axis=-1)
        cos_lamp = np.clip(d[..., 2] / dist, 0.0, 1.0)
        return (w[:, None] * S * 0.25 * cos_lamp / dist * cos_floor(p[:, None, :], x[None])).sum(0)

    ref = expected(term)
    check(img, ref)


def test_spot_lamp():
    S, L, size, blend = 40.0, np.array([0.0, 0.0, 1.2]), 0.8, 0.1
    img = render(floor_scene("spot", scenes.spot_light(scenes.light_info(1), L, (0.0, 0.0, 1.0), 0.0, (S, S, S),
                                                       size, blend)))

    def term(x):
        to_lamp = L - x
        d2 = (to_lamp ** 2).sum(-1)
        c = to_lamp[..., 2] / np.sqrt(d2)
        t = np.clip((c - size) / blend, 0.0, 1.0)
        att = np.where(c <= size, 0.0, np.where(c - size < blend, c * t * t * (3.0 - 2.0 * t), c))
        return S * att / (4.0 * math.pi * d2) * cos_floor(L, x)

    ref = expected(term)
    assert (ref == 0.0).any() and (ref > 0.0).any()  # the cone's edge crosses the image
    img_out = img[ref == 0.0]
    assert np.all(img_out == 0.0)
    check(img, ref)
