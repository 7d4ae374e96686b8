"""Lens camera of the shared device library (camera.h, spectral.h) on the CPU.

The reference traces camera rays through the un-vendored "lore" lens library
(generate.metal:16-70,122-163); its surface model and refraction code are not
in the reference tree, so this library's lens model is its own (radius,
thickness, aperture, n_d, V_d per surface) and is pinned here against optics
and colorimetry rather than against lore's outputs (lore itself stays parity
unpinned):

* the stop's opening: the unit circle when fully open, and for a small stop a
  regular polygon with one edge per blade whose area is that of the circle of
  the stop's radius;
* refraction: Snell's law, coplanarity and total internal reflection;
* dispersion: the Cauchy fit returns n_d at the d line and the Abbe number;
* a flat system of apertures: rays pass unbent and the apertures clip them;
* a thick biconvex singlet: rays from its rear focal point leave collimated,
  at the slope -y / f of the thick-lens effective focal length (paraxial
  optics, textbook form);
* the camera weight: its mean over the rear element equals the projected
  solid angle pi a^2 / (h^2 + a^2) of a disc of radius a at distance h, and
  the spectral weight has mean 1 per channel;
* the colour tables: published CIE 1931 values, equal sums of the three
  matching functions, D65 white to sRGB (1, 1, 1).
Planted mistakes show that each check catches a plausible error.
"""
import numpy as np
import pytest

import raymond_amd.scenes as scenes
from tests import devlib

D_LINE = 0.58756  # micrometres


# ---------------------------------------------------------------------------
# Stop

def boundary_radius(phi, relative_stop, blades, rmax, L=None):
    """Radius of the opening's edge along directions phi, by bisection."""
    lo = np.zeros(phi.size)
    hi = np.full(phi.size, rmax)
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        inside = devlib.lens_aperture(np.stack([mid * np.cos(phi), mid * np.sin(phi)], 1), relative_stop, blades, L)
        lo = np.where(inside, mid, lo)
        hi = np.where(inside, hi, mid)
    return 0.5 * (lo + hi)


def check_aperture(L=None):
    rng = np.random.default_rng(1)
    # fully open: the unit circle
    p = rng.uniform(-1.2, 1.2, (200000, 2))
    r = np.hypot(p[:, 0], p[:, 1])
    inside = devlib.lens_aperture(p, 1.0, 7, L)
    clear = np.abs(r - 1.0) > 1e-5
    assert np.array_equal(inside[clear], r[clear] <= 1.0)
    # a small stop: a regular polygon of the circle's area, one edge per blade
    for blades in (5, 6, 9):
        s = 0.002
        phi = np.linspace(0.0, 2 * np.pi, 3600, endpoint=False)
        rb = boundary_radius(phi, s, blades, 3 * s, L)
        area = 0.5 * np.sum(rb ** 2) * (2 * np.pi / phi.size)
        assert abs(area / (np.pi * s * s) - 1.0) < 2e-3, (blades, area / (np.pi * s * s))
        # n-fold symmetry, and edges that are straight: inradius / circumradius = cos(pi / n)
        step = phi.size // blades if phi.size % blades == 0 else None
        if step:
            assert np.allclose(rb, np.roll(rb, step), rtol=1e-4)
        assert abs(rb.min() / rb.max() - np.cos(np.pi / blades)) < 2e-3, blades
        # the edge points between two corners lie on one line
        k = np.argsort(rb)[:1][0]  # an edge midpoint
        half = int(np.pi / blades / (2 * np.pi / phi.size) * 0.9)
        idx = (k + np.arange(-half, half + 1)) % phi.size
        pts = np.stack([rb[idx] * np.cos(phi[idx]), rb[idx] * np.sin(phi[idx])], 1)
        normal = pts[half] / np.linalg.norm(pts[half])
        assert np.ptp(pts @ normal) < 2e-3 * s, blades


def test_stop_opening():
    check_aperture()


# ---------------------------------------------------------------------------
# Refraction and dispersion

def test_refraction_follows_snell():
    rng = np.random.default_rng(2)
    n = 20000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    eta = rng.uniform(0.5, 1.8, n)
    out, ok = devlib.lens_refract(d, nrm, eta)
    d64, n64, o64 = d.astype(np.float32).astype(np.float64), nrm.astype(np.float32).astype(np.float64), \
        out.astype(np.float64)
    eta = eta.astype(np.float32).astype(np.float64)
    # |d x n| keeps its precision near normal incidence
    sin_i = np.linalg.norm(np.cross(d64, n64), axis=1) / (np.linalg.norm(d64, axis=1) * np.linalg.norm(n64, axis=1))
    tir = eta * sin_i > 1.0
    clear = np.abs(eta * sin_i - 1.0) > 1e-4
    assert np.array_equal(ok[clear], ~tir[clear])
    t = ok & clear
    assert np.allclose(np.linalg.norm(o64[t], axis=1), 1.0, atol=1e-6)
    sin_t = np.linalg.norm(np.cross(o64[t], n64[t]), axis=1)
    assert np.allclose(sin_t, eta[t] * sin_i[t], atol=2e-6)
    # in the plane of incidence, on the far side of the surface
    assert np.allclose(np.sum(np.cross(d64[t], n64[t]) * o64[t], axis=1), 0.0, atol=2e-6)
    assert np.all(np.sign(np.sum(o64[t] * n64[t], axis=1)) == np.sign(np.sum(d64[t] * n64[t], axis=1)))


def test_dispersion_returns_abbe_number():
    for nd, vd in ((1.5168, 64.17), (1.7847, 25.72), (1.4585, 67.82)):
        n_d, n_F, n_C = devlib.lens_ior(nd, vd, [D_LINE, 0.48613, 0.65627])
        assert abs(n_d - nd) < 2e-7
        assert abs((n_d - 1.0) / (n_F - n_C) / vd - 1.0) < 1e-4
        assert n_F > n_d > n_C  # normal dispersion
    assert np.all(devlib.lens_ior(1.6, 0.0, [0.4, 0.55, 0.7]) == np.float32(1.6))


# ---------------------------------------------------------------------------
# Geometry

def test_flat_system_passes_rays_unbent_and_clips_at_apertures():
    # object medium, plane (aperture 8) 10 mm before a plane (aperture 5) 20 mm before the image plane
    surf = np.array([[0, 0, 100, 1, 0], [0, 10, 8, 1, 0], [0, 20, 5, 1, 0], [0, 0, 100, 1, 0]], np.float32)
    rng = np.random.default_rng(3)
    n = 50000
    o = np.zeros((n, 3))
    o[:, :2] = rng.uniform(-3, 3, (n, 2))
    d = np.stack([rng.uniform(-0.4, 0.4, n), rng.uniform(-0.4, 0.4, n), -np.ones(n)], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    oo, dd, ok = devlib.lens_trace(surf, devlib.lens_settings(), o, d, D_LINE)
    o64, d64 = o.astype(np.float32).astype(np.float64), d.astype(np.float32).astype(np.float64)
    at = lambda z: o64[:, :2] + d64[:, :2] * (z / -d64[:, 2])[:, None]
    r20, r30 = np.linalg.norm(at(20.0), axis=1), np.linalg.norm(at(30.0), axis=1)
    expect = (r20 <= 5.0) & (r30 <= 8.0)
    clear = (np.abs(r20 - 5.0) > 1e-4) & (np.abs(r30 - 8.0) > 1e-4)
    assert np.array_equal(ok[clear], expect[clear])
    assert 0.2 < expect.mean() < 0.9
    assert np.allclose(dd[ok], d64[ok], atol=1e-6)
    # the exit point lies on the first surface (its vertex is the frame origin)
    assert np.allclose(oo[ok][:, :2], at(30.0)[ok], atol=1e-4)
    assert np.allclose(oo[ok][:, 2], 0.0, atol=1e-4)


def thick_lens(surf):
    """Effective focal length and back focal distance of the singlet from
    paraxial optics: P = P1 + P2 - (d / n) P1 P2, BFD = f (1 - d P1 / n),
    with R1 = +R (centre toward the image) and R2 = -R."""
    n, R, d = float(surf[1, 3]), float(surf[1, 0]), float(surf[1, 1])
    P1 = (n - 1.0) / R
    P2 = (1.0 - n) / float(surf[2, 0])
    P = P1 + P2 - d / n * P1 * P2
    f = 1.0 / P
    return f, f * (1.0 - d * P1 / n)


def check_singlet(L=None):
    surf = scenes.singlet_lens(50.0, 4.0)
    f_eff, bfd = thick_lens(surf)
    # the image plane 2 mm before the rear focal point (BFD behind the back
    # vertex, the stop 5 mm behind it) and the sensor moved back onto it
    # (focus > 0 moves the sensor away from the lens)
    surf[3, 1] = bfd - 5.0 - 2.0
    z_focus = 2.0
    z_stop = -float(surf[3, 1])
    ang = np.linspace(0, 2 * np.pi, 16, endpoint=False)
    for ys in (0.0, 0.4, -0.7):
        aim = np.stack([0.25 * np.cos(ang), 0.25 * np.sin(ang), np.full(ang.size, z_stop)], 1)
        o = np.tile([0.0, ys, z_focus], (ang.size, 1))
        d = aim - o
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        _, dd, ok = devlib.lens_trace(surf, devlib.lens_settings(stop_index=3), o, d, D_LINE, L)
        assert ok.all()
        dd = dd.astype(np.float64)
        slope = dd[:, :2] / -dd[:, 2:3]
        # collimated, at the slope of the ray through the principal points
        assert np.ptp(slope[:, 0]) < 2e-5 and np.ptp(slope[:, 1]) < 2e-5, (ys, np.ptp(slope, axis=0))
        assert abs(slope[:, 1].mean() - (-ys / f_eff)) < 2e-5, (ys, slope[:, 1].mean(), -ys / f_eff)
        assert abs(slope[:, 0].mean()) < 2e-5


def test_singlet_collimates_from_its_rear_focal_point():
    check_singlet()


# ---------------------------------------------------------------------------
# Camera weight

def flat_camera(a=10.0, h=20.0):
    return np.array([[0, 0, 100, 1, 0], [0, h, a, 1, 0], [0, 0, 1e4, 1, 0]], np.float32)


def check_camera_weight(L=None):
    a, h = 10.0, 20.0
    surf = flat_camera(a, h)
    n = 200000
    o, d, w, ok = devlib.lens_camera_ray(surf, devlib.lens_settings(camera_scale=0.5), 1001, 1001, 500, 500, 7, n, L)
    assert ok.mean() > 0.999
    w = np.where(ok[:, None], w, 0.0).astype(np.float64)
    expect = np.pi * a * a / (h * h + a * a)
    se = w[:, 0].std() / np.sqrt(n)
    assert abs(w[:, 0].mean() - expect) < 4 * se + 1e-4, (w[:, 0].mean(), expect)
    assert np.all(w[:, 0] == w[:, 1]) and np.all(w[:, 0] == w[:, 2])
    # rays leave the rear element's plane, in scene units (camera_scale 0.5),
    # toward the object side
    assert np.allclose(o[ok][:, 2], 0.0, atol=1e-5) and np.all(np.hypot(o[ok][:, 0], o[ok][:, 1]) <= 0.5 * a + 1e-4)
    assert np.all(d[ok][:, 2] < 0)


def test_camera_weight_is_projected_solid_angle():
    check_camera_weight()


def test_spectral_weight_has_unit_mean():
    surf = flat_camera()
    n = 400000
    _, _, w, ok = devlib.lens_camera_ray(surf, devlib.lens_settings(spectral=1), 1001, 1001, 500, 500, 8, n)
    _, _, w0, _ = devlib.lens_camera_ray(surf, devlib.lens_settings(spectral=0), 1001, 1001, 500, 500, 8, n)
    ratio = w.astype(np.float64) / w0[:, :1].astype(np.float64)
    for c in range(3):
        m, se = ratio[ok, c].mean(), ratio[ok, c].std() / np.sqrt(ok.sum())
        assert abs(m - 1.0) < 4 * se, (c, m, se)


# ---------------------------------------------------------------------------
# Colour tables

# CIE 1931 2-degree observer, 5 nm table (CIE 15:2004), 4 decimals
CIE_PUBLISHED = {
    450: (0.3362, 0.0380, 1.7721), 500: (0.0049, 0.3230, 0.2720), 555: (0.5121, 1.0000, 0.0057),
    600: (1.0622, 0.6310, 0.0008), 650: (0.2835, 0.1070, 0.0000), 700: (0.0114, 0.0041, 0.0000),
}


def test_colour_tables():
    nm = np.array(sorted(CIE_PUBLISHED), np.float32)
    xyz, _, norm = devlib.spectral(nm)
    assert np.allclose(xyz, np.array([CIE_PUBLISHED[int(v)] for v in nm]), atol=1e-6)
    grid = np.arange(380, 781, 5).astype(np.float32)
    tab = devlib.spectral(grid[:-1])[0].astype(np.float64)
    sums = tab.sum(axis=0)
    # equal-energy white has x = y = z
    assert np.ptp(sums) / sums.mean() < 2e-4, sums
    assert abs(sums[1] - 21.371) < 2e-3, sums
    # between table rows: linear interpolation; outside 380-780 nm: zero
    mid = devlib.spectral(grid[:-1] + 2.5)[0].astype(np.float64)
    assert np.allclose(mid[:-1], 0.5 * (tab[:-1] + tab[1:]), atol=1e-6)
    assert np.all(devlib.spectral([780.0, 800.0, 370.0])[0] == 0.0)
    # D65 white to linear sRGB (1, 1, 1)
    assert np.allclose(devlib.xyz_to_rgb([0.95047, 1.0, 1.08883]), 1.0, atol=2e-4)
    # white spectrum: mean over 380-780 nm of rgb * norm * 400 nm is 1 per channel
    fine = (380.0 + (np.arange(400000) + 0.5) * 1e-3).astype(np.float32)
    rgb = devlib.spectral(fine)[1].astype(np.float64)
    assert np.allclose(rgb.mean(axis=0) * norm * 400.0, 1.0, atol=2e-4), rgb.mean(axis=0) * norm * 400.0


# ---------------------------------------------------------------------------
# Planted mistakes

PLANTS = [
    ("refraction with the index ratio inverted", "if (!lens_refract(d, normal, n2 / n1)) return false;",
     "if (!lens_refract(d, normal, n1 / n2)) return false;", check_singlet),
    ("sphere cap on the far half", "t = ((z0 - s.radius) * s.radius <= 0.0f) ? t0 : t1;",
     "t = ((z0 - s.radius) * s.radius <= 0.0f) ? t1 : t0;", check_singlet),
    ("thickness of the surface left behind", "o.z += s.thickness;", "o.z += L.surfaces[i - 1].thickness;",
     check_singlet),
    ("sensor pdf without the cosine", "const float inv_pdf = (abs_(dir.z) * inv_d2) *",
     "const float inv_pdf = (inv_d2) *", check_camera_weight),
    ("blades without the area normalisation", "const float norm = sqrt_(tan_(x) / x);", "const float norm = 1.0f;",
     check_aperture),
]


@pytest.mark.parametrize("name,old,new,check", PLANTS, ids=[p[0] for p in PLANTS])
def test_planted_mistakes_fail(name, old, new, check):
    L = devlib.lib(plant=("camera.h", old, new))
    with pytest.raises(AssertionError):
        check(L)
