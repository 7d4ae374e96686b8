"""GPU parity against the CPU oracle: traversal hits bit-exact, per-sample
radiance within 1e-4 relative per pixel (tests/parity.py), identical per-depth
ray counts."""
import numpy as np
import pytest

import raymond_amd.api as api
import raymond_amd.scenes as scenes
from oracle import binding
from tests import parity

pytestmark = pytest.mark.gpu


def random_rays(n, lo, hi, seed, tmin=0.0, tmax=np.inf):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


def edge_rays(desc, n, seed):
    """Rays aimed at triangle edges and vertices of the scene (watertightness, ties)."""
    rng = np.random.default_rng(seed)
    verts, _, _, faces, _, inst = desc.arrays()
    targets = []
    for d in inst:
        m = np.array(d["transform"], np.float64).reshape(4, 4).T
        f = faces[d["face_offset"]:d["face_offset"] + d["face_count"]].astype(np.int64) + d["vertex_offset"]
        v = verts[f]  # (F, 3, 3)
        vw = v @ m[:3, :3].T + m[:3, 3]
        k = rng.integers(0, f.shape[0], n // len(inst) + 1)
        a = rng.integers(0, 3, k.size)
        t = rng.uniform(0, 1, k.size)
        p0 = vw[k, a]
        p1 = vw[k, (a + 1) % 3]
        targets.append(p0 * (1 - t[:, None]) + p1 * t[:, None])
        targets.append(vw[k, a])
    tgt = np.concatenate(targets)[:n]
    o = tgt + rng.normal(size=tgt.shape) * 0.7
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((tgt.shape[0], 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = 0.0
    r[:, 4:7] = d
    r[:, 7] = np.inf
    return r


assert_hits_equal = parity.assert_hits_equal


@pytest.fixture(scope="module")
def cornell(device):
    desc = scenes.cornell_box(64, 64)
    desc.envmap_resolution = 256
    scene = scenes.build_device_scene(device, desc)
    orc = binding.OracleScene(desc)
    orc.build_envmap()
    return desc, scene, orc


def test_device_is_gfx950(device):
    info = device.info()
    assert info.arch.decode().startswith("gfx950"), info.arch
    assert info.wavefront_size == 64


def test_closest_cornell_matches_brute_force(cornell):
    desc, scene, orc = cornell
    rays = np.concatenate([random_rays(20000, 0.02, 0.98, 1), edge_rays(desc, 20000, 2)])
    gpu = scene.trace_closest(rays)
    ref = orc.trace_closest(rays, brute_force=True)
    assert_hits_equal(gpu, ref)
    assert np.mean(gpu[0] >= 0) > 0.5


def axis_rays(n, seed):
    """Rays with zero direction components: the six axis directions and
    directions in the coordinate planes, from inside the box and from points
    exactly on its wall planes (rays that lie in a wall or graze an edge)."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(0.02, 0.98, (n, 3))
    on_wall = rng.random(n) < 0.5
    axis = rng.integers(0, 3, n)
    side = rng.integers(0, 2, n).astype(np.float64)
    o[on_wall, axis[on_wall]] = side[on_wall]
    d = np.zeros((n, 3))
    kind = rng.integers(0, 2, n)
    a1 = rng.integers(0, 3, n)
    sign = np.where(rng.random(n) < 0.5, -1.0, 1.0)
    d[np.arange(n), a1] = sign                              # kind 0: an axis
    a2 = (a1 + 1 + rng.integers(0, 2, n)) % 3              # kind 1: a coordinate plane
    ang = rng.uniform(0, 2 * np.pi, n)
    pl = kind == 1
    d[pl] = 0.0
    d[pl, a1[pl]] = np.cos(ang[pl])
    d[pl, a2[pl]] = np.sin(ang[pl])
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = 0.0
    r[:, 4:7] = d
    r[:, 7] = np.inf
    return r


def test_axis_aligned_rays_match_brute_force(cornell):
    desc, scene, orc = cornell
    rays = axis_rays(20000, 7)
    assert (rays[:, 4:7] == 0.0).any(axis=1).all()
    gpu = scene.trace_closest(rays)
    ref = orc.trace_closest(rays, brute_force=True)
    assert_hits_equal(gpu, ref)
    assert np.mean(gpu[0] >= 0) > 0.5
    short = rays.copy()
    short[:, 7] = 0.3
    assert np.array_equal(scene.trace_any(short), orc.trace_any(short, brute_force=True))


def test_any_cornell_matches_brute_force(cornell):
    desc, scene, orc = cornell
    rays = random_rays(20000, 0.02, 0.98, 3, tmin=1e-3, tmax=0.5)
    assert np.array_equal(scene.trace_any(rays), orc.trace_any(rays, brute_force=True))


def test_render_cornell_per_sample(cornell):
    desc, scene, orc = cornell
    W, H = 64, 64
    r = api.Renderer(scene, W, H, max_depth=8)
    ref_total = np.zeros((H, W, 4), np.float32)
    r.set_uniforms(accumulate=0)  # each call's image holds that call's samples
    total = np.zeros((H, W, 4), np.float32)
    for frame in range(3):
        st = r.execute(1)
        img = r.image()
        sample = img
        total += img
        ref, res = orc.render(W, H, 8, frames=1, frame_begin=frame)
        ref_total += ref
        assert [st.rays[d] for d in range(8)] == [res.rays[d] for d in range(8)]
        assert [st.shadow_rays[d] for d in range(7)] == [res.shadow_rays[d] for d in range(7)]
        parity.assert_radiance_close(sample, ref)


def test_render_cornell_config_mode_rr3(cornell):
    desc, scene, orc = cornell
    W, H = 64, 64
    r = api.Renderer(scene, W, H, max_depth=8)
    r.set_uniforms(rr_depth=3)
    st = r.execute(1)
    ref, res = orc.render(W, H, 8, frames=1, rr_depth=3)
    assert st.total_rays == res.total_rays
    parity.assert_radiance_close(r.image(), ref)


@pytest.fixture(scope="module")
def small_spheres(device):
    desc = scenes.sphere_field(96, 54, grid=2, level=3)
    desc.envmap_resolution = 256
    scene = scenes.build_device_scene(device, desc)
    orc = binding.OracleScene(desc)
    orc.build_envmap()
    return desc, scene, orc


def test_envmap_bit_exact(small_spheres):
    desc, scene, orc = small_spheres
    gp, gm = scene.read_envmap()
    op, om = orc.read_envmap()
    assert np.array_equal(gp.view(np.uint32), op.view(np.uint32))
    assert np.array_equal(gm.view(np.uint32), om.view(np.uint32))


def test_closest_spheres_matches_brute_force(small_spheres):
    desc, scene, orc = small_spheres
    rays = np.concatenate([random_rays(20000, -2.0, 2.0, 5), edge_rays(desc, 20000, 6)])
    assert_hits_equal(scene.trace_closest(rays), orc.trace_closest(rays, brute_force=True))


def test_render_spheres_per_sample(small_spheres):
    desc, scene, orc = small_spheres
    W, H = 96, 54
    r = api.Renderer(scene, W, H, max_depth=12)
    st = r.execute(1)
    ref, res = orc.render(W, H, 12, frames=1)
    assert st.total_rays == res.total_rays
    parity.assert_radiance_close(r.image(), ref)


@pytest.mark.parametrize("spl,samples,graph", [(1, 3, 0), (2, 5, 0), (2, 5, 1)])
def test_regenerating_stream_matches_frames(cornell, spl, samples, graph):
    """A stream of `samples` spp through a queue of spl frames (paths are
    regenerated as others terminate) equals the sum of per-frame oracle
    renders, launched directly or replayed from a captured graph."""
    desc, scene, orc = cornell
    W, H = 64, 64
    r = api.Renderer(scene, W, H, max_depth=8, samples_per_launch=spl)
    r.set_uniforms(use_graph=graph)
    st = r.execute(samples)
    ref, res = orc.render(W, H, 8, frames=samples, frame_begin=0)
    assert st.total_rays == res.total_rays
    assert st.iterations > 8
    parity.assert_radiance_close(r.image(), ref)


@pytest.fixture(scope="module")
def graphs(device):
    """The shader-graph scene (random 20-60 node graphs per triangle) with 8 graphs."""
    desc = scenes.shader_graph_scene(48, 48, materials=8, seed=3)
    desc.envmap_resolution = 256
    scene = scenes.build_device_scene(device, desc)
    orc = binding.OracleScene(desc)
    orc.build_envmap()
    return desc, scene, orc


@pytest.mark.parametrize("sort,grid,shift", [(0, 0, 0), (1, 0, 0), (1, 1, 3)])
def test_render_graphs_per_sample(graphs, sort, grid, shift):
    """Per sample against the oracle, unsorted, sorted by material, and with
    the tuner's other knobs: one persistent trace workgroup per CU and 8
    materials per sort bin."""
    desc, scene, orc = graphs
    W, H = 48, 48
    r = api.Renderer(scene, W, H, max_depth=8)
    r.set_uniforms(material_sort=sort, trace_blocks_per_cu=grid, sort_key_shift=shift)
    r.set_uniforms(accumulate=0)  # each call's image holds that call's samples
    total = np.zeros((H, W, 4), np.float32)
    for frame in range(2):
        st = r.execute(1)
        img = r.image()
        sample = img
        total += img
        ref, res = orc.render(W, H, 8, frames=1, frame_begin=frame)
        assert [st.rays[d] for d in range(8)] == [res.rays[d] for d in range(8)]
        assert [st.shadow_rays[d] for d in range(7)] == [res.shadow_rays[d] for d in range(7)]
        parity.assert_radiance_close(sample, ref)


def test_material_sort_matches_unsorted(graphs):
    """Sorting only permutes the shading order: every path does the same work,
    so ray counts are equal and the images differ only by the order of the
    floating-point accumulation (relative 1e-5)."""
    desc, scene, orc = graphs
    W, H = 48, 48
    out = []
    for sort in (0, 1):
        r = api.Renderer(scene, W, H, max_depth=8, samples_per_launch=2)
        r.set_uniforms(material_sort=sort)
        st = r.execute(6)
        out.append((st.total_rays, st.iterations, r.image()))
    assert out[0][0] == out[1][0]
    assert out[0][1] == out[1][1]
    parity.assert_radiance_close(out[1][2], out[0][2], rtol=1e-5, what="sorted vs unsorted")


def test_graph_replay_matches_direct_launches(cornell):
    """Several streams, each captured once and replayed chunk by chunk with
    host work between the replays, match direct launches iteration by
    iteration (ray counts) and in the image."""
    desc, scene, orc = cornell
    W, H = 64, 64
    out = []
    for graph in (0, 1):
        r = api.Renderer(scene, W, H, max_depth=8, samples_per_launch=1)
        r.set_uniforms(use_graph=graph, rr_depth=3)
        counts = []
        for _ in range(3):
            st = r.execute(4)
            counts.append([st.rays[d] for d in range(33)] + [st.shadow_rays[d] for d in range(32)])
        out.append((counts, r.image()))
    assert out[0][0] == out[1][0]
    parity.assert_radiance_close(out[1][1], out[0][1], rtol=1e-5, what="graph replay vs direct")


@pytest.fixture(scope="module")
def lamps(device):
    """The Cornell box lit by a rectangular and a circular area lamp, a point
    lamp with a radius, a spot lamp and a sun, through two lamp shaders."""
    desc = scenes.lamp_box(48, 48)
    desc.envmap_resolution = 64
    scene = scenes.build_device_scene(device, desc)
    orc = binding.OracleScene(desc)
    orc.build_envmap()
    return desc, scene, orc


@pytest.mark.parametrize("mode", [2, 1, 0])
def test_render_lamps_per_sample(lamps, mode):
    """Every analytic lamp type (sampling, lamp shader, shadow flag, MIS flag)
    under MIS, light sampling only and BSDF sampling only."""
    desc, scene, orc = lamps
    W, H = 48, 48
    r = api.Renderer(scene, W, H, max_depth=6)
    r.set_uniforms(sampling_mode=mode)
    r.set_uniforms(accumulate=0)  # each call's image holds that call's samples
    total = np.zeros((H, W, 4), np.float32)
    for frame in range(2):
        st = r.execute(1)
        img = r.image()
        sample = img
        total += img
        ref, res = orc.render(W, H, 6, frames=1, frame_begin=frame, sampling_mode=mode)
        assert [st.rays[d] for d in range(6)] == [res.rays[d] for d in range(6)]
        assert [st.shadow_rays[d] for d in range(5)] == [res.shadow_rays[d] for d in range(5)]
        parity.assert_radiance_close(sample, ref)
        if mode != 0:
            assert ref.max() > 0.0


def test_batch_of_lamp_scenes_matches_solo(device, lamps):
    """Each job of a batch keeps its own lamps."""
    desc, scene, orc = lamps
    d2 = scenes.lamp_box(32, 24)
    d2.point_lights = []
    d2.envmap_resolution = 64
    s2 = scenes.build_device_scene(device, d2)
    jobs = [(scene, 48, 48, 2, 6, 0), (s2, 32, 24, 3, 5, 0)]
    b = api.Batch(device, jobs)
    b.render(queue_paths=4096)
    for j, (s, w, h, spp, depth, _) in enumerate(jobs):
        r = api.Renderer(s, w, h, max_depth=depth)
        r.execute(spp)
        ref = r.image()
        assert np.abs(b.image(j) - ref).max() <= 1e-5 * max(np.abs(ref).max(), 1e-6), j


def _hable_partial(x):
    A, B, C, D, E, F = (np.float32(v) for v in (0.15, 0.50, 0.10, 0.20, 0.02, 0.30))
    return (x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F) - E / F


@pytest.mark.parametrize("tonemap", [0, 1, 2])
def test_resolve_matches_normalize_image(cornell, tonemap):
    """rm_resolve against a float32 restatement of normalizeImage
    (device/kernels/utils/normalizeImage.metal:11-59): exposure, division by the
    accumulated sample count, then linear / Hable / ACES."""
    desc, scene, orc = cornell
    r = api.Renderer(scene, 64, 64, max_depth=4)
    r.set_uniforms(exposure=0.7, tonemapping=tonemap)
    r.execute(3)
    raw = r.image()[..., :3].astype(np.float32)
    c = (np.float32(0.7) * raw) / np.float32(3.0)
    if tonemap == 1:
        c = _hable_partial(c * np.float32(2.0)) * (np.float32(1.0) / _hable_partial(np.float32(11.2)))
    elif tonemap == 2:
        a, b, cc, d, e = (np.float32(v) for v in (2.51, 0.03, 2.43, 0.59, 0.14))
        c = np.clip((c * (a * c + b)) / (c * (cc * c + d) + e), 0.0, 1.0)
    out = r.resolve()
    assert np.all(out[..., 3] == 1.0)
    assert np.allclose(out[..., :3], c, rtol=2e-6, atol=1e-7)
    assert c.max() > 0.05


def test_stage_timers(cornell):
    """timing=1 records per-stage GPU times (the reference's stage timers,
    Renderer.swift) without changing what is rendered."""
    desc, scene, orc = cornell
    out = []
    for timing in (0, 1):
        r = api.Renderer(scene, 64, 64, max_depth=8, samples_per_launch=2)
        r.set_uniforms(timing=timing)
        st = r.execute(4)
        out.append((st, r.image()))
    (st0, img0), (st1, img1) = out
    assert st0.total_rays == st1.total_rays and st0.iterations == st1.iterations
    assert np.abs(img0 - img1).max() <= 1e-5 * max(np.abs(img0).max(), 1e-6)
    assert st0.ms_trace_closest == 0.0 and st0.ms_shade == 0.0
    for v in (st1.ms_trace_closest, st1.ms_shade, st1.ms_trace_any):
        assert v > 0.0
    assert st1.ms_trace_closest + st1.ms_shade + st1.ms_trace_any <= st1.ms_total
    assert st1.trace_closest_launches == st1.iterations
