"""Batched jobs on the GPU: every job of a batch renders what it renders alone
(same seeds, same ray counts; images equal up to the order of the
floating-point accumulation), with and without the material sort; per-job
completion is recorded; the autotuner times its candidates and caches the
choice."""
import os

import numpy as np
import pytest

import raymond_amd.api as api
import raymond_amd.scenes as scenes

pytestmark = pytest.mark.gpu

JOBS = [  # (triangles, width, height, samples, depth)
    (0, 48, 32, 3, 5),
    (2000, 32, 32, 5, 8),
    (300, 24, 40, 2, 12),
]


@pytest.fixture(scope="module")
def jobs(device):
    out = []
    for tris, w, h, spp, depth in JOBS:
        d = scenes.cornell_box(w, h) if tris == 0 else scenes.tessellated_cornell_box(w, h, tris)
        d.envmap_resolution = 64
        out.append((scenes.build_device_scene(device, d), w, h, spp, depth))
    return out


def solo(scene, w, h, spp, depth):
    r = api.Renderer(scene, w, h, max_depth=depth)
    st = r.execute(spp)
    return r.image(), st.total_rays


@pytest.mark.parametrize("sort,graph,order", [(0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 0, 0), (1, 1, 0)])
def test_batch_matches_solo_renders(device, jobs, sort, graph, order):
    b = api.Batch(device, [(s, w, h, spp, depth, 0) for s, w, h, spp, depth in jobs])
    st = b.render(queue_paths=4096, material_sort=sort, use_graph=graph, order=order)
    total = 0
    for j, (s, w, h, spp, depth) in enumerate(jobs):
        ref, rays = solo(s, w, h, spp, depth)
        total += rays
        img = b.image(j)
        scale = max(np.abs(ref).max(), 1e-6)
        assert np.abs(img - ref).max() <= 1e-5 * scale, j
    assert st.total_rays == total
    t = b.job_times()
    assert np.all(t >= 0.0) and st.makespan_ms == pytest.approx(t.max())


def test_batch_splits_into_bounce_loops(device, jobs):
    """A queue much smaller than the batch forces several bounce loops; the
    result is the same."""
    b = api.Batch(device, [(s, w, h, spp, depth, 0) for s, w, h, spp, depth in jobs])
    st = b.render(queue_paths=16)
    assert st.loops > 1
    s, w, h, spp, depth = jobs[1]
    ref, _ = solo(s, w, h, spp, depth)
    assert np.abs(b.image(1) - ref).max() <= 1e-5 * max(np.abs(ref).max(), 1e-6)


def test_autotune_times_candidates_and_caches(device, jobs, tmp_path):
    os.environ["RM_TUNE_CACHE"] = str(tmp_path / "tune.txt")
    try:
        b = api.Batch(device, [(s, w, h, spp, depth, 0) for s, w, h, spp, depth in jobs])
        cfg, tried, hit = b.autotune()
        # stage 1: 4 queue depths x sort off/on; stage 2: trace grids and sort bins from the best
        assert not hit and len(tried) >= 8 and all(t[4] > 0 for t in tried)
        assert [(t[0], t[1]) for t in tried[:8]] == [(q, m) for q in (1 << 21, 1 << 22, 1 << 23, 1 << 24) for m in (0, 1)]
        knobs = lambda c: (c.queue_paths, c.material_sort, c.trace_blocks_per_cu, c.sort_key_shift)
        assert knobs(cfg) in [tuple(t[:4]) for t in tried]
        assert max(tried, key=lambda t: t[4])[:4] == knobs(cfg)
        cfg2, tried2, hit2 = b.autotune()
        assert hit2 and not tried2 and knobs(cfg2) == knobs(cfg)
        # the calibration renders leave no trace in the images
        st = b.render(queue_paths=cfg.queue_paths, material_sort=cfg.material_sort,
                      trace_blocks_per_cu=cfg.trace_blocks_per_cu, sort_key_shift=cfg.sort_key_shift)
        s, w, h, spp, depth = jobs[0]
        ref, _ = solo(s, w, h, spp, depth)
        assert np.abs(b.image(0) - ref).max() <= 1e-5 * max(np.abs(ref).max(), 1e-6)
        assert st.total_rays > 0
    finally:
        del os.environ["RM_TUNE_CACHE"]


def test_shortest_first_completes_small_jobs_first(device, jobs):
    """With shortest-job-first slices the jobs complete in order of their
    path counts (each job's last paths end before the next job's start)."""
    b = api.Batch(device, [(s, w, h, spp, depth, 0) for s, w, h, spp, depth in jobs])
    b.render(queue_paths=64, order=api.BATCH_SHORTEST_FIRST)
    t = b.job_times()
    paths = [w * h * spp for _, w, h, spp, _ in jobs]
    order = np.argsort(paths, kind="stable")
    assert np.all(np.diff(t[order]) >= 0.0)
