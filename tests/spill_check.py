"""Parity run of the library built with one LDS row of traversal stack
(raymond_amd/libraymond_amd_spill.so), in a process of its own: the library
is chosen when raymond_amd.api is imported.  With one LDS row every stack
entry below the first lives in the global spill rows, so this exercises
their indexing (row stride = the persistent grid's thread count) and the
sizing of the spill buffer for scenes, raw ray batches, renders, batches
and two-level scenes.

Run by tests/test_full_size_gpu.py as `python -m tests.spill_check`; prints
one JSON line.
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main():
    import raymond_amd.api as api
    import raymond_amd.scenes as scenes
    from oracle import binding
    from tests import parity
    from tests.test_gpu_parity import axis_rays, edge_rays, random_rays

    assert os.path.basename(api.LIB_PATH) == "libraymond_amd_spill.so", api.LIB_PATH
    dev = api.Device(0)
    out = {}
    kept = []
    # BVH8 depth >= 3: with one LDS row, the second stack entry is the first
    # in the spill rows (a depth-2 tree pushes at its root level only)
    for name, desc, lo, hi in (("tessellated", scenes.tessellated_cornell_box(48, 48, 2000), 0.02, 0.98),
                               ("spheres", scenes.sphere_field(64, 36, grid=3, level=4), -2.5, 2.5)):
        desc.envmap_resolution = 64
        scene = scenes.build_device_scene(dev, desc)
        orc = binding.OracleScene(desc)
        orc.build_envmap(threads=parity.ORACLE_THREADS)
        depth = scene.stats().bvh_depth
        rays = np.concatenate([random_rays(20000, lo, hi, 1), edge_rays(desc, 20000, 2), axis_rays(4000, 3)])
        parity.assert_hits_equal(scene.trace_closest(rays), orc.trace_closest(rays, brute_force=True))
        short = random_rays(20000, lo, hi, 4, tmin=1e-3, tmax=0.7)
        assert np.array_equal(scene.trace_any(short), orc.trace_any(short, brute_force=True)), name
        W, H = desc.width, desc.height
        r = api.Renderer(scene, W, H, max_depth=8, samples_per_launch=2)
        r.set_uniforms(accumulate=0)
        st = r.execute(3)
        ref, res = orc.render(W, H, 8, frames=3, threads=parity.ORACLE_THREADS)
        assert st.total_rays == res.total_rays, (name, st.total_rays, res.total_rays)
        worst = parity.assert_radiance_close(r.image(), ref, what=name)
        out[name] = dict(bvh_depth=int(depth), rays=int(rays.shape[0]), worst_rel=worst)
        kept.append((name, desc, scene, orc))
    # two-level: the marker entries and both trees' groups in the spill rows
    inst = scenes.instanced_field(48, 36, grid=4, shapes=2, level=2)
    inst.envmap_resolution = 64
    scene = scenes.build_device_scene(dev, inst, two_level=True)
    orc = binding.OracleScene(inst)
    orc.build_envmap(threads=parity.ORACLE_THREADS)
    orc.set_two_level(True)
    cam = parity.camera_rays(inst, 48, 36, seed=5)
    ref = orc.trace_closest(cam, threads=parity.ORACLE_THREADS)
    parity.assert_hits_equal(scene.trace_closest(cam), ref)
    shadow = parity.bounce_rays(cam, ref[0], seed=6, count=4096, tmax=2.0)
    assert np.array_equal(scene.trace_any(shadow), orc.trace_any(shadow, threads=parity.ORACLE_THREADS))
    r = api.Renderer(scene, 48, 36, max_depth=6)
    r.set_uniforms(accumulate=0)
    st = r.execute(1)
    img, res = orc.render(48, 36, 6, frames=1, threads=parity.ORACLE_THREADS)
    assert st.total_rays == res.total_rays, ("two-level", st.total_rays, res.total_rays)
    out["two_level"] = dict(bvh_depth=int(scene.stats().bvh_depth), rays=int(cam.shape[0] + shadow.shape[0]),
                            worst_rel=parity.assert_radiance_close(r.image(), img, what="two-level"))
    # a batch keeps its own spill rows for the merged scene; its jobs share
    # one material library, so the second job is the plain Cornell box
    plain = scenes.cornell_box(32, 40)
    plain.envmap_resolution = 64
    plain_orc = binding.OracleScene(plain)
    plain_orc.build_envmap(threads=parity.ORACLE_THREADS)
    batch = [kept[0], ("cornell", plain, scenes.build_device_scene(dev, plain), plain_orc)]
    jobs = [(scene, desc.width, desc.height, 2, 6, 0) for (_, desc, scene, _) in batch]
    b = api.Batch(dev, jobs)
    b.render(queue_paths=4096)
    for j, (name, desc, scene, orc) in enumerate(batch):
        ref, _ = orc.render(desc.width, desc.height, 6, frames=2, threads=parity.ORACLE_THREADS)
        out.setdefault(name, {})["batch_worst_rel"] = parity.assert_radiance_close(b.image(j), ref,
                                                                                   what="batch " + name)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
