"""Radiance cache on the GPU: the MFMA forward and backward kernels against the
double-precision oracle, the Adam step against its closed form, online
training, and the cache inside the renderer."""
import numpy as np
import pytest

import raymond_amd.api as api
import raymond_amd.scenes as scenes
from oracle import binding
from tests.test_rcache_cpu import numpy_gradients, random_features, random_params

pytestmark = pytest.mark.gpu

LOG2 = 14


@pytest.fixture(scope="module")
def cache(device):
    return api.RadianceCache(device, log2_table_size=LOG2, seed=1)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_infer_matches_oracle(cache):
    rng = np.random.default_rng(10)
    mlp, table = random_params(rng, LOG2)
    cache.write_params(mlp, table)
    f = random_features(rng, 1000)
    got = cache.infer(f)
    ref = binding.rc_forward(LOG2, mlp, table, f)
    # FP16 encodings and activations, FP32 accumulation
    assert rel_l2(got, ref) < 1e-2
    assert np.abs(got - ref).max() < 2e-2 * np.abs(ref).max()


def test_gradients_match_oracle(cache):
    rng = np.random.default_rng(11)
    mlp, table = random_params(rng, LOG2)
    cache.write_params(mlp, table)
    f = random_features(rng, 256)
    y = binding.rc_forward(LOG2, mlp, table, f)
    t = (y + rng.normal(size=y.shape) * 0.5).astype(np.float32)
    t[7] = np.nan  # unusable sample: contributes nothing
    gm, gt, loss = cache.gradients(f, t)
    rm, rt, rloss = binding.rc_gradients(LOG2, mlp, table, f, t)
    assert abs(loss - rloss) < 1e-2 * rloss
    w = np.ones(9216, bool)
    w3 = np.zeros((64, 32), bool)
    w3[:, :3] = True
    w[7168:] = w3.ravel()
    assert np.all(gm[~w] == 0.0)
    check_gradients(gm, gt, rm, rt, numpy_gradients(LOG2, mlp, table, f, t, fp16=True))
    assert np.array_equal(gt != 0.0, rt != 0.0) or (np.sum((gt != 0) != (rt != 0)) < 1e-3 * gt.size)


def check_gradients(gm, gt, rm, rt, half, log=print):
    """GPU gradients against the numpy restatement that rounds where the
    kernels hold binary16 values (bound 5e-3: accumulation order and the odd
    binary16 rounding that float32 and float64 sums put on different sides),
    and against the double oracle within that restatement's own distance
    from it plus the same 5e-3.  Relative L2 per block."""
    hm, ht, _ = half
    blocks = (("W1", slice(0, 3072)), ("W2", slice(3072, 7168)), ("W3", slice(7168, 9216)))
    for name, got, half_ref, ref in [(n, gm[sl], hm[sl], rm[sl]) for n, sl in blocks] + [("table", gt, ht, rt)]:
        e_half, e_ref, e_model = rel_l2(got, half_ref), rel_l2(got, ref), rel_l2(half_ref, ref)
        log("%s gradient: relative L2 %.3g to the binary16 restatement, %.3g to the oracle (restatement %.3g)"
            % (name, e_half, e_ref, e_model))
        assert e_half < 5e-3, name
        assert e_ref < e_model + 5e-3, name


def test_adam_first_step_closed_form(cache):
    """First step from zero moments: m^ = g, v^ = g^2, so p1 = p0 - lr g / (|g| + eps)."""
    rng = np.random.default_rng(12)
    dev = cache.device
    c = api.RadianceCache(dev, log2_table_size=LOG2, seed=2, learning_rate=1e-2)
    mlp, table = random_params(rng, LOG2)
    c.write_params(mlp, table)
    f = random_features(rng, 512)
    t = rng.uniform(0, 1, (512, 3)).astype(np.float32)
    gm, gt, _ = c.gradients(f, t)
    p0m, p0t = c.read_params()
    c.train(f, t, steps=1)
    p1m, p1t = c.read_params()
    for p0, p1, g in ((p0m, p1m, gm), (p0t, p1t, gt)):
        big = np.abs(g) > 1e-6
        expect = p0[big] - 1e-2 * g[big] / (np.abs(g[big]) + 1e-15)
        assert np.abs(p1[big] - expect).max() < 1e-6
        assert np.array_equal(p1[g == 0.0], p0[g == 0.0])
    assert c.stats().steps == 1


def test_training_fits_a_smooth_function(device):
    rng = np.random.default_rng(13)
    c = api.RadianceCache(device, log2_table_size=LOG2, seed=3, batch_size=4096)
    f = random_features(rng, 4096)
    p = f[:, 0:3]
    t = np.stack([0.5 + 0.4 * np.sin(6 * p[:, 0]), 0.5 + 0.4 * np.cos(5 * p[:, 1]), p[:, 2] ** 2], 1)
    t = t.astype(np.float32)
    first = c.train(f, t, steps=1)
    last = c.train(f, t, steps=300)
    assert last < 0.05 * first, (first, last)
    y = c.infer(f)
    assert np.sqrt(np.mean((y - t) ** 2)) < 0.05


@pytest.fixture(scope="module")
def cornell_scene(device):
    desc = scenes.cornell_box(64, 64)
    desc.envmap_resolution = 64
    return scenes.build_device_scene(device, desc)


def test_render_with_cache(device, cornell_scene):
    W = H = 64
    plain = api.Renderer(cornell_scene, W, H, max_depth=12)
    plain.set_uniforms(rr_depth=3)
    sp = plain.execute(4)
    r = api.Renderer(cornell_scene, W, H, max_depth=12)
    r.set_uniforms(rr_depth=3)
    c = api.RadianceCache(device, seed=4, batch_size=1024)
    r.set_cache(c, mode=1)
    st = r.execute(4)
    cs = c.stats()
    img = r.image()
    assert np.all(np.isfinite(img))
    assert cs.queries > 0 and cs.train_records > 0 and cs.steps == 4 * 2
    # the last in-render step's relative L2 loss, read back by the stats call
    assert np.isfinite(cs.last_loss) and cs.last_loss > 0.0
    # paths end at their first diffuse vertex, so far fewer rays are traced
    assert st.total_rays < 0.7 * sp.total_rays
    # frozen cache: no training vertices, no steps
    r2 = api.Renderer(cornell_scene, W, H, max_depth=12)
    r2.set_cache(c, mode=2)
    r2.execute(1)
    assert c.stats().steps == cs.steps and c.stats().train_records == cs.train_records


def test_cache_reduces_error_at_equal_samples(device, cornell_scene):
    """What the trained cache gains: after training, a few samples per pixel
    with the cache have a lower error against a converged plain render than
    the same number of plain samples."""
    W = H = 64
    ref_r = api.Renderer(cornell_scene, W, H, max_depth=12, samples_per_launch=4)
    ref_r.set_uniforms(rr_depth=3, accumulate_f64=1)
    ref_r.execute(4096)
    ref = ref_r.image_f64()[..., :3] / 4096.0
    n = 4
    plain = api.Renderer(cornell_scene, W, H, max_depth=12)
    plain.set_uniforms(rr_depth=3)
    plain.execute(n)
    err_plain = np.sqrt(np.mean((plain.image()[..., :3] / n - ref) ** 2))
    c = api.RadianceCache(device, seed=5, batch_size=2048)
    tr = api.Renderer(cornell_scene, W, H, max_depth=12)
    tr.set_uniforms(rr_depth=3)
    tr.set_cache(c, mode=1)
    tr.execute(256)  # online training
    r = api.Renderer(cornell_scene, W, H, max_depth=12)
    r.set_uniforms(rr_depth=3)
    r.set_cache(c, mode=2)
    r.execute(n)
    err_cache = np.sqrt(np.mean((r.image()[..., :3] / n - ref) ** 2))
    assert err_cache < err_plain, (err_cache, err_plain)
