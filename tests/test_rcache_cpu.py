"""Radiance-cache oracle on the CPU: FP16 rounding, grid geometry and index
rule against independent numpy restatements, the encoding at grid vertices,
the forward and backward passes against numpy matrix products, and the
analytic gradients against central differences of the oracle's own loss.
The numpy restatement also rounds, on request, where the GPU kernels hold
binary16 values; the GPU tests compare with that."""
import numpy as np
import pytest

from oracle import binding

LOG2 = 12


def test_round_half_matches_numpy():
    rng = np.random.default_rng(1)
    vals = np.concatenate([
        rng.normal(size=2000) * 10.0 ** rng.integers(-9, 5, 2000),
        [0.0, -0.0, 65504.0, 65519.0, 65520.0, 1e-8, 3e-8, 6e-8, 2.0 ** -24, 2.0 ** -25, 2.0 ** -14, 1.0 + 2.0 ** -11,
         1.0 + 3 * 2.0 ** -11, -1.0 - 2.0 ** -11],
    ]).astype(np.float32)
    with np.errstate(over="ignore"):
        ref = vals.astype(np.float16).astype(np.float32)
    got = binding.rc_round_half(vals)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("log2", [12, 19])
def test_grid_levels(log2):
    res, dense = binding.rc_grid(log2)
    assert res[0] == 16 and res[-1] == 2048
    ratio = res[1:] / res[:-1]
    b = (2048 / 16) ** (1 / 15)
    assert np.all(np.abs(ratio - b) < 0.05)
    assert np.array_equal(dense, ((res.astype(np.uint64) + 1) ** 3 <= 2 ** log2).astype(np.uint32))


@pytest.mark.parametrize("log2", [12, 19])
def test_index_rule(log2):
    res, dense = binding.rc_grid(log2)
    rng = np.random.default_rng(2)
    for level in range(16):
        for _ in range(40):
            x, y, z = (int(v) for v in rng.integers(0, res[level] + 1, 3))
            if dense[level]:
                n = int(res[level]) + 1
                ref = x + n * (y + n * z)
            else:
                ref = ((x * 1) ^ ((y * 2654435761) & 0xFFFFFFFF) ^ ((z * 805459861) & 0xFFFFFFFF)) & ((1 << log2) - 1)
            assert binding.rc_index(log2, level, x, y, z) == ref


def random_params(rng, log2, table_scale=0.5):
    """FP16-exact masters so the oracle's rounding is the identity."""
    mlp = np.zeros(9216, np.float32)
    mlp[:42 * 64] = rng.uniform(-0.4, 0.4, 42 * 64)
    mlp[3072:3072 + 4096] = rng.uniform(-0.3, 0.3, 4096)
    w3 = np.zeros((64, 32), np.float32)
    w3[:, :3] = rng.uniform(-0.3, 0.3, (64, 3))
    mlp[7168:] = w3.ravel()
    table = rng.uniform(-table_scale, table_scale, 16 * 2 ** log2 * 2).astype(np.float32)
    return mlp.astype(np.float16).astype(np.float32), table.astype(np.float16).astype(np.float32)


def random_features(rng, n):
    f = np.zeros((n, 12), np.float32)
    f[:, 0:3] = rng.uniform(0, 1, (n, 3))
    f[:, 3:12] = rng.uniform(0, 1, (n, 9))
    return f


def numpy_encode(log2, table, f):
    """Independent restatement of the input layer in float64: the trilinear
    hash-grid encoding of 16 levels x 2 features, the 9 further features and
    a constant 1.  Also returns each (level, corner)'s table entry and weight,
    shape (16, 8, n), for the table gradient."""
    res, dense = binding.rc_grid(log2)
    T = 2 ** log2
    tab = table.reshape(16, T, 2).astype(np.float64)
    n = f.shape[0]
    x = np.zeros((n, 48))
    idxs = np.zeros((16, 8, n), np.int64)
    ws = np.zeros((16, 8, n))
    for l in range(16):
        p = f[:, 0:3].astype(np.float32) * np.float32(res[l])
        c = np.clip(np.floor(p), 0, res[l] - 1)
        fr = (p - c).astype(np.float64)
        c = c.astype(np.int64)
        for k in range(8):
            b = np.array([k & 1, (k >> 1) & 1, (k >> 2) & 1])
            g = c + b
            if dense[l]:
                m = int(res[l]) + 1
                idx = g[:, 0] + m * (g[:, 1] + m * g[:, 2])
            else:
                idx = (g[:, 0] ^ ((g[:, 1] * 2654435761) & 0xFFFFFFFF) ^ ((g[:, 2] * 805459861) & 0xFFFFFFFF)) & (T - 1)
            w = np.prod(np.where(b[None, :] == 1, fr, 1.0 - fr), axis=1)
            x[:, 2 * l:2 * l + 2] += w[:, None] * tab[l, idx]
            idxs[l, k], ws[l, k] = idx, w
    x[:, 32:41] = f[:, 3:12]
    x[:, 41] = 1.0
    return x, idxs, ws


def numpy_weights(mlp):
    W1 = mlp[:3072].reshape(48, 64).astype(np.float64)
    W2 = mlp[3072:7168].reshape(64, 64).astype(np.float64)
    W3 = mlp[7168:].reshape(64, 32)[:, :3].astype(np.float64)
    return W1, W2, W3


def numpy_forward(log2, mlp, table, f):
    """Independent restatement: trilinear encoding + three matrix products."""
    W1, W2, W3 = numpy_weights(mlp)
    x = numpy_encode(log2, table, f)[0]
    h1 = np.maximum(x @ W1, 0)
    h2 = np.maximum(h1 @ W2, 0)
    return h2 @ W3


def to_half(a):
    return np.asarray(a).astype(np.float16).astype(np.float64)


def numpy_gradients(log2, mlp, table, f, t, fp16=False):
    """Mean relative-L2 loss sum_c (y - t)^2 / (y^2 + 0.01) and its gradients
    by numpy matrix products, laid out as the oracle's (mlp, table) arrays.

    With fp16 the values that the GPU kernels hold in binary16 are rounded
    where they do: the encoded inputs, both hidden activations, dY, dZ2 and
    dZ1 (the MFMA operands, products summed exactly), and each table-gradient
    contribution, scaled by 4096 / batch, clamped to +-2048 and summed in
    binary16 in sample order.  Samples with a non-finite target contribute
    nothing."""
    W1, W2, W3 = numpy_weights(mlp)
    r = to_half if fp16 else (lambda a: a)
    n = f.shape[0]
    x, idxs, ws = numpy_encode(log2, table, f)
    x = r(x)
    h1 = r(np.maximum(x @ W1, 0))
    h2 = r(np.maximum(h1 @ W2, 0))
    y = h2 @ W3
    t = np.asarray(t, np.float64)
    ok = np.all(np.isfinite(t), axis=1)
    tt = np.where(ok[:, None], t, 0.0)
    den = y * y + 0.01
    loss = float(np.sum(np.where(ok[:, None], (y - tt) ** 2 / den, 0.0)) / n)
    dy = r(np.where(ok[:, None], 2.0 * (y - tt) / den, 0.0))
    dz2 = r(dy @ W3.T) * (h2 > 0)
    dz1 = r(dz2 @ W2.T) * (h1 > 0)
    gm = np.zeros(9216)
    gm[:3072] = (x.T @ dz1).ravel() / n
    gm[3072:7168] = (h1.T @ dz2).ravel() / n
    g3 = np.zeros((64, 32))
    g3[:, :3] = h2.T @ dy / n
    gm[7168:] = g3.ravel()
    dx = dz1 @ W1[:32].T  # encoding features
    T = 2 ** log2
    if fp16:
        acc = np.zeros(16 * T * 2, np.float16)
        scale = np.float32(1.0 / n) * np.float32(4096.0)
        for l in range(16):
            for c in range(2):
                g = (dx[:, 2 * l + c] * scale).astype(np.float32)
                for k in range(8):
                    v = np.clip(ws[l, k].astype(np.float32) * g, -2048.0, 2048.0).astype(np.float16)
                    np.add.at(acc, (l * T + idxs[l, k]) * 2 + c, v)
        gt = acc.astype(np.float64) / 4096.0
    else:
        gt = np.zeros(16 * T * 2)
        for l in range(16):
            for c in range(2):
                for k in range(8):
                    np.add.at(gt, (l * T + idxs[l, k]) * 2 + c, ws[l, k] * dx[:, 2 * l + c] / n)
    return gm, gt, loss


def test_encoding_at_grid_vertices():
    """At a vertex of the level-0 grid the interpolation weights collapse onto
    one table entry."""
    rng = np.random.default_rng(3)
    mlp, table = random_params(rng, LOG2)
    res, dense = binding.rc_grid(LOG2)
    # identity-like first layer: output o = input o for o < 32, rest zero
    mlp_id = np.zeros(9216, np.float32)
    w1 = np.zeros((48, 64), np.float32)
    w1[np.arange(32), np.arange(32)] = 1.0
    mlp_id[:3072] = w1.ravel()
    w2 = np.eye(64, dtype=np.float32)
    mlp_id[3072:7168] = w2.ravel()
    w3 = np.zeros((64, 32), np.float32)
    w3[0, 0] = 1.0  # y0 = relu(encoded feature 0 of level 0)
    w3[1, 1] = 1.0
    mlp_id[7168:] = w3.ravel()
    tab = np.abs(table)  # positive so ReLU passes it
    f = random_features(rng, 64)
    cells = rng.integers(0, res[0], (64, 3))
    f[:, 0:3] = (cells / res[0]).astype(np.float32)
    out = binding.rc_forward(LOG2, mlp_id, tab, f)
    idx = np.array([binding.rc_index(LOG2, 0, int(a), int(b), int(c)) for a, b, c in cells])
    t0 = tab.reshape(16, -1, 2)[0]
    assert np.allclose(out[:, 0], t0[idx, 0], rtol=0, atol=1e-12)
    assert np.allclose(out[:, 1], t0[idx, 1], rtol=0, atol=1e-12)


def test_forward_matches_numpy():
    rng = np.random.default_rng(4)
    mlp, table = random_params(rng, LOG2)
    f = random_features(rng, 200)
    out = binding.rc_forward(LOG2, mlp, table, f)
    ref = numpy_forward(LOG2, mlp, table, f)
    assert np.allclose(out, ref, rtol=1e-9, atol=1e-9)


def test_gradients_match_numpy():
    """The oracle's backward pass against the numpy restatement (double),
    with an unusable sample; and the binary16 restatement's distance from it
    stays at the size that the binary16 operands explain."""
    rng = np.random.default_rng(6)
    mlp, table = random_params(rng, LOG2)
    f = random_features(rng, 300)
    y = binding.rc_forward(LOG2, mlp, table, f)
    t = (y + rng.normal(size=y.shape) * 0.5).astype(np.float32)
    t[7] = np.nan
    gm, gt, loss = binding.rc_gradients(LOG2, mlp, table, f, t)
    nm, nt, nloss = numpy_gradients(LOG2, mlp, table, f, t)
    assert abs(loss - nloss) < 1e-12 * nloss
    assert np.allclose(gm, nm, rtol=1e-9, atol=1e-12 * np.abs(nm).max())
    assert np.allclose(gt, nt, rtol=1e-9, atol=1e-12 * np.abs(nt).max())
    hm, ht, _ = numpy_gradients(LOG2, mlp, table, f, t, fp16=True)
    e = np.linalg.norm(hm - nm) / np.linalg.norm(nm)
    assert 1e-4 < e < 0.1


def test_gradients_match_central_differences():
    """Perturbations by 4 FP16 ulps of parameters in [0.25, 0.5) are exact, so
    central differences of the oracle's double-precision loss check its
    analytic gradient.  The gradient holds the denominator y^2 + 0.01 at the
    base point's y, so the differenced loss does too."""
    rng = np.random.default_rng(5)
    mlp, table = random_params(rng, LOG2)
    f = random_features(rng, 16)
    y0 = binding.rc_forward(LOG2, mlp, table, f)
    t = (y0 + rng.normal(size=y0.shape) * 0.3).astype(np.float32)
    h = 4 * 2.0 ** -12

    def check(m, tb, which, k):
        den = binding.rc_forward(LOG2, m, tb, f) ** 2 + 0.01

        def loss(mm, tt):
            y = binding.rc_forward(LOG2, mm, tt, f)
            return float(np.mean(np.sum((y - t) ** 2 / den, axis=1)))

        gm, gt, l0 = binding.rc_gradients(LOG2, m, tb, f, t)
        assert abs(l0 - loss(m, tb)) < 1e-12
        if which == "mlp":
            mp, mm = m.copy(), m.copy()
            mp[k] += h
            mm[k] -= h
            g_ref = (loss(mp, tb) - loss(mm, tb)) / (2 * h)
            g = gm[k]
        else:
            tp, tm = tb.copy(), tb.copy()
            tp[k] += h
            tm[k] -= h
            g_ref = (loss(m, tp) - loss(m, tm)) / (2 * h)
            g = gt[k]
        assert abs(g - g_ref) <= 1e-7 + 1e-4 * abs(g_ref), (which, k, g, g_ref)
        return g

    # weights of all three layers (input rows 5, 10 and the constant input 41)
    nonzero = 0
    for i in [5, 64 * 10 + 3, 64 * 41 + 7, 3072 + 130, 3072 + 4000, 7168 + 32 * 5 + 1, 7168 + 32 * 60 + 2]:
        m = mlp.copy()
        m[i] = 0.375
        nonzero += check(m, table, "mlp", i) != 0.0
    assert nonzero >= 5
    # table entries under sample 0 at a dense-sized and a hashed level
    res, dense = binding.rc_grid(LOG2)
    for level in (0, 15):
        p = f[0, 0:3] * np.float32(res[level])
        c = np.clip(np.floor(p), 0, res[level] - 1).astype(np.int64)
        idx = binding.rc_index(LOG2, level, int(c[0]), int(c[1]), int(c[2]))
        k = ((level * 2 ** LOG2) + idx) * 2
        tb = table.copy()
        tb[k] = 0.375
        assert check(mlp, tb, "table", k) != 0.0


def test_feature_atan2_and_acos_against_numpy():
    """The float atan2/acos of the vertex features (Abramowitz and Stegun
    4.4.49 after octant reduction) against numpy over all quadrants, axes and
    the clamp of acos."""
    rng = np.random.default_rng(5)
    y = np.concatenate([rng.normal(size=4000), [0.0, 0.0, 1.0, -1.0, 0.0, 1e-30]]).astype(np.float32)
    x = np.concatenate([rng.normal(size=4000), [1.0, -1.0, 0.0, 0.0, 0.0, -1.0]]).astype(np.float32)
    got = np.array([binding.rc_atan2(float(a), float(b)) for a, b in zip(y, x)])
    assert np.abs(got - np.arctan2(y.astype(np.float64), x.astype(np.float64))).max() < 2e-6
    z = np.concatenate([np.linspace(-1, 1, 2001), [1.5, -1.5]]).astype(np.float32)
    got = np.array([binding.rc_acos(float(v)) for v in z])
    assert np.abs(got - np.arccos(np.clip(z.astype(np.float64), -1, 1))).max() < 2e-3
