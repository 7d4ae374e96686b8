"""Shared checks of the GPU parity tests.

Radiance is compared per pixel and channel, |gpu - oracle| <= rtol |oracle| +
atol, with rtol = 1e-4 (the per-sample bound) and a small absolute floor for
float32 accumulation noise; the worst relative error is returned so that a
test can print it.  Hits are compared bit for bit.
"""
import os

import numpy as np

RTOL = 1e-4
ATOL = 1e-6

# host threads the oracle may use: the GPU box's CPU share for one GPU
ORACLE_THREADS = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", "16") or 16)))


def assert_radiance_close(got, ref, rtol=RTOL, atol=ATOL, what="radiance"):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    assert got.shape == ref.shape, (got.shape, ref.shape)
    assert np.all(np.isfinite(got)), "%s: non-finite values" % what
    d = np.abs(got - ref)
    bound = rtol * np.abs(ref) + atol
    bad = d > bound
    if bad.any():
        k = np.unravel_index(np.argmax(d - bound), d.shape)
        raise AssertionError("%s: %d of %d values outside |d| <= %g |ref| + %g; worst at %s: gpu %r oracle %r"
                             % (what, int(bad.sum()), d.size, rtol, atol, k, got[k], ref[k]))
    lit = np.abs(ref) > 1e-3
    return float((d[lit] / np.abs(ref[lit])).max()) if lit.any() else 0.0


def describe_rays(rays, idx, gpu, ref, limit=8):
    """The rays at idx as exact float32 bits with both sides' results."""
    lines = []
    for i in idx[:limit]:
        bits = " ".join("%08x" % b for b in np.ascontiguousarray(rays[i], np.float32).view(np.uint32))
        lines.append("ray %d [%s] gpu t=%r inst=%d prim=%d oracle t=%r inst=%d prim=%d"
                     % (i, bits, float(gpu[0][i]), int(gpu[1][i]), int(gpu[2][i]), float(ref[0][i]), int(ref[1][i]),
                        int(ref[2][i])))
    return "\n".join(lines)


def assert_hits_equal(gpu, ref, rays=None):
    t, inst, prim, bary = gpu
    rt, rinst, rprim, rbary = ref
    hit = rt >= 0
    if rays is not None:
        both = hit & (t >= 0)
        bad = np.flatnonzero(((t >= 0) != hit) | (both & ((t.view(np.uint32) != rt.view(np.uint32)) |
                                                          (inst != rinst) | (prim != rprim) |
                                                          (bary.view(np.uint32) != rbary.view(np.uint32)).any(1))))
        if bad.size:
            raise AssertionError("%d rays differ (%d hit/miss)\n%s" % (bad.size, np.sum((t >= 0) != hit),
                                                                      describe_rays(rays, bad, gpu, ref)))
    assert np.array_equal(t >= 0, hit), "hit/miss mismatch: %d rays" % np.sum((t >= 0) != hit)
    assert np.array_equal(t.view(np.uint32)[hit], rt.view(np.uint32)[hit]), \
        "distance bits differ on %d rays" % np.sum(t.view(np.uint32)[hit] != rt.view(np.uint32)[hit])
    assert np.array_equal(inst[hit], rinst[hit]), "instance ids differ"
    assert np.array_equal(prim[hit], rprim[hit]), "primitive ids differ"
    assert np.array_equal(bary.view(np.uint32)[hit], rbary.view(np.uint32)[hit]), "barycentric bits differ"


def camera_rays(desc, width, height, seed, count=None):
    """Pinhole rays through jittered pixel positions of the scene camera (the
    camera looks down its -z axis, +y up, focal length over half the film
    width), one per pixel or `count` at random pixels."""
    rng = np.random.default_rng(seed)
    m = np.asarray(desc.camera["transform"], np.float64)
    f = float(desc.camera["focal"])
    if count is None:
        y, x = np.mgrid[0:height, 0:width]
        x, y = x.ravel().astype(np.float64), y.ravel().astype(np.float64)
    else:
        x = rng.integers(0, width, count).astype(np.float64)
        y = rng.integers(0, height, count).astype(np.float64)
    u = ((x + rng.random(x.size)) / width) * 2.0 - 1.0
    v = -(((y + rng.random(y.size)) / height) * 2.0 - 1.0) * (height / width)
    d = np.stack([u, v, np.full_like(u, -f)], 1) @ m[:3, :3].T
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((x.size, 8), np.float32)
    r[:, 0:3] = m[:3, 3]
    r[:, 3] = 1e-3
    r[:, 4:7] = d
    r[:, 7] = np.inf
    return r


def bounce_rays(rays, t, seed, count, tmax=np.inf):
    """Rays from `count` hit points of `rays` in uniformly random directions."""
    rng = np.random.default_rng(seed)
    hit = np.flatnonzero(t >= 0)
    k = rng.choice(hit, size=min(count, hit.size), replace=hit.size < count)
    p = rays[k, 0:3].astype(np.float64) + rays[k, 4:7].astype(np.float64) * t[k, None]
    d = rng.normal(size=(k.size, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((k.size, 8), np.float32)
    r[:, 0:3] = p
    r[:, 3] = 1e-4
    r[:, 4:7] = d
    r[:, 7] = tmax if np.isscalar(tmax) else tmax
    return r


def pixel_subset(width, height, count, seed):
    """A fixed random subset of pixel indices y * width + x, sorted."""
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(width * height, size=count, replace=False)).astype(np.uint32)


def subset_of(image, pixels):
    h, w = image.shape[:2]
    return image.reshape(h * w, -1)[pixels]
