"""Image files for texture slots: PNG (8 and 16 bit, grey, grey+alpha, RGB,
RGBA and palette, not interlaced), OpenEXR (scanline, uncompressed or ZIP,
half or float channels) and PFM.  Decoded with the standard library's zlib.

Texels come back as float32 arrays of shape (height, width, 1) for grey
images and (height, width, 4) otherwise, row 0 at the top of the picture,
integer formats scaled to [0, 1] with no colour-space conversion: the
reference loads its textures with sRGB decoding off and lets the image node
convert (MaterialBuilder.swift:719-737, nodes.hpp:537-552).
"""
import os
import struct
import zlib

import numpy as np


def _png(data):
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG file")
    pos, chunks = 8, []
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        chunks.append((kind, data[pos + 8:pos + 8 + n]))
        pos += 12 + n
    ihdr = chunks[0][1]
    w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", ihdr)
    if interlace:
        raise ValueError("interlaced PNG is not supported")
    channels = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    if depth not in (8, 16) and not (ctype == 3 and depth in (1, 2, 4, 8)):
        raise ValueError("PNG bit depth %d is not supported" % depth)
    raw = zlib.decompress(b"".join(c for k, c in chunks if k == b"IDAT"))
    bits = channels * depth
    stride = (w * bits + 7) // 8
    bpp = max(1, bits // 8)
    out = np.zeros((h, stride), np.uint8)
    prev = np.zeros(stride, np.int32)
    p = 0
    for y in range(h):
        ftype = raw[p]
        line = np.frombuffer(raw, np.uint8, stride, p + 1).astype(np.int32)
        p += 1 + stride
        if ftype == 0:
            cur = line
        elif ftype == 2:
            cur = (line + prev) & 0xff
        else:
            cur = line.copy()
            for i in range(stride):  # sub, average, Paeth depend on the decoded left neighbour
                a = cur[i - bpp] if i >= bpp else 0
                b = prev[i]
                c = prev[i - bpp] if i >= bpp else 0
                if ftype == 1:
                    pred = a
                elif ftype == 3:
                    pred = (a + b) >> 1
                elif ftype == 4:
                    pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                    pred = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
                else:
                    raise ValueError("bad PNG filter %d" % ftype)
                cur[i] = (cur[i] + pred) & 0xff
        out[y] = cur
        prev = cur
    if depth == 16:
        v = out.view(">u2").reshape(h, w, channels).astype(np.float32) / 65535.0
    elif depth == 8:
        v = out.reshape(h, w, channels).astype(np.float32)
        v = v if ctype == 3 else v / 255.0
    else:
        v = np.unpackbits(out, axis=1)[:, :w * depth].reshape(h, w, depth)
        v = (v * (1 << np.arange(depth - 1, -1, -1))).sum(2)[..., None].astype(np.float32)
    if ctype == 3:
        plte = np.frombuffer(dict(chunks)[b"PLTE"], np.uint8).reshape(-1, 3).astype(np.float32) / 255.0
        idx = v[..., 0].astype(np.int64)
        rgba = np.ones((h, w, 4), np.float32)
        rgba[..., :3] = plte[idx]
        if b"tRNS" in dict(chunks):
            trns = np.frombuffer(dict(chunks)[b"tRNS"], np.uint8).astype(np.float32) / 255.0
            a = np.ones(256, np.float32)
            a[:trns.size] = trns
            rgba[..., 3] = a[idx]
        return rgba
    if channels == 1:
        return v
    rgba = np.ones((h, w, 4), np.float32)
    if channels == 2:
        rgba[..., :3] = v[..., :1]
        rgba[..., 3] = v[..., 1]
    else:
        rgba[..., :channels] = v
    return rgba


def _exr(data):
    if struct.unpack("<I", data[:4])[0] != 20000630:
        raise ValueError("not an OpenEXR file")
    if struct.unpack("<I", data[4:8])[0] & 0x200:
        raise ValueError("tiled OpenEXR is not supported")
    pos, attrs = 8, {}
    while data[pos] != 0:
        end = data.index(b"\0", pos)
        name = data[pos:end].decode()
        end2 = data.index(b"\0", end + 1)
        typ = data[end + 1:end2].decode()
        n = struct.unpack("<i", data[end2 + 1:end2 + 5])[0]
        attrs[name] = (typ, data[end2 + 5:end2 + 5 + n])
        pos = end2 + 5 + n
    pos += 1
    chans, p, blob = [], 0, attrs["channels"][1]
    while blob[p] != 0:
        e = blob.index(b"\0", p)
        name = blob[p:e].decode()
        ptype = struct.unpack("<i", blob[e + 1:e + 5])[0]
        chans.append((name, ptype))
        p = e + 17
    comp = attrs["compression"][1][0]
    if comp not in (0, 2, 3):
        raise ValueError("OpenEXR compression %d is not supported" % comp)
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    lines_per_block = {0: 1, 2: 1, 3: 16}[comp]
    nblocks = (h + lines_per_block - 1) // lines_per_block
    offsets = struct.unpack("<%dQ" % nblocks, data[pos:pos + 8 * nblocks])
    chans = sorted(chans)  # channels are stored alphabetically
    size = {1: 2, 2: 4, 0: 4}
    out = {name: np.zeros((h, w), np.float32) for name, _ in chans}
    for b, off in enumerate(offsets):
        yline, n = struct.unpack("<ii", data[off:off + 8])
        block = data[off + 8:off + 8 + n]
        rows = min(lines_per_block, y1 + 1 - yline)
        expect = rows * w * sum(size[t] for _, t in chans)
        if comp != 0 and n < expect:
            z = np.frombuffer(zlib.decompress(block), np.uint8).astype(np.int32)
            # undo the predictor, then interleave the two halves
            z = np.cumsum(np.concatenate([z[:1], z[1:] - 128])) & 0xff
            z = z.astype(np.uint8)
            half = (z.size + 1) // 2
            t = np.empty(z.size, np.uint8)
            t[0::2] = z[:half]
            t[1::2] = z[half:]
            block = t.tobytes()
        q = 0
        for r in range(rows):
            for name, ptype in chans:
                k = size[ptype] * w
                dt = {1: "<f2", 2: "<f4", 0: "<u4"}[ptype]
                out[name][yline - y0 + r] = np.frombuffer(block, dt, w, q).astype(np.float32)
                q += k
    names = [c for c, _ in chans]
    if set(names) >= {"R", "G", "B"}:
        rgba = np.ones((h, w, 4), np.float32)
        for i, c in enumerate("RGB"):
            rgba[..., i] = out[c]
        if "A" in out:
            rgba[..., 3] = out["A"]
        return rgba
    return out[names[0]][..., None]


def _pfm(data):
    header = data.split(b"\n", 3)
    kind = header[0].strip()
    w, h = (int(v) for v in header[1].split())
    scale = float(header[2])
    chans = 3 if kind == b"PF" else 1
    v = np.frombuffer(header[3], "<f4" if scale < 0 else ">f4", w * h * chans).reshape(h, w, chans)
    v = v[::-1].astype(np.float32)  # PFM stores the bottom row first
    if chans == 1:
        return v
    rgba = np.ones((h, w, 4), np.float32)
    rgba[..., :3] = v
    return rgba


def load(path):
    """Texels of an image file, (h, w, 1) or (h, w, 4) float32."""
    with open(path, "rb") as f:
        data = f.read()
    ext = os.path.splitext(path)[1].lower()
    if data[:8] == b"\x89PNG\r\n\x1a\n":
        return _png(data)
    if data[:4] == struct.pack("<I", 20000630):
        return _exr(data)
    if ext == ".pfm" or data[:2] in (b"PF", b"Pf"):
        return _pfm(data)
    raise ValueError("unsupported image format: %s" % path)


def placeholder():
    """The texture of an image that cannot be loaded: one black texel (the
    reference substitutes a 1x1 zero texture, MaterialBuilder.swift:224-233)."""
    return np.zeros((1, 1, 1), np.float32)
