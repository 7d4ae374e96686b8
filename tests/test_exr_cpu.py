"""OpenEXR output (rmh_save_exr; the reference's saveFrame,
host/Renderer.swift:804-807 and utils/Metal.swift:75-124), read back by a
decoder written here from the OpenEXR file-format document: magic number and
version, the attribute header, the line offset table, and scanline blocks of
(y, byte count, per-channel little-endian floats)."""
import struct

import numpy as np
import pytest

import raymond_amd.api as api


def read_exr(path):
    data = open(path, "rb").read()
    magic, version = struct.unpack_from("<ii", data, 0)
    assert magic == 20000630 and version & 0xff == 2 and not version & 0x200  # scanline, single part
    pos, attrs = 8, {}
    while data[pos] != 0:
        end = data.index(b"\0", pos)
        name = data[pos:end].decode()
        tend = data.index(b"\0", end + 1)
        typ = data[end + 1:tend].decode()
        size, = struct.unpack_from("<i", data, tend + 1)
        attrs[name] = (typ, data[tend + 5:tend + 5 + size])
        pos = tend + 5 + size
    pos += 1
    chans, raw = [], attrs["channels"][1]
    q = 0
    while raw[q] != 0:
        end = raw.index(b"\0", q)
        ptype, = struct.unpack_from("<i", raw, end + 1)
        xs, ys = struct.unpack_from("<ii", raw, end + 9)
        chans.append((raw[q:end].decode(), ptype, xs, ys))
        q = end + 17
    assert attrs["compression"] == ("compression", b"\0")
    assert attrs["lineOrder"] == ("lineOrder", b"\0")
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    assert attrs["displayWindow"][1] == attrs["dataWindow"][1]
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from("<%dQ" % h, data, pos)
    out = {c[0]: np.zeros((h, w), np.float32) for c in chans}
    for off in offsets:
        y, n = struct.unpack_from("<ii", data, off)
        assert n == 4 * w * len(chans)
        p = off + 8
        for name, ptype, xs, ys in chans:
            assert ptype == 2 and xs == 1 and ys == 1  # FLOAT, full resolution
            out[name][y - y0] = np.frombuffer(data, "<f4", w, p)
            p += 4 * w
    return [c[0] for c in chans], out


def test_exr_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.uniform(-1.0, 50.0, (7, 13, 4)).astype(np.float32)
    norm = np.float32(1.0 / 7.0)
    path = str(tmp_path / "frame.exr")
    api.save_exr(path, img, float(norm))
    names, ch = read_exr(path)
    assert names == ["A", "B", "G", "R"]  # channel list sorted by name
    for k, n in enumerate("RGB"):
        assert np.array_equal(ch[n].view(np.uint32), (img[..., k] * norm).astype(np.float32).view(np.uint32))
    assert np.all(ch["A"] == 1.0)


def test_exr_rejects_bad_arguments(tmp_path):
    with pytest.raises(api.RmError):
        api.save_exr(str(tmp_path / "none" / "x.exr"), np.zeros((2, 2, 4), np.float32))
    with pytest.raises(api.RmError):
        api.save_exr(str(tmp_path / "e.exr"), np.zeros((0, 2, 4), np.float32))
    with pytest.raises(ValueError):
        api.save_exr(str(tmp_path / "x.exr"), np.zeros((2, 2, 3), np.float32))
