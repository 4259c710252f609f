"""Output files as the reference's host program writes them (HostProgram/main.cpp:118-170):
the PNG's tone mapping and sRGB encoding at known values, the PNG read back by PIL, and the EXR
read back by a minimal scanline parser (magic, version, header attributes, offsets, channels)."""
import math
import struct

import numpy as np

from vlr_amd import image_io


def test_tone_mapping_known_values():
    # 1 - exp(-v) = 0.5 at v = ln 2; sRGB(0.5) = 0.7353570; * 256 -> 188
    rgba = np.zeros((1, 4, 4), np.float32)
    rgba[0, 0, :3] = math.log(2.0)
    rgba[0, 1, :3] = 1e6          # saturates to 255
    rgba[0, 2, :3] = -1.0         # clamped to 0
    rgba[0, 3, :3] = [0.0, math.log(2.0) / 2, 0.0]
    out = image_io.tone_map_srgb8(rgba, brightness=1.0)
    assert out[0, 0].tolist() == [188, 188, 188, 255]
    assert out[0, 1].tolist() == [255, 255, 255, 255]
    assert out[0, 2].tolist() == [0, 0, 0, 255]
    # brightness 2 doubles the argument of the tone curve
    assert image_io.tone_map_srgb8(rgba[:, 3:4], brightness=2.0)[0, 0, 1] == 188
    # debug renders skip the tone mapping: 0.5 linear -> int(128)
    dbg = image_io.tone_map_srgb8(np.full((1, 1, 4), 0.5, np.float32), debug=True)
    assert dbg[0, 0, 0] == 128


def test_png_round_trip(tmp_path):
    from PIL import Image
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (7, 5, 4), dtype=np.uint8)
    image_io.write_png(str(tmp_path / "a.png"), img)
    back = np.asarray(Image.open(tmp_path / "a.png").convert("RGBA"))
    assert np.array_equal(back, img)


def read_exr(path):
    data = open(path, "rb").read()
    magic, version = struct.unpack_from("<ii", data, 0)
    assert magic == 20000630 and version == 2
    pos, attrs = 8, {}
    while data[pos] != 0:
        name_end = data.index(b"\x00", pos)
        name = data[pos:name_end].decode()
        type_end = data.index(b"\x00", name_end + 1)
        size = struct.unpack_from("<i", data, type_end + 1)[0]
        attrs[name] = data[type_end + 5:type_end + 5 + size]
        pos = type_end + 5 + size
    pos += 1
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    assert attrs["compression"] == b"\x00"
    chans, p = [], 0
    cl = attrs["channels"]
    while cl[p] != 0:
        e = cl.index(b"\x00", p)
        chans.append((cl[p:e].decode(), struct.unpack_from("<i", cl, e + 1)[0]))
        p = e + 1 + 16
    assert [c for c, _ in chans] == ["A", "B", "G", "R"] and all(t == 2 for _, t in chans)
    offsets = struct.unpack_from("<%dQ" % h, data, pos)
    out = np.zeros((h, w, 4), np.float32)
    for y, off in enumerate(offsets):
        yy, size = struct.unpack_from("<ii", data, off)
        assert yy == y and size == 16 * w
        planes = np.frombuffer(data, "<f4", 4 * w, off + 8).reshape(4, w)
        out[y, :, 3], out[y, :, 2], out[y, :, 1], out[y, :, 0] = planes
    return out


def test_exr_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    img = (rng.normal(size=(6, 9, 4)) * 100).astype(np.float32)
    image_io.write_exr(str(tmp_path / "a.exr"), img)
    assert np.array_equal(read_exr(str(tmp_path / "a.exr")), img)


def test_save_output_writes_both(tmp_path):
    img = np.full((3, 2, 4), 0.25, np.float32)
    image_io.save_output(img, str(tmp_path / "out"))
    assert (tmp_path / "out.png").exists() and (tmp_path / "out.exr").exists()
    np.testing.assert_array_equal(read_exr(str(tmp_path / "out.exr")), img)
