"""Writing a rendered output buffer to disk, as the reference's host program does
(HostProgram/main.cpp:118-170, saveOutputBufferAsImageFile):

  PNG  8-bit sRGB after the simple tone mapping 1 - exp(-k * rgb) (k the brightness
       coefficient; skipped for debug renders), negative values clamped to 0, each channel
       min(v * 256, 255), alpha 255
  EXR  the linear RGBA32F output as it is: scanline, uncompressed, 32-bit float channels

Both are written with the standard library only (zlib for PNG's deflate stream).
"""
import struct
import zlib

import numpy as np

__all__ = ["tone_map_srgb8", "write_png", "write_exr", "save_output"]


def _srgb_encode(v):
    v = np.asarray(v, np.float32)
    return np.where(v <= 0.0031308, 12.92 * v, 1.055 * np.power(np.maximum(v, 0.0031308), 1 / 2.4) - 0.055)


def tone_map_srgb8(rgba: np.ndarray, brightness: float = 1.0, debug: bool = False) -> np.ndarray:
    """(H, W, 4) float linear RGBA -> (H, W, 4) uint8 as main.cpp:128-160."""
    rgb = np.maximum(np.asarray(rgba, np.float32)[..., :3], 0.0)
    if not debug:
        rgb = rgb * np.float32(brightness)
        rgb = 1.0 - np.exp(-rgb)
        rgb = _srgb_encode(rgb)
    q = np.minimum((rgb * 256).astype(np.uint32), 255).astype(np.uint8)
    out = np.empty(rgb.shape[:2] + (4,), np.uint8)
    out[..., :3] = q
    out[..., 3] = 255
    return out


def write_png(path: str, rgba8: np.ndarray):
    """(H, W, 4) uint8, rows top to bottom."""
    img = np.ascontiguousarray(rgba8, np.uint8)
    h, w = img.shape[:2]
    raw = b"".join(b"\x00" + img[y].tobytes() for y in range(h))   # filter type 0 per row

    def chunk(kind, data):
        body = kind + data
        return struct.pack(">I", len(data)) + body + struct.pack(">I", zlib.crc32(body) & 0xFFFFFFFF)

    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n")
        f.write(chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)))
        f.write(chunk(b"IDAT", zlib.compress(raw, 6)))
        f.write(chunk(b"IEND", b""))


def write_exr(path: str, rgba: np.ndarray):
    """(H, W, 4) float32 linear RGBA as a scanline OpenEXR file without compression."""
    img = np.ascontiguousarray(rgba, np.float32)
    h, w = img.shape[:2]

    def attr(name, kind, data):
        return name.encode() + b"\x00" + kind.encode() + b"\x00" + struct.pack("<i", len(data)) + data

    chans = b"".join(c.encode() + b"\x00" + struct.pack("<iBBBBii", 2, 0, 0, 0, 0, 1, 1) for c in "ABGR") + b"\x00"
    header = (struct.pack("<ii", 20000630, 2) +
              attr("channels", "chlist", chans) +
              attr("compression", "compression", b"\x00") +
              attr("dataWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1)) +
              attr("displayWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1)) +
              attr("lineOrder", "lineOrder", b"\x00") +
              attr("pixelAspectRatio", "float", struct.pack("<f", 1.0)) +
              attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0)) +
              attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\x00")
    line_bytes = 4 * w * 4
    table_end = len(header) + 8 * h
    offsets = [table_end + y * (8 + line_bytes) for y in range(h)]
    with open(path, "wb") as f:
        f.write(header)
        f.write(struct.pack("<%dQ" % h, *offsets))
        for y in range(h):
            f.write(struct.pack("<ii", y, line_bytes))
            for c in (3, 2, 1, 0):   # A, B, G, R: channels in name order
                f.write(img[y, :, c].astype("<f4").tobytes())


def save_output(rgba: np.ndarray, path_without_ext: str, brightness: float = 1.0, debug: bool = False):
    """Both files of the reference's saveOutputBufferAsImageFile for an output buffer read with
    vlrContextReadOutputBuffer."""
    write_png(path_without_ext + ".png", tone_map_srgb8(rgba, brightness, debug))
    write_exr(path_without_ext + ".exr", rgba)
