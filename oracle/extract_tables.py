"""Write the published colour-science tables the spectral path needs into a
binary data file, read from the reference's source files where they lie.

The tables are data, not code:
  * CIE 1931 2-degree colour-matching functions, 360-830 nm at 1 nm
    (libVLR/shared/spectrum_base.cpp:6-190, the CIE's published observer);
  * the Meng et al. 2015 spectral-upsampling grid (12x14 cells) and its data
    points (xy*, uv, 95-sample spectra) for float precision
    (libVLR/shared/spectrum_types.cpp:687-1050).

Output layout (little endian), shared by the product and the oracle:
  char[8]  magic "VLRSPEC1"
  u32      n_cmf (471);  f32 xbar[n], ybar[n], zbar[n]
  u32      grid_w, grid_h; per cell: u8 inside, u8 num_points, u8 idx[6]
  u32      n_points;     per point: f32 xystar[2], uv[2], spectrum[95]

Run only where the reference checkout exists:
  python oracle/extract_tables.py /root/reference vlr_amd/data/spectral_tables.bin
"""
import re
import struct
import sys
from pathlib import Path

FLOAT = r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?f?"


def _floats(text):
    out = []
    for tok in re.findall(FLOAT, text):
        out.append(float(tok.rstrip("fF")))
    return out


def _array_body(src, name):
    i = src.index(name)
    j = src.index("{", i)
    depth, k = 0, j
    while True:
        c = src[k]
        if c == "{":
            depth += 1
        elif c == "}":
            depth -= 1
            if depth == 0:
                return src[j + 1:k]
        k += 1


def extract(ref_root: Path):
    base = (ref_root / "libVLR/shared/spectrum_base.cpp").read_text(encoding="utf-8-sig")
    cmf = []
    for nm in ("xbar_CIE1931_2deg[]", "ybar_CIE1931_2deg[]", "zbar_CIE1931_2deg[]"):
        vals = _floats(_array_body(base, "const float " + nm))
        assert len(vals) == 471, (nm, len(vals))
        cmf.append(vals)

    spec = (ref_root / "libVLR/shared/spectrum_types.cpp").read_text(encoding="utf-8-sig")
    # float-precision tables come first in the file
    grid_txt = _array_body(spec, "UpsampledSpectrumTemplate<float, NumSpectralSamples>::spectrum_grid[]")
    cells = []
    for m in re.finditer(r"\{\s*(\d+)\s*,\s*(\d+)\s*,\s*\{([^}]*)\}\s*\}", grid_txt):
        inside, npts = int(m.group(1)), int(m.group(2))
        idx = [255 if t.strip() == "UINT8_MAX" else int(t) for t in m.group(3).split(",")]
        assert len(idx) == 6
        cells.append((inside, npts, idx))
    assert len(cells) == 12 * 14, len(cells)

    pts_txt = _array_body(spec, "UpsampledSpectrumTemplate<float, NumSpectralSamples>::spectrum_data_points[]")
    points = []
    for m in re.finditer(r"\{\s*\{([^}]*)\}\s*,\s*\{([^}]*)\}\s*,\s*\{([^}]*)\}\s*\}", pts_txt):
        xy, uv, sp = _floats(m.group(1)), _floats(m.group(2)), _floats(m.group(3))
        assert len(xy) == 2 and len(uv) == 2 and len(sp) == 95
        points.append(xy + uv + sp)
    assert len(points) == 186, len(points)
    return cmf, cells, points


def write(out: Path, cmf, cells, points):
    b = bytearray(b"VLRSPEC1")
    b += struct.pack("<I", 471)
    for arr in cmf:
        b += struct.pack("<471f", *arr)
    b += struct.pack("<II", 12, 14)
    for inside, npts, idx in cells:
        b += struct.pack("<BB6B", inside, npts, *idx)
    b += struct.pack("<I", len(points))
    for p in points:
        b += struct.pack("<99f", *p)
    out.parent.mkdir(parents=True, exist_ok=True)
    out.write_bytes(bytes(b))


if __name__ == "__main__":
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    dst = Path(sys.argv[2] if len(sys.argv) > 2 else Path(__file__).resolve().parent.parent / "vlr_amd/data/spectral_tables.bin")
    write(dst, *extract(ref))
    print("wrote", dst, dst.stat().st_size, "bytes")
