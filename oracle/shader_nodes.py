"""CPU restatement of the shader-node programs (test infrastructure only).

This is the checker for the device's node programs (vlr_amd/csrc/shader_nodes.h), evaluated
directly from the node graph, recursively, as the reference's calcNode does
(shared/kernel_common.h:658-737), with:
  conversions       NodeTypeInfo<T>::convertFrom (shared/shared.h:390-470); a plug whose value
                    cannot convert to the wanted type gives the default
  node programs     GPU_kernels/shader_nodes.cu:20-483 (EnvironmentTexture outside the
                    environment emitter: :457-483)
  sampled spectra   RegularSampledSpectrum / IrregularSampledSpectrum::evaluate
                    (shared/spectrum_types.cpp:394-412, 496-518); in the RGB build toXYZ
                    (spectrum_types.cpp:414-591) and transformToRenderingRGB
  images            LinearImage2D's conversions at creation (image.cpp:410-790): in the
                    spectral build, spectra from RGB sources stored as quantised Meng grid
                    coordinates (uvsA8x4 from 8-bit sources, uvsA16Fx4 from half / float ones);
                    8-bit sRGB texels degammed when read (the hardware sRGB path of
                    image.h:74, shader_nodes.cpp:1665-1683), half / float sRGB ones at creation
  sampling          CUDA's normalised-coordinate texture fetch: linear filtering at
                    u * width - 0.5 with full-precision weights (the hardware's 8-bit weights
                    are not reproduced), wrap modes on texel indices, tex2Dgather order
                    (i0, j1), (i1, j1), (i1, j0), (i0, j0); a one-channel texel reads
                    (v, 0, 0, 1) and a two-channel one (r, g, 0, 1)
Spectra in the spectral build go through the oracle's Meng upsampling (oracle/ov_spectrum.h).
Arithmetic is float32 where the device's is.
"""
import math

import numpy as np

from oracle import binding as ob

f32 = np.float32
PLUG = {"float1": 0, "float2": 1, "float3": 2, "float4": 3, "Point3D": 4, "Vector3D": 5, "Normal3D": 6,
        "Spectrum": 7, "Alpha": 8, "TextureCoordinates": 9}
# the C API's spectrum-type strings (queryable.cpp:34-37) and values (spectrum_base.h:20-27)
SPEC_TYPES = {"Reflectance": 0, "Transmittance": 0, "Light Source": 1, "NA": 3}
COLOR_SPACES = {"Rec709(D65) sRGB Gamma": 0, "Rec709(D65)": 1, "XYZ": 2, "xyY": 3}
# value classes: the C++ type of a plug's program
F1, F2, F3, F4, P3, V3, N3, SPEC = range(8)
CLASS_OF = {0: F1, 1: F2, 2: F3, 3: F4, 4: P3, 5: V3, 6: N3, 7: SPEC, 8: F1, 9: P3}
EER = f32(0.009355121400914532)
# spectrum_base.h:109-128 (column-major)
MAT_D65_TO_XYZ = np.array([0.4124564, 0.2126729, 0.0193339, 0.3575761, 0.7151522, 0.1191920, 0.1804375, 0.0721750,
                           0.9503041], f32)
MAT_E_TO_XYZ = np.array([0.4969, 0.2562, 0.0233, 0.3391, 0.6782, 0.1130, 0.1640, 0.0656, 0.8637], f32)


# spectrum_base.h:109-128, XYZ to Rec.709 (D65 white for light sources, E for the rest)
MAT_XYZ_TO_D65 = np.array([3.2404542, -0.9692660, 0.0556434, -1.5371385, 1.8760108, -0.2040259, -0.4985314,
                           0.0415560, 1.0572252], f32)
MAT_XYZ_TO_E = np.array([2.6897, -1.0221, 0.0612, -1.2759, 1.9783, -0.2245, -0.4138, 0.0438, 1.1633], f32)

_CMF = None


def cmf_table():
    """The CIE 1931 2-degree CMF at 1 nm, 360-830 nm (xbar, ybar, zbar), from the data file
    the product reads (extracted from shared/spectrum_base.cpp:6-190)."""
    global _CMF
    if _CMF is None:
        raw = ob.TABLES.read_bytes()
        assert raw[:8] == b"VLRSPEC1"
        n = int(np.frombuffer(raw, np.uint32, 1, 8)[0])
        t = np.frombuffer(raw, np.float32, 3 * n, 12).reshape(3, n)
        _CMF = (t[0].copy(), t[1].copy(), t[2].copy())
    return _CMF


class _Kahan:
    """CompensatedSum (shared/common_internal.h:351-368) in float32."""

    def __init__(self):
        self.s, self.c = f32(0), f32(0)

    def add(self, v):
        inp = f32(f32(v) - self.c)
        t = f32(self.s + inp)
        self.c = f32(f32(t - self.s) - inp)
        self.s = t


def sampled_rgb(spectrum_type, regular, lo, hi, lambdas, values):
    """The RGB build's triplet of a sampled spectrum: Regular/IrregularSampledSpectrum::toXYZ
    (shared/spectrum_types.cpp:414-488, 521-591) over the 1 nm CMF table and the spectrum's
    samples, divided by integralCMF (shared/spectrum_base.cpp:196-205), then
    transformToRenderingRGB (spectrum_base.h:138-152) with negative channels clamped to zero
    (shader_nodes.cpp:1058-1063).  A one-sample spectrum is constant (the reference's bin
    width would be 0 / 0 there)."""
    xb, yb, zb = cmf_table()
    cm = (xb, yb, zb)
    ncmf = len(yb)
    low, binw = f32(360), f32(1)
    values = [f32(v) for v in values]
    lambdas = [f32(v) for v in lambdas]
    n = len(values)
    yint = _Kahan()
    for i in range(1, ncmf):
        yint.add(f32(f32(f32(yb[i - 1] + yb[i]) * f32(1)) * f32(0.5)))
    sums = [_Kahan(), _Kahan(), _Kahan()]
    if n == 1:
        for i in range(1, ncmf):
            for k in range(3):
                sums[k].add(f32(f32(values[0] * f32(cm[k][i - 1] + cm[k][i])) * f32(0.5)))
    else:
        lo, hi = f32(lo), f32(hi)
        first = lo if regular else lambdas[0]
        last = hi if regular else lambdas[n - 1]
        bw = f32(f32(hi - lo) / f32(n - 1)) if regular else f32(0)
        at = (lambda k: f32(lo + f32(f32(k) * bw))) if regular else (lambda k: lambdas[k])
        ci, bi = 0, 0
        wl = min(low, first)
        prev = [f32(0)] * 3
        prev_v, half = f32(0), f32(0)
        while True:
            if wl < low:
                c = [f32(0)] * 3
            elif wl == f32(low + f32(f32(ci) * binw)):
                c = [cm[k][ci] for k in range(3)]
                ci += 1
            else:
                idx = min(int(f32(f32(wl - low) / binw)), ncmf - 1)
                t = f32(f32(wl - f32(low + f32(f32(idx) * binw))) / binw)
                c = [f32(f32(f32(f32(1) - t) * cm[k][idx]) + f32(t * cm[k][idx + 1])) for k in range(3)]
            if wl < first:
                v = values[0]
            elif wl > last:
                v = values[n - 1]
            elif wl == at(bi):
                v = values[bi]
                bi += 1
            elif regular:
                idx = min(int(f32(f32(wl - lo) / bw)), n - 2)
                t = f32(f32(wl - f32(lo + f32(f32(idx) * bw))) / bw)
                v = f32(f32(f32(f32(1) - t) * values[idx]) + f32(t * values[idx + 1]))
            else:
                j = max(bi - 1, 0)
                while j < n and lambdas[j] < wl:   # std::lower_bound from max(bi - 1, 0)
                    j += 1
                idx = max(j - 1, 0)
                t = f32(f32(wl - lambdas[idx]) / f32(lambdas[idx + 1] - lambdas[idx]))
                v = f32(f32(f32(f32(1) - t) * values[idx]) + f32(t * values[idx + 1]))
            avg = f32(f32(prev_v + v) * f32(0.5))
            for k in range(3):
                sums[k].add(f32(f32(avg * f32(prev[k] + c[k])) * half))
            prev = c
            prev_v = v
            prev_wl = wl
            wl = min(f32(low + f32(f32(ci) * binw)), at(bi) if bi < n else f32(np.inf))
            half = f32(f32(wl - prev_wl) * f32(0.5))
            if ci == ncmf:
                break
    xyz = [f32(sums[k].s / yint.s) for k in range(3)]
    rgb = tristim(MAT_XYZ_TO_D65 if spectrum_type == 1 else MAT_XYZ_TO_E, xyz)
    return np.array([max(f32(0), x) for x in rgb], f32)


def wavelengths(offset):
    """WavelengthSamples::createWithEqualOffsets (spectrum_types.h:55-65), 4 samples."""
    return np.array([f32(360.0) + f32(470.0) * (f32(i) + f32(offset)) / f32(4) for i in range(4)], f32)


def degamma(v):
    v = f32(v)
    return f32(v / f32(12.92)) if v <= f32(0.04045) else f32(math.pow((v + f32(0.055)) / f32(1.055), 2.4))


def tristim(m, rgb):
    r, g, b = (f32(x) for x in rgb)
    return [f32(m[0] * r + m[3] * g + m[6] * b), f32(m[1] * r + m[4] * g + m[7] * b),
            f32(m[2] * r + m[5] * g + m[8] * b)]


def xy_to_uv(x, y):
    return (f32(f32(16.730260708356887) * x + f32(7.7801960340706) * y - f32(2.170152247475828)),
            f32(f32(-7.530081094743006) * x + f32(16.192422314095225) * y + f32(1.1125529268825947)))


# ---- block compression (Direct3D block-compression specification) ----
def _rgb565(c):
    return [f32((c >> 11) & 31) / f32(31), f32((c >> 5) & 63) / f32(63), f32(c & 31) / f32(31)]


def _color_block(b, three_color_allowed):
    c0, c1 = int(b[0]) | int(b[1]) << 8, int(b[2]) | int(b[3]) << 8
    idx = int.from_bytes(bytes(b[4:8]), "little")
    p0, p1 = _rgb565(c0), _rgb565(c1)
    if c0 > c1 or not three_color_allowed:
        pal = [p0 + [f32(1)], p1 + [f32(1)],
               [f32((f32(2) * p0[k] + p1[k]) / f32(3)) for k in range(3)] + [f32(1)],
               [f32((p0[k] + f32(2) * p1[k]) / f32(3)) for k in range(3)] + [f32(1)]]
    else:
        pal = [p0 + [f32(1)], p1 + [f32(1)], [f32((p0[k] + p1[k]) / f32(2)) for k in range(3)] + [f32(1)],
               [f32(0)] * 4]
    return [pal[(idx >> (2 * t)) & 3] for t in range(16)]


def _channel_block(b, signed):
    if signed:
        s0, s1 = int(np.int8(b[0])), int(np.int8(b[1]))
        eight = s0 > s1
        e0, e1 = f32(max(s0, -127)) / f32(127), f32(max(s1, -127)) / f32(127)
    else:
        eight = int(b[0]) > int(b[1])
        e0, e1 = f32(int(b[0])) / f32(255), f32(int(b[1])) / f32(255)
    p = [e0, e1]
    if eight:
        p += [f32((f32(8 - i) * e0 + f32(i - 1) * e1) / f32(7)) for i in range(2, 8)]
    else:
        p += [f32((f32(6 - i) * e0 + f32(i - 1) * e1) / f32(5)) for i in range(2, 6)]
        p += [f32(-1) if signed else f32(0), f32(1)]
    idx = int.from_bytes(bytes(b[2:8]), "little")
    return [p[(idx >> (3 * t)) & 7] for t in range(16)]


# BC6H and BC7 (Direct3D 11 specification: "BC6H Format", "BC7 Format"): partitions of a 4x4
# block into 2 or 3 subsets (texel t's subset: bit t of the 2-subset mask, digit t of the
# 3-subset row), the anchor texel of each further subset (whose index has one bit fewer), and
# the interpolation weights out of 64 for 2-, 3- and 4-bit indices
BC_PART2 = [0xCCCC, 0x8888, 0xEEEE, 0xECC8, 0xC880, 0xFEEC, 0xFEC8, 0xEC80, 0xC800, 0xFFEC, 0xFE80, 0xE800, 0xFFE8,
            0xFF00, 0xFFF0, 0xF000, 0xF710, 0x008E, 0x7100, 0x08CE, 0x008C, 0x7310, 0x3100, 0x8CCE, 0x088C, 0x3110,
            0x6666, 0x366C, 0x17E8, 0x0FF0, 0x718E, 0x399C, 0xAAAA, 0xF0F0, 0x5A5A, 0x33CC, 0x3C3C, 0x55AA, 0x9696,
            0xA55A, 0x73CE, 0x13C8, 0x324C, 0x3BDC, 0x6996, 0xC33C, 0x9966, 0x0660, 0x0272, 0x04E4, 0x4E40, 0x2720,
            0xC936, 0x936C, 0x39C6, 0x639C, 0x9336, 0x9CC6, 0x817E, 0xE718, 0xCCF0, 0x0FCC, 0x7744, 0xEE22]
BC_PART3 = [
    "0011001102212222", "0001001122112221", "0000200122112211", "0222002200110111",
    "0000000011221122", "0011001100220022", "0022002211111111", "0011001122112211",
    "0000000011112222", "0000111111112222", "0000111122222222", "0012001200120012",
    "0112011201120112", "0122012201220122", "0011011211221222", "0011200122002220",
    "0001001101121122", "0111001120012200", "0000112211221122", "0022002200221111",
    "0111011102220222", "0001000122212221", "0000001101220122", "0000110022102210",
    "0122012200110000", "0012001211222222", "0110122112210110", "0000011012211221",
    "0022110211020022", "0110011020022222", "0011012201220011", "0000200022112221",
    "0000000211221222", "0222002200120011", "0011001200220222", "0120012001200120",
    "0000111122220000", "0120120120120120", "0120201212010120", "0011220011220011",
    "0011112222000011", "0101010122222222", "0000000021212121", "0022112200221122",
    "0022001100220011", "0220122102201221", "0101222222220101", "0000212121212121",
    "0101010101012222", "0222011102220111", "0002111200021112", "0000211221122112",
    "0222011101110222", "0002111211120002", "0110011001102222", "0000000021122112",
    "0110011022222222", "0022001100110022", "0022112211220022", "0000000000002112",
    "0002000100020001", "0222122202221222", "0101222222222222", "0111201122012220",
]
BC_ANCHOR2 = [15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 15, 2, 8, 2, 2, 8, 8, 15, 2, 8, 2, 2, 8, 8,
              2, 2, 15, 15, 6, 8, 2, 8, 15, 15, 2, 8, 2, 2, 2, 15, 15, 6, 6, 2, 6, 8, 15, 15, 2, 2, 15, 15, 15, 15, 15,
              2, 2, 15]
BC_ANCHOR3A = [3, 3, 15, 15, 8, 3, 15, 15, 8, 8, 6, 6, 6, 5, 3, 3, 3, 3, 8, 15, 3, 3, 6, 10, 5, 8, 8, 6, 8, 5, 15, 15,
               8, 15, 3, 5, 6, 10, 8, 15, 15, 3, 15, 5, 15, 15, 15, 15, 3, 15, 5, 5, 5, 8, 5, 10, 5, 10, 8, 13, 15, 12,
               3, 3]
BC_ANCHOR3B = [15, 8, 8, 3, 15, 15, 3, 8, 15, 15, 15, 15, 15, 15, 15, 8, 15, 8, 15, 3, 15, 8, 15, 8, 3, 15, 6, 10, 15,
               15, 10, 8, 15, 3, 15, 10, 10, 8, 9, 10, 6, 15, 8, 15, 3, 6, 6, 8, 15, 3, 15, 15, 15, 15, 15, 15, 15, 15,
               15, 15, 3, 15, 15, 8]
BC_WEIGHTS = {2: [0, 21, 43, 64], 3: [0, 9, 18, 27, 37, 46, 55, 64],
              4: [0, 4, 9, 13, 17, 21, 26, 30, 34, 38, 43, 47, 51, 55, 60, 64]}


class _BitReader:
    """The 128 bits of a block, read from bit 0 up."""

    def __init__(self, b):
        self.v, self.p = int.from_bytes(bytes(b), "little"), 0

    def get(self, n):
        r = (self.v >> self.p) & ((1 << n) - 1)
        self.p += n
        return r


# BC7 modes: subsets, partition bits, rotation bits, index-selection bits, colour bits, alpha
# bits, per-endpoint p-bits, per-subset shared p-bits, index bits, secondary index bits
BC7_MODES = [(3, 4, 0, 0, 4, 0, 1, 0, 3, 0), (2, 6, 0, 0, 6, 0, 0, 1, 3, 0), (3, 6, 0, 0, 5, 0, 0, 0, 2, 0),
             (2, 6, 0, 0, 7, 0, 1, 0, 2, 0), (1, 0, 2, 1, 5, 6, 0, 0, 2, 3), (1, 0, 2, 0, 7, 8, 0, 0, 2, 2),
             (1, 0, 0, 0, 7, 7, 1, 0, 4, 0), (2, 6, 0, 0, 5, 5, 1, 0, 2, 0)]


def _bc7_block(b):
    """16 RGBA8 texels; the reserved mode (no set bit in the first byte) gives zeros."""
    br = _BitReader(b)
    mode = 0
    while mode < 8 and br.get(1) == 0:
        mode += 1
    if mode == 8:
        return [[0, 0, 0, 0]] * 16
    ns, pb, rb, isb, cb, ab, epb, spb, ib, ib2 = BC7_MODES[mode]
    part, rot, isel = br.get(pb), br.get(rb), br.get(isb)
    ep = [[[0] * 4 for _ in range(2)] for _ in range(ns)]
    for c in range(3):
        for s in range(ns):
            for e in range(2):
                ep[s][e][c] = br.get(cb)
    for s in range(ns if ab else 0):
        for e in range(2):
            ep[s][e][3] = br.get(ab)
    cprec, aprec = cb, ab
    if epb or spb:
        pbits = [[br.get(1), br.get(1)] for _ in range(ns)] if epb else [[p, p] for p in (br.get(1) for _ in range(ns))]
        for s in range(ns):
            for e in range(2):
                for c in range(4):
                    ep[s][e][c] = (ep[s][e][c] << 1) | pbits[s][e]
        cprec, aprec = cb + 1, (ab + 1 if ab else 0)

    def expand(v, prec):              # to 8 bits, the top bits repeated below
        v <<= 8 - prec
        return v | (v >> prec)
    for s in range(ns):
        for e in range(2):
            for c in range(3):
                ep[s][e][c] = expand(ep[s][e][c], cprec)
            ep[s][e][3] = expand(ep[s][e][3], aprec) if ab else 255
    anchors = [0] if ns == 1 else ([0, BC_ANCHOR2[part]] if ns == 2 else [0, BC_ANCHOR3A[part], BC_ANCHOR3B[part]])
    idx1 = [br.get(ib - 1 if t in anchors else ib) for t in range(16)]
    idx2 = [br.get(ib2 - 1 if t == 0 else ib2) for t in range(16)] if ib2 else idx1
    out = []
    for t in range(16):
        s = 0 if ns == 1 else ((BC_PART2[part] >> t) & 1 if ns == 2 else int(BC_PART3[part][t]))
        if ib2:
            ci, cbits, ai, abits = (idx2[t], ib2, idx1[t], ib) if isel else (idx1[t], ib, idx2[t], ib2)
        else:
            ci, cbits, ai, abits = idx1[t], ib, idx1[t], ib
        px = []
        for c in range(4):
            w = BC_WEIGHTS[cbits][ci] if c < 3 else BC_WEIGHTS[abits][ai]
            px.append(((64 - w) * ep[s][0][c] + w * ep[s][1][c] + 32) >> 6)
        if rot:
            px[rot - 1], px[3] = px[3], px[rot - 1]
        out.append(px)
    return out


# BC6H modes: mode value and its field width, transformed (delta) endpoints, endpoint bits,
# delta bits (R, G, B), regions, and where each endpoint bit lies after the mode field (in
# stream order; x[a:b] puts bit b first, then on towards a; endpoints 0-3 are region 0's two,
# then region 1's two)
BC6H_MODES = [
    (0b00, 2, True, 10, (5, 5, 5), 2, "g2[4] b2[4] b3[4] r0[9:0] g0[9:0] b0[9:0] r1[4:0] g3[4] g2[3:0] g1[4:0] b3[0] g3[3:0] b1[4:0] b3[1] b2[3:0] r2[4:0] b3[2] r3[4:0] b3[3]"),
    (0b01, 2, True, 7, (6, 6, 6), 2, "g2[5] g3[4] g3[5] r0[6:0] b3[0] b3[1] b2[4] g0[6:0] b2[5] b3[2] g2[4] b0[6:0] b3[3] b3[5] b3[4] r1[5:0] g2[3:0] g1[5:0] g3[3:0] b1[5:0] b2[3:0] r2[5:0] r3[5:0]"),
    (0b00010, 5, True, 11, (5, 4, 4), 2, "r0[9:0] g0[9:0] b0[9:0] r1[4:0] r0[10] g2[3:0] g1[3:0] g0[10] b3[0] g3[3:0] b1[3:0] b0[10] b3[1] b2[3:0] r2[4:0] b3[2] r3[4:0] b3[3]"),
    (0b00110, 5, True, 11, (4, 5, 4), 2, "r0[9:0] g0[9:0] b0[9:0] r1[3:0] r0[10] g3[4] g2[3:0] g1[4:0] g0[10] g3[3:0] b1[3:0] b0[10] b3[1] b2[3:0] r2[3:0] b3[0] b3[2] r3[3:0] g2[4] b3[3]"),
    (0b01010, 5, True, 11, (4, 4, 5), 2, "r0[9:0] g0[9:0] b0[9:0] r1[3:0] r0[10] b2[4] g2[3:0] g1[3:0] g0[10] b3[0] g3[3:0] b1[4:0] b0[10] b2[3:0] r2[3:0] b3[1] b3[2] r3[3:0] b3[4] b3[3]"),
    (0b01110, 5, True, 9, (5, 5, 5), 2, "r0[8:0] b2[4] g0[8:0] g2[4] b0[8:0] b3[4] r1[4:0] g3[4] g2[3:0] g1[4:0] b3[0] g3[3:0] b1[4:0] b3[1] b2[3:0] r2[4:0] b3[2] r3[4:0] b3[3]"),
    (0b10010, 5, True, 8, (6, 5, 5), 2, "r0[7:0] g3[4] b2[4] g0[7:0] b3[2] g2[4] b0[7:0] b3[3] b3[4] r1[5:0] g2[3:0] g1[4:0] b3[0] g3[3:0] b1[4:0] b3[1] b2[3:0] r2[5:0] r3[5:0]"),
    (0b10110, 5, True, 8, (5, 6, 5), 2, "r0[7:0] b3[0] b2[4] g0[7:0] g2[5] g2[4] b0[7:0] g3[5] b3[4] r1[4:0] g3[4] g2[3:0] g1[5:0] g3[3:0] b1[4:0] b3[1] b2[3:0] r2[4:0] b3[2] r3[4:0] b3[3]"),
    (0b11010, 5, True, 8, (5, 5, 6), 2, "r0[7:0] b3[1] b2[4] g0[7:0] b2[5] g2[4] b0[7:0] b3[5] b3[4] r1[4:0] g3[4] g2[3:0] g1[4:0] b3[0] g3[3:0] b1[5:0] b2[3:0] r2[4:0] b3[2] r3[4:0] b3[3]"),
    (0b11110, 5, False, 6, (6, 6, 6), 2, "r0[5:0] g3[4] b3[0] b3[1] b2[4] g0[5:0] g2[5] b2[5] b3[2] g2[4] b0[5:0] g3[5] b3[3] b3[5] b3[4] r1[5:0] g2[3:0] g1[5:0] g3[3:0] b1[5:0] b2[3:0] r2[5:0] r3[5:0]"),
    (0b00011, 5, False, 10, (10, 10, 10), 1, "r0[9:0] g0[9:0] b0[9:0] r1[9:0] g1[9:0] b1[9:0]"),
    (0b00111, 5, True, 11, (9, 9, 9), 1, "r0[9:0] g0[9:0] b0[9:0] r1[8:0] r0[10] g1[8:0] g0[10] b1[8:0] b0[10]"),
    (0b01011, 5, True, 12, (8, 8, 8), 1, "r0[9:0] g0[9:0] b0[9:0] r1[7:0] r0[10:11] g1[7:0] g0[10:11] b1[7:0] b0[10:11]"),
    (0b01111, 5, True, 16, (4, 4, 4), 1, "r0[9:0] g0[9:0] b0[9:0] r1[3:0] r0[10:15] g1[3:0] g0[10:15] b1[3:0] b0[10:15]"),
]


def _bc6h_layout(text):
    out = []
    for seg in text.split():
        rng = seg[3:-1]
        a, b = (int(x) for x in rng.split(":")) if ":" in rng else (int(rng), int(rng))
        step = 1 if a >= b else -1
        for k in range(b, a + step, step):
            out.append(("rgb".index(seg[0]), int(seg[1]), k))
    return out


BC6H_LAYOUTS = [_bc6h_layout(m[6]) for m in BC6H_MODES]


def _sext(v, n):
    v &= (1 << n) - 1
    return v - (1 << n) if v & (1 << (n - 1)) else v


def _bc6h_block(b, signed):
    """16 RGB texels as float32 (half-float values); reserved modes give zeros."""
    br = _BitReader(b)
    m = br.get(2)
    if m & 2:
        m |= br.get(3) << 2
    width = 5 if m & 2 else 2
    mode = next((i for i, md in enumerate(BC6H_MODES) if md[0] == m and md[1] == width), None)
    if mode is None:
        return [[f32(0)] * 3] * 16
    _, _, transformed, eb, db, regions, _ = BC6H_MODES[mode]
    e = [[0, 0, 0] for _ in range(4)]
    for c, i, k in BC6H_LAYOUTS[mode]:
        e[i][c] |= br.get(1) << k
    part = br.get(5) if regions == 2 else 0
    n_ep = 2 * regions
    for c in range(3):
        if signed:
            e[0][c] = _sext(e[0][c], eb)
        for i in range(1, n_ep):
            if transformed:     # endpoint = base + sign-extended delta, wrapped to eb bits
                v = (e[0][c] + _sext(e[i][c], db[c])) & ((1 << eb) - 1)
                e[i][c] = _sext(v, eb) if signed else v
            elif signed:
                e[i][c] = _sext(e[i][c], eb)

    def unquantize(c):
        if not signed:
            if eb >= 15:
                return c
            if c == 0:
                return 0
            if c == (1 << eb) - 1:
                return 0xFFFF
            return ((c << 16) + 0x8000) >> eb
        if eb >= 16:
            return c
        neg = c < 0
        c = -c if neg else c
        q = 0 if c == 0 else (0x7FFF if c >= (1 << (eb - 1)) - 1 else ((c << 15) + 0x4000) >> (eb - 1))
        return -q if neg else q

    ue = [[unquantize(e[i][c]) for c in range(3)] for i in range(n_ep)]
    anchors = [0, BC_ANCHOR2[part]] if regions == 2 else [0]
    ib = 3 if regions == 2 else 4
    idx = [br.get(ib - 1 if t in anchors else ib) for t in range(16)]
    out = []
    for t in range(16):
        s = (BC_PART2[part] >> t) & 1 if regions == 2 else 0
        w = BC_WEIGHTS[ib][idx[t]]
        px = []
        for c in range(3):
            q = (ue[2 * s][c] * (64 - w) + ue[2 * s + 1][c] * w + 32) >> 6
            if signed:   # to half-float bits: magnitude scaled by 31/32, sign in bit 15
                h = (0x8000 | (((-q) * 31) >> 5)) if q < 0 else (q * 31) >> 5
            else:
                h = (q * 31) >> 6
            px.append(f32(np.array([h], np.uint16).view(np.float16)[0]))
        out.append(px)
    return out


def decode_bc(fmt, data, width, height):
    """Mip level 0 of a block-compressed image as (height, width, 4) float32 texels: BC1-BC5 as
    floats in [0, 1] ([-1, 1] signed), BC7 as 8-bit unorm values / 255, BC6H as the half-float
    values with alpha 1."""
    data = np.ascontiguousarray(data, np.uint8).reshape(-1)
    bb = 8 if fmt in ("BC1", "BC4", "BC4_Signed") else 16
    bw, bh = (width + 3) // 4, (height + 3) // 4
    out = np.zeros((height, width, 4), f32)
    out[..., 3] = 1
    for by in range(bh):
        for bx in range(bw):
            b = data[(by * bw + bx) * bb:(by * bw + bx + 1) * bb]
            if fmt == "BC1":
                t = _color_block(b, True)
            elif fmt == "BC2":
                t = _color_block(b[8:], False)
                a = int.from_bytes(bytes(b[:8]), "little")
                t = [c[:3] + [f32((a >> (4 * k)) & 15) / f32(15)] for k, c in enumerate(t)]
            elif fmt == "BC3":
                t = _color_block(b[8:], False)
                a = _channel_block(b[:8], False)
                t = [c[:3] + [a[k]] for k, c in enumerate(t)]
            elif fmt in ("BC4", "BC4_Signed"):
                r = _channel_block(b, fmt == "BC4_Signed")
                t = [[r[k], f32(0), f32(0), f32(1)] for k in range(16)]
            elif fmt == "BC7":
                t = [[f32(v) / f32(255) for v in px] for px in _bc7_block(b)]
            elif fmt in ("BC6H", "BC6H_Signed"):
                t = [px + [f32(1)] for px in _bc6h_block(b, fmt == "BC6H_Signed")]
            else:
                r = _channel_block(b[:8], fmt == "BC5_Signed")
                g = _channel_block(b[8:], fmt == "BC5_Signed")
                t = [[r[k], g[k], f32(0), f32(1)] for k in range(16)]
            for k in range(16):
                x, y = bx * 4 + (k & 3), by * 4 + (k >> 2)
                if x < width and y < height:
                    out[y, x] = t[k]
    return out


# ---- images ----
class Texture:
    def __init__(self, img, spectral, env=False):
        """env: the texels as EnvironmentTextureShaderNode's sampler reads them (its read mode is
        plain normalised float, shader_nodes.cpp:1915-1919: no hardware sRGB read), and for a
        missing image its null image, 1 x 1 RGBA16F (1, 0, 1, 1) light source in Rec.709 D65
        (shader_nodes.cpp:1880-1881)."""
        if img is None:
            from types import SimpleNamespace
            px = np.array([1.0, 0.0, 1.0, 1.0], np.float16).view(np.uint8).reshape(1, 1, 8)
            img = SimpleNamespace(data=px, width=1, height=1, format="RGBA16Fx4", spectrum_type="Light Source",
                                  color_space="Rec709(D65)")
        self.width, self.height = img.width, img.height
        fmt = img.format
        if fmt.startswith("BC"):
            # decoded on the host; the texture unit's sRGB read of BC1-BC4 and BC7 degammas the colour
            # channels per texel (image.cpp:336-346)
            cs = COLOR_SPACES[img.color_space]
            self.uvs, self.kind, self.gray = False, "rgba", False
            self.spectrum_type = SPEC_TYPES[img.spectrum_type]
            t = decode_bc(fmt, img.data, img.width, img.height)
            degam = not env and cs == 0 and fmt in ("BC1", "BC2", "BC3", "BC4", "BC7")
            if degam:
                for y in range(img.height):
                    for x in range(img.width):
                        for k in ((0,) if fmt == "BC4" else (0, 1, 2)):
                            t[y, x, k] = degamma(t[y, x, k])
            self.color_space = 1 if degam else cs
            self.texels = t
            return
        st = SPEC_TYPES[img.spectrum_type]
        cs = COLOR_SPACES[img.color_space]
        srgb = cs == 0
        rgb8 = fmt in ("RGB8x3", "RGB_8x4", "RGBA8x4")
        half, full = fmt == "RGBA16Fx4", fmt == "RGBA32Fx4"
        self.uvs = spectral and st != 3 and (rgb8 or half or full)
        self.kind = ("uvs8" if rgb8 else "uvs16") if self.uvs else "rgba"
        self.gray = fmt in ("Gray32F", "Gray8", "GrayA8x2")
        self.spectrum_type = st
        hw = not env and srgb and not self.uvs and fmt in ("RGB8x3", "RGB_8x4", "RGBA8x4", "Gray8")
        create = srgb and not self.uvs and ((not spectral and (half or full)) or fmt in ("RG32Fx2", "Gray32F"))
        self.color_space = 1 if (hw or create) else cs
        raw = np.ascontiguousarray(img.data, np.uint8).reshape(img.height, img.width, -1)
        t = np.zeros((img.height, img.width, 4), f32)
        for y in range(img.height):
            for x in range(img.width):
                px = raw[y, x]
                c = [f32(0), f32(0), f32(0), f32(1)]
                if rgb8:
                    c[:3] = [f32(px[k]) / f32(255) for k in range(3)]
                    c[3] = f32(px[3]) / f32(255) if fmt == "RGBA8x4" else f32(1)
                elif half:
                    c = list(px[:8].view(np.float16).astype(f32))
                elif full:
                    c = list(px[:16].view(f32))
                elif fmt == "RG32Fx2":
                    c[:2] = list(px[:8].view(f32))
                elif fmt == "Gray32F":
                    c[0] = px[:4].view(f32)[0]
                elif fmt == "Gray8":
                    c[0] = f32(px[0]) / f32(255)
                elif fmt == "GrayA8x2":
                    c[0], c[1] = f32(px[0]) / f32(255), f32(px[1]) / f32(255)
                if self.uvs:
                    rgb = [degamma(v) if srgb else f32(v) for v in c[:3]]
                    X, Y, Z = tristim(MAT_E_TO_XYZ if st in (0, 2) else MAT_D65_TO_XYZ, rgb)
                    b = f32(X + Y + Z)
                    xy = (f32(X / b), f32(Y / b)) if b > 0 else (f32(1.0 / 3.0), f32(1.0 / 3.0))
                    u, v = xy_to_uv(*xy)
                    if rgb8:
                        q = lambda a: f32(int(f32(255) * min(max(a, f32(0)), f32(1)))) / f32(255)
                        t[y, x] = [q(u / f32(12)), q(v / f32(14)), q(b / f32(3)), c[3]]
                    else:
                        h = lambda a: f32(np.float16(a))
                        t[y, x] = [h(u), h(v), h(b), c[3] if half else h(c[3])]
                    continue
                if hw:
                    c[:3] = [degamma(v) for v in c[:3]]
                    if fmt == "Gray8":
                        c[1] = c[2] = f32(0)
                elif create:
                    n = 2 if fmt == "RG32Fx2" else (1 if fmt == "Gray32F" else 3)
                    for k in range(n):
                        c[k] = degamma(c[k])
                        if half:
                            c[k] = f32(np.float16(c[k]))
                elif srgb and fmt == "GrayA8x2":
                    c[0] = f32(int(min(255.0, math.floor(256.0 * float(degamma(c[0])))))) / f32(255)
                t[y, x] = c
        self.texels = t

    def texel(self, x, y, wu, wv):
        def wrap(i, n, mode):
            if mode == 0:
                return i % n, False
            if mode == 2:
                m = i % (2 * n)
                return (m if m < n else 2 * n - 1 - m), False
            if mode == 3 and (i < 0 or i >= n):
                return 0, True
            return min(max(i, 0), n - 1), False
        ix, ox = wrap(x, self.width, wu)
        iy, oy = wrap(y, self.height, wv)
        return np.zeros(4, f32) if (ox or oy) else self.texels[iy, ix]

    def sample(self, u, v, linear, wu, wv):
        u, v = f32(u), f32(v)
        if not linear:
            return self.texel(int(math.floor(u * f32(self.width))), int(math.floor(v * f32(self.height))), wu, wv)
        fx, fy = f32(u * f32(self.width) - f32(0.5)), f32(v * f32(self.height) - f32(0.5))
        x0, y0 = math.floor(fx), math.floor(fy)
        a, b = f32(fx - f32(x0)), f32(fy - f32(y0))
        w = [f32((1 - a) * (1 - b)), f32(a * (1 - b)), f32((1 - a) * b), f32(a * b)]
        t = [self.texel(x0, y0, wu, wv), self.texel(x0 + 1, y0, wu, wv), self.texel(x0, y0 + 1, wu, wv),
             self.texel(x0 + 1, y0 + 1, wu, wv)]
        return np.array([f32(w[0] * t[0][c] + w[1] * t[1][c] + w[2] * t[2][c] + w[3] * t[3][c]) for c in range(4)], f32)

    def gather(self, u, v, comp, wu, wv):
        x0 = math.floor(f32(f32(u) * f32(self.width) - f32(0.5)))
        y0 = math.floor(f32(f32(v) * f32(self.height) - f32(0.5)))
        return np.array([self.texel(x0, y0 + 1, wu, wv)[comp], self.texel(x0 + 1, y0 + 1, wu, wv)[comp],
                         self.texel(x0 + 1, y0, wu, wv)[comp], self.texel(x0, y0, wu, wv)[comp]], f32)


# ---- evaluation ----
class Point:
    """A surface point: position, geometric normal, shading frame, texture coordinates, and the
    wavelength offset of the spectral build."""

    def __init__(self, row):
        r = np.asarray(row, f32)
        self.p, self.gn = r[0:3], r[3:6]
        self.tx, self.ty, self.tz = r[6:9], r[9:12], r[12:15]
        self.uv = r[15:17]
        self.wl_offset = r[17]


def convert(want, src, v):
    """NodeTypeInfo<want>::convertFrom<src>; None where it is not defined."""
    v = np.asarray(v, f32)
    if want == F1:
        return v.copy() if src == F1 else None
    if want == F2:
        return v.copy() if src == F2 else (np.array([v[0], v[0], 0, 0], f32) if src == F1 else None)
    if want == F3:
        if src in (F3, P3, V3, N3):
            return np.array([v[0], v[1], v[2], 0], f32)
        return np.array([v[0], v[0], v[0], 0], f32) if src == F1 else None
    if want == F4:
        return v.copy() if src == F4 else (np.full(4, v[0], f32) if src == F1 else None)
    if want in (P3, V3):
        return np.array([v[0], v[1], v[2], 0], f32) if src in (P3, V3, N3, F3) else None
    if want == N3:
        if src == N3:
            return np.array([v[0], v[1], v[2], 0], f32)
        if src in (F3, P3, V3):
            ln = f32(math.sqrt(float(f32(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]))))
            return np.array([v[0] / ln, v[1] / ln, v[2] / ln, 0], f32)
    return None


class Evaluator:
    def __init__(self, spectral):
        self.spectral = spectral
        self.textures = {}

    def texture(self, img, env=False):
        key = (id(img), env)
        if key not in self.textures:
            self.textures[key] = Texture(img, self.spectral, env)
        return self.textures[key]

    def calc(self, plug, want, default, pt):
        """calcNode for a numeric type."""
        if plug is None:
            return np.asarray(default, f32)
        node, ptype, opt = plug
        src = CLASS_OF[ptype]
        if src == SPEC:
            return np.asarray(default, f32)
        out = convert(want, src, self.value(node, ptype, opt, pt))
        return np.asarray(default, f32) if out is None else out

    def value(self, node, ptype, opt, pt):
        z = np.zeros(4, f32)
        k = node.type
        fl = lambda name, i, d: f32(node.floats.get(name, [d] * (i + 1))[i])
        inp = lambda name: node.plugs.get(name)
        if k == "Geometry":
            if ptype == PLUG["Point3D"]:
                return np.array([*pt.p, 0], f32)
            if ptype == PLUG["Normal3D"]:
                return np.array([*(pt.gn if opt == 0 else pt.tz if opt == 1 else np.zeros(3, f32)), 0], f32)
            if ptype == PLUG["Vector3D"]:
                return np.array([*(pt.tx if opt == 0 else pt.ty if opt == 1 else np.zeros(3, f32)), 0], f32)
            if ptype == PLUG["TextureCoordinates"]:
                return np.array([pt.uv[0], pt.uv[1], 0, 0], f32)
            return z
        if k == "Tangent":
            # shader_nodes.cu:56-87 with the identity instance transform
            tt = node.enums.get("tangent type", "TC0 Direction")
            x, y, zz = pt.p
            if tt == "TC0 Direction":
                return np.array([*pt.tx, 0], f32)
            t = {"Radial X": (0, -zz, y), "Radial Y": (zz, 0, -x), "Radial Z": (-y, x, 0)}[tt]
            return np.array([*t, 0], f32)
        if k in ("Float2", "Float3", "Float4"):
            n = {"Float2": 2, "Float3": 3, "Float4": 4}[k]
            width = {0: 1, 1: 2, 2: 3, 3: 4}.get(ptype, 0)
            if opt + width > n:
                return z
            c = [self.calc(inp(str(i)), F1, [fl(str(i), 0, 0.0), 0, 0, 0], pt)[0] for i in range(opt, opt + width)]
            return np.array(c + [0] * (4 - width), f32)
        if k == "ScaleAndOffsetFloat":
            value = self.calc(inp("value"), F1, z, pt)[0]
            scale = self.calc(inp("scale"), F1, [fl("scale", 0, 1.0), 0, 0, 0], pt)[0]
            offset = self.calc(inp("offset"), F1, [fl("offset", 0, 0.0), 0, 0, 0], pt)[0]
            return np.array([f32(scale * value) + offset, 0, 0, 0], f32)
        if k == "ScaleAndOffsetUVTextureMap2D":
            return np.array([f32(fl("scale", 0, 1) * pt.uv[0]) + fl("offset", 0, 0),
                             f32(fl("scale", 1, 1) * pt.uv[1]) + fl("offset", 1, 0), 0, 0], f32)
        if k == "Image2DTexture":
            tex, wu, wv, linear, bump, coeff = self.sampler(node)
            tc = self.calc(inp("texcoord"), P3, [pt.uv[0], pt.uv[1], 0, 0], pt)
            if ptype == PLUG["Normal3D"]:
                # shader_nodes.cu:367-411
                n = np.array([0, 0, 1], f32)
                if bump != 2:
                    t = tex.sample(tc[0], tc[1], linear, wu, wv)
                    if opt < 2:
                        n = (t[0:3] if opt == 0 else t[1:4]).astype(f32)
                        n = (f32(2) * n - f32(1)).astype(f32)
                        if bump == 0:
                            n[1] = -n[1]
                else:
                    t = tex.gather(tc[0], tc[1], opt, wu, wv)
                    c = f32(5.0 / 1024)
                    dhdu = f32(f32(c * f32(tex.width)) * f32(t[1] - t[0]))
                    dhdv = f32(f32(c * f32(tex.height)) * f32(t[0] - t[3]))
                    n = np.array([-dhdu, dhdv, 1], f32)
                n[0] *= coeff
                n[1] *= coeff
                ln = f32(math.sqrt(float(f32(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]))))
                return np.array([n[0] / ln, n[1] / ln, n[2] / ln, 0], f32)
            t = tex.sample(tc[0], tc[1], linear, wu, wv)
            if ptype in (PLUG["float1"], PLUG["Alpha"]) and opt < 4:
                return np.array([t[opt], 0, 0, 0], f32)
            if ptype == PLUG["float2"] and opt < 3:
                return np.array([t[opt], t[opt + 1], 0, 0], f32)
            if ptype == PLUG["float3"] and opt < 2:
                return np.array([t[opt], t[opt + 1], t[opt + 2], 0], f32)
            if ptype == PLUG["float4"]:
                return t.copy()
            return z
        return z

    def sampler(self, node):
        wraps = ["Repeat", "Clamp to Edge", "Mirror", "Clamp to Border"]
        bumps = ["Normal Map (DirectX)", "Normal Map (OpenGL)", "Height Map"]
        wu = wraps.index(node.enums.get("wrap u", "Repeat"))
        wv = wraps.index(node.enums.get("wrap v", "Repeat"))
        linear = node.enums.get("mag filter", "Linear") == "Linear"
        bump = bumps.index(node.enums.get("bump type", "Normal Map (DirectX)"))
        # 5-bit coefficient (shader_nodes.cpp:1685-1688, shared.h:743-746)
        q = min(int(round(31 * float(node.floats.get("bump coeff", [1.0])[0]) * 0.5)), 31)
        return self.texture(node.image), wu, wv, linear, bump, f32(q + 1) / f32(16)

    def spectrum(self, plug, default, pt):
        """calcNode for a spectrum: the values at the point's 4 wavelengths (spectral build) or
        RGB (RGB build).  default: a callable giving the immediate's spectrum."""
        if plug is None or plug[1] != PLUG["Spectrum"]:
            return default()
        node, _, _ = plug
        lam = wavelengths(pt.wl_offset)
        st = lambda: SPEC_TYPES[node.enums.get("spectrum type", "Reflectance")]
        cs = lambda: COLOR_SPACES[node.enums.get("color space", "Rec709(D65)")]
        if node.type == "TripletSpectrum":
            tr = node.floats.get("triplet", [0.18, 0.18, 0.18])
            return self.triplet(st(), cs(), tr, lam)
        if node.type in ("RegularSampledSpectrum", "IrregularSampledSpectrum"):
            if not self.spectral:
                return self.sampled_rgb(node, st())
            return self.sampled(node, lam)
        if node.type == "Float3ToSpectrum":
            # shader_nodes.cu:254-274
            v = self.calc(node.plugs.get("value"), F3, [*node.floats.get("value", [0, 0, 0]), 0], pt)
            c = [min(max(f32(f32(0.5) * v[i] + f32(0.5)), f32(0)), f32(1)) for i in range(3)]
            if not self.spectral:
                return np.array(c, f32)
            return ob.meng_eval(st(), cs(), c, lam)
        if node.type == "Image2DTexture":
            # shader_nodes.cu:413-452
            tex, wu, wv, linear, _, _ = self.sampler(node)
            tc = self.calc(node.plugs.get("texcoord"), P3, [pt.uv[0], pt.uv[1], 0, 0], pt)
            t = tex.sample(tc[0], tc[1], linear, wu, wv).copy()
            if tex.gray:
                t[1] = t[2] = t[0]
            if not self.spectral:
                return t[:3].copy()
            if tex.uvs:
                u, v, s = t[0], t[1], t[2]
                if tex.kind == "uvs8":
                    u, v, s = f32(u * f32(12)), f32(v * f32(14)), f32(s * f32(3))
                return ob.meng_uvs_eval(u, v, f32(s / EER), lam)
            return ob.meng_eval(tex.spectrum_type, tex.color_space, t[:3], lam)
        if node.type == "EnvironmentTexture":
            # shader_nodes.cu:457-483: sampled at the surface's texture coordinates with repeat
            # wrapping; only uvsA16Fx4 texels are Meng coordinates, other texels a light-source
            # triplet in the image's colour space; the RGB build takes the texel as it is
            tex = self.texture(node.image, env=True)
            linear = node.enums.get("mag filter", "Linear") == "Linear"
            t = tex.sample(pt.uv[0], pt.uv[1], linear, 0, 0).copy()
            if not self.spectral:
                return t[:3].copy()
            if tex.kind == "uvs16":
                return ob.meng_uvs_eval(t[0], t[1], f32(t[2] / EER), lam)
            return ob.meng_eval(1, tex.color_space, t[:3], lam)
        return default()

    def triplet(self, st, cs, tr, lam):
        if self.spectral:
            return ob.meng_eval(st, cs, tr, lam)
        return ob.triplet_rgb(st, cs, tr)

    def sampled_rgb(self, node, spectrum_type):
        values = node.floats.get("values", [0.0]) or [0.0]
        if node.type == "RegularSampledSpectrum":
            return sampled_rgb(spectrum_type, True, node.floats.get("min wavelength", [360.0])[0],
                               node.floats.get("max wavelength", [830.0])[0], [], values)
        lam = list(node.floats.get("wavelengths", []))[:len(values)]
        lam += [0.0] * (len(values) - len(lam))
        return sampled_rgb(spectrum_type, False, 0, 0, lam, values)

    def sampled(self, node, lam):
        values = [f32(v) for v in node.floats["values"]]
        n = len(values)
        out = np.zeros(4, f32)
        if node.type == "RegularSampledSpectrum":
            lo, hi = f32(node.floats["min wavelength"][0]), f32(node.floats["max wavelength"][0])
            for i, wl in enumerate(lam):
                binf = f32(f32(f32(wl - lo) / f32(hi - lo)) * f32(n - 1))
                if binf <= 0:
                    out[i] = values[0]
                elif binf >= n - 1:
                    out[i] = values[n - 1]
                else:
                    b = int(binf)
                    t = f32(binf - f32(b))
                    out[i] = f32((f32(1) - t) * values[b]) + f32(t * values[b + 1])
            return out
        lambdas = [f32(v) for v in node.floats["wavelengths"]]
        base = 0
        for i, wl in enumerate(lam):
            p2 = 1
            while p2 < n - base:
                p2 <<= 1
            d = p2 >> 1
            while d > 0:
                if base + d < n and lambdas[base + d] <= wl:
                    base += d
                d >>= 1
            if base >= n - 1:
                out[i] = values[n - 1]
                continue
            t = f32(f32(wl - lambdas[base]) / f32(lambdas[base + 1] - lambdas[base]))
            out[i] = values[base] if t <= 0 else f32(f32((f32(1) - t) * values[base]) + f32(t * values[base + 1]))
        return out

    def evaluate(self, plug, pt):
        """What the device's entry point returns for one plug at one point (4 floats)."""
        node, ptype, opt = plug
        if ptype == PLUG["Spectrum"]:
            s = self.spectrum(plug, lambda: np.zeros(4 if self.spectral else 3, f32), pt)
            return np.array(list(s) + [0] * (4 - len(s)), f32)
        cls = CLASS_OF[ptype]
        return self.calc(plug, cls, np.zeros(4, f32), pt)
