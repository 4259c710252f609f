"""Pins of the shader-node oracle (oracle/shader_nodes.py) by closed forms and invariants:
bilinear filtering reproduces affine texel fields, the wrap modes' periodicity and symmetry,
the gather order against the filter's weights, the plug conversions, sampled spectra against
numpy's piecewise-linear interpolation, the RGB build's triplet of a sampled spectrum against
the equal-energy white and the CIE table, sRGB decoding at known values, the Meng texel
encodings against the direct upsampling of the same colour, normal and height maps of known
fields, and the radial tangents."""
import math

import numpy as np
import pytest

from oracle import binding as ob
from oracle import shader_nodes as osn
from vlr_amd.scenes import ImageDesc, NodeDesc, REC709, REC709_SRGB

P = osn.PLUG


def rgba32f(arr):
    a = np.ascontiguousarray(arr, np.float32)
    h, w = a.shape[:2]
    return ImageDesc(a.view(np.uint8).reshape(h, w, 16), w, h, "RGBA32Fx4", "NA", REC709)


def point(u=0.5, v=0.5, p=(0.0, 0.0, 0.0), offset=0.25):
    row = np.zeros(20, np.float32)
    row[0:3] = p
    row[3:6] = (0, 0, 1)
    row[6:9] = (1, 0, 0)
    row[9:12] = (0, 1, 0)
    row[12:15] = (0, 0, 1)
    row[15:17] = (u, v)
    row[17] = offset
    return osn.Point(row)


def tex_node(img, **enums):
    e = {"wrap u": "Clamp to Edge", "wrap v": "Clamp to Edge"}
    e.update(enums)
    return NodeDesc("Image2DTexture", enums=e, image=img)


def test_wavelengths():
    np.testing.assert_allclose(osn.wavelengths(0.0), [360.0, 477.5, 595.0, 712.5])
    np.testing.assert_allclose(osn.wavelengths(0.5), [418.75, 536.25, 653.75, 771.25])


def test_bilinear_reproduces_affine_fields():
    W, H = 7, 5
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float32)
    field = np.stack([1 + 2 * xs, 3 - ys, 0.5 * xs + 0.25 * ys, np.ones_like(xs)], -1)
    tex = osn.Texture(rgba32f(field), False)
    rng = np.random.default_rng(1)
    for _ in range(50):
        # interior: texel-centre coordinates between 0 and W-1 / H-1
        x, y = rng.uniform(0, W - 1), rng.uniform(0, H - 1)
        u, v = (x + 0.5) / W, (y + 0.5) / H
        got = tex.sample(u, v, True, 1, 1)
        fx, fy = np.float32(u) * W - 0.5, np.float32(v) * H - 0.5
        want = [1 + 2 * fx, 3 - fy, 0.5 * fx + 0.25 * fy, 1]
        np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5)


def test_nearest_and_wrap_modes():
    W, H = 4, 3
    rng = np.random.default_rng(2)
    field = rng.uniform(0, 1, (H, W, 4)).astype(np.float32)
    tex = osn.Texture(rgba32f(field), False)
    for j in range(H):
        for i in range(W):
            u, v = (i + 0.5) / W, (j + 0.5) / H
            np.testing.assert_array_equal(tex.sample(u, v, False, 0, 0), field[j, i])
            # repeat: period 1 in both coordinates
            np.testing.assert_array_equal(tex.sample(u + 2, v - 1, False, 0, 0), field[j, i])
            # mirror: texel i at u and at 2 - u (and at -u)
            np.testing.assert_array_equal(tex.sample(2 - u, v, False, 2, 0), field[j, i])
            np.testing.assert_array_equal(tex.sample(-u, v, False, 2, 0), field[j, i])
            # clamp to edge: beyond the right edge reads the last column
            np.testing.assert_array_equal(tex.sample(1.7, v, False, 1, 1), field[j, W - 1])
    # clamp to border: outside reads zero, linear filtering at the edge blends with zero
    np.testing.assert_array_equal(tex.sample(-0.2, 0.5, False, 3, 3), np.zeros(4))
    edge = tex.sample(0.0, 0.5 / H, True, 3, 3)
    np.testing.assert_allclose(edge, 0.5 * field[0, 0], rtol=1e-6)


def test_gather_order_matches_the_filter_weights():
    W, H = 5, 4
    rng = np.random.default_rng(3)
    field = rng.uniform(0, 1, (H, W, 4)).astype(np.float32)
    tex = osn.Texture(rgba32f(field), False)
    for _ in range(40):
        u, v = rng.uniform(0, 1), rng.uniform(0, 1)
        fx, fy = np.float32(u) * W - 0.5, np.float32(v) * H - 0.5
        a, b = fx - math.floor(fx), fy - math.floor(fy)
        for comp in range(4):
            g = tex.gather(u, v, comp, 0, 0)
            # CUDA's tex2Dgather: x (i0, j1), y (i1, j1), z (i1, j0), w (i0, j0)
            blend = (1 - a) * (1 - b) * g[3] + a * (1 - b) * g[2] + (1 - a) * b * g[0] + a * b * g[1]
            assert abs(blend - tex.sample(u, v, True, 0, 0)[comp]) < 1e-5


def test_conversions():
    ev = osn.Evaluator(False)
    f3 = NodeDesc("Float3", floats={"0": [3.0], "1": [0.0], "2": [4.0]})
    # Normal3D from float3: normalised
    n = osn.convert(osn.N3, osn.F3, ev.value(f3, P["float3"], 0, point()))
    np.testing.assert_allclose(n[:3], [0.6, 0.0, 0.8], rtol=1e-6)
    # float2 from float1: broadcast; float4 from float1: broadcast
    np.testing.assert_array_equal(osn.convert(osn.F2, osn.F1, [2, 0, 0, 0])[:2], [2, 2])
    np.testing.assert_array_equal(osn.convert(osn.F4, osn.F1, [2, 0, 0, 0]), [2, 2, 2, 2])
    # float1 from float3 is not defined: the consumer's default is used
    assert osn.convert(osn.F1, osn.F3, [1, 2, 3, 0]) is None
    scale = NodeDesc("ScaleAndOffsetFloat", floats={"scale": [2.0], "offset": [0.5]},
                     plugs={"value": (f3, P["float3"], 0)})
    np.testing.assert_array_equal(ev.value(scale, P["float1"], 0, point())[0], 0.5)
    scale.plugs["value"] = (f3, P["float1"], 2)
    np.testing.assert_array_equal(ev.value(scale, P["float1"], 0, point())[0], 8.5)
    # a component window past the node's width gives zero
    np.testing.assert_array_equal(ev.value(f3, P["float2"], 2, point()), np.zeros(4))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_sampled_spectra_match_linear_interpolation(seed):
    rng = np.random.default_rng(seed)
    ev = osn.Evaluator(True)
    n = int(rng.integers(2, 40))
    values = rng.uniform(0, 2, n).astype(np.float32)
    lo, hi = float(rng.uniform(350, 450)), float(rng.uniform(650, 850))
    reg = NodeDesc("RegularSampledSpectrum", floats={"min wavelength": [lo], "max wavelength": [hi],
                                                     "values": list(values)})
    lam = np.sort(rng.uniform(300, 900, n)).astype(np.float32)
    irr = NodeDesc("IrregularSampledSpectrum", floats={"wavelengths": list(lam), "values": list(values)})
    for off in (0.0, 0.3, 0.99):
        pt = point(offset=off)
        wl = osn.wavelengths(off)
        grid = np.linspace(lo, hi, n)
        np.testing.assert_allclose(ev.spectrum((reg, P["Spectrum"], 0), None, pt), np.interp(wl, grid, values),
                                   rtol=2e-5, atol=2e-5)
        np.testing.assert_allclose(ev.spectrum((irr, P["Spectrum"], 0), None, pt), np.interp(wl, lam, values),
                                   rtol=2e-5, atol=2e-5)


def test_srgb_texels_decoded():
    px = np.array([[[0, 128, 255, 255]]], np.uint8)
    tex = osn.Texture(ImageDesc(px, 1, 1, "RGBA8x4", "NA", REC709_SRGB), False)
    # sRGB 128 / 255 in linear light
    np.testing.assert_allclose(tex.texels[0, 0], [0.0, 0.21586050, 1.0, 1.0], rtol=1e-6)
    assert tex.color_space == osn.COLOR_SPACES[REC709]
    # spectrum type NA keeps RGBA in the spectral build too
    assert osn.Texture(ImageDesc(px, 1, 1, "RGBA8x4", "NA", REC709_SRGB), True).kind == "rgba"


def test_environment_texture_node_reads():
    """EnvironmentTexture on a surface (shader_nodes.cu:457-483, shader_nodes.cpp:1862-1921): the
    sampler reads without the hardware sRGB conversion, the RGB build returns the texel as it is,
    a missing image reads the null image (1, 0, 1), and in the spectral build an 8-bit image (stored
    as uvsA8x4) is taken as a light-source triplet while a half-float one is Meng coordinates."""
    px = np.array([[[0, 128, 255, 255]]], np.uint8)
    img = ImageDesc(px, 1, 1, "RGBA8x4", "Reflectance", REC709_SRGB)
    node = NodeDesc("EnvironmentTexture", image=img, enums={"mag filter": "Nearest"})
    rgb = osn.Evaluator(False).spectrum((node, P["Spectrum"], 0), None, point())
    np.testing.assert_array_equal(rgb, np.float32([0, 128, 255]) / np.float32(255))
    null = NodeDesc("EnvironmentTexture")
    np.testing.assert_array_equal(osn.Evaluator(False).spectrum((null, P["Spectrum"], 0), None, point()), [1, 0, 1])
    ev = osn.Evaluator(True)
    pt = point(offset=0.3)
    lam = osn.wavelengths(0.3)
    t8 = ev.texture(img, env=True)
    assert t8.kind == "uvs8"
    want8 = ob.meng_eval(1, t8.color_space, t8.texels[0, 0, :3], lam)
    np.testing.assert_array_equal(ev.spectrum((node, P["Spectrum"], 0), None, pt), want8)
    t16 = ev.texture(None, env=True)
    assert t16.kind == "uvs16"
    u, v, b = t16.texels[0, 0, :3]
    np.testing.assert_array_equal(ev.spectrum((null, P["Spectrum"], 0), None, pt),
                                  ob.meng_uvs_eval(u, v, np.float32(b / osn.EER), lam))
    # the null image is magenta: its upsampled spectrum is high at both ends and low in the middle
    s = ob.meng_uvs_eval(u, v, np.float32(b / osn.EER), np.float32([420, 550, 680, 700]))
    assert s[0] > 3 * s[1] and s[2] > 3 * s[1]


@pytest.mark.parametrize("fmt,tol", [("RGBA8x4", 0.03), ("RGBA16Fx4", 2e-3), ("RGBA32Fx4", 2e-3)])
def test_meng_texel_encodings_match_direct_upsampling(fmt, tol):
    rng = np.random.default_rng(4)
    rgb = rng.uniform(0.05, 0.95, (6, 3)).astype(np.float32)
    for c in rgb:
        if fmt == "RGBA8x4":
            q = np.round(c * 255).astype(np.uint8)
            data = np.array([[[q[0], q[1], q[2], 255]]], np.uint8)
            lin = q.astype(np.float32) / 255
        elif fmt == "RGBA16Fx4":
            h = np.array([c[0], c[1], c[2], 1], np.float16)
            data = h.view(np.uint8).reshape(1, 1, 8)
            lin = h[:3].astype(np.float32)
        else:
            data = np.array([c[0], c[1], c[2], 1], np.float32).view(np.uint8).reshape(1, 1, 16)
            lin = c
        img = ImageDesc(data, 1, 1, fmt, "Reflectance", REC709)
        node = tex_node(img, **{"mag filter": "Nearest"})
        ev = osn.Evaluator(True)
        assert ev.texture(img).kind == ("uvs8" if fmt == "RGBA8x4" else "uvs16")
        pt = point(offset=0.4)
        got = ev.spectrum((node, P["Spectrum"], 0), None, pt)
        want = ob.meng_eval(0, 1, lin, osn.wavelengths(0.4))
        np.testing.assert_allclose(got, want, rtol=tol, atol=tol)


def test_normal_maps():
    flat = rgba32f(np.array([[[0.5, 0.5, 1.0, 1.0]]], np.float32))
    ev = osn.Evaluator(False)
    for bump in ("Normal Map (DirectX)", "Normal Map (OpenGL)"):
        n = ev.value(tex_node(flat, **{"bump type": bump}), P["Normal3D"], 0, point())
        np.testing.assert_array_equal(n[:3], [0, 0, 1])
    tilted = rgba32f(np.array([[[0.75, 0.75, 1.0, 1.0]]], np.float32))
    # the default coefficient 1 is stored in 5 bits as round(31 * 1 / 2) = 16 and read as 17 / 16
    c = 17 / 16
    s = 1 / math.sqrt(0.5 * c * c + 1)
    dx = ev.value(tex_node(tilted, **{"bump type": "Normal Map (DirectX)"}), P["Normal3D"], 0, point())
    gl = ev.value(tex_node(tilted, **{"bump type": "Normal Map (OpenGL)"}), P["Normal3D"], 0, point())
    np.testing.assert_allclose(dx[:3], [0.5 * c * s, -0.5 * c * s, s], rtol=1e-6)
    np.testing.assert_allclose(gl[:3], [0.5 * c * s, 0.5 * c * s, s], rtol=1e-6)
    # coefficient 2 (stored as 31, read as 32 / 16) doubles the tilt
    node2 = NodeDesc("Image2DTexture", floats={"bump coeff": [2.0]}, enums={"bump type": "Normal Map (DirectX)"},
                     image=tilted)
    dx2 = ev.value(node2, P["Normal3D"], 0, point())
    np.testing.assert_allclose(dx2[:3] / dx2[2], [1.0, -1.0, 1.0], rtol=1e-6)


def test_height_map_of_a_ramp():
    W, H, k = 16, 8, 0.01
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float32)
    field = np.stack([k * xs, np.zeros_like(xs), np.zeros_like(xs), np.ones_like(xs)], -1)
    img = rgba32f(field)
    node = tex_node(img, **{"bump type": "Height Map"})
    n = osn.Evaluator(False).value(node, P["Normal3D"], 0, point(u=0.4, v=0.6))
    # slope k per texel: dh/du = (5 / 1024) * W * k, tilted by the stored coefficient 17 / 16;
    # the normal leans against it
    dhdu = 5.0 / 1024 * W * k * (17 / 16)
    want = np.array([-dhdu, 0.0, 1.0]) / math.hypot(dhdu, 1.0)
    np.testing.assert_allclose(n[:3], want, rtol=1e-5, atol=1e-7)


def test_radial_tangents():
    ev = osn.Evaluator(False)
    cases = {"Radial X": ((0, 1, 0), (0, 0, 1)), "Radial Y": ((0, 0, 1), (1, 0, 0)),
             "Radial Z": ((1, 0, 0), (0, 1, 0))}
    for tt, (p, t) in cases.items():
        node = NodeDesc("Tangent", enums={"tangent type": tt})
        got = ev.value(node, P["Vector3D"], 0, point(p=p))
        np.testing.assert_array_equal(got[:3], t)
        # perpendicular to the position and to the axis
        assert abs(float(np.dot(got[:3], p))) == 0.0


def test_float3_to_spectrum_rgb_maps_minus_one_to_one_onto_zero_to_one():
    ev = osn.Evaluator(False)
    f3 = NodeDesc("Float3", floats={"0": [-1.0], "1": [0.0], "2": [3.0]})
    node = NodeDesc("Float3ToSpectrum", plugs={"value": (f3, P["float3"], 0)})
    np.testing.assert_array_equal(ev.spectrum((node, P["Spectrum"], 0), None, point()), [0.0, 0.5, 1.0])


# ---- the RGB build's triplet of a sampled spectrum ----
# XYZ -> linear sRGB (D65), Lindbloom's published matrix (rows)
XYZ_TO_SRGB_D65 = np.array([[3.2404542, -1.5371385, -0.4985314], [-0.9692660, 1.8760108, 0.0415560],
                            [0.0556434, -0.2040259, 1.0572252]])


def cmf64():
    return [np.asarray(t, np.float64) for t in osn.cmf_table()]


def test_sampled_rgb_of_a_flat_spectrum():
    """A flat reflectance of 1 is the equal-energy white: XYZ = (int xbar, int ybar, int zbar) /
    int ybar, which the E-white matrix takes to (1, 1, 1) within the CIE table's imbalance; a
    flat light source goes through the D65 matrix. Both also as a regular spectrum over
    400-700 nm, held constant beyond its ends, and at half the value (linearity)."""
    x, y, z = cmf64()
    xyz = np.array([np.trapezoid(x), np.trapezoid(y), np.trapezoid(z)]) / np.trapezoid(y)
    refl = osn.sampled_rgb(0, False, 0, 0, [360.0], [1.0])
    np.testing.assert_allclose(refl, [1, 1, 1], atol=2e-3)
    light = osn.sampled_rgb(1, False, 0, 0, [360.0], [1.0])
    np.testing.assert_allclose(light, XYZ_TO_SRGB_D65 @ xyz, rtol=2e-5)
    for st, want in ((0, refl), (1, light)):
        np.testing.assert_allclose(osn.sampled_rgb(st, True, 400.0, 700.0, [], [1.0] * 7), want, rtol=2e-6)
        np.testing.assert_allclose(osn.sampled_rgb(st, False, 0, 0, [412.5, 700.25], [0.5, 0.5]), 0.5 * want,
                                   rtol=2e-6)


@pytest.mark.parametrize("lam", [455, 500, 555, 610])
def test_sampled_rgb_of_a_one_nanometre_triangle(lam):
    """Values 0, 1, 0 at lam - 1, lam, lam + 1 nm: each interval adds the mean value times the
    mean CMF times its width (toXYZ's rule), so XYZ = (c(lam-1) + 2 c(lam) + c(lam+1)) / 4 over
    int ybar. A trapezoid of the products, or a 1 nm sum, would give c(lam) instead."""
    cm = cmf64()
    i = lam - 360
    xyz = np.array([(c[i - 1] + 2 * c[i] + c[i + 1]) / 4 for c in cm]) / np.trapezoid(cm[1])
    got = osn.sampled_rgb(1, False, 0, 0, [lam - 1.0, float(lam), lam + 1.0], [0.0, 1.0, 0.0])
    want = np.maximum(XYZ_TO_SRGB_D65 @ xyz, 0)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-9)
    plain = np.maximum(XYZ_TO_SRGB_D65 @ (np.array([c[i] for c in cm]) / np.trapezoid(cm[1])), 0)
    assert np.max(np.abs(got - plain)) > 1e-4 * np.max(np.abs(plain))


def test_sampled_rgb_between_grid_points():
    """Samples off the 1 nm grid: a linear ramp sampled at 3 irregular points equals the same
    ramp given as a regular spectrum of 2 samples (interpolation between samples is linear),
    and the result is linear in the values."""
    lo, hi = 380.5, 779.25
    mid = 512.75
    f = lambda l: 0.2 + 1.3 * (l - lo) / (hi - lo)
    for st in (0, 1):
        a = osn.sampled_rgb(st, False, 0, 0, [lo, mid, hi], [f(lo), f(mid), f(hi)])
        b = osn.sampled_rgb(st, True, lo, hi, [], [f(lo), f(hi)])
        np.testing.assert_allclose(a, b, rtol=2e-5)
        c = osn.sampled_rgb(st, True, lo, hi, [], [2 * f(lo), 2 * f(hi)])
        np.testing.assert_allclose(c, 2 * b, rtol=2e-6)


# ---- block-compressed images: invariants of the Direct3D block formats ----
def bc1_block(c0, c1, indices):
    idx = 0
    for t, i in enumerate(indices):
        idx |= i << (2 * t)
    return np.frombuffer(np.array([c0, c1], "<u2").tobytes() + int(idx).to_bytes(4, "little"), np.uint8)


def bc4_block(e0, e1, indices, signed=False):
    idx = 0
    for t, i in enumerate(indices):
        idx |= i << (3 * t)
    head = np.array([e0, e1], np.int8 if signed else np.uint8).tobytes()
    return np.frombuffer(head + int(idx).to_bytes(6, "little"), np.uint8)


def test_bc1_endpoints_and_thirds():
    white, black = 0xFFFF, 0x0000
    t = osn.decode_bc("BC1", bc1_block(white, black, [0, 1, 2, 3] * 4), 4, 4)
    np.testing.assert_array_equal(t[0, 0], [1, 1, 1, 1])      # index 0: first endpoint exactly
    np.testing.assert_array_equal(t[0, 1], [0, 0, 0, 1])      # index 1: second endpoint
    np.testing.assert_allclose(t[0, 2, :3], 2 / 3, rtol=1e-6)  # thirds between them
    np.testing.assert_allclose(t[0, 3, :3], 1 / 3, rtol=1e-6)
    # c0 <= c1: three colours and transparent black
    t = osn.decode_bc("BC1", bc1_block(black, white, [2, 3] * 8), 4, 4)
    np.testing.assert_allclose(t[0, 0], [0.5, 0.5, 0.5, 1], rtol=1e-6)
    np.testing.assert_array_equal(t[0, 1], [0, 0, 0, 0])
    # a solid block is constant whatever its indices; 565 channels expand to 1 at full scale
    red = 0xF800
    rng = np.random.default_rng(0)
    t = osn.decode_bc("BC1", bc1_block(red, red, rng.integers(0, 3, 16)), 4, 4)
    assert np.all(t == np.array([1, 0, 0, 1], np.float32))


def test_bc_block_layout_and_partial_blocks():
    # 6 x 5 texels: 2 x 2 blocks, texels past the edge dropped; block k is solid grey level k
    greys = [0x0000, 0x4208, 0x8410, 0xFFFF]
    data = np.concatenate([bc1_block(g, g, [0] * 16) for g in greys])
    t = osn.decode_bc("BC1", data, 6, 5)
    assert t.shape == (5, 6, 4)
    for k, (bx, by) in enumerate([(0, 0), (1, 0), (0, 1), (1, 1)]):
        blk = t[by * 4:by * 4 + 4, bx * 4:bx * 4 + 4]
        assert np.all(blk == blk[0, 0])
    assert t[0, 0, 0] == 0 and t[4, 5, 0] == 1


def test_bc4_and_bc5_palettes():
    t = osn.decode_bc("BC4", bc4_block(255, 0, list(range(8)) * 2), 4, 4)
    r = t[..., 0].reshape(-1)[:8]
    np.testing.assert_allclose(r, [1, 0, 6 / 7, 5 / 7, 4 / 7, 3 / 7, 2 / 7, 1 / 7], rtol=1e-6, atol=1e-7)
    # e0 <= e1: fifths, then the constants 0 and 1
    t = osn.decode_bc("BC4", bc4_block(0, 255, list(range(8)) * 2), 4, 4)
    np.testing.assert_allclose(t[..., 0].reshape(-1)[:8], [0, 1, 1 / 5, 2 / 5, 3 / 5, 4 / 5, 0, 1], atol=1e-7)
    # signed: -128 reads as -127 (-1), constants -1 and 1
    t = osn.decode_bc("BC4_Signed", bc4_block(-128, 127, [0, 1, 6, 7] * 4, signed=True), 4, 4)
    np.testing.assert_allclose(t[..., 0].reshape(-1)[:4], [-1, 1, -1, 1], atol=1e-7)
    # BC5 = two BC4 blocks in red and green
    d = np.concatenate([bc4_block(255, 0, [0] * 16), bc4_block(255, 0, [1] * 16)])
    t = osn.decode_bc("BC5", d, 4, 4)
    assert np.all(t[..., 0] == 1) and np.all(t[..., 1] == 0) and np.all(t[..., 3] == 1)


def test_bc2_and_bc3_alpha():
    a = 0
    for k in range(16):
        a |= k << (4 * k)
    d = np.concatenate([np.frombuffer(int(a).to_bytes(8, "little"), np.uint8), bc1_block(0xFFFF, 0xFFFF, [0] * 16)])
    t = osn.decode_bc("BC2", d, 4, 4)
    np.testing.assert_allclose(t[..., 3].reshape(-1), np.arange(16) / 15, rtol=1e-6)
    # BC3: the colour block is always four-colour, alpha from a BC4 block
    d = np.concatenate([bc4_block(0, 255, [7] * 16), bc1_block(0x0000, 0xFFFF, [3] * 16)])
    t = osn.decode_bc("BC3", d, 4, 4)
    np.testing.assert_allclose(t[0, 0], [2 / 3, 2 / 3, 2 / 3, 1], rtol=1e-6)


# ---- BC6H and BC7 ----
def _pillow_bcn(data, n, mode, fmt=None):
    from PIL import Image
    args = (n,) if fmt is None else (n, fmt)
    return np.asarray(Image.frombytes(mode, (4, 4), bytes(data), "bcn", args))


def _random_blocks(rng, count, mode_bits, width):
    for _ in range(count):
        v = int.from_bytes(rng.integers(0, 256, 16, dtype=np.uint8).tobytes(), "little")
        yield ((v & ~((1 << width) - 1)) | mode_bits).to_bytes(16, "little")


def test_bc7_matches_pillows_decoder():
    """Pillow's BCn decoder (an implementation of the same Direct3D specification) decodes 300
    random blocks of each of BC7's eight modes (every partition, rotation, index selection and
    p-bit) to the same RGBA8 texels as the oracle.  The reserved mode (no bit set in the first
    byte) decodes to zeros, as the specification says (Pillow returns opaque black)."""
    rng = np.random.default_rng(71)
    for mode in range(8):
        for b in _random_blocks(rng, 300, 1 << mode, mode + 1):
            want = _pillow_bcn(b, 7, "RGBA").reshape(16, 4)
            assert np.array_equal(np.array(osn._bc7_block(b)), want), (mode, b.hex())
    assert osn._bc7_block(bytes(16)) == [[0, 0, 0, 0]] * 16
    # decode_bc: values / 255, partial blocks
    data = np.frombuffer(b"".join(_random_blocks(rng, 4, 1 << 6, 7)), np.uint8)
    t = osn.decode_bc("BC7", data, 6, 5)
    assert t.shape == (5, 6, 4)
    np.testing.assert_array_equal(t[:4, :4], _pillow_bcn(data[:16], 7, "RGBA").astype(np.float32) / 255)


def test_bc6h_unsigned_matches_pillows_decoder():
    """200 random blocks of each of BC6H's fourteen modes (every endpoint-bit layout, delta
    width and partition): the oracle's half-float texels, read as Pillow reads them
    (clamp to [0, 1], times 255, truncated), equal Pillow's 8-bit texels to within one step
    (Pillow's values differ by one half-float ulp in about 0.1 % of texels); reserved modes
    decode to zeros."""
    rng = np.random.default_rng(72)
    worst, count = 0, 0
    for mode_bits, width, *_ in osn.BC6H_MODES:
        for b in _random_blocks(rng, 200, mode_bits, width):
            want = _pillow_bcn(b, 6, "RGB", "uf").reshape(16, 3).astype(np.float64)
            got = np.array(osn._bc6h_block(b, False), np.float64)
            assert np.all(np.isfinite(got)) and got.min() >= 0 and got.max() <= 65504
            worst = max(worst, float(np.abs(np.floor(np.clip(got, 0, 1) * 255) - want).max()))
            count += 1
    assert worst <= 1, worst
    for reserved in (0b10011, 0b10111, 0b11011, 0b11111):
        blk = (reserved | (0xABCDEF << 5)).to_bytes(16, "little")
        assert osn._bc6h_block(blk, False) == [[0.0] * 3] * 16


def _bc6h_mode11(e0, e1, index_bits):
    """A BC6H mode-11 block (one region, raw 10-bit endpoints) with the given RGB endpoints and
    16 4-bit indices (texel 0's index has 3 bits)."""
    v, p = 0b00011, 5
    for e in (e0, e1):
        for c in range(3):
            v |= (e[c] & 0x3FF) << p
            p += 10
    for t, i in enumerate(index_bits):
        n = 3 if t == 0 else 4
        v |= (i & ((1 << n) - 1)) << p
        p += n
    assert p == 128
    return v.to_bytes(16, "little")


def test_bc6h_signed_closed_forms():
    """The signed variant (Pillow decodes it as unsigned, so it pins nothing there) at values
    the specification fixes: the largest 10-bit magnitude 511 unquantizes to 0x7FFF and ends as
    the largest half, +-65504; zero stays zero; weight 64 picks the second endpoint; -1
    unquantizes to -((2^15 + 2^14) >> 9) = -96, which becomes the half -(96 * 31 >> 5) ulps =
    -93 * 2^-24 (at 11 bits: -48, then -46 * 2^-24); a transformed mode adds a sign-extended
    delta to the base (mode 12: base -1, delta +1 gives endpoint 0)."""
    top = osn._bc6h_block(_bc6h_mode11((511, 511, 511), (-511, 0, 511), [0] * 16), True)
    assert top[0] == [65504.0, 65504.0, 65504.0]
    blk = osn._bc6h_block(_bc6h_mode11((0, 511, -511), (-511, 0, 511), [15] * 16), True)
    assert blk[5] == [-65504.0, 0.0, 65504.0]
    blk = osn._bc6h_block(_bc6h_mode11((-1, 0, 1), (0, 0, 0), [0] * 16), True)
    assert blk[3][0] == np.float32(-93 * 2.0 ** -24) and blk[3][1] == 0.0 and blk[3][2] == np.float32(93 * 2.0 ** -24)
    # unsigned: the largest 10-bit value 1023 becomes 0xFFFF and the half 65504 too
    assert osn._bc6h_block(_bc6h_mode11((1023, 0, 1), (0, 0, 0), [0] * 16), False)[0][0] == 65504.0
    # mode 12 (11-bit base, 9-bit deltas): r0 = -1 (all ones), r1 delta = +1 -> endpoint 0
    v, p = 0b00111, 5
    fields = [(0x3FF, 10), (0, 10), (0, 10), (1, 9), (1, 1), (0, 9), (0, 1), (0, 9), (0, 1)]
    for val, n in fields:
        v |= val << p
        p += n
    # indices: texel 0 (3 bits) 0, texel 1 picks weight 64 (index 15)
    v |= 0 << p
    v |= 15 << (p + 3)
    b = v.to_bytes(16, "little")
    t = osn._bc6h_block(b, True)
    # endpoint 0: -1 at 11 bits unquantizes to -((2^15 + 2^14) >> 10) = -48, the half -(48 * 31 >> 5)
    assert t[0][0] == np.float32(-46 * 2.0 ** -24)
    assert t[1][0] == 0.0                                 # endpoint 1: -1 + 1 = 0
