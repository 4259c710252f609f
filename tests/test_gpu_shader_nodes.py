"""Shader-node parity: the device's compiled node programs against the oracle's recursive
restatement (oracle/shader_nodes.py), plug by plug at random surface points, in the RGB and the
spectral build; then renders whose node-driven parameters have an equivalent node-free scene
that the oracle path tracer renders (a two-texel albedo texture against a floor split in two,
constant normal maps against tilted vertex normals, a radial tangent against vertex tangents).

Tolerance: float32 arithmetic contracted differently on the device (fused multiply-adds) moves
results by a few ulps, so every value must agree within 2e-4 relative plus 2e-5 of the case's
scale.  A point whose texture coordinate lands within float rounding of a nearest-filter texel
edge (or whose wavelength offset lands on a wavelength-bin edge) may fall on the other side on
the device; such a value is accepted only if it matches, at the same tolerance, the oracle
evaluated at the point with its texture coordinates or wavelength offset moved by 2e-6 (a few
float32 ulps of the inputs).  No other value is exempt."""
import math

import numpy as np
import pytest

from oracle import binding as ob
from oracle import shader_nodes as osn
from vlr_amd import api, scenes
from vlr_amd.scenes import ImageDesc, NodeDesc, REC709, REC709_SRGB, XYZ

pytestmark = pytest.mark.gpu

P = osn.PLUG
NPTS = 192


def random_points(n, seed):
    rng = np.random.default_rng(seed)
    pts = np.zeros((n, 20), np.float32)
    pts[:, 0:3] = rng.uniform(-2, 2, (n, 3))
    z = rng.normal(size=(n, 3))
    z /= np.linalg.norm(z, axis=1, keepdims=True)
    a = np.where(np.abs(z[:, :1]) < 0.9, [[1.0, 0, 0]], [[0, 1.0, 0]])
    x = np.cross(a, z)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    y = np.cross(z, x)
    g = z + rng.normal(scale=0.1, size=(n, 3))
    pts[:, 3:6] = g / np.linalg.norm(g, axis=1, keepdims=True)
    pts[:, 6:9], pts[:, 9:12], pts[:, 12:15] = x, y, z
    pts[:, 15:17] = rng.uniform(-0.25, 1.25, (n, 2))
    pts[:, 17] = rng.uniform(0, 1, n)
    return pts


def image(fmt, w, h, spectrum_type, color_space, seed):
    rng = np.random.default_rng(seed)
    if fmt.startswith("BC"):
        bb = 8 if fmt in ("BC1", "BC4", "BC4_Signed") else 16
        data = rng.integers(0, 256, ((w + 3) // 4) * ((h + 3) // 4) * bb, dtype=np.uint8)
        return ImageDesc(data, w, h, fmt, spectrum_type, color_space)
    if fmt in ("RGBA8x4", "RGB8x3", "Gray8", "GrayA8x2"):
        ch = {"RGBA8x4": 4, "RGB8x3": 3, "Gray8": 1, "GrayA8x2": 2}[fmt]
        data = rng.integers(0, 256, (h, w, ch), dtype=np.uint8)
    elif fmt == "RGBA16Fx4":
        data = rng.uniform(0, 1.2, (h, w, 4)).astype(np.float16).view(np.uint8).reshape(h, w, 8)
    else:
        ch = {"RGBA32Fx4": 4, "RG32Fx2": 2, "Gray32F": 1}[fmt]
        data = rng.uniform(0, 1, (h, w, ch)).astype(np.float32).view(np.uint8).reshape(h, w, 4 * ch)
    return ImageDesc(data, w, h, fmt, spectrum_type, color_space)


IMAGES = {
    "rgba8_srgb_refl": image("RGBA8x4", 5, 3, "Reflectance", REC709_SRGB, 1),
    "rgb8_linear_na": image("RGB8x3", 4, 4, "NA", REC709, 2),
    "rgba16f_srgb_light": image("RGBA16Fx4", 3, 5, "Light Source", REC709_SRGB, 3),
    "rgba32f_refl": image("RGBA32Fx4", 6, 2, "Reflectance", REC709, 4),
    "rgba32f_xyz": image("RGBA32Fx4", 2, 2, "Reflectance", XYZ, 5),
    "gray8_srgb": image("Gray8", 4, 3, "NA", REC709_SRGB, 6),
    "gray32f_srgb_refl": image("Gray32F", 3, 3, "Reflectance", REC709_SRGB, 7),
    "graya8_srgb_refl": image("GrayA8x2", 5, 2, "Reflectance", REC709_SRGB, 8),
    "rg32f_srgb": image("RG32Fx2", 2, 5, "NA", REC709_SRGB, 9),
    # block-compressed: partial blocks, hardware-style sRGB decode of BC1-BC4, signed BC5
    "bc1_srgb_refl": image("BC1", 6, 5, "Reflectance", REC709_SRGB, 10),
    "bc2_linear_light": image("BC2", 4, 4, "Light Source", REC709, 11),
    "bc3_srgb_refl": image("BC3", 8, 4, "Reflectance", REC709_SRGB, 12),
    "bc4_srgb_na": image("BC4", 8, 8, "NA", REC709_SRGB, 13),
    "bc5s_na": image("BC5_Signed", 4, 8, "NA", REC709, 14),
    # BC7 (sRGB read like BC1-BC4) and BC6H half floats, unsigned and signed
    "bc7_srgb_refl": image("BC7", 6, 7, "Reflectance", REC709_SRGB, 15),
    "bc6h_light": image("BC6H", 5, 4, "Light Source", REC709, 16),
    "bc6hs_na": image("BC6H_Signed", 4, 4, "NA", REC709, 17),
}
WRAPS = ["Repeat", "Clamp to Edge", "Mirror", "Clamp to Border"]


def uv_map():
    return NodeDesc("ScaleAndOffsetUVTextureMap2D", floats={"scale": [1.7, -0.8], "offset": [0.1, 0.3]})


def texture_cases():
    cases = []
    for k, (name, img) in enumerate(IMAGES.items()):
        for filt in ("Nearest", "Linear"):
            node = NodeDesc("Image2DTexture", image=img,
                            enums={"mag filter": filt, "wrap u": WRAPS[k % 4], "wrap v": WRAPS[(k + 1 + (filt == "Linear")) % 4]})
            if (k + (filt == "Linear")) % 2:
                node.plugs["texcoord"] = (uv_map(), P["TextureCoordinates"], 0)
            tag = f"{name}-{filt}"
            cases.append((f"tex-spectrum-{tag}", (node, P["Spectrum"], 0)))
            cases.append((f"tex-float4-{tag}", (node, P["float4"], 0)))
            cases.append((f"tex-float3-{tag}", (node, P["float3"], 1)))
            cases.append((f"tex-float2-{tag}", (node, P["float2"], 2)))
            cases.append((f"tex-float1-{tag}", (node, P["float1"], k % 4)))
            cases.append((f"tex-alpha-{tag}", (node, P["Alpha"], 3)))
    # EnvironmentTexture on a surface: read at the surface's texture coordinates, repeat wrap,
    # no hardware sRGB read; no image reads the node's magenta null image
    for k, name in enumerate(("rgba16f_srgb_light", "rgba8_srgb_refl", "rgb8_linear_na", "gray8_srgb",
                              "rgba32f_xyz", "bc1_srgb_refl", "bc6h_light", None)):
        filt = ("Nearest", "Linear")[k % 2]
        node = NodeDesc("EnvironmentTexture", image=IMAGES[name] if name else None, enums={"mag filter": filt})
        cases.append((f"envtex-{name}-{filt}", (node, P["Spectrum"], 0)))
    for bump in ("Normal Map (DirectX)", "Normal Map (OpenGL)", "Height Map"):
        for coeff in (0.7, 1.0, 1.5):
            img = IMAGES["rgba32f_refl"] if coeff != 1.0 else IMAGES["rgba8_srgb_refl"]
            node = NodeDesc("Image2DTexture", image=img, floats={"bump coeff": [coeff]},
                            enums={"bump type": bump, "mag filter": "Linear", "wrap u": "Repeat",
                                   "wrap v": "Mirror" if coeff > 1 else "Clamp to Edge"})
            for opt in (range(4) if bump == "Height Map" else range(3)):
                cases.append((f"normal-{bump}-{coeff}-{opt}", (node, P["Normal3D"], opt)))
    return cases


def value_cases():
    cases = []
    geo = NodeDesc("Geometry")
    cases += [("geo-point", (geo, P["Point3D"], 0)), ("geo-normal0", (geo, P["Normal3D"], 0)),
              ("geo-normal1", (geo, P["Normal3D"], 1)), ("geo-normal2", (geo, P["Normal3D"], 2)),
              ("geo-vector0", (geo, P["Vector3D"], 0)), ("geo-vector1", (geo, P["Vector3D"], 1)),
              ("geo-texcoord", (geo, P["TextureCoordinates"], 0))]
    for tt in ("TC0 Direction", "Radial X", "Radial Y", "Radial Z"):
        cases.append((f"tangent-{tt}", (NodeDesc("Tangent", enums={"tangent type": tt}), P["Vector3D"], 0)))
    cases.append(("uvmap", (uv_map(), P["TextureCoordinates"], 0)))
    tex = NodeDesc("Image2DTexture", image=IMAGES["rgba32f_refl"], enums={"mag filter": "Linear"})
    f2 = NodeDesc("Float2", floats={"0": [0.25], "1": [-1.5]})
    sof = NodeDesc("ScaleAndOffsetFloat", floats={"scale": [1.5], "offset": [-0.25]},
                   plugs={"value": (tex, P["float1"], 2), "offset": (f2, P["float1"], 1)})
    f3 = NodeDesc("Float3", floats={"0": [0.5], "1": [2.0], "2": [-3.0]},
                  plugs={"1": (sof, P["float1"], 0)})
    f4 = NodeDesc("Float4", floats={"0": [1.0], "1": [2.0], "2": [3.0], "3": [4.0]},
                  plugs={"0": (tex, P["float1"], 0), "3": (f2, P["float1"], 0)})
    for node, width in ((f2, 2), (f3, 3), (f4, 4)):
        for ptype, w in (("float1", 1), ("float2", 2), ("float3", 3), ("float4", 4)):
            if w > width:
                continue
            for opt in range(width - w + 1):
                cases.append((f"{node.type}-{ptype}-{opt}", (node, P[ptype], opt)))
    cases.append(("scale-offset", (sof, P["float1"], 0)))
    # conversions at inputs
    tc3 = NodeDesc("Image2DTexture", image=IMAGES["rgba16f_srgb_light"],
                   plugs={"texcoord": (f3, P["float3"], 0)})        # Point3D from float3
    cases.append(("conv-p3-from-f3", (tc3, P["float4"], 0)))
    tcn = NodeDesc("Image2DTexture", image=IMAGES["rgba16f_srgb_light"],
                   plugs={"texcoord": (geo, P["Normal3D"], 1)})     # Point3D from Normal3D
    cases.append(("conv-p3-from-n3", (tcn, P["float2"], 1)))
    return cases


def spectrum_cases(spectral):
    cases = []
    for st in ("Reflectance", "Light Source", "NA"):
        for cs in (REC709_SRGB, REC709, XYZ, "xyY"):
            tr = [0.3, 0.35, 0.4] if cs == "xyY" else [0.8, 0.35, 0.1]
            node = NodeDesc("TripletSpectrum", floats={"triplet": tr}, enums={"spectrum type": st, "color space": cs})
            cases.append((f"triplet-{st}-{cs}", (node, P["Spectrum"], 0)))
    geo = NodeDesc("Geometry")
    f1 = NodeDesc("Float2", floats={"0": [0.4], "1": [-0.2]})
    for name, src in (("n3", (geo, P["Normal3D"], 0)), ("p3", (geo, P["Point3D"], 0)),
                      ("f1", (f1, P["float1"], 1)), ("none", None)):
        node = NodeDesc("Float3ToSpectrum", floats={"value": [0.2, -0.6, 0.9]},
                        enums={"spectrum type": "Reflectance", "color space": REC709_SRGB})
        if src is not None:
            node.plugs["value"] = src
        cases.append((f"float3-to-spectrum-{name}", (node, P["Spectrum"], 0)))
    # sampled spectra: per-wavelength values in the spectral build, one triplet in the RGB build
    rng = np.random.default_rng(11)
    vals = list(rng.uniform(0, 1.5, 17))
    cases.append(("regular", (NodeDesc("RegularSampledSpectrum", floats={
        "min wavelength": [380.0], "max wavelength": [780.0], "values": vals}), P["Spectrum"], 0)))
    cases.append(("regular-narrow", (NodeDesc("RegularSampledSpectrum", floats={
        "min wavelength": [500.0], "max wavelength": [600.0], "values": vals[:5]}), P["Spectrum"], 0)))
    lam = list(np.sort(rng.uniform(340, 860, 23)))
    cases.append(("irregular", (NodeDesc("IrregularSampledSpectrum", floats={
        "wavelengths": lam, "values": list(rng.uniform(0, 2, 23))}), P["Spectrum"], 0)))
    cases.append(("irregular-light", (NodeDesc("IrregularSampledSpectrum", floats={
        "wavelengths": [450.25, 520.0, 610.5], "values": [0.5, 2.0, 1.0]},
        enums={"spectrum type": "Light Source"}), P["Spectrum"], 0)))
    return cases


def all_cases(spectral):
    return value_cases() + spectrum_cases(spectral) + texture_cases()


@pytest.fixture(scope="module")
def contexts(torch_gpu):
    ctxs = {s: api.Context(spectral=s) for s in (False, True)}
    yield ctxs
    for c in ctxs.values():
        c.close()


def gpu_eval(torch, ctx, plug_desc, spectral, pts, cache, keep):
    node, ptype, opt = plug_desc
    plug = api.node_plug(ctx, node, ptype, opt, cache, keep)
    dev = torch.device("cuda")
    p = torch.from_numpy(pts).to(dev)
    out = torch.zeros((len(pts), 4), dtype=torch.float32, device=dev)
    ctx.evaluate_node_plug(plug, spectral, p.data_ptr(), len(pts), out.data_ptr())
    torch.cuda.synchronize()
    return out.cpu().numpy()


EDGE_STEP = 2e-6


def nudged_points(row):
    """The point with its texture coordinates and wavelength offset moved by EDGE_STEP."""
    out = []
    for du in (-EDGE_STEP, 0.0, EDGE_STEP):
        for dv in (-EDGE_STEP, 0.0, EDGE_STEP):
            if du or dv:
                r = row.copy()
                r[15] += du
                r[16] += dv
                out.append(r)
    for dw in (-EDGE_STEP, EDGE_STEP):
        r = row.copy()
        r[17] = np.float32(min(max(r[17] + dw, 0.0), np.nextafter(np.float32(1), np.float32(0))))
        out.append(r)
    return out


def check(name, got, want, pts=None, oracle_eval=None):
    scale = max(1.0, float(np.abs(want).max()))

    def close(a, b):
        return np.all(np.abs(a - b) <= 2e-4 * np.maximum(np.abs(a), np.abs(b)) + 2e-5 * scale, axis=-1)

    tight = close(got, want)
    edge = []
    for i in np.flatnonzero(~tight):
        if oracle_eval is not None and any(close(got[i], oracle_eval(r)) for r in nudged_points(pts[i])):
            edge.append(int(i))
            continue
        raise AssertionError((name, int(i), got[i].tolist(), want[i].tolist()))
    if edge:
        print(f"  {name}: {len(edge)} of {len(got)} points next to an edge, matched by the nudged oracle")


@pytest.mark.parametrize("spectral", [False, True])
def test_node_plugs_match_oracle(torch_gpu, contexts, spectral):
    ctx = contexts[spectral]
    pts = random_points(NPTS, 100 + spectral)
    points = [osn.Point(r) for r in pts]
    ev = osn.Evaluator(spectral)
    cache, keep = {}, []
    cases = all_cases(spectral)
    for name, plug in cases:
        print("case", name, flush=True)
        got = gpu_eval(torch_gpu, ctx, plug, spectral, pts, cache, keep)
        want = np.stack([ev.evaluate(plug, p) for p in points])
        check(name, got, want, pts, lambda r, plug=plug: ev.evaluate(plug, osn.Point(r)))
    print(len(cases), "plugs checked")


def test_sampled_spectrum_rgb_build_is_a_constant_triplet(torch_gpu, contexts):
    """The RGB build reduces a sampled spectrum to one triplet (its value against the oracle is
    checked with the other plugs); it must not depend on the point."""
    ctx = contexts[False]
    pts = random_points(64, 7)
    node = NodeDesc("RegularSampledSpectrum", floats={"min wavelength": [400.0], "max wavelength": [700.0],
                                                      "values": [0.2, 0.9, 0.4]})
    got = gpu_eval(torch_gpu, ctx, (node, P["Spectrum"], 0), False, pts, {}, [])
    assert np.all(got == got[0]) and np.all(got[0, :3] > 0) and got[0, 3] == 0


# ---- renders ----
def render(desc, W, H, frames, max_len, spectral=False):
    with api.Context(spectral=spectral, max_path_length=max_len) as ctx:
        bs = api.build_scene(ctx, desc)
        ctx.bind_output(W, H)
        for f in range(frames):
            ctx.render(bs.camera, f == 0, frames)
        return ctx.read_output()[..., :3], ctx.read_rng()


def oracle_render(desc, W, H, frames, max_len):
    ref, _, rng, _ = ob.Scene(desc, spectral=False).render(W, H, frames, max_path_length=max_len,
                                                           rng_state=ob.seed_states(W, H))
    return ref[..., :3], rng


def compare_where_paths_match(img, rng, ref, rng_ref, min_same=0.9, rtol=1e-3):
    same = rng == rng_ref
    print("pixels with identical RNG state:", same.mean())
    assert same.mean() > min_same
    eq = same.reshape(img.shape[:2])
    a, b = img[eq].astype(np.float64), ref[eq].astype(np.float64)
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-3 * max(1e-12, np.abs(b).mean()))
    q = np.quantile(rel, 0.999)
    print("relative difference, 99.9th percentile:", q)
    assert q < rtol


def cornell_with_floor_group(floor_vertices=None):
    """The Cornell box with its floor as a group of its own (material 0, the white walls)."""
    d = scenes.cornell_box()
    room = d.meshes[0]
    wi = scenes.quad_indices(5)
    if floor_vertices is not None:
        room.vertices = room.vertices.copy()
        room.vertices[0:4] = floor_vertices
    room.groups = [(wi[0:6].copy(), 0), (wi[6:18].copy(), 0), (wi[18:24].copy(), 1), (wi[24:30].copy(), 2)]
    return d


def test_two_texel_albedo_matches_split_floor(torch_gpu):
    rgb0, rgb1 = (0.7, 0.2, 0.1), (0.1, 0.3, 0.8)
    # texels hold the rendering RGB that the immediate colours convert to
    c0, c1 = ob.triplet_rgb(0, 1, rgb0), ob.triplet_rgb(0, 1, rgb1)
    texels = np.array([[[*c0, 1.0], [*c1, 1.0]]], np.float32)
    img = ImageDesc(texels.view(np.uint8).reshape(1, 2, 16), 2, 1, "RGBA32Fx4", "NA", REC709)
    tex = NodeDesc("Image2DTexture", image=img,
                   enums={"mag filter": "Nearest", "wrap u": "Clamp to Edge", "wrap v": "Clamp to Edge"})
    textured = cornell_with_floor_group()
    textured.materials.append(scenes.MaterialDesc(scenes.MATTE, plugs={"albedo": (tex, P["Spectrum"], 0)}))
    textured.meshes[0].groups[0] = (textured.meshes[0].groups[0][0], len(textured.materials) - 1)
    # the floor's u runs along z (u = (z + 1.5) / 3): texel 0 for z < 0, texel 1 for z > 0
    split = cornell_with_floor_group()
    split.materials += [scenes.MaterialDesc(scenes.MATTE, {"albedo": (REC709, *rgb0)}),
                        scenes.MaterialDesc(scenes.MATTE, {"albedo": (REC709, *rgb1)})]
    room = split.meshes[0]
    halves = np.concatenate([
        scenes._quad((-1.5, 0, -1.5), (-1.5, 0, 0), (1.5, 0, 0), (1.5, 0, -1.5), (0, 1, 0), (1, 0, 0)),
        scenes._quad((-1.5, 0, 0), (-1.5, 0, 1.5), (1.5, 0, 1.5), (1.5, 0, 0), (0, 1, 0), (1, 0, 0))])
    base = len(room.vertices)
    room.vertices = np.concatenate([room.vertices, halves])
    q = scenes.quad_indices(2) + base
    room.groups = [(q[0:6].copy(), len(split.materials) - 2), (q[6:12].copy(), len(split.materials) - 1)] + \
        room.groups[1:]
    W = H = 64
    img_gpu, rng_gpu = render(textured, W, H, 8, 8)
    ref, rng_ref = oracle_render(split, W, H, 8, 8)
    compare_where_paths_match(img_gpu, rng_gpu, ref, rng_ref)


def test_node_free_scenes_run_the_kernels_without_node_evaluation(torch_gpu):
    """The environment texture is read by the environment emitter itself: the icosphere grid
    sets up no node program, so its frames run the kernel variants without node evaluation."""
    desc = scenes.icosphere_grid(grid=2, env_size=(64, 32))
    with api.Context(spectral=True, max_path_length=4) as ctx:
        bs = api.build_scene(ctx, desc)
        ctx.bind_output(16, 16)
        ctx.render(bs.camera, True, 1)
        assert ctx.accel_stats()["numNodeProgramSteps"] == 0
    textured = cornell_with_floor_group()
    tex = NodeDesc("Image2DTexture", image=flat_map([0.2, 0.4, 0.6, 1.0]))
    textured.materials[0] = scenes.MaterialDesc(scenes.MATTE, plugs={"albedo": (tex, P["Spectrum"], 0)})
    with api.Context(spectral=False, max_path_length=4) as ctx:
        bs = api.build_scene(ctx, textured)
        ctx.bind_output(16, 16)
        ctx.render(bs.camera, True, 1)
        assert ctx.accel_stats()["numNodeProgramSteps"] == 1


def flat_map(rgba):
    t = np.array([[rgba]], np.float32)
    return ImageDesc(t.view(np.uint8).reshape(1, 1, 16), 1, 1, "RGBA32Fx4", "NA", REC709)


def test_flat_normal_map_and_tc0_tangent_leave_the_image_unchanged(torch_gpu):
    W = H = 48
    plain = cornell_with_floor_group()
    img0, rng0 = render(plain, W, H, 4, 8)
    mapped = cornell_with_floor_group()
    # nearest filtering reads the texel exactly (linear weights need not sum to exactly 1)
    nmap = NodeDesc("Image2DTexture", image=flat_map([0.5, 0.5, 1.0, 1.0]), enums={"mag filter": "Nearest"})
    tan = NodeDesc("Tangent", enums={"tangent type": "TC0 Direction"})
    mapped.meshes[0].group_nodes = {0: ((nmap, P["Normal3D"], 0), (tan, P["Vector3D"], 0)),
                                    1: ((nmap, P["Normal3D"], 0), None)}
    img1, rng1 = render(mapped, W, H, 4, 8)
    assert np.array_equal(rng0, rng1)
    np.testing.assert_allclose(img1, img0, rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("bump", ["Normal Map (DirectX)", "Normal Map (OpenGL)"])
@pytest.mark.parametrize("axis", ["x", "y"])
def test_constant_normal_map_matches_tilted_vertex_normals(torch_gpu, bump, axis):
    # texel 0.75 on one axis: local normal (0.5 c, 0, 1) or (0, +-0.5 c, 1) with c = 17 / 16
    texel = [0.75, 0.5, 1.0, 1.0] if axis == "x" else [0.5, 0.75, 1.0, 1.0]
    c = 17 / 16
    if axis == "x":
        local = np.array([0.5 * c, 0.0, 1.0])
    else:
        local = np.array([0.0, (-0.5 if bump == "Normal Map (DirectX)" else 0.5) * c, 1.0])
    local /= np.linalg.norm(local)
    # the floor's shading frame: x = (1, 0, 0), y = z cross x = (0, 0, -1), z = (0, 1, 0)
    world = local[0] * np.array([1.0, 0, 0]) + local[1] * np.array([0, 0, -1.0]) + local[2] * np.array([0, 1.0, 0])
    floor = scenes._quad((-1.5, 0, -1.5), (-1.5, 0, 1.5), (1.5, 0, 1.5), (1.5, 0, -1.5), world, (1, 0, 0))
    floor[:, 3:6] = world.astype(np.float32)
    mapped = cornell_with_floor_group()
    nmap = NodeDesc("Image2DTexture", image=flat_map(texel), enums={"bump type": bump, "mag filter": "Nearest"})
    mapped.meshes[0].group_nodes = {0: ((nmap, P["Normal3D"], 0), None)}
    tilted = cornell_with_floor_group(floor)
    W = H = 64
    img, rng = render(mapped, W, H, 8, 8)
    ref, rng_ref = oracle_render(tilted, W, H, 8, 8)
    compare_where_paths_match(img, rng, ref, rng_ref)


def test_radial_tangent_matches_vertex_tangents(torch_gpu):
    # Radial Y at (x, 0, z) is (z, 0, -x): affine in position, so vertex tangents interpolate to it
    floor = scenes._quad((-1.5, 0, -1.5), (-1.5, 0, 1.5), (1.5, 0, 1.5), (1.5, 0, -1.5), (0, 1, 0), (1, 0, 0))
    floor[:, 6:9] = np.stack([floor[:, 2], np.zeros(4, np.float32), -floor[:, 0]], 1)
    mapped = cornell_with_floor_group()
    tan = NodeDesc("Tangent", enums={"tangent type": "Radial Y"})
    mapped.meshes[0].group_nodes = {0: (None, (tan, P["Vector3D"], 0))}
    tangents = cornell_with_floor_group(floor)
    W = H = 64
    img, rng = render(mapped, W, H, 8, 8)
    ref, rng_ref = oracle_render(tangents, W, H, 8, 8)
    compare_where_paths_match(img, rng, ref, rng_ref)
