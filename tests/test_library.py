"""The C-ABI library without a GPU: it loads, exports every function the headers declare, its
headers compile as C and C++, and the host side of the API (object model, parameters, error
codes of vlr.cpp:23-34) behaves like the reference's.  No kernel is launched here."""
import ctypes as C
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from vlr_amd import api

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include" / "vlr").glob("*.h"))

# vlr.cpp:870-906 and 955-991 (names the reference's factories accept)
MATERIAL_NAMES = ["Matte", "SpecularReflection", "SpecularScattering", "MicrofacetReflection",
                  "MicrofacetScattering", "LambertianScattering", "UE4", "OldStyle", "DiffuseEmitter",
                  "DirectionalEmitter", "PointEmitter", "Multi", "EnvironmentEmitter"]
NODE_NAMES = ["Geometry", "Tangent", "Float2", "Float3", "Float4", "ScaleAndOffsetFloat", "TripletSpectrum",
              "RegularSampledSpectrum", "IrregularSampledSpectrum", "Float3ToSpectrum",
              "ScaleAndOffsetUVTextureMap2D", "Image2DTexture", "EnvironmentTexture"]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = h.read_text()
        names.update(re.findall(r"VLR_API\s+[\w\s\*]+?\b(vlr\w+)\s*\(", text))
    return names


def test_every_declared_function_is_exported():
    names = declared_functions()
    assert len(names) > 100
    L = C.CDLL(str(api.LIB_PATH))
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_every_declared_function():
    names = declared_functions() - {"vlrGetErrorMessage"}
    assert names == set(api.PROTOTYPES), sorted(names ^ set(api.PROTOTYPES))


@pytest.mark.parametrize("lang,compiler", [("c", "gcc"), ("c++", "g++")])
def test_headers_compile(tmp_path, lang, compiler):
    if shutil.which(compiler) is None:
        pytest.skip(f"{compiler} not available")
    src = tmp_path / ("t.c" if lang == "c" else "t.cpp")
    src.write_text('#include "vlr/vlr.h"\nint main(void) { VLRContext c = 0; (void)c; return 0; }\n')
    r = subprocess.run([compiler, "-fsyntax-only", "-x", lang, "-D__HIP_PLATFORM_AMD__", f"-I{ROOT / 'include'}",
                        "-I/opt/rocm/include", str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_cpp_wrapper_compiles(tmp_path):
    """vlrcpp.h (the vlr::Context wrapper) with a caller in the style of the reference's
    HostProgram/cpp_compile_test.cpp: scene, mesh with a material group, nodes, camera, render."""
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = tmp_path / "t.cpp"
    src.write_text(
        '#include "vlr/vlrcpp.h"\n'
        "using namespace vlr;\n"
        "int main() {\n"
        "  ContextRef context = Context::create(nullptr, true);\n"
        "  SceneRef scene = context->createScene();\n"
        '  TriangleMeshSurfaceNodeRef mesh = context->createTriangleMeshSurfaceNode("mesh");\n'
        "  Vertex v[] = {Vertex{Point3D(-1, 0, -1), Normal3D(0, 1, 0), Vector3D(1, 0, 0), TexCoord2D(0, 1)}};\n"
        "  mesh->setVertices(v, 1);\n"
        '  ShaderNodeRef tex = context->createShaderNode("Image2DTexture");\n'
        '  tex->set("min filter", "Nearest");\n'
        '  SurfaceMaterialRef mat = context->createSurfaceMaterial("Matte");\n'
        '  mat->set("albedo", tex->getPlug(VLRShaderNodePlugType_Spectrum, 0));\n'
        "  uint32_t group[] = {0, 0, 0};\n"
        "  mesh->addMaterialGroup(group, 3, mat, ShaderNodePlug(), ShaderNodePlug(), ShaderNodePlug());\n"
        '  InternalNodeRef node = context->createInternalNode("trf", context->createStaticTransform(scale(2.0f)));\n'
        "  node->addChild(mesh);\n"
        "  scene->addChild(node);\n"
        '  CameraRef camera = context->createCamera("Perspective");\n'
        '  camera->set("position", Point3D(0, 1.5f, 6.0f));\n'
        '  camera->set("orientation", qRotateY(3.14159265f));\n'
        "  context->bindOutputBuffer(64, 64, 0);\n"
        "  context->setScene(scene);\n"
        "  uint32_t n;\n"
        "  context->render(nullptr, camera, false, 1, true, 1, &n);\n"
        "  hipArray_t array;\n"
        "  context->getOutputBuffer(&array);\n"
        "  return 0;\n"
        "}\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-D__HIP_PLATFORM_AMD__",
                        f"-I{ROOT / 'include'}", "-I/opt/rocm/include", str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_example_callers_are_built():
    """build() compiles the C and C++ callers of vlr_amd/examples against the library."""
    bin_dir = ROOT / "vlr_amd" / "examples" / "bin"
    for name in ("cornell", "cornell_cpp"):
        exe = bin_dir / name
        assert exe.exists(), f"{exe} missing: run vlr_amd/build.py"
        r = subprocess.run(["ldd", str(exe)], capture_output=True, text=True)
        assert "libvlr_amd.so" in r.stdout and "not found" not in r.stdout.split("libvlr_amd.so")[1].split("\n")[0]


# ---------------------------------------------------------------------------------------------
@pytest.fixture()
def ctx():
    c = api.Context()
    yield c
    c.close()


def code_of(name, *args):
    return api.try_call(name, *args)


def test_error_codes_follow_the_reference(ctx):
    h = api.H()
    assert code_of("vlrSurfaceMaterialCreate", ctx.h, b"NoSuchMaterial", C.byref(h)) == 3
    assert code_of("vlrSurfaceMaterialCreate", None, b"Matte", C.byref(h)) == 3   # null handle
    assert code_of("vlrSurfaceMaterialCreate", ctx.h, b"Matte", None) == 3
    m = ctx.material("Matte")
    # a material handle where a node is expected -> InvalidInstance
    assert code_of("vlrTriangleMeshSurfaceNodeSetVertices", m, None, 0) == 2
    msg = api.lib().vlrGetErrorMessage(2)
    assert b"Instance" in msg or b"instance" in msg


@pytest.mark.parametrize("name", MATERIAL_NAMES)
def test_every_material_type_is_creatable(ctx, name):
    m = ctx.material(name)
    t = api.CP()
    api.call("vlrObjectGetType", m, C.byref(t))
    assert t.value
    n = api.U32()
    api.call("vlrQueryableGetNumParameters", m, C.byref(n))
    for i in range(n.value):
        info = api.H()
        api.call("vlrQueryableGetParameterInfo", m, i, C.byref(info))
        pname = api.CP()
        api.call("vlrParameterInfoGetName", info, C.byref(pname))
        assert pname.value


@pytest.mark.parametrize("name", NODE_NAMES)
def test_every_shader_node_type_is_creatable(ctx, name):
    ctx.shader_node(name)


def test_parameters_round_trip(ctx):
    m = ctx.material("UE4")
    api.set_float(m, "roughness", 0.37)
    f = api.F()
    api.call("vlrQueryableGetFloat", m, b"roughness", C.byref(f))
    assert abs(f.value - 0.37) < 1e-7
    api.set_spectrum(m, "base color", "Rec709(D65)", 0.1, 0.2, 0.3)
    s = api.VLRImmediateSpectrum()
    api.call("vlrQueryableGetImmediateSpectrum", m, b"base color", C.byref(s))
    assert (s.e0, s.e1, s.e2) == pytest.approx((0.1, 0.2, 0.3))
    assert s.colorSpace == b"Rec709(D65)"
    # parameter names are matched without regard to case and spaces (queryable.cpp)
    api.call("vlrQueryableSetFloat", m, b"Roughness", 0.5)
    api.call("vlrQueryableGetFloat", m, b"roughness", C.byref(f))
    assert abs(f.value - 0.5) < 1e-7
    assert code_of("vlrQueryableSetFloat", m, b"no such parameter", 1.0) == 3


def test_plug_types_follow_the_conversion_rules(ctx):
    """A plug connects where its value converts to the socket's type (shared.h:390-470);
    set() refuses the rest with InvalidArgument (vlr.cpp:705-717)."""
    def plug(node, ptype, option=0):
        p = api.VLRShaderNodePlug()
        api.call("vlrShaderNodeGetPlug", node, ptype, option, C.byref(p))
        return p
    geo = ctx.shader_node("Geometry")
    f3 = ctx.shader_node("Float3")
    tex = ctx.shader_node("Image2DTexture")
    sof = ctx.shader_node("ScaleAndOffsetFloat")
    f2s = ctx.shader_node("Float3ToSpectrum")
    mat = ctx.material("Matte")
    ue4 = ctx.material("UE4")
    setp = lambda q, name, p: code_of("vlrQueryableSetShaderNodePlug", q, name, p)
    assert setp(sof, b"value", plug(f3, api.PLUG_FLOAT1, 2)) == 0
    assert setp(sof, b"value", plug(tex, api.PLUG_ALPHA, 3)) == 0
    assert setp(sof, b"value", plug(f3, api.PLUG_FLOAT3)) == 3          # float from float3
    assert setp(tex, b"texcoord", plug(geo, api.PLUG_NORMAL3D, 1)) == 0  # Point3D from Normal3D
    assert setp(tex, b"texcoord", plug(f3, api.PLUG_FLOAT3)) == 0
    assert setp(tex, b"texcoord", plug(f3, api.PLUG_FLOAT2)) == 3
    assert setp(f2s, b"value", plug(f3, api.PLUG_FLOAT1)) == 0          # float3 from float
    assert setp(f2s, b"value", plug(tex, api.PLUG_FLOAT4)) == 3
    assert setp(mat, b"albedo", plug(tex, api.PLUG_SPECTRUM)) == 0
    assert setp(mat, b"albedo", plug(tex, api.PLUG_FLOAT3)) == 3
    assert setp(ue4, b"occlusion/roughness/metallic", plug(tex, api.PLUG_FLOAT3, 1)) == 0
    assert setp(ue4, b"occlusion/roughness/metallic", plug(tex, api.PLUG_FLOAT2)) == 3


def test_scene_graph(ctx):
    mesh = api.H()
    api.call("vlrTriangleMeshSurfaceNodeCreate", ctx.h, b"quad", C.byref(mesh))
    verts = np.zeros((4, 11), np.float32)
    verts[:, 0:3] = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]]
    verts[:, 3:6] = [0, 0, 1]
    verts[:, 6:9] = [1, 0, 0]
    api.call("vlrTriangleMeshSurfaceNodeSetVertices", mesh, verts.ctypes.data, 4)
    idx = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    mat = ctx.material("Matte")
    api.call("vlrTriangleMeshSurfaceNodeAddMaterialGroup", mesh, idx.ctypes.data, 6, mat, api.NULL_PLUG,
             api.NULL_PLUG, api.NULL_PLUG)
    xf = ctx.transform([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 2, 3, 4, 1])
    node = api.H()
    api.call("vlrInternalNodeCreate", ctx.h, b"inst", xf, C.byref(node))
    api.call("vlrParentNodeAddChild", node, mesh)
    n = api.U32()
    api.call("vlrParentNodeGetNumChildren", node, C.byref(n))
    assert n.value == 1
    child = api.H()
    api.call("vlrParentNodeGetChildAt", node, 0, C.byref(child))
    assert child.value == mesh.value
    name = api.CP()
    api.call("vlrNodeGetName", node, C.byref(name))
    assert name.value == b"inst"
    m, inv = (C.c_float * 16)(), (C.c_float * 16)()
    api.call("vlrStaticTransformGetArrays", xf, m, inv)
    M = np.array(m[:]).reshape(4, 4).T
    I = np.array(inv[:]).reshape(4, 4).T
    np.testing.assert_allclose(M @ I, np.eye(4), atol=1e-6)
    api.call("vlrParentNodeRemoveChild", node, mesh)
    api.call("vlrParentNodeGetNumChildren", node, C.byref(n))
    assert n.value == 0


def test_camera_parameters(ctx):
    cam = api.H()
    api.call("vlrCameraCreate", ctx.h, b"Perspective", C.byref(cam))
    api.call("vlrQueryableSetFloat", cam, b"fovy", 0.7)
    f = api.F()
    api.call("vlrQueryableGetFloat", cam, b"fovy", C.byref(f))
    assert abs(f.value - 0.7) < 1e-7
    assert code_of("vlrCameraCreate", ctx.h, b"Orthographic", C.byref(cam)) == 3


def test_knobs(ctx):
    api.call("vlrContextSetMaxPathLength", ctx.h, 7)
    n = api.U32()
    api.call("vlrContextGetMaxPathLength", ctx.h, C.byref(n))
    assert n.value == 7
    api.call("vlrContextSetSpectralRendering", ctx.h, True)
    b = C.c_bool()
    api.call("vlrContextGetSpectralRendering", ctx.h, C.byref(b))
    assert b.value
    # light paths of VLRRenderer_BPT: 500,000 by default (context.h:303), at most 2^20
    assert ctx.num_light_paths() == 500000
    ctx.set_num_light_paths(921600)
    assert ctx.num_light_paths() == 921600
    assert code_of("vlrContextSetNumLightPaths", ctx.h, 0) == 3
    assert code_of("vlrContextSetNumLightPaths", ctx.h, (1 << 20) + 1) == 3
    assert ctx.num_light_paths() == 921600


def test_block_compressed_formats(ctx):
    """Every block-compressed format of the reference (BC1-BC7, signed BC4/BC5/BC6H) is accepted;
    a format name that is not one is refused with InvalidArgument."""
    data = np.zeros(16, np.uint8)
    ptrs = (C.c_void_p * 1)(data.ctypes.data)
    sizes = (C.c_size_t * 1)(data.nbytes)
    for fmt, ok in (("BC1", True), ("BC3", True), ("BC4_Signed", True), ("BC5", True), ("BC6H", True),
                    ("BC6H_Signed", True), ("BC7", True), ("RGBA8x4", False), ("BC8", False)):
        h = api.H()
        rc = code_of("vlrBlockCompressedImage2DCreate", ctx.h, C.cast(ptrs, C.c_void_p), C.cast(sizes, C.c_void_p), 1,
                     4, 4, fmt.encode(), b"Reflectance", b"Rec709(D65)", C.byref(h))
        assert (rc == 0) == ok, (fmt, rc)


def test_material_group_plugs(ctx):
    """Normal, tangent and alpha plugs are taken; an alpha plug that does not convert to a float
    is ignored, as scene.cpp:225-246 ignores it (the call still succeeds)."""
    mesh = api.H()
    api.call("vlrTriangleMeshSurfaceNodeCreate", ctx.h, b"quad", C.byref(mesh))
    verts = np.zeros((4, 11), np.float32)
    api.call("vlrTriangleMeshSurfaceNodeSetVertices", mesh, verts.ctypes.data, 4)
    idx = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    mat = ctx.material("Matte")
    f2 = ctx.shader_node("Float2")
    tex = ctx.shader_node("Image2DTexture")

    def plug(node, ptype, option=0):
        p = api.VLRShaderNodePlug()
        api.call("vlrShaderNodeGetPlug", node, ptype, option, C.byref(p))
        return p
    assert code_of("vlrTriangleMeshSurfaceNodeAddMaterialGroup", mesh, idx.ctypes.data, 6, mat,
                   plug(tex, api.PLUG_NORMAL3D), api.NULL_PLUG, plug(f2, api.PLUG_FLOAT1)) == 0
    assert code_of("vlrTriangleMeshSurfaceNodeAddMaterialGroup", mesh, idx.ctypes.data, 6, mat,
                   api.NULL_PLUG, api.NULL_PLUG, plug(tex, api.PLUG_SPECTRUM)) == 0
