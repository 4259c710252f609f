"""ctypes binding of libvlr_amd.so (the C API of include/vlr/vlr.h and vlr_amd.h).

The library is the product: this module only declares prototypes, checks VLRResult codes
and offers :func:`build_scene`, which creates a vlr_amd.scenes.SceneDesc through the C API
(TriangleMeshSurfaceNode / InternalNode / Scene / SurfaceMaterial / Camera), the same calls a
C++ caller of vlrcpp.h would make.  There is no fallback: when the library is missing,
:func:`lib` raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Dict, List

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libvlr_amd.so"

VLRResult_NoError = 0
RESULT_NAMES = ["NoError", "InvalidContext", "InvalidInstance", "InvalidArgument", "IncompatibleNodeType",
                "InternalError"]

VLRRenderer_PathTracing = 0
VLRRenderer_LightTracing = 1
VLRRenderer_BPT = 2
VLRRenderer_DebugRendering = 3

PLUG_FLOAT1, PLUG_FLOAT2, PLUG_FLOAT3, PLUG_FLOAT4, PLUG_POINT3D, PLUG_VECTOR3D, PLUG_NORMAL3D = 0, 1, 2, 3, 4, 5, 6
PLUG_SPECTRUM = 7
PLUG_ALPHA, PLUG_TEXCOORD = 8, 9


class VLRError(RuntimeError):
    def __init__(self, fn, code):
        name = RESULT_NAMES[code] if 0 <= code < len(RESULT_NAMES) else str(code)
        super().__init__(f"{fn} -> VLRResult_{name}")
        self.code = code


class VLRPoint3D(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class VLRQuaternion(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class VLRImmediateSpectrum(C.Structure):
    _fields_ = [("colorSpace", C.c_char_p), ("e0", C.c_float), ("e1", C.c_float), ("e2", C.c_float)]


# VLRRenderer (public_types.h:37-42)
RENDERER_PATH_TRACING, RENDERER_LIGHT_TRACING, RENDERER_BPT, RENDERER_DEBUG = 0, 1, 2, 3


class VLRShaderNodePlug(C.Structure):
    _fields_ = [("nodeRef", C.c_size_t), ("info", C.c_uint32)]


class VLRAccelStats(C.Structure):
    _fields_ = [("numBlas", C.c_uint32), ("numTlasInstances", C.c_uint32), ("numTriangles", C.c_uint64),
                ("numBlasNodes", C.c_uint32), ("numTlasNodes", C.c_uint32), ("msBlas", C.c_float),
                ("msTlas", C.c_float), ("blasReused", C.c_uint32), ("numNodeProgramSteps", C.c_uint32),
                ("tlasDepth", C.c_uint32), ("blasDepth", C.c_uint32), ("tlasRefit", C.c_uint32)]


# 64-byte quantised 4-wide BVH node: child k's box on axis a is
# origin[a] + byte_k(qlo[a] / qhi[a]) * 2^(byte_a(scaleExp) - 127)
BVH4_NODE = np.dtype([("origin", "<f4", 3), ("scaleExp", "<u4"), ("qlo", "<u4", 3), ("qhi", "<u4", 3),
                      ("child", "<i4", 4), ("pad", "<u4", 2)])
TRI_RECORD = np.dtype([("v0", "<f4", 3), ("geomInst", "<u4"), ("v1", "<f4", 3), ("prim", "<u4"),
                       ("v2", "<f4", 3), ("pad", "<u4")])
ACCEL_TLAS, ACCEL_BLAS, ACCEL_TRIS, ACCEL_INSTANCES = 0, 1, 2, 3
ACCEL_BUILDER_DEVICE, ACCEL_BUILDER_HOST = 0, 1


def decode_bvh4(nodes: np.ndarray):
    """Child boxes of quantised nodes: (lo, hi), each (n, 4, 3) float32, computed as the traversal
    does (origin + float(q) * step, an exact product and one rounding)."""
    e = ((nodes["scaleExp"][:, None] >> (8 * np.arange(3, dtype=np.uint32))) & 0xFF).astype(np.int32) - 127
    step = np.ldexp(np.float32(1.0), e).astype(np.float32)                       # (n, 3)
    shift = (8 * np.arange(4, dtype=np.uint32))[None, :, None]
    qlo = ((nodes["qlo"][:, None, :] >> shift) & 0xFF).astype(np.float32)     # (n, 4, 3)
    qhi = ((nodes["qhi"][:, None, :] >> shift) & 0xFF).astype(np.float32)
    org = nodes["origin"][:, None, :].astype(np.float32)
    lo = (org + (qlo * step[:, None, :]).astype(np.float32)).astype(np.float32)
    hi = (org + (qhi * step[:, None, :]).astype(np.float32)).astype(np.float32)
    return lo, hi


class VLRRenderStats(C.Structure):
    _fields_ = [("numFrames", C.c_uint64), ("paths", C.c_uint64), ("closestRays", C.c_uint64),
                ("shadowRays", C.c_uint64), ("droppedNonFinite", C.c_uint64), ("msGenerate", C.c_double),
                ("msIntersect", C.c_double), ("msShade", C.c_double), ("msShadow", C.c_double),
                ("msAccumulate", C.c_double), ("msConvert", C.c_double), ("launchesIntersect", C.c_uint64),
                ("launchesShadow", C.c_uint64), ("launchesShade", C.c_uint64), ("nodesClosest", C.c_uint64),
                ("trianglesClosest", C.c_uint64), ("instancesClosest", C.c_uint64), ("nodesShadow", C.c_uint64),
                ("trianglesShadow", C.c_uint64), ("instancesShadow", C.c_uint64), ("msTail", C.c_double),
                ("launchesTail", C.c_uint64), ("closestRaysWavefront", C.c_uint64),
                ("shadowRaysWavefront", C.c_uint64), ("msLightPaths", C.c_double), ("msEyePaths", C.c_double),
                ("msCacheQuery", C.c_double), ("msCacheTrain", C.c_double)]


class VLRBatchJob(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("camera", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32),
                ("numSamples", C.c_uint32), ("maxPathLength", C.c_uint32)]


class VLRRadianceCacheSettings(C.Structure):
    _fields_ = [("learningRate", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("epsilon", C.c_float),
                ("trainingPaths", C.c_uint32), ("terminationThreshold", C.c_float), ("warmupSteps", C.c_uint32),
                ("seed", C.c_uint32), ("fullRadiance", C.c_uint32)]


RAY_DTYPE = np.dtype([("org", np.float32, 3), ("tmin", np.float32), ("dir", np.float32, 3), ("tmax", np.float32)])
HIT_DTYPE = np.dtype([("inst", np.int32), ("geomInst", np.int32), ("prim", np.int32), ("t", np.float32),
                      ("b1", np.float32), ("b2", np.float32)])
BSDF_QUERY_DTYPE = np.dtype([("dirLocal", np.float32, 3), ("geomNormalLocal", np.float32, 3),
                             ("wlOffset", np.float32), ("wlSelect", np.float32), ("uComponent", np.float32),
                             ("uDir", np.float32, 2), ("evalDir", np.float32, 3), ("singleIsSelected", np.int32),
                             ("transportMode", np.int32)])
BSDF_RESULT_DTYPE = np.dtype([("sampledDir", np.float32, 3), ("sampledPDF", np.float32),
                              ("sampledType", np.uint32), ("sampledValue", np.float32, 4),
                              ("evalValue", np.float32, 4), ("evalPDF", np.float32), ("revValue", np.float32, 4),
                              ("revPDF", np.float32), ("sampledRevValue", np.float32, 4),
                              ("sampledRevPDF", np.float32)])

H = C.c_void_p
U32 = C.c_uint32
F = C.c_float
P = C.c_void_p
CP = C.c_char_p
B = C.c_bool

# name -> argtypes (all return VLRResult)
PROTOTYPES: Dict[str, list] = {
    "vlrCreateContext": [H, B, U32, C.POINTER(H)],
    "vlrDestroyContext": [H],
    "vlrContextGetCUcontext": [H, C.POINTER(H)],
    "vlrContextBindOutputBuffer": [H, U32, U32, U32],
    "vlrContextGetOutputBufferSize": [H, C.POINTER(U32), C.POINTER(U32)],
    "vlrContextGetOutputBuffer": [H, C.POINTER(H)],
    "vlrContextReadOutputBuffer": [H, P],
    "vlrContextSetScene": [H, H],
    "vlrContextSetRenderer": [H, C.c_int],
    "vlrContextSetDebugRenderingAttribute": [H, C.c_int],
    "vlrContextSetProbePixel": [H, C.c_int32, C.c_int32],
    "vlrContextRender": [H, H, H, B, U32, B, U32, C.POINTER(U32)],
    "vlrObjectGetType": [H, C.POINTER(CP)],
    "vlrParameterInfoGetName": [H, C.POINTER(CP)],
    "vlrParameterInfoGetSocketForm": [H, C.POINTER(C.c_int)],
    "vlrParameterInfoGetType": [H, C.POINTER(CP)],
    "vlrParameterInfoGetTupleSize": [H, C.POINTER(U32)],
    "vlrGetNumEnumMembers": [CP, C.POINTER(U32)],
    "vlrGetEnumMember": [CP, U32, C.POINTER(CP)],
    "vlrQueryableGetNumParameters": [H, C.POINTER(U32)],
    "vlrQueryableGetParameterInfo": [H, U32, C.POINTER(H)],
    "vlrQueryableGetEnumValue": [H, CP, C.POINTER(CP)],
    "vlrQueryableGetPoint3D": [H, CP, C.POINTER(VLRPoint3D)],
    "vlrQueryableGetVector3D": [H, CP, C.POINTER(VLRPoint3D)],
    "vlrQueryableGetNormal3D": [H, CP, C.POINTER(VLRPoint3D)],
    "vlrQueryableGetQuaternion": [H, CP, C.POINTER(VLRQuaternion)],
    "vlrQueryableGetFloat": [H, CP, C.POINTER(F)],
    "vlrQueryableGetFloatTuple": [H, CP, P, U32],
    "vlrQueryableGetFloatArray": [H, CP, C.POINTER(P), C.POINTER(U32)],
    "vlrQueryableGetImage2D": [H, CP, C.POINTER(H)],
    "vlrQueryableGetImmediateSpectrum": [H, CP, C.POINTER(VLRImmediateSpectrum)],
    "vlrQueryableGetSurfaceMaterial": [H, CP, C.POINTER(H)],
    "vlrQueryableGetShaderNodePlug": [H, CP, C.POINTER(VLRShaderNodePlug)],
    "vlrQueryableSetEnumValue": [H, CP, CP],
    "vlrQueryableSetPoint3D": [H, CP, C.POINTER(VLRPoint3D)],
    "vlrQueryableSetVector3D": [H, CP, C.POINTER(VLRPoint3D)],
    "vlrQueryableSetNormal3D": [H, CP, C.POINTER(VLRPoint3D)],
    "vlrQueryableSetQuaternion": [H, CP, C.POINTER(VLRQuaternion)],
    "vlrQueryableSetFloat": [H, CP, F],
    "vlrQueryableSetFloatTuple": [H, CP, P, U32],
    "vlrQueryableSetImage2D": [H, CP, H],
    "vlrQueryableSetImmediateSpectrum": [H, CP, C.POINTER(VLRImmediateSpectrum)],
    "vlrQueryableSetSurfaceMaterial": [H, CP, H],
    "vlrQueryableSetShaderNodePlug": [H, CP, VLRShaderNodePlug],
    "vlrImage2DGetWidth": [H, C.POINTER(U32)],
    "vlrImage2DGetHeight": [H, C.POINTER(U32)],
    "vlrImage2DGetStride": [H, C.POINTER(U32)],
    "vlrImage2DGetOriginalDataFormat": [H, C.POINTER(CP)],
    "vlrImage2DOriginalHasAlpha": [H, C.POINTER(B)],
    "vlrLinearImage2DCreate": [H, P, U32, U32, CP, CP, CP, C.POINTER(H)],
    "vlrLinearImage2DDestroy": [H, H],
    "vlrBlockCompressedImage2DCreate": [H, P, P, U32, U32, U32, CP, CP, CP, C.POINTER(H)],
    "vlrBlockCompressedImage2DDestroy": [H, H],
    "vlrShaderNodeCreate": [H, CP, C.POINTER(H)],
    "vlrShaderNodeDestroy": [H, H],
    "vlrShaderNodeGetPlug": [H, C.c_int, U32, C.POINTER(VLRShaderNodePlug)],
    "vlrSurfaceMaterialCreate": [H, CP, C.POINTER(H)],
    "vlrSurfaceMaterialDestroy": [H, H],
    "vlrStaticTransformCreate": [H, P, C.POINTER(H)],
    "vlrStaticTransformDestroy": [H, H],
    "vlrStaticTransformGetArrays": [H, P, P],
    "vlrNodeSetName": [H, CP],
    "vlrNodeGetName": [H, C.POINTER(CP)],
    "vlrTriangleMeshSurfaceNodeCreate": [H, CP, C.POINTER(H)],
    "vlrTriangleMeshSurfaceNodeDestroy": [H, H],
    "vlrTriangleMeshSurfaceNodeSetVertices": [H, P, U32],
    "vlrTriangleMeshSurfaceNodeAddMaterialGroup": [H, P, U32, H, VLRShaderNodePlug, VLRShaderNodePlug,
                                                   VLRShaderNodePlug],
    "vlrPointSurfaceNodeCreate": [H, CP, C.POINTER(H)],
    "vlrPointSurfaceNodeDestroy": [H, H],
    "vlrPointSurfaceNodeSetVertices": [H, P, U32],
    "vlrPointSurfaceNodeAddMaterialGroup": [H, P, U32, H],
    "vlrParentNodeSetTransform": [H, H],
    "vlrParentNodeGetTransform": [H, C.POINTER(H)],
    "vlrParentNodeAddChild": [H, H],
    "vlrParentNodeRemoveChild": [H, H],
    "vlrParentNodeGetNumChildren": [H, C.POINTER(U32)],
    "vlrParentNodeGetChildren": [H, P],
    "vlrParentNodeGetChildAt": [H, U32, C.POINTER(H)],
    "vlrInternalNodeCreate": [H, CP, H, C.POINTER(H)],
    "vlrInternalNodeDestroy": [H, H],
    "vlrSceneCreate": [H, H, C.POINTER(H)],
    "vlrSceneDestroy": [H, H],
    "vlrSceneSetEnvironment": [H, H],
    "vlrSceneSetEnvironmentRotation": [H, F],
    "vlrCameraCreate": [H, CP, C.POINTER(H)],
    "vlrCameraDestroy": [H, H],
    # extensions
    "vlrContextSetMaxPathLength": [H, U32],
    "vlrContextGetMaxPathLength": [H, C.POINTER(U32)],
    "vlrContextSetSpectralRendering": [H, B],
    "vlrContextGetSpectralRendering": [H, C.POINTER(B)],
    "vlrContextGetRenderStats": [H, C.POINTER(VLRRenderStats)],
    "vlrContextResetRenderStats": [H],
    "vlrContextSetKernelTiming": [H, B],
    "vlrContextSetDenoiserAuxBuffers": [H, B],
    "vlrContextReadDenoiserAuxBuffers": [H, P, P],
    "vlrContextSetTraversalStatistics": [H, B],
    "vlrContextGetQueueCounts": [H, P, P, U32],
    "vlrContextSetWavefrontDepth": [H, U32],
    "vlrContextGetWavefrontDepth": [H, C.POINTER(U32)],
    "vlrContextSetScheduleAutotuning": [H, B],
    "vlrContextGetSchedule": [H, C.POINTER(U32), C.POINTER(U32), C.POINTER(B)],
    "vlrContextReadRNGStates": [H, P],
    "vlrContextWriteRNGStates": [H, P],
    "vlrContextReadLightRNGStates": [H, P],
    "vlrContextWriteLightRNGStates": [H, P],
    "vlrContextSetNumLightPaths": [H, U32],
    "vlrContextSetSplitFrames": [H, B],
    "vlrContextSetShadeSorting": [H, B],
    "vlrContextGetShadeSorting": [H, C.POINTER(B)],
    "vlrContextRenderBatch": [H, H, P, U32, P],
    "vlrContextReadBatchOutput": [H, U32, P],
    "vlrContextSetBatchGraphs": [H, B],
    "vlrContextGetBatchSchedule": [H, C.POINTER(U32), C.POINTER(U32)],
    "vlrContextSetRadianceCache": [H, B],
    "vlrContextGetRadianceCache": [H, C.POINTER(B)],
    "vlrContextSetRadianceCacheSettings": [H, C.POINTER(VLRRadianceCacheSettings)],
    "vlrContextGetRadianceCacheSettings": [H, C.POINTER(VLRRadianceCacheSettings)],
    "vlrContextResetRadianceCache": [H],
    "vlrRadianceCacheNumParameters": [C.POINTER(U32)],
    "vlrContextReadRadianceCacheWeights": [H, P],
    "vlrContextWriteRadianceCacheWeights": [H, P],
    "vlrContextReadRadianceCacheGradient": [H, P],
    "vlrContextGetRadianceCacheLoss": [H, C.POINTER(C.c_float), C.POINTER(U32)],
    "vlrContextRadianceCacheInfer": [H, H, P, U32, P],
    "vlrContextRadianceCacheTrain": [H, H, P, P, U32, U32],
    "vlrContextGetNumLightPaths": [H, C.POINTER(U32)],
    "vlrContextGetAccumulationChannels": [H, C.POINTER(U32)],
    "vlrContextReadAccumulation": [H, P, C.c_size_t],
    "vlrContextWriteAccumulation": [H, P, C.c_size_t, U32],
    "vlrContextGetBptWavelengthFrames": [H, C.POINTER(C.c_uint64)],
    "vlrContextSetBptWavelengthFrames": [H, C.c_uint64],
    "vlrContextTraceRays": [H, H, H, P, U32, P, B],
    "vlrContextEvaluateBSDFs": [H, H, H, B, P, U32, P],
    "vlrContextEvaluateShaderNodePlug": [H, H, VLRShaderNodePlug, B, P, U32, P],
    "vlrContextSetAccelBuilder": [H, U32],
    "vlrContextGetAccelBuilder": [H, C.POINTER(U32)],
    "vlrContextGetAccelStats": [H, C.POINTER(VLRAccelStats)],
    "vlrContextSetTlasRefit": [H, B],
    "vlrContextReadAccel": [H, H, U32, P, C.c_size_t, C.POINTER(C.c_size_t)],
    "vlrBvh4NodeCapacity": [U32, C.POINTER(U32)],
    "vlrBuildBvh4Host": [P, U32, U32, P, P, U32, P, C.POINTER(U32)],
    "vlrContextBuildBvh4": [H, H, P, U32, U32, P, P, U32, P, C.POINTER(U32)],
    "vlrBuildBvh4NoAlloc": [H, P, U32, U32, P, P, U32, P, C.POINTER(U32)],
}

_lib = None


def lib():
    """Load libvlr_amd.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m vlr_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        for name, args in PROTOTYPES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.vlrGetErrorMessage.argtypes = [C.c_int]
        L.vlrGetErrorMessage.restype = C.c_char_p
        _lib = L
    return _lib


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != VLRResult_NoError:
        raise VLRError(name, rc)
    return rc


def try_call(name, *args) -> int:
    return getattr(lib(), name)(*args)


def _b(s):
    return s.encode() if isinstance(s, str) else s


NULL_PLUG = VLRShaderNodePlug(0, 0)


class Context:
    """Owns a VLRContext (vlrcpp.h:684-749 equivalent) and the objects created through it."""

    def __init__(self, spectral: bool = False, max_path_length: int = 25):
        h = H()
        call("vlrCreateContext", None, False, 8, C.byref(h))
        self.h = h
        self.spectral = spectral
        if spectral:
            call("vlrContextSetSpectralRendering", self.h, True)
        call("vlrContextSetMaxPathLength", self.h, max_path_length)
        self.width = self.height = 0
        self.num_accum = U32(0)
        self._keep: List = []

    def close(self):
        if self.h:
            call("vlrDestroyContext", self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- object creation ----
    def material(self, type_name: str):
        m = H()
        call("vlrSurfaceMaterialCreate", self.h, _b(type_name), C.byref(m))
        return m

    def shader_node(self, type_name: str):
        n = H()
        call("vlrShaderNodeCreate", self.h, _b(type_name), C.byref(n))
        return n

    def transform(self, mat16):
        t = H()
        arr = (C.c_float * 16)(*[float(x) for x in mat16])
        call("vlrStaticTransformCreate", self.h, arr, C.byref(t))
        return t

    def bind_output(self, width: int, height: int):
        call("vlrContextBindOutputBuffer", self.h, width, height, 0)
        self.width, self.height = width, height

    def set_scene(self, scene):
        call("vlrContextSetScene", self.h, scene)

    def render(self, camera, first_frame: bool, limit: int, stream=None, shrink: int = 1) -> int:
        call("vlrContextRender", self.h, stream, camera, False, shrink, first_frame, limit, C.byref(self.num_accum))
        return self.num_accum.value

    def read_output(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 4), np.float32)
        call("vlrContextReadOutputBuffer", self.h, out.ctypes.data)
        return out

    def read_rng(self) -> np.ndarray:
        out = np.zeros(self.width * self.height, np.uint64)
        call("vlrContextReadRNGStates", self.h, out.ctypes.data)
        return out

    def write_rng(self, states: np.ndarray):
        s = np.ascontiguousarray(states, np.uint64)
        call("vlrContextWriteRNGStates", self.h, s.ctypes.data)

    # ---- LVC-BPT ----
    def set_renderer(self, renderer: int):
        call("vlrContextSetRenderer", self.h, renderer)

    def set_num_light_paths(self, n: int):
        call("vlrContextSetNumLightPaths", self.h, n)

    def num_light_paths(self) -> int:
        n = U32()
        call("vlrContextGetNumLightPaths", self.h, C.byref(n))
        return n.value

    def read_light_rng(self) -> np.ndarray:
        out = np.zeros(self.num_light_paths(), np.uint64)
        call("vlrContextReadLightRNGStates", self.h, out.ctypes.data)
        return out

    def write_light_rng(self, states: np.ndarray):
        s = np.ascontiguousarray(states, np.uint64)
        assert s.size == self.num_light_paths()
        call("vlrContextWriteLightRNGStates", self.h, s.ctypes.data)

    # ---- batches ----
    def render_batch(self, jobs, stream=None):
        """jobs: list of (scene handle, camera handle, width, height, samples, max path length).
        Returns the per-job completion times in ms."""
        arr = (VLRBatchJob * len(jobs))()
        for k, (scene, cam, w, h, spp, depth) in enumerate(jobs):
            arr[k] = VLRBatchJob(C.cast(scene, C.c_void_p).value if not isinstance(scene, int) else scene,
                                 C.cast(cam, C.c_void_p).value if not isinstance(cam, int) else cam, w, h, spp, depth)
        ms = np.zeros(len(jobs), np.float32)
        call("vlrContextRenderBatch", self.h, stream, C.cast(arr, C.c_void_p), len(jobs), ms.ctypes.data)
        self._batch_sizes = [(w, h) for (_, _, w, h, _, _) in jobs]
        return ms

    def read_batch_output(self, job: int) -> np.ndarray:
        w, h = self._batch_sizes[job]
        out = np.zeros((h, w, 4), np.float32)
        call("vlrContextReadBatchOutput", self.h, job, out.ctypes.data)
        return out

    def set_batch_graphs(self, on: bool):
        call("vlrContextSetBatchGraphs", self.h, on)

    def batch_schedule(self):
        wd, tb = U32(), U32()
        call("vlrContextGetBatchSchedule", self.h, C.byref(wd), C.byref(tb))
        return wd.value, tb.value

    # ---- radiance cache ----
    def set_radiance_cache(self, on: bool):
        call("vlrContextSetRadianceCache", self.h, on)

    def cache_settings(self) -> dict:
        st = VLRRadianceCacheSettings()
        call("vlrContextGetRadianceCacheSettings", self.h, C.byref(st))
        return {f: getattr(st, f) for f, _ in VLRRadianceCacheSettings._fields_}

    def set_cache_settings(self, **kw):
        cur = self.cache_settings()
        cur.update(kw)
        st = VLRRadianceCacheSettings(**cur)
        call("vlrContextSetRadianceCacheSettings", self.h, C.byref(st))

    def reset_radiance_cache(self):
        call("vlrContextResetRadianceCache", self.h)

    def read_cache_weights(self) -> np.ndarray:
        out = np.zeros(cache_num_parameters(), np.float32)
        call("vlrContextReadRadianceCacheWeights", self.h, out.ctypes.data)
        return out

    def write_cache_weights(self, params: np.ndarray):
        a = np.ascontiguousarray(params, np.float32)
        assert a.size == cache_num_parameters()
        call("vlrContextWriteRadianceCacheWeights", self.h, a.ctypes.data)

    def read_cache_gradient(self) -> np.ndarray:
        out = np.zeros(cache_num_parameters(), np.float32)
        call("vlrContextReadRadianceCacheGradient", self.h, out.ctypes.data)
        return out

    def cache_loss(self):
        loss, steps = C.c_float(), U32()
        call("vlrContextGetRadianceCacheLoss", self.h, C.byref(loss), C.byref(steps))
        return loss.value, steps.value

    def cache_infer(self, records_ptr: int, n: int, out_ptr: int, stream=None):
        call("vlrContextRadianceCacheInfer", self.h, stream, C.c_void_p(records_ptr), n, C.c_void_p(out_ptr))

    def cache_train(self, records_ptr: int, targets_ptr: int, n: int, num_out: int, stream=None):
        call("vlrContextRadianceCacheTrain", self.h, stream, C.c_void_p(records_ptr), C.c_void_p(targets_ptr), n,
             num_out)

    def accumulation_channels(self) -> int:
        ch = U32()
        call("vlrContextGetAccumulationChannels", self.h, C.byref(ch))
        return ch.value

    def read_accumulation(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, self.accumulation_channels(), 2), np.float32)
        call("vlrContextReadAccumulation", self.h, out.ctypes.data, out.size)
        return out

    def write_accumulation(self, acc: np.ndarray, frames: int):
        a = np.ascontiguousarray(acc, np.float32)
        expect = (self.height, self.width, self.accumulation_channels(), 2)
        if a.shape != expect:
            raise ValueError(f"accumulation of shape {a.shape}, the context expects {expect}")
        call("vlrContextWriteAccumulation", self.h, a.ctypes.data, a.size, frames)

    def bpt_wavelength_frames(self) -> int:
        v = C.c_uint64()
        call("vlrContextGetBptWavelengthFrames", self.h, C.byref(v))
        return v.value

    def set_bpt_wavelength_frames(self, frames: int):
        call("vlrContextSetBptWavelengthFrames", self.h, frames)

    def stats(self) -> dict:
        s = VLRRenderStats()
        call("vlrContextGetRenderStats", self.h, C.byref(s))
        return {f: getattr(s, f) for f, _ in VLRRenderStats._fields_}

    def reset_stats(self):
        call("vlrContextResetRenderStats", self.h)

    def set_denoiser_aux_buffers(self, on: bool):
        call("vlrContextSetDenoiserAuxBuffers", self.h, on)

    def read_denoiser_aux_buffers(self):
        """(albedo, normal): H x W x 4 float32 each, converted as the reference's copyBuffers."""
        alb = np.zeros((self.height, self.width, 4), np.float32)
        nrm = np.zeros((self.height, self.width, 4), np.float32)
        call("vlrContextReadDenoiserAuxBuffers", self.h, alb.ctypes.data, nrm.ctypes.data)
        return alb, nrm

    def set_kernel_timing(self, on: bool):
        call("vlrContextSetKernelTiming", self.h, on)

    def queue_counts(self, max_depth: int = 64):
        c = np.zeros(max_depth, np.uint32)
        s = np.zeros(max_depth, np.uint32)
        call("vlrContextGetQueueCounts", self.h, c.ctypes.data, s.ctypes.data, max_depth)
        return c, s

    def set_wavefront_depth(self, depth: int):
        call("vlrContextSetWavefrontDepth", self.h, depth)

    def set_split_frames(self, on: bool):
        call("vlrContextSetSplitFrames", self.h, on)

    def set_shade_sorting(self, on: bool):
        call("vlrContextSetShadeSorting", self.h, on)

    def shade_sorting(self) -> bool:
        b = B()
        call("vlrContextGetShadeSorting", self.h, C.byref(b))
        return bool(b.value)

    def set_autotune(self, on: bool):
        call("vlrContextSetScheduleAutotuning", self.h, on)

    def schedule(self):
        wd, tb, tuned = U32(), U32(), B()
        call("vlrContextGetSchedule", self.h, C.byref(wd), C.byref(tb), C.byref(tuned))
        return wd.value, tb.value, bool(tuned.value)

    def set_traversal_stats(self, on: bool):
        call("vlrContextSetTraversalStatistics", self.h, on)

    def trace_rays(self, scene, rays_dev_ptr: int, n: int, hits_dev_ptr: int, any_hit: bool, stream=None):
        call("vlrContextTraceRays", self.h, stream, scene, C.c_void_p(rays_dev_ptr), n, C.c_void_p(hits_dev_ptr),
             any_hit)

    def evaluate_bsdfs(self, material, spectral: bool, q_ptr: int, n: int, r_ptr: int, stream=None):
        call("vlrContextEvaluateBSDFs", self.h, stream, material, spectral, C.c_void_p(q_ptr), n, C.c_void_p(r_ptr))

    def evaluate_node_plug(self, plug, spectral: bool, points_ptr: int, n: int, values_ptr: int, stream=None):
        """One shader-node plug at n device-resident surface points (20 floats each)."""
        call("vlrContextEvaluateShaderNodePlug", self.h, stream, plug, spectral, C.c_void_p(points_ptr), n,
             C.c_void_p(values_ptr))

    def set_accel_builder(self, builder: int):
        call("vlrContextSetAccelBuilder", self.h, builder)

    def accel_builder(self) -> int:
        b = U32()
        call("vlrContextGetAccelBuilder", self.h, C.byref(b))
        return b.value

    def set_tlas_refit(self, on: bool):
        call("vlrContextSetTlasRefit", self.h, on)

    def accel_stats(self) -> dict:
        s = VLRAccelStats()
        call("vlrContextGetAccelStats", self.h, C.byref(s))
        return {f: getattr(s, f) for f, _ in VLRAccelStats._fields_}

    def read_accel(self, scene, which: int) -> np.ndarray:
        """One part of the scene's acceleration structure (ACCEL_TLAS / _BLAS / _TRIS /
        _INSTANCES) as a numpy array."""
        n = C.c_size_t()
        call("vlrContextReadAccel", self.h, scene, which, None, 0, C.byref(n))
        buf = np.zeros(n.value, np.uint8)
        call("vlrContextReadAccel", self.h, scene, which, buf.ctypes.data, n.value, C.byref(n))
        if which in (ACCEL_TLAS, ACCEL_BLAS):
            return buf.view(BVH4_NODE)
        if which == ACCEL_TRIS:
            return buf.view(TRI_RECORD)
        return buf.view(np.uint32).reshape(-1, 2)

    def build_bvh4(self, boxes_ptr: int, n: int, max_leaf: int, remap_ptr, nodes_ptr: int, capacity: int,
                   order_ptr: int, stream=None) -> int:
        """Device build over device-resident boxes (6 floats each); returns the node count."""
        cnt = U32()
        call("vlrContextBuildBvh4", self.h, stream, C.c_void_p(boxes_ptr), n, max_leaf,
             C.c_void_p(remap_ptr) if remap_ptr else None, C.c_void_p(nodes_ptr), capacity, C.c_void_p(order_ptr),
             C.byref(cnt))
        return cnt.value


def bvh4_node_capacity(n: int) -> int:
    c = U32()
    call("vlrBvh4NodeCapacity", n, C.byref(c))
    return c.value


def build_bvh4_host(boxes: np.ndarray, max_leaf: int, remap=None):
    """Host build over (n, 6) boxes; returns (nodes, order)."""
    b = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
    n = len(b)
    cap = bvh4_node_capacity(n)
    nodes = np.zeros(cap, BVH4_NODE)
    order = np.zeros(max(n, 1), np.uint32)
    r = None if remap is None else np.ascontiguousarray(remap, np.uint32)
    cnt = U32()
    call("vlrBuildBvh4Host", b.ctypes.data, n, max_leaf, None if r is None else r.ctypes.data, nodes.ctypes.data, cap,
         order.ctypes.data, C.byref(cnt))
    return nodes[:cnt.value].copy(), order[:n].copy()


def cache_num_parameters() -> int:
    n = U32()
    call("vlrRadianceCacheNumParameters", C.byref(n))
    return n.value


def set_spectrum(obj, name: str, cs: str, e0: float, e1: float, e2: float):
    s = VLRImmediateSpectrum(_b(cs), e0, e1, e2)
    call("vlrQueryableSetImmediateSpectrum", obj, _b(name), C.byref(s))


def set_float(obj, name: str, v: float):
    call("vlrQueryableSetFloat", obj, _b(name), float(v))


def build_image(ctx: Context, img, keep: list):
    data = np.ascontiguousarray(img.data, np.uint8)
    keep.append(data)
    h = H()
    if img.format.startswith("BC"):
        # one mip level of blocks (vlrBlockCompressedImage2DCreate, vlr.h)
        ptrs = (C.c_void_p * 1)(data.ctypes.data)
        sizes = (C.c_size_t * 1)(data.nbytes)
        keep += [ptrs, sizes]
        call("vlrBlockCompressedImage2DCreate", ctx.h, C.cast(ptrs, C.c_void_p), C.cast(sizes, C.c_void_p), 1,
             img.width, img.height, _b(img.format), _b(img.spectrum_type), _b(img.color_space), C.byref(h))
        return h
    call("vlrLinearImage2DCreate", ctx.h, data.ctypes.data, img.width, img.height, _b(img.format),
         _b(img.spectrum_type), _b(img.color_space), C.byref(h))
    return h


def build_node(ctx: Context, node, cache: dict, keep: list):
    """Creates a scenes.NodeDesc graph through the C API (each node once); returns its handle."""
    if id(node) in cache:
        return cache[id(node)]
    h = ctx.shader_node(node.type)
    cache[id(node)] = h
    for name, vals in node.floats.items():
        arr = (C.c_float * len(vals))(*[float(v) for v in vals])
        call("vlrQueryableSetFloatTuple", h, _b(name), arr, len(vals))
    for name, value in node.enums.items():
        call("vlrQueryableSetEnumValue", h, _b(name), _b(value))
    if node.image is not None:
        call("vlrQueryableSetImage2D", h, b"image", build_image(ctx, node.image, keep))
    for name, (child, ptype, option) in node.plugs.items():
        call("vlrQueryableSetShaderNodePlug", h, _b(name), node_plug(ctx, child, ptype, option, cache, keep))
    return h


def node_plug(ctx: Context, node, ptype: int, option: int, cache: dict, keep: list) -> VLRShaderNodePlug:
    plug = VLRShaderNodePlug()
    call("vlrShaderNodeGetPlug", build_node(ctx, node, cache, keep), ptype, option, C.byref(plug))
    return plug


def build_material(ctx: Context, m, env_texture_node=None, node_cache=None, keep=None):
    h = ctx.material(m.type)
    node_cache = {} if node_cache is None else node_cache
    keep = [] if keep is None else keep
    for name, (node, ptype, option) in getattr(m, "plugs", {}).items():
        call("vlrQueryableSetShaderNodePlug", h, _b(name), node_plug(ctx, node, ptype, option, node_cache, keep))
    for name, (cs, e0, e1, e2) in m.spectra.items():
        if m.use_env_texture and name == "emittance" and env_texture_node is not None:
            continue
        set_spectrum(h, name, cs, e0, e1, e2)
    for name, v in m.floats.items():
        set_float(h, name, v)
    if m.use_env_texture and env_texture_node is not None:
        plug = VLRShaderNodePlug()
        call("vlrShaderNodeGetPlug", env_texture_node, PLUG_SPECTRUM, 0, C.byref(plug))
        call("vlrQueryableSetShaderNodePlug", h, b"emittance", plug)
    return h


class BuiltScene:
    def __init__(self):
        self.scene = None
        self.camera = None
        self.materials = []
        self.meshes = []
        self.instances = []
        self.keep = []


def build_scene(ctx: Context, desc) -> BuiltScene:
    """Create the objects of a SceneDesc through the C API and make it the context's scene."""
    bs = BuiltScene()
    env_node = None
    if desc.env is not None and desc.env.texture_half is not None:
        tex = np.ascontiguousarray(desc.env.texture_half, np.uint16)
        bs.keep.append(tex)
        img = H()
        call("vlrLinearImage2DCreate", ctx.h, tex.ctypes.data, tex.shape[1], tex.shape[0], b"RGBA16Fx4",
             b"Light Source", b"Rec709(D65)", C.byref(img))
        env_node = ctx.shader_node("EnvironmentTexture")
        call("vlrQueryableSetImage2D", env_node, b"image", img)
    node_cache = {}
    for m in desc.materials:
        bs.materials.append(build_material(ctx, m, env_node, node_cache, bs.keep))
    for mesh in desc.meshes:
        n = H()
        v = np.ascontiguousarray(mesh.vertices, np.float32)
        assert v.shape[1] == 11
        if mesh.points:
            call("vlrPointSurfaceNodeCreate", ctx.h, _b(mesh.name), C.byref(n))
            call("vlrPointSurfaceNodeSetVertices", n, v.ctypes.data, len(v))
            for idx, mi in mesh.groups:
                ix = np.ascontiguousarray(idx, np.uint32)
                call("vlrPointSurfaceNodeAddMaterialGroup", n, ix.ctypes.data, len(ix), bs.materials[mi])
        else:
            call("vlrTriangleMeshSurfaceNodeCreate", ctx.h, _b(mesh.name), C.byref(n))
            call("vlrTriangleMeshSurfaceNodeSetVertices", n, v.ctypes.data, len(v))
            for gi, (idx, mi) in enumerate(mesh.groups):
                ix = np.ascontiguousarray(idx, np.uint32)
                plugs = [NULL_PLUG, NULL_PLUG, NULL_PLUG]
                for k, p in enumerate(getattr(mesh, "group_nodes", {}).get(gi, ())):
                    if p is not None:
                        plugs[k] = node_plug(ctx, p[0], p[1], p[2], node_cache, bs.keep)
                call("vlrTriangleMeshSurfaceNodeAddMaterialGroup", n, ix.ctypes.data, len(ix), bs.materials[mi],
                     plugs[0], plugs[1], plugs[2])
        bs.meshes.append(n)
    ident = ctx.transform(np.eye(4, dtype=np.float32).reshape(16, order="F"))
    scene = H()
    call("vlrSceneCreate", ctx.h, ident, C.byref(scene))
    for inst in desc.instances:
        xf = ctx.transform(inst.transform)
        node = H()
        call("vlrInternalNodeCreate", ctx.h, _b(inst.name), xf, C.byref(node))
        for mi in inst.meshes:
            call("vlrParentNodeAddChild", node, bs.meshes[mi])
        call("vlrParentNodeAddChild", scene, node)
        bs.instances.append(node)
    if desc.env is not None:
        call("vlrSceneSetEnvironment", scene, bs.materials[desc.env.material])
        call("vlrSceneSetEnvironmentRotation", scene, float(desc.env.rotation))
    cam = H()
    call("vlrCameraCreate", ctx.h, b"Perspective", C.byref(cam))
    c = desc.camera
    call("vlrQueryableSetPoint3D", cam, b"position", C.byref(VLRPoint3D(*c.position)))
    call("vlrQueryableSetQuaternion", cam, b"orientation", C.byref(VLRQuaternion(*c.orientation)))
    set_float(cam, "aspect", c.aspect)
    set_float(cam, "sensitivity", c.sensitivity)
    set_float(cam, "fovy", c.fovy)
    set_float(cam, "lens radius", c.lens_radius)
    set_float(cam, "op distance", c.op_distance)
    ctx.set_scene(scene)
    bs.scene = scene
    bs.camera = cam
    return bs
