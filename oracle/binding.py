"""Oracle (test infrastructure only): ctypes binding of oracle/build/liboracle.so.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, and only there;
the product (vlr_amd) never imports it.  Converts a vlr_amd.scenes.SceneDesc into the
oracle's flat C description.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "liboracle.so"
TABLES = HERE.parent / "vlr_amd" / "data" / "spectral_tables.bin"


class OracleVertex(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("normal", C.c_float * 3), ("tc0Direction", C.c_float * 3),
                ("texCoord", C.c_float * 2)]


class OracleMaterial(C.Structure):
    _fields_ = [("type", C.c_int32), ("colorSpace", C.c_int32 * 3), ("spectra", (C.c_float * 3) * 3),
                ("scalars", C.c_float * 4), ("useEnvTexture", C.c_int32)]


class OracleGeomInst(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("numVertices", C.c_uint32), ("indices", C.c_void_p),
                ("numIndices", C.c_uint32), ("material", C.c_int32), ("isPoints", C.c_int32),
                ("alpha", C.c_float)]


class OracleInstance(C.Structure):
    _fields_ = [("transform", C.c_float * 16), ("geomInsts", C.c_void_p), ("numGeomInsts", C.c_uint32)]


class OracleCamera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("orientation", C.c_float * 4), ("sensitivity", C.c_float),
                ("aspect", C.c_float), ("fovY", C.c_float), ("lensRadius", C.c_float), ("opDistance", C.c_float)]


class OracleScene(C.Structure):
    _fields_ = [("materials", C.c_void_p), ("numMaterials", C.c_uint32),
                ("geomInsts", C.c_void_p), ("numGeomInsts", C.c_uint32),
                ("instances", C.c_void_p), ("numInstances", C.c_uint32),
                ("envMaterial", C.c_int32), ("envTextureHalf", C.c_void_p), ("envWidth", C.c_uint32),
                ("envHeight", C.c_uint32), ("envRotation", C.c_float), ("camera", OracleCamera)]


class OracleRenderStats(C.Structure):
    _fields_ = [("closestRays", C.c_uint64), ("shadowRays", C.c_uint64), ("paths", C.c_uint64),
                ("droppedNonFinite", C.c_uint64)]


class OracleBSDFQuery(C.Structure):
    _fields_ = [("dirLocal", C.c_float * 3), ("geomNormalLocal", C.c_float * 3), ("wlOffset", C.c_float),
                ("wlSelect", C.c_float), ("uComponent", C.c_float), ("uDir", C.c_float * 2),
                ("evalDir", C.c_float * 3), ("singleIsSelected", C.c_int32)]


class OracleBSDFResult(C.Structure):
    _fields_ = [("sampledDir", C.c_float * 3), ("sampledPDF", C.c_float), ("sampledType", C.c_uint32),
                ("sampledValue", C.c_float * 4), ("evalValue", C.c_float * 4), ("evalPDF", C.c_float),
                ("revValue", C.c_float * 4), ("revPDF", C.c_float)]


RAY_DTYPE = np.dtype([("org", np.float32, 3), ("tmin", np.float32), ("dir", np.float32, 3), ("tmax", np.float32)])
HIT_DTYPE = np.dtype([("inst", np.int32), ("geomInst", np.int32), ("prim", np.int32), ("t", np.float32),
                      ("b1", np.float32), ("b2", np.float32)])

MATERIAL_CODES = {"Matte": 0, "SpecularReflection": 1, "SpecularScattering": 2, "MicrofacetReflection": 3,
                  "UE4": 4, "DiffuseEmitter": 5, "PointEmitter": 6, "EnvironmentEmitter": 7,
                  "MicrofacetScattering": 8}
# slot order of each type's immediate spectra / floats (names as in the reference's materials.cpp)
SPECTRUM_SLOTS = {"Matte": ["albedo"], "SpecularReflection": ["coeff", "eta", "k"],
                  "SpecularScattering": ["coeff", "eta ext", "eta int"], "MicrofacetReflection": ["eta", "k"],
                  "UE4": ["base color"], "DiffuseEmitter": ["emittance"], "PointEmitter": ["intensity"],
                  "EnvironmentEmitter": ["emittance"],
                  "MicrofacetScattering": ["coeff", "eta ext", "eta int"]}
FLOAT_SLOTS = {"MicrofacetReflection": ["roughness", "anisotropy", "rotation"],
               "MicrofacetScattering": ["roughness", "anisotropy", "rotation"],
               "UE4": ["occlusion", "roughness", "metallic"], "DiffuseEmitter": ["scale"],
               "PointEmitter": ["scale"], "EnvironmentEmitter": ["scale"]}
# defaults of the reference's material constructors (materials.cpp:273, 385-386, 533-534, 684-685,
# 1262, 1615, 1941, 2205)
SPECTRUM_DEFAULTS = {("Matte", "albedo"): (1, 0.18, 0.18, 0.18),
                     ("SpecularReflection", "coeff"): (1, 0.8, 0.8, 0.8),
                     ("SpecularReflection", "eta"): (1, 1.0, 1.0, 1.0),
                     ("SpecularReflection", "k"): (1, 0.0, 0.0, 0.0),
                     ("SpecularScattering", "coeff"): (1, 0.8, 0.8, 0.8),
                     ("SpecularScattering", "eta ext"): (1, 1.0, 1.0, 1.0),
                     ("SpecularScattering", "eta int"): (1, 1.5, 1.5, 1.5),
                     ("MicrofacetReflection", "eta"): (1, 1.0, 1.0, 1.0),
                     ("MicrofacetReflection", "k"): (1, 0.0, 0.0, 0.0),
                     ("MicrofacetScattering", "coeff"): (1, 0.8, 0.8, 0.8),
                     ("MicrofacetScattering", "eta ext"): (1, 1.0, 1.0, 1.0),
                     ("MicrofacetScattering", "eta int"): (1, 1.5, 1.5, 1.5),
                     ("UE4", "base color"): (1, 0.18, 0.18, 0.18),
                     ("DiffuseEmitter", "emittance"): (1, np.pi, np.pi, np.pi),
                     ("PointEmitter", "intensity"): (1, np.pi, np.pi, np.pi),
                     ("EnvironmentEmitter", "emittance"): (1, np.pi, np.pi, np.pi)}
FLOAT_DEFAULTS = {("MicrofacetScattering", "roughness"): 0.1, ("MicrofacetScattering", "anisotropy"): 0.0,
                  ("MicrofacetScattering", "rotation"): 0.0,
                  ("MicrofacetReflection", "roughness"): 0.1, ("MicrofacetReflection", "anisotropy"): 0.0,
                  ("MicrofacetReflection", "rotation"): 0.0, ("UE4", "occlusion"): 0.0,
                  ("UE4", "roughness"): 0.1, ("UE4", "metallic"): 0.0, ("DiffuseEmitter", "scale"): 1.0,
                  ("PointEmitter", "scale"): 1.0, ("EnvironmentEmitter", "scale"): 1.0}
COLOR_SPACE_CODES = {"Rec709(D65) sRGB Gamma": 0, "Rec709(D65)": 1, "XYZ": 2, "xyY": 3}

_lib = None


def build(force: bool = False) -> Path:
    """Compile liboracle.so with the committed Makefile."""
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(HERE)] + (["-B"] if force else []), check=True)
    return LIB_PATH


def build_native(out_dir) -> Path:
    """Compile a copy tuned for this machine's CPU (-march=native) into out_dir."""
    subprocess.run(["make", "-s", "-C", str(HERE), "-B", "ARCH=-march=native", f"OUT={out_dir}"], check=True)
    return Path(out_dir) / "liboracle.so"


def lib(path=None):
    """The oracle library: the default build, or `path` (e.g. a build_native copy) when given
    on the first call."""
    global _lib
    if _lib is None:
        if path is None and not LIB_PATH.exists():
            build()
        L = C.CDLL(str(path or LIB_PATH))
        L.oracle_load_tables.argtypes = [C.c_char_p]
        L.oracle_scene_create.restype = C.c_void_p
        L.oracle_scene_create.argtypes = [C.POINTER(OracleScene), C.c_int]
        L.oracle_scene_destroy.argtypes = [C.c_void_p]
        L.oracle_scene_bounds.argtypes = [C.c_void_p, C.c_void_p]
        L.oracle_trace_closest.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_int, C.c_int]
        L.oracle_trace_any.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_int, C.c_int]
        L.oracle_render.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(OracleRenderStats)]
        L.oracle_render_guides.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int,
                                           C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.POINTER(OracleRenderStats)]
        L.oracle_bsdf_queries.argtypes = [C.POINTER(OracleMaterial), C.c_int, C.c_void_p, C.c_uint32, C.c_void_p]
        L.oracle_pcg32_stream.restype = C.c_uint32
        L.oracle_pcg32_stream.argtypes = [C.c_uint64, C.c_uint32, C.c_void_p, C.c_void_p]
        L.oracle_seed_states.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p]
        L.oracle_render_bpt.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.POINTER(OracleRenderStats)]
        L.oracle_render_lt.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                       C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(OracleRenderStats)]
        L.oracle_sample_lights.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]
        L.oracle_env_lookup.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]
        L.oracle_pcg32_floats.argtypes = [C.c_uint64, C.c_uint32, C.c_void_p]
        L.oracle_mt19937_64.argtypes = [C.c_uint64, C.c_uint32, C.c_void_p]
        L.oracle_meng_eval.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_void_p, C.c_void_p]
        L.oracle_meng_uvs_eval.argtypes = [C.c_float, C.c_float, C.c_float, C.c_void_p, C.c_void_p]
        L.oracle_cmf_bins.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_triplet_rgb.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_void_p]
        if L.oracle_load_tables(str(TABLES).encode()) != 0:
            raise RuntimeError(f"oracle: cannot load {TABLES}")
        _lib = L
    return _lib


def material_struct(m) -> OracleMaterial:
    om = OracleMaterial()
    om.type = MATERIAL_CODES[m.type]
    for k, name in enumerate(SPECTRUM_SLOTS.get(m.type, [])):
        if name in m.spectra:
            cs, e0, e1, e2 = m.spectra[name]
            cs = COLOR_SPACE_CODES[cs] if isinstance(cs, str) else int(cs)
        else:
            cs, e0, e1, e2 = SPECTRUM_DEFAULTS[(m.type, name)]
        om.colorSpace[k] = cs
        om.spectra[k][0], om.spectra[k][1], om.spectra[k][2] = e0, e1, e2
    for k, name in enumerate(FLOAT_SLOTS.get(m.type, [])):
        om.scalars[k] = m.floats.get(name, FLOAT_DEFAULTS[(m.type, name)])
    om.useEnvTexture = 1 if m.use_env_texture else 0
    return om


LIGHT_SAMPLE_DTYPE = np.dtype([("position", np.float32, 3), ("normal", np.float32, 3), ("areaPDF", np.float32),
                               ("lightProb", np.float32), ("inst", np.int32), ("prim", np.int32),
                               ("atInfinity", np.int32), ("valid", np.int32)])


class Scene:
    """Oracle scene built from a SceneDesc (keeps the numpy buffers alive)."""

    def __init__(self, desc, spectral=None):
        L = lib()
        self.desc = desc
        spectral = desc.spectral if spectral is None else spectral
        self._keep = []
        mats = (OracleMaterial * len(desc.materials))(*[material_struct(m) for m in desc.materials])
        geoms = []
        mesh_to_geoms = []
        for mesh in desc.meshes:
            v = np.ascontiguousarray(mesh.vertices, np.float32)
            self._keep.append(v)
            ids = []
            for gi, (idx, mat) in enumerate(mesh.groups):
                ix = np.ascontiguousarray(idx, np.uint32)
                self._keep.append(ix)
                alpha = float(getattr(mesh, "group_alpha", {}).get(gi, -1.0))
                g = OracleGeomInst(v.ctypes.data, len(v), ix.ctypes.data, len(ix), mat, 1 if mesh.points else 0, alpha)
                ids.append(len(geoms))
                geoms.append(g)
            mesh_to_geoms.append(ids)
        garr = (OracleGeomInst * max(1, len(geoms)))(*geoms)
        insts = []
        for inst in desc.instances:
            gi = np.asarray([g for mi in inst.meshes for g in mesh_to_geoms[mi]], np.uint32)
            self._keep.append(gi)
            oi = OracleInstance()
            for k in range(16):
                oi.transform[k] = float(inst.transform[k])
            oi.geomInsts = gi.ctypes.data
            oi.numGeomInsts = len(gi)
            insts.append(oi)
        iarr = (OracleInstance * max(1, len(insts)))(*insts)
        sd = OracleScene()
        sd.materials, sd.numMaterials = C.cast(mats, C.c_void_p), len(desc.materials)
        sd.geomInsts, sd.numGeomInsts = C.cast(garr, C.c_void_p), len(geoms)
        sd.instances, sd.numInstances = C.cast(iarr, C.c_void_p), len(insts)
        sd.envMaterial = -1
        if desc.env is not None:
            sd.envMaterial = desc.env.material
            if desc.env.texture_half is not None:
                tex = np.ascontiguousarray(desc.env.texture_half, np.uint16)
                self._keep.append(tex)
                sd.envTextureHalf = tex.ctypes.data
                sd.envHeight, sd.envWidth = tex.shape[0], tex.shape[1]
            sd.envRotation = desc.env.rotation
        cam = desc.camera
        sd.camera.position[:] = list(cam.position)
        sd.camera.orientation[:] = list(cam.orientation)
        sd.camera.sensitivity, sd.camera.aspect, sd.camera.fovY = cam.sensitivity, cam.aspect, cam.fovy
        sd.camera.lensRadius, sd.camera.opDistance = cam.lens_radius, cam.op_distance
        self._keep += [mats, garr, iarr, sd]
        self.mesh_to_geoms = mesh_to_geoms
        self.spectral = bool(spectral)
        self.handle = L.oracle_scene_create(C.byref(sd), 1 if spectral else 0)
        if not self.handle:
            raise RuntimeError("oracle_scene_create failed")

    def close(self):
        if self.handle:
            lib().oracle_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bounds(self):
        out = np.zeros(5, np.float32)
        lib().oracle_scene_bounds(self.handle, out.ctypes.data)
        return out

    def trace_closest(self, rays: np.ndarray, use_bvh=True, threads=0) -> np.ndarray:
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        lib().oracle_trace_closest(self.handle, rays.ctypes.data, len(rays), hits.ctypes.data, int(use_bvh), threads)
        return hits

    def trace_any(self, rays: np.ndarray, use_bvh=True, threads=0) -> np.ndarray:
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        occ = np.zeros(len(rays), np.uint8)
        lib().oracle_trace_any(self.handle, rays.ctypes.data, len(rays), occ.ctypes.data, int(use_bvh), threads)
        return occ

    def render_bpt(self, width, height, frames, max_path_length=25, num_light_paths=500000, threads=0,
                   rng_state=None, light_rng_state=None, want_accum=False):
        """LVC-BPT render; returns (image, stats, rng states, light-path rng states, accumulation)."""
        out = np.zeros((height, width, 4), np.float32)
        st = OracleRenderStats()
        rs = None if rng_state is None else np.ascontiguousarray(rng_state, np.uint64)
        ls = None if light_rng_state is None else np.ascontiguousarray(light_rng_state, np.uint64)
        ch = 16 if self.spectral else 3
        acc = np.zeros((height, width, ch, 2), np.float32) if want_accum else None
        rc = lib().oracle_render_bpt(self.handle, width, height, frames, max_path_length, num_light_paths, threads,
                                     out.ctypes.data, None if rs is None else rs.ctypes.data,
                                     None if ls is None else ls.ctypes.data,
                                     None if acc is None else acc.ctypes.data, C.byref(st))
        if rc != 0:
            raise RuntimeError("oracle_render_bpt failed")
        stats = {"closest_rays": st.closestRays, "shadow_rays": st.shadowRays, "paths": st.paths,
                 "dropped_non_finite": st.droppedNonFinite}
        return out, stats, rs, ls, acc

    def render_lt(self, width, height, frames, max_path_length=25, num_light_paths=500000, threads=0,
                  light_rng_state=None, want_accum=False):
        """Light tracing render; returns (image, stats, light-path rng states, accumulation)."""
        out = np.zeros((height, width, 4), np.float32)
        st = OracleRenderStats()
        ls = None if light_rng_state is None else np.ascontiguousarray(light_rng_state, np.uint64)
        ch = 16 if self.spectral else 3
        acc = np.zeros((height, width, ch, 2), np.float32) if want_accum else None
        rc = lib().oracle_render_lt(self.handle, width, height, frames, max_path_length, num_light_paths, threads,
                                    out.ctypes.data, None if ls is None else ls.ctypes.data,
                                    None if acc is None else acc.ctypes.data, C.byref(st))
        if rc != 0:
            raise RuntimeError("oracle_render_lt failed")
        stats = {"closest_rays": st.closestRays, "shadow_rays": st.shadowRays, "paths": st.paths,
                 "dropped_non_finite": st.droppedNonFinite}
        return out, stats, ls, acc

    def render_guides(self, width, height, frames, max_path_length=25, threads=0, rng_state=None):
        """Path-traced render with the denoiser guides; returns (image, albedo, normal, rng states,
        stats), the guides as W x H RGBA after copyBuffers' conversion."""
        out = np.zeros((height, width, 4), np.float32)
        alb = np.zeros((height, width, 4), np.float32)
        nrm = np.zeros((height, width, 4), np.float32)
        st = OracleRenderStats()
        rs = None if rng_state is None else np.ascontiguousarray(rng_state, np.uint64)
        rc = lib().oracle_render_guides(self.handle, width, height, frames, max_path_length, threads, out.ctypes.data,
                                        None if rs is None else rs.ctypes.data, alb.ctypes.data, nrm.ctypes.data,
                                        C.byref(st))
        if rc != 0:
            raise RuntimeError("oracle_render_guides failed")
        stats = {"closest_rays": st.closestRays, "shadow_rays": st.shadowRays, "paths": st.paths,
                 "dropped_non_finite": st.droppedNonFinite}
        return out, alb, nrm, rs, stats

    def sample_lights(self, u: np.ndarray) -> np.ndarray:
        """u: (n, 3) uniform numbers -> LIGHT_SAMPLE_DTYPE records."""
        u = np.ascontiguousarray(u, np.float32)
        out = np.zeros(len(u), LIGHT_SAMPLE_DTYPE)
        lib().oracle_sample_lights(self.handle, len(u), u.ctypes.data, out.ctypes.data)
        return out

    def env_lookup(self, dirs: np.ndarray) -> np.ndarray:
        """dirs: (n, 3) world directions -> (n, 4): tex u, tex v, solid-angle pdf, texel channel 0."""
        d = np.ascontiguousarray(dirs, np.float32)
        out = np.zeros((len(d), 4), np.float32)
        lib().oracle_env_lookup(self.handle, len(d), d.ctypes.data, out.ctypes.data)
        return out

    def render(self, width, height, frames, max_path_length=25, threads=0, rng_state=None, want_accum=False,
               pixel_range=None):
        out = np.zeros((height, width, 4), np.float32)
        st = OracleRenderStats()
        rs = None
        if rng_state is not None:
            rs = np.ascontiguousarray(rng_state, np.uint64)
        ch = 16 if self.spectral else 3
        acc = np.zeros((height, width, ch, 2), np.float32) if want_accum else None
        pr = None if pixel_range is None else np.asarray(pixel_range, np.int32)
        rc = lib().oracle_render(self.handle, width, height, frames, max_path_length, threads, out.ctypes.data,
                                 None if rs is None else rs.ctypes.data,
                                 None if acc is None else acc.ctypes.data,
                                 None if pr is None else pr.ctypes.data, C.byref(st))
        if rc != 0:
            raise RuntimeError("oracle_render failed")
        stats = {"closest_rays": st.closestRays, "shadow_rays": st.shadowRays, "paths": st.paths,
                 "dropped_non_finite": st.droppedNonFinite}
        return out, stats, rs, acc


def bsdf_queries(material, spectral: bool, queries: np.ndarray) -> np.ndarray:
    """queries: structured array of OracleBSDFQuery fields (see query_dtype())."""
    q = np.ascontiguousarray(queries, query_dtype())
    out = np.zeros(len(q), result_dtype())
    om = material_struct(material)
    rc = lib().oracle_bsdf_queries(C.byref(om), 1 if spectral else 0, q.ctypes.data, len(q), out.ctypes.data)
    if rc != 0:
        raise RuntimeError("oracle_bsdf_queries failed")
    return out


def query_dtype():
    return np.dtype([("dirLocal", np.float32, 3), ("geomNormalLocal", np.float32, 3), ("wlOffset", np.float32),
                     ("wlSelect", np.float32), ("uComponent", np.float32), ("uDir", np.float32, 2),
                     ("evalDir", np.float32, 3), ("singleIsSelected", np.int32), ("transportMode", np.int32)])


def result_dtype():
    return np.dtype([("sampledDir", np.float32, 3), ("sampledPDF", np.float32), ("sampledType", np.uint32),
                     ("sampledValue", np.float32, 4), ("evalValue", np.float32, 4), ("evalPDF", np.float32),
                     ("revValue", np.float32, 4), ("revPDF", np.float32), ("sampledRevValue", np.float32, 4),
                     ("sampledRevPDF", np.float32)])


def pcg32_stream(state: int, n: int):
    out = np.zeros(n, np.uint32)
    fin = np.zeros(1, np.uint64)
    lib().oracle_pcg32_stream(C.c_uint64(state), n, out.ctypes.data, fin.ctypes.data)
    return out, int(fin[0])


def pcg32_floats(state: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.float32)
    lib().oracle_pcg32_floats(C.c_uint64(state), n, out.ctypes.data)
    return out


def mt19937_64(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint64)
    lib().oracle_mt19937_64(C.c_uint64(seed), n, out.ctypes.data)
    return out


def seed_states(width: int, height: int) -> np.ndarray:
    out = np.zeros(width * height, np.uint64)
    lib().oracle_seed_states(width, height, out.ctypes.data)
    return out


def meng_eval(spectrum_type: int, color_space: int, rgb, lambdas) -> np.ndarray:
    lam = np.asarray(lambdas, np.float32)
    out = np.zeros(4, np.float32)
    lib().oracle_meng_eval(spectrum_type, color_space, *map(float, rgb), lam.ctypes.data, out.ctypes.data)
    return out


def meng_uvs_eval(u: float, v: float, scale: float, lambdas) -> np.ndarray:
    lam = np.ascontiguousarray(lambdas, np.float32)
    out = np.zeros(4, np.float32)
    lib().oracle_meng_uvs_eval(C.c_float(u), C.c_float(v), C.c_float(scale), lam.ctypes.data, out.ctypes.data)
    return out


def cmf_bins():
    x, y, z = (np.zeros(16, np.float32) for _ in range(3))
    integ = np.zeros(1, np.float32)
    lib().oracle_cmf_bins(x.ctypes.data, y.ctypes.data, z.ctypes.data, integ.ctypes.data)
    return x, y, z, float(integ[0])


def triplet_rgb(spectrum_type: int, color_space: int, rgb) -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib().oracle_triplet_rgb(spectrum_type, color_space, *map(float, rgb), out.ctypes.data)
    return out
