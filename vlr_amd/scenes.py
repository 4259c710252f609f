"""Procedural scene descriptions for the light-transport workloads.

A :class:`SceneDesc` is plain numpy data: materials with the reference's immediate
parameters, triangle/point meshes with material groups, instances (column-major 4x4
local-to-world transforms over lists of meshes), an optional environment and a perspective
camera.  :func:`vlr_amd.api.build_scene` turns one into objects of the C API
(``vlrTriangleMeshSurfaceNodeCreate`` ... ``vlrSceneSetEnvironment``); the oracle in
``oracle/binding.py`` reads the same description.

Generators:
  * :func:`cornell_box` -- the RGB Cornell box (5 wall quads, two rotated boxes, one
    0.25 x 0.25 light quad of radiance (17, 12, 4)), camera as in
    HostProgram/scene.cpp:789-801 of the reference.
  * :func:`icosphere_grid` -- a grid of 1,280-triangle icospheres over a ground plane with
    GGX (UE4, metallic 1) and smooth dielectric materials and an exp-mapped value-noise
    environment.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

# Material type names are those of vlrSurfaceMaterialCreate (reference vlr.cpp:955-991).
MATTE = "Matte"
SPECULAR_REFLECTION = "SpecularReflection"
SPECULAR_SCATTERING = "SpecularScattering"
MICROFACET_REFLECTION = "MicrofacetReflection"
MICROFACET_SCATTERING = "MicrofacetScattering"
UE4 = "UE4"
DIFFUSE_EMITTER = "DiffuseEmitter"
POINT_EMITTER = "PointEmitter"
ENVIRONMENT_EMITTER = "EnvironmentEmitter"

# Colour-space enum strings of the reference (queryable.cpp:34-45).
REC709_SRGB = "Rec709(D65) sRGB Gamma"
REC709 = "Rec709(D65)"
XYZ = "XYZ"
XYY = "xyY"
COLOR_SPACES = [REC709_SRGB, REC709, XYZ, XYY]


@dataclass
class ImageDesc:
    """A LinearImage2D: raw texel bytes in one of the C API's formats."""
    data: np.ndarray                 # uint8, height x width x bytes per texel
    width: int
    height: int
    format: str                      # "RGBA8x4", "RGB8x3", "RGBA16Fx4", "RGBA32Fx4", "Gray8", "Gray32F", ...
    spectrum_type: str = "Reflectance"
    color_space: str = REC709


@dataclass
class NodeDesc:
    """A shader node: its type name (vlrShaderNodeCreate), immediate parameters and inputs."""
    type: str
    floats: dict = field(default_factory=dict)      # name -> list of floats
    enums: dict = field(default_factory=dict)       # name -> enum string
    plugs: dict = field(default_factory=dict)       # name -> (NodeDesc, plug type, option)
    image: Optional[ImageDesc] = None


@dataclass
class MaterialDesc:
    type: str
    # immediate spectra: parameter name -> (colour space, e0, e1, e2)
    spectra: dict = field(default_factory=dict)
    # float parameters: name -> value
    floats: dict = field(default_factory=dict)
    # environment emitter: emittance from the scene's environment texture
    use_env_texture: bool = False
    # parameters connected to shader nodes: name -> (NodeDesc, plug type, option)
    plugs: dict = field(default_factory=dict)


@dataclass
class MeshDesc:
    name: str
    vertices: np.ndarray            # float32 [N, 11]: position(3) normal(3) tc0Direction(3) texCoord(2)
    groups: List[Tuple[np.ndarray, int]]  # (uint32 indices, material index)
    points: bool = False
    # shader nodes of a triangle group: group index -> (normal plug, tangent plug[, alpha plug]),
    # each a (NodeDesc, plug type, option) or None
    group_nodes: dict = field(default_factory=dict)
    # constant alpha of a triangle group for the oracle (group index -> alpha); the C API takes
    # the same value as an alpha node in group_nodes
    group_alpha: dict = field(default_factory=dict)


@dataclass
class InstanceDesc:
    name: str
    transform: np.ndarray           # float32 [16], column-major local-to-world
    meshes: List[int]


@dataclass
class CameraDesc:
    position: Tuple[float, float, float]
    orientation: Tuple[float, float, float, float]  # quaternion x, y, z, w
    aspect: float
    fovy: float
    sensitivity: float = 1.0
    lens_radius: float = 0.0
    op_distance: float = 1.0


@dataclass
class EnvDesc:
    material: int
    texture_half: Optional[np.ndarray] = None   # uint16 [H, W, 4] RGBA16F
    rotation: float = 0.0


@dataclass
class SceneDesc:
    materials: List[MaterialDesc]
    meshes: List[MeshDesc]
    instances: List[InstanceDesc]
    camera: CameraDesc
    env: Optional[EnvDesc] = None
    spectral: bool = False
    name: str = ""

    @property
    def num_triangles(self) -> int:
        per_mesh = [sum(len(ix) // 3 for ix, _ in m.groups) for m in self.meshes]
        return int(sum(per_mesh[i] for inst in self.instances for i in inst.meshes if not self.meshes[i].points))


# --------------------------------------------------------------------------------------------
# helpers

def identity4() -> np.ndarray:
    return np.eye(4, dtype=np.float32).reshape(16, order="F").copy()


def to_colmajor(m: np.ndarray) -> np.ndarray:
    return np.asarray(m, dtype=np.float32).reshape(4, 4).reshape(16, order="F").copy()


def translate(t) -> np.ndarray:
    m = np.eye(4, dtype=np.float64)
    m[:3, 3] = t
    return m


def scale(s) -> np.ndarray:
    m = np.eye(4, dtype=np.float64)
    m[0, 0] = m[1, 1] = m[2, 2] = s
    return m


def rotate_y(angle: float) -> np.ndarray:
    c, s = math.cos(angle), math.sin(angle)
    m = np.eye(4, dtype=np.float64)
    m[0, 0], m[0, 2], m[2, 0], m[2, 2] = c, s, -s, c
    return m


def q_rotate_y(angle: float):
    """Quaternion of a rotation about +Y (basic_types.h:1417-1432 of the reference)."""
    s, c = math.sin(angle / 2), math.cos(angle / 2)
    return (0.0, s, 0.0, c)


def q_look_at(position, target, up=(0.0, 1.0, 0.0)):
    """Camera orientation whose local +Z looks from ``position`` to ``target``."""
    z = np.asarray(target, np.float64) - np.asarray(position, np.float64)
    z /= np.linalg.norm(z)
    x = np.cross(np.asarray(up, np.float64), z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    m = np.stack([x, y, z], axis=1)  # columns = local axes in world space
    tr = m[0, 0] + m[1, 1] + m[2, 2]
    if tr > 0:
        s = math.sqrt(tr + 1.0) * 2
        w = 0.25 * s
        qx = (m[2, 1] - m[1, 2]) / s
        qy = (m[0, 2] - m[2, 0]) / s
        qz = (m[1, 0] - m[0, 1]) / s
    elif m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
        s = math.sqrt(1.0 + m[0, 0] - m[1, 1] - m[2, 2]) * 2
        w = (m[2, 1] - m[1, 2]) / s
        qx = 0.25 * s
        qy = (m[0, 1] + m[1, 0]) / s
        qz = (m[0, 2] + m[2, 0]) / s
    elif m[1, 1] > m[2, 2]:
        s = math.sqrt(1.0 + m[1, 1] - m[0, 0] - m[2, 2]) * 2
        w = (m[0, 2] - m[2, 0]) / s
        qx = (m[0, 1] + m[1, 0]) / s
        qy = 0.25 * s
        qz = (m[1, 2] + m[2, 1]) / s
    else:
        s = math.sqrt(1.0 + m[2, 2] - m[0, 0] - m[1, 1]) * 2
        w = (m[1, 0] - m[0, 1]) / s
        qx = (m[0, 2] + m[2, 0]) / s
        qy = (m[1, 2] + m[2, 1]) / s
        qz = 0.25 * s
    return (float(qx), float(qy), float(qz), float(w))


def _quad(p0, p1, p2, p3, n, t):
    v = np.zeros((4, 11), np.float32)
    for i, p in enumerate((p0, p1, p2, p3)):
        v[i, 0:3] = p
        v[i, 3:6] = n
        v[i, 6:9] = t
    v[:, 9:11] = [[0, 1], [1, 1], [1, 0], [0, 0]]
    return v


def box_mesh(sx: float, sy: float, sz: float) -> np.ndarray:
    """Axis-aligned box [−sx/2, sx/2] x [0, sy] x [−sz/2, sz/2] as 6 quads (24 vertices, 12 tris)."""
    hx, hz = sx / 2, sz / 2
    quads = [
        _quad((-hx, sy, -hz), (-hx, sy, hz), (hx, sy, hz), (hx, sy, -hz), (0, 1, 0), (1, 0, 0)),     # top
        _quad((-hx, 0, hz), (-hx, 0, -hz), (hx, 0, -hz), (hx, 0, hz), (0, -1, 0), (1, 0, 0)),        # bottom
        _quad((-hx, 0, hz), (hx, 0, hz), (hx, sy, hz), (-hx, sy, hz), (0, 0, 1), (1, 0, 0)),         # +z
        _quad((hx, 0, -hz), (-hx, 0, -hz), (-hx, sy, -hz), (hx, sy, -hz), (0, 0, -1), (-1, 0, 0)),   # -z
        _quad((hx, 0, hz), (hx, 0, -hz), (hx, sy, -hz), (hx, sy, hz), (1, 0, 0), (0, 0, -1)),        # +x
        _quad((-hx, 0, -hz), (-hx, 0, hz), (-hx, sy, hz), (-hx, sy, -hz), (-1, 0, 0), (0, 0, 1)),    # -x
    ]
    return np.concatenate(quads, axis=0)


def quad_indices(num_quads: int) -> np.ndarray:
    idx = []
    for q in range(num_quads):
        b = 4 * q
        idx += [b, b + 1, b + 2, b, b + 2, b + 3]
    return np.asarray(idx, np.uint32)


# --------------------------------------------------------------------------------------------
# Cornell box (RGB, Lambert)

def cornell_box(seed: int = 20261202, light_size: float = 0.25, angles=None) -> SceneDesc:
    """The Cornell box; the two boxes turn about +Y by seeded angles in [0, 30 degrees], or by
    `angles` (radians: short box, tall box) when given."""
    rng = np.random.default_rng(seed)
    mats = [
        MaterialDesc(MATTE, {"albedo": (REC709, 0.73, 0.73, 0.73)}),
        MaterialDesc(MATTE, {"albedo": (REC709, 0.65, 0.05, 0.05)}),
        MaterialDesc(MATTE, {"albedo": (REC709, 0.12, 0.45, 0.15)}),
        MaterialDesc(DIFFUSE_EMITTER, {"emittance": (REC709, math.pi * 17, math.pi * 12, math.pi * 4)},
                     {"scale": 1.0}),
    ]
    walls = np.concatenate([
        _quad((-1.5, 0, -1.5), (-1.5, 0, 1.5), (1.5, 0, 1.5), (1.5, 0, -1.5), (0, 1, 0), (1, 0, 0)),      # floor
        _quad((-1.5, 0, -1.5), (1.5, 0, -1.5), (1.5, 3, -1.5), (-1.5, 3, -1.5), (0, 0, 1), (1, 0, 0)),    # back
        _quad((-1.5, 3, -1.5), (1.5, 3, -1.5), (1.5, 3, 1.5), (-1.5, 3, 1.5), (0, -1, 0), (1, 0, 0)),     # ceiling
        _quad((-1.5, 0, 1.5), (-1.5, 0, -1.5), (-1.5, 3, -1.5), (-1.5, 3, 1.5), (1, 0, 0), (0, 0, -1)),   # left
        _quad((1.5, 0, -1.5), (1.5, 0, 1.5), (1.5, 3, 1.5), (1.5, 3, -1.5), (-1, 0, 0), (0, 0, 1)),       # right
    ], axis=0)
    wi = quad_indices(5)
    room = MeshDesc("CornellBox", walls, [(wi[0:18].copy(), 0), (wi[18:24].copy(), 1), (wi[24:30].copy(), 2)])
    h = light_size / 2
    light_v = _quad((-h, 2.99, -h), (h, 2.99, -h), (h, 2.99, h), (-h, 2.99, h), (0, -1, 0), (1, 0, 0))
    light = MeshDesc("Light", light_v, [(quad_indices(1), 3)])
    short_box = MeshDesc("ShortBox", box_mesh(0.6, 0.6, 0.6), [(quad_indices(6), 0)])
    tall_box = MeshDesc("TallBox", box_mesh(0.6, 1.2, 0.6), [(quad_indices(6), 0)])
    a0, a1 = rng.uniform(0.0, math.radians(30.0), size=2)
    if angles is not None:
        a0, a1 = float(angles[0]), float(angles[1])
    insts = [
        InstanceDesc("Room", identity4(), [0, 1]),
        InstanceDesc("ShortBoxInst", to_colmajor(translate((0.55, 0.0, 0.45)) @ rotate_y(-a0)), [2]),
        InstanceDesc("TallBoxInst", to_colmajor(translate((-0.5, 0.0, -0.4)) @ rotate_y(a1)), [3]),
    ]
    cam = CameraDesc((0.0, 1.5, 6.0), q_rotate_y(math.pi), 1.0, math.radians(40.0))
    return SceneDesc(mats, [room, light, short_box, tall_box], insts, cam, None, False, "cornell_box")


def cornell_many_lights(seed: int = 20261202, num_points: int = 64, num_quads: int = 8) -> SceneDesc:
    """The Cornell box with its ceiling light plus `num_points` point lights of seeded intensity
    U[1, 20] (one PointSurfaceNode, one PointEmitter material per point) and `num_quads`
    0.2 x 0.2 ceiling quad lights of seeded colour."""
    desc = cornell_box(seed)
    rng = np.random.default_rng(seed + 7)
    mats, meshes, insts = list(desc.materials), list(desc.meshes), list(desc.instances)
    pts = np.zeros((num_points, 11), np.float32)
    pts[:, 0] = rng.uniform(-1.2, 1.2, num_points)
    pts[:, 1] = rng.uniform(0.3, 2.7, num_points)
    pts[:, 2] = rng.uniform(-1.2, 1.2, num_points)
    pts[:, 3:6] = (0.0, 1.0, 0.0)
    pts[:, 6:9] = (1.0, 0.0, 0.0)
    groups = []
    for k in range(num_points):
        v = float(rng.uniform(1.0, 20.0))
        mats.append(MaterialDesc(POINT_EMITTER, {"intensity": (REC709, v, v, v)}, {"scale": 1.0}))
        groups.append((np.array([k], np.uint32), len(mats) - 1))
    meshes.append(MeshDesc("PointLights", pts, groups, points=True))
    quad_v, quad_groups = [], []
    for k in range(num_quads):
        cx, cz = rng.uniform(-1.2, 1.2, 2)
        col = rng.uniform(2.0, 8.0, 3)
        mats.append(MaterialDesc(DIFFUSE_EMITTER, {"emittance": (REC709, *map(float, math.pi * col))},
                                 {"scale": 1.0}))
        quad_v.append(_quad((cx - 0.1, 2.985, cz - 0.1), (cx + 0.1, 2.985, cz - 0.1), (cx + 0.1, 2.985, cz + 0.1),
                            (cx - 0.1, 2.985, cz + 0.1), (0, -1, 0), (1, 0, 0)))
        quad_groups.append((quad_indices(1) + 4 * k, len(mats) - 1))
    meshes.append(MeshDesc("QuadLights", np.concatenate(quad_v), quad_groups))
    insts.append(InstanceDesc("Lights", identity4(), [len(meshes) - 2, len(meshes) - 1]))
    return SceneDesc(mats, meshes, insts, desc.camera, None, False, "cornell_many_lights")


def tessellated_cornell(seed: int, target_tris: int) -> SceneDesc:
    """The Cornell box holding unit icospheres of jittered radius, each its own mesh (unique
    triangles, no instancing), subdivided and counted so the scene has about `target_tris`
    triangles; matte spheres of seeded albedo."""
    desc = cornell_box(seed)
    rng = np.random.default_rng(seed + 11)
    level = 3
    while level < 7 and 20 * 4 ** (level + 1) * 4 <= target_tris:
        level += 1
    per = 20 * 4 ** level
    count = max(1, int(round((target_tris - desc.num_triangles) / per)))
    sv, si = _icosphere_cached(level)
    mats, meshes, insts = list(desc.materials), list(desc.meshes), list(desc.instances)
    first = len(meshes)
    for k in range(count):
        alb = rng.uniform(0.1, 0.8, size=3)
        mats.append(MaterialDesc(MATTE, {"albedo": (REC709, *map(float, alb))}))
        r = float(rng.uniform(0.08, 0.3))
        c = rng.uniform((-1.2, 0.3, -1.2), (1.2, 2.5, 1.2))
        v = sv.copy()
        v[:, 0:3] = v[:, 0:3] * r + c.astype(np.float32)
        meshes.append(MeshDesc(f"Blob{k}", v, [(si, len(mats) - 1)]))
    insts.append(InstanceDesc("Blobs", identity4(), list(range(first, len(meshes)))))
    return SceneDesc(mats, meshes, insts, desc.camera, None, False, f"tessellated_cornell_{target_tris}")


_ICO_CACHE = {}


def _icosphere_cached(level):
    if level not in _ICO_CACHE:
        _ICO_CACHE[level] = icosphere(level)
    return _ICO_CACHE[level]


# --------------------------------------------------------------------------------------------
# Icosphere grid (spectral, GGX + dielectric, environment light)

def icosphere(subdivisions: int = 3):
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.asarray(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        cache = {}
        new_faces = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new_faces += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = new_faces
    p = np.asarray(verts, np.float64)
    v = np.zeros((len(p), 11), np.float32)
    v[:, 0:3] = p
    v[:, 3:6] = p
    up = np.tile(np.array([0.0, 1.0, 0.0]), (len(p), 1))
    tang = np.cross(up, p)
    lens = np.linalg.norm(tang, axis=1)
    pole = lens < 1e-6
    tang[pole] = np.array([1.0, 0.0, 0.0])
    lens[pole] = 1.0
    v[:, 6:9] = tang / lens[:, None]
    v[:, 9] = 0.5 + np.arctan2(p[:, 0], p[:, 2]) / (2 * math.pi)
    v[:, 10] = np.arccos(np.clip(p[:, 1], -1, 1)) / math.pi
    return v, np.asarray(faces, np.uint32).reshape(-1)


def flat_shaded(vertices: np.ndarray, indices: np.ndarray):
    """Unshares the vertices of a triangle list so that every vertex normal is its face normal."""
    tri = indices.reshape(-1, 3)
    v = vertices[tri.reshape(-1)].copy()
    p = v[:, 0:3].astype(np.float64).reshape(-1, 3, 3)
    n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    v[:, 3:6] = np.repeat(n, 3, axis=0)
    t = p[:, 1] - p[:, 0]
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    v[:, 6:9] = np.repeat(t, 3, axis=0)
    return v, np.arange(len(v), dtype=np.uint32)


def value_noise_env(width: int, height: int, seed: int, lo: float = 0.01, hi: float = 50.0) -> np.ndarray:
    """RGBA16F equirectangular map of smooth value noise mapped exponentially onto [lo, hi]."""
    rng = np.random.default_rng(seed)
    ys = (np.arange(height) + 0.5) / height
    xs = (np.arange(width) + 0.5) / width
    out = np.zeros((height, width, 3), np.float64)
    for c in range(3):
        acc = np.zeros((height, width))
        amp, tot = 1.0, 0.0
        for octave, cells in enumerate((4, 8, 16)):
            g = rng.uniform(0, 1, size=(cells + 1, 2 * cells + 1))
            g[:, -1] = g[:, 0]  # wrap in phi
            gy = ys * cells
            gx = xs * 2 * cells
            iy = np.minimum(gy.astype(int), cells - 1)
            ix = np.minimum(gx.astype(int), 2 * cells - 1)
            fy = gy - iy
            fx = gx - ix
            sy = fy * fy * (3 - 2 * fy)
            sx = fx * fx * (3 - 2 * fx)
            v00 = g[iy][:, ix]
            v01 = g[iy][:, ix + 1]
            v10 = g[iy + 1][:, ix]
            v11 = g[iy + 1][:, ix + 1]
            top = v00 * (1 - sx)[None, :] + v01 * sx[None, :]
            bot = v10 * (1 - sx)[None, :] + v11 * sx[None, :]
            acc += amp * (top * (1 - sy)[:, None] + bot * sy[:, None])
            tot += amp
            amp *= 0.5
        out[:, :, c] = acc / tot
    lum = out.mean(axis=2, keepdims=True)
    tint = 0.75 + 0.5 * (out - lum)
    val = lo * np.power(hi / lo, lum) * tint
    rgba = np.ones((height, width, 4), np.float16)
    rgba[:, :, :3] = np.clip(val, lo, hi).astype(np.float16)
    return rgba.view(np.uint16)


def icosphere_grid(seed: int = 20261202, grid: int = 32, spacing: float = 2.2, env_size=(1024, 512),
                   aspect: float = 16.0 / 9.0, dielectric_fraction: float = 0.1) -> SceneDesc:
    rng = np.random.default_rng(seed)
    sphere_v, sphere_i = icosphere(3)
    mats: List[MaterialDesc] = []
    meshes: List[MeshDesc] = []
    insts: List[InstanceDesc] = []
    half = (grid - 1) * spacing / 2
    ground_mat = len(mats)
    mats.append(MaterialDesc(MATTE, {"albedo": (REC709, 0.5, 0.5, 0.5)}))
    ext = half + spacing * 4
    gv = _quad((-ext, 0, -ext), (-ext, 0, ext), (ext, 0, ext), (ext, 0, -ext), (0, 1, 0), (1, 0, 0))
    meshes.append(MeshDesc("Ground", gv, [(quad_indices(1), ground_mat)]))
    insts.append(InstanceDesc("GroundInst", identity4(), [0]))
    for gz in range(grid):
        for gx in range(grid):
            s = float(rng.uniform(0.3, 1.0))
            if rng.uniform() < dielectric_fraction:
                mats.append(MaterialDesc(SPECULAR_SCATTERING, {
                    "coeff": (REC709, 1.0, 1.0, 1.0), "eta ext": (REC709, 1.0, 1.0, 1.0),
                    "eta int": (REC709, 1.5, 1.5, 1.5)}))
            else:
                base = rng.uniform(0.0, 1.0, size=3)
                rough = float(rng.uniform(0.05, 0.8))
                mats.append(MaterialDesc(UE4, {"base color": (REC709, *map(float, base))},
                                         {"roughness": rough, "metallic": 1.0, "occlusion": 0.0}))
            mi = len(meshes)
            meshes.append(MeshDesc(f"Sphere{gz}_{gx}", sphere_v, [(sphere_i, len(mats) - 1)]))
            xf = translate((gx * spacing - half, s, gz * spacing - half)) @ scale(s)
            insts.append(InstanceDesc(f"SphereInst{gz}_{gx}", to_colmajor(xf), [mi]))
    env_mat = len(mats)
    mats.append(MaterialDesc(ENVIRONMENT_EMITTER, {"emittance": (REC709, 1.0, 1.0, 1.0)}, {"scale": 1.0},
                             use_env_texture=True))
    env = EnvDesc(env_mat, value_noise_env(env_size[0], env_size[1], seed + 1), 0.0)
    cam_pos = (0.0, half * 0.55, half * 1.9)
    cam = CameraDesc(cam_pos, q_look_at(cam_pos, (0.0, 0.0, -half * 0.15)), aspect, math.radians(45.0))
    return SceneDesc(mats, meshes, insts, cam, env, True, "icosphere_grid")


def white_furnace(spectral: bool = False, env_size=(64, 32), flat: bool = False) -> SceneDesc:
    """A unit Lambert sphere of albedo 1 inside a constant environment of radiance 1.

    With flat shading (shading normal = geometric normal) every pixel of a converged render
    equals the environment radiance times the camera weight (pixel-averaged cos^4 times
    4 tan^2(fovy / 2) aspect): the sphere is invisible, a closed-form check of energy
    conservation.  Smooth vertex normals make the shading-normal correction lose a little energy."""
    v, i = icosphere(3)
    if flat:
        v, i = flat_shaded(v, i)
    mats = [MaterialDesc(MATTE, {"albedo": (REC709, 1.0, 1.0, 1.0)}),
            MaterialDesc(ENVIRONMENT_EMITTER, {"emittance": (REC709, 1.0, 1.0, 1.0)}, {"scale": 1.0},
                         use_env_texture=True)]
    tex = np.ones((env_size[1], env_size[0], 4), np.float16).view(np.uint16)
    cam = CameraDesc((0.0, 0.0, 4.0), q_rotate_y(math.pi), 1.0, math.radians(40.0))
    return SceneDesc(mats, [MeshDesc("Sphere", v, [(i, 0)])], [InstanceDesc("SphereInst", identity4(), [0])],
                     cam, EnvDesc(1, tex, 0.0), spectral, "white_furnace")
