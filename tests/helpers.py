"""Shared helpers for the parity tests: ray sets and device round trips."""
import numpy as np

from vlr_amd import api


def random_rays_in_box(n, lo, hi, seed, tmax=np.float32(3.4028235e38)):
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, api.RAY_DTYPE)
    rays["org"] = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["dir"] = d.astype(np.float32)
    rays["tmin"] = 0.0
    rays["tmax"] = tmax
    return rays


def rays_towards(n, origin_lo, origin_hi, target_lo, target_hi, seed):
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, api.RAY_DTYPE)
    o = rng.uniform(origin_lo, origin_hi, size=(n, 3))
    t = rng.uniform(target_lo, target_hi, size=(n, 3))
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["org"] = o.astype(np.float32)
    rays["dir"] = d.astype(np.float32)
    rays["tmax"] = np.float32(3.4028235e38)
    return rays


def gpu_trace(torch, ctx, scene, rays, any_hit=False):
    dev = torch.device("cuda")
    r = torch.from_numpy(rays.view(np.uint8).copy()).to(dev)
    h = torch.zeros(len(rays) * api.HIT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    ctx.trace_rays(scene, r.data_ptr(), len(rays), h.data_ptr(), any_hit)
    torch.cuda.synchronize()
    return h.cpu().numpy().view(api.HIT_DTYPE)


def gpu_bsdf(torch, ctx, material, spectral, queries):
    dev = torch.device("cuda")
    q = torch.from_numpy(np.ascontiguousarray(queries, api.BSDF_QUERY_DTYPE).view(np.uint8).copy()).to(dev)
    r = torch.zeros(len(queries) * api.BSDF_RESULT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    ctx.evaluate_bsdfs(material, spectral, q.data_ptr(), len(queries), r.data_ptr())
    torch.cuda.synchronize()
    return r.cpu().numpy().view(api.BSDF_RESULT_DTYPE)

