"""Denoiser guides (first-hit albedo and normal, path_tracing.cu:95-108, post_process.cu:70-90)
of the GPU path tracer against the oracle's.

Guides draw no random numbers, so pixels whose PCG32 states agree after the render saw the
same first hits in every frame: there the GPU's guides equal the oracle's up to float
rounding.  The guides leave the image unchanged, and restart with the accumulation."""
import numpy as np
import pytest

from oracle import binding as ob
from vlr_amd import api, scenes

pytestmark = pytest.mark.gpu


def render(desc, W, H, frames, depth, spectral, guides=True):
    with api.Context(spectral=spectral, max_path_length=depth) as ctx:
        bs = api.build_scene(ctx, desc)
        ctx.bind_output(W, H)
        ctx.set_denoiser_aux_buffers(guides)
        for f in range(frames):
            ctx.render(bs.camera, f == 0, frames)
        img = ctx.read_output()
        alb, nrm = ctx.read_denoiser_aux_buffers() if guides else (None, None)
        rng = ctx.read_rng()
    return img, alb, nrm, rng


@pytest.mark.parametrize("scene,spectral", [("cornell", False), ("cornell", True), ("grid", True)])
def test_guides_match_the_oracle(torch_gpu, scene, spectral):
    desc = scenes.cornell_box() if scene == "cornell" else scenes.icosphere_grid(grid=3, env_size=(64, 32))
    W, H, frames, depth = 64, 48, 6, 8
    img, alb, nrm, rng = render(desc, W, H, frames, depth, spectral)
    _, oalb, onrm, orng, _ = ob.Scene(desc, spectral=spectral).render_guides(W, H, frames, depth,
                                                                             rng_state=ob.seed_states(W, H))
    eq = (rng == orng).reshape(H, W)
    print("identical RNG states:", eq.mean())
    assert eq.mean() > 0.95
    a, b = alb[eq][:, :3].astype(np.float64), oalb[eq][:, :3].astype(np.float64)
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-3)
    print("albedo relative difference: max", rel.max(), "99.9%", np.quantile(rel, 0.999))
    assert np.quantile(rel, 0.999) < 1e-5 and rel.max() < 1e-3
    dn = np.abs(nrm[eq][:, :3] - onrm[eq][:, :3])
    print("normal difference: max", dn.max())
    assert dn.max() < 1e-5
    assert np.all(alb[..., 3] == 1.0) and np.all(nrm[..., 3] == 1.0)
    # guides are not all trivial: misses (zero) and hits (unit normals) both occur in the grid
    unit = np.abs(np.linalg.norm(nrm[..., :3], axis=-1) - 1.0) < 1e-5
    assert unit.mean() > 0.5


def test_guides_leave_the_image_unchanged_and_restart(torch_gpu):
    desc = scenes.cornell_box()
    W, H, depth = 48, 32, 8
    plain, _, _, _ = render(desc, W, H, 4, depth, False, guides=False)
    with_guides, alb4, nrm4, _ = render(desc, W, H, 4, depth, False)
    assert np.array_equal(plain.view(np.uint32), with_guides.view(np.uint32))
    # a first frame restarts the guides: 3 frames, then 2 from firstFrame, equal the guides of
    # the last 2 frames of an uninterrupted 5-frame render's generator sequence
    with api.Context(spectral=False, max_path_length=depth) as ctx:
        bs = api.build_scene(ctx, desc)
        ctx.bind_output(W, H)
        ctx.set_denoiser_aux_buffers(True)
        for f in range(3):
            ctx.render(bs.camera, f == 0, 10)
        seeds = ctx.read_rng()
        for f in range(2):
            ctx.render(bs.camera, f == 0, 10)
        alb_restart, nrm_restart = ctx.read_denoiser_aux_buffers()
    with api.Context(spectral=False, max_path_length=depth) as ctx:
        bs = api.build_scene(ctx, desc)
        ctx.bind_output(W, H)
        ctx.write_rng(seeds)
        ctx.set_denoiser_aux_buffers(True)
        for f in range(2):
            ctx.render(bs.camera, f == 0, 10)
        alb_fresh, nrm_fresh = ctx.read_denoiser_aux_buffers()
    assert np.array_equal(alb_restart, alb_fresh) and np.array_equal(nrm_restart, nrm_fresh)


def render_guides(desc, W, H, frames, spectral, renderer):
    with api.Context(spectral=spectral, max_path_length=8) as ctx:
        ctx.set_renderer(renderer)
        bs = api.build_scene(ctx, desc)
        ctx.bind_output(W, H)
        ctx.set_denoiser_aux_buffers(True)
        for f in range(frames):
            ctx.render(bs.camera, f == 0, frames)
        alb, nrm = ctx.read_denoiser_aux_buffers()
        return alb, nrm, ctx.read_rng()


@pytest.mark.parametrize("spectral,renderer", [(False, api.RENDERER_LIGHT_TRACING), (True, api.RENDERER_LIGHT_TRACING),
                                               (False, api.RENDERER_BPT)])
def test_light_tracing_and_bpt_guides_equal_the_path_tracers(torch_gpu, spectral, renderer):
    """Light tracing's guide pass (aux_buffer_generator.cu) draws each pixel's primary ray as the
    path tracer does (jitter, wavelengths, lens point), and BPT's eye paths draw jitter then lens
    point (the pinhole lens point does not move the ray; RGB albedos do not depend on the
    wavelengths): after one frame their guides equal the path tracer's, whose are pinned by the
    oracle.  Over several frames they stay finite, with unit or zero normals."""
    desc = scenes.cornell_box()
    W, H = 48, 32
    pa, pn, _ = render_guides(desc, W, H, 1, spectral, api.RENDERER_PATH_TRACING)
    la, ln, lrng = render_guides(desc, W, H, 1, spectral, renderer)
    np.testing.assert_allclose(la, pa, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(ln, pn, rtol=1e-6, atol=1e-7)
    if renderer == api.RENDERER_LIGHT_TRACING:
        assert not np.array_equal(lrng, ob.seed_states(W, H))     # the guide pass drew from the pixels
    a4, n4, _ = render_guides(desc, W, H, 4, spectral, renderer)
    assert np.all(np.isfinite(a4)) and np.all(np.isfinite(n4))
    norms = np.linalg.norm(n4[..., :3], axis=-1)
    assert np.all((np.abs(norms - 1.0) < 1e-4) | (norms == 0.0))
