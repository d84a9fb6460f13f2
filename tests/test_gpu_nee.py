"""Environment NEE (MRT_MODE_NEE) against the reference estimator (BSDF-only pt).

NEE has no counterpart in the reference's pt; it is judged by what it must preserve: the same
expectation as the BSDF-only estimator (whose frames are bit-exact to the CPU oracle), with lower
error at equal sample count. Standard errors come from the per-pixel second moment the renderer
writes in .a (mean of yuv_luminance^2, render_pt.slang:119-127).
"""
import math

import numpy as np
import pytest

from merian_amd import abi, scenes
from merian_amd.runtime import pt_params

pytestmark = pytest.mark.gpu

W, H = 96, 54


def lum(img):
    return img[..., 0] * 0.299 + img[..., 1] * 0.587 + img[..., 2] * 0.114


def rel_mse(x, r, eps=1e-2):
    return float(np.mean((x[..., :3] - r[..., :3]) ** 2 / (r[..., :3] ** 2 + eps)))


@pytest.fixture(scope="module")
def sky(ctx):
    sc = scenes.full_scene(seed=5, instances=32, env_size=(256, 128))
    ds = ctx.scene(sc)
    cam_spec = scenes.CameraSpec(sc.camera.eye, sc.camera.target, sc.camera.up, sc.camera.vfov, W, H)
    yield sc, ds, cam_spec
    ds.destroy()


def _render(ctx, ds, spec, mode, spp, frame, L=6):
    cam = ctx.camera(spec, 0)  # fixed primary hit: compare estimators of the same pixel integral
    img, _, _, st = ctx.render(ds, pt_params(spp=spp, max_path_length=L, mode=mode, record_signature=False), cam,
                               frame, W, H, signature=False)
    return img, st


def test_nee_is_unbiased_and_reduces_error(ctx, sky):
    sc, ds, spec = sky
    ref, _ = _render(ctx, ds, spec, abi.MODE_BSDF, 4096, 100)
    nee, st = _render(ctx, ds, spec, abi.MODE_NEE, 1024, 200)
    assert st.shadow_rays > 0
    # block means (6 x 6 pixel blocks) agree within their standard errors
    def blocks(img, spp):
        l = lum(img).astype(np.float64)
        var = np.maximum(img[..., 3].astype(np.float64) - l * l, 0) / spp
        b = lambda a: a.reshape(H // 6, 6, W // 6, 6).mean((1, 3))
        return b(l), np.sqrt(b(var) / 36)
    m_ref, s_ref = blocks(ref, 4096)
    m_nee, s_nee = blocks(nee, 1024)
    z = (m_nee - m_ref) / np.sqrt(s_ref ** 2 + s_nee ** 2 + 1e-12)
    assert np.abs(z).max() < 6.0, f"max |z| = {np.abs(z).max():.2f}"
    assert abs(np.mean(z)) < 0.5
    # equal spp: NEE's error against the reference is lower
    b256, _ = _render(ctx, ds, spec, abi.MODE_BSDF, 64, 300)
    n256, _ = _render(ctx, ds, spec, abi.MODE_NEE, 64, 400)
    e_b, e_n = rel_mse(b256, ref), rel_mse(n256, ref)
    assert e_n < e_b, (e_n, e_b)


def test_nee_without_environment_equals_bsdf_mode(ctx):
    # an emissive quad over a diffuse floor, no environment: NEE samples nothing and draws no
    # random numbers, so both modes produce the identical frame
    b = scenes.SceneBuilder("lamp")
    lamp = scenes.pbrt_diffuse((0.0, 0.0, 0.0))
    lamp["emission"] = (5.0, 4.0, 3.0)
    m_lamp = b.add_material(lamp)
    m_floor = b.add_material(scenes.pbrt_diffuse((0.7, 0.7, 0.7)))
    P, N, T = scenes.quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1))  # front = -y (faces down)
    b.add_mesh(P, N, T, m_lamp)
    P, N, T = scenes.quad((-4, 0, -4), (-4, 0, 4), (4, 0, 4), (4, 0, -4))
    b.add_mesh(P, N, T, m_floor)
    sc = b.build()
    spec = scenes.CameraSpec((0.0, 1.0, 5.0), (0.0, 0.5, 0.0), (0.0, 1.0, 0.0), math.radians(50), 64, 48)
    ds = ctx.scene(sc)
    cam = ctx.camera(spec, 0)
    a, _, _, _ = ctx.render(ds, pt_params(spp=8, max_path_length=5), cam, 0, 64, 48)
    n, _, _, st = ctx.render(ds, pt_params(spp=8, max_path_length=5, mode=abi.MODE_NEE), cam, 0, 64, 48)
    ds.destroy()
    assert st.shadow_rays == 0
    assert a.tobytes() == n.tobytes()
    assert a[..., :3].max() > 0
