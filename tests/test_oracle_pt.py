"""End-to-end known answers of the oracle's pt (render_pt.slang:24-130), CPU only.

A convex, flat-shaded Lambert object (DiffuseMaterial, albedo a) under a constant environment L:
every continuation ray leaves the object and escapes, and the Lambert sample weight is exactly
a, so each object pixel is a * L (max path length >= 2), 0 (max path length 1) and background
pixels are L (or 0 without "emission on primary"). The second moment is the mean of
yuv_luminance(contribution)^2. An emissive OpenPBR quad emits only from its front side
(openpbr-material.slang:93-104). A lat-long map varying only with v pins the
v = acos(y) / pi mapping and the bilinear filter (utils/math.slang:243-249).
"""
import math

import numpy as np

from merian_amd import abi, scenes
from oracle import oracle as O


def convex_scene(albedo=(0.6, 0.4, 0.2), L=(2.0, 1.5, 0.5)):
    b = scenes.SceneBuilder("convex")
    P, T = scenes.icosphere(2)
    N = scenes.vertex_normals(P, T)
    m = b.add_material(scenes.diffuse_material(albedo))
    b.add_mesh(P * 1.0, N, T, m, flags=abi.GEOM_FLAT_SHADING)
    sc = b.build(**scenes.constant_env(L))
    sc.camera = scenes.CameraSpec((0.0, 0.0, 4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.radians(50), 64, 64,
                                  jitter=abi.JITTER_NONE)
    return sc


def render(sc, spp, L, emission_on_primary=True, frame=0):
    osc = O.OracleScene(sc)
    cam = O.camera_look_at(sc.camera, frame)
    W, H = sc.camera.width, sc.camera.height
    gb = osc.gbuffer(cam, W, H)
    img, sig, segs = osc.render(O.pt_params(spp, L, emission_on_primary), cam, frame, W, H, gb)
    return img, gb.reshape(H, W), segs


def test_convex_lambert_under_constant_env():
    albedo, Lenv = (0.6, 0.4, 0.2), (2.0, 1.5, 0.5)
    sc = convex_scene(albedo, Lenv)
    aL = (np.float32(albedo) * np.float32(Lenv)).astype(np.float32)
    for spp, L in [(1, 2), (2, 2), (2, 5)]:
        img, gb, segs = render(sc, spp, L)
        obj = gb["instance_id"] != 0xFFFFFFFF
        assert 0.2 < obj.mean() < 0.8
        assert np.array_equal(img[~obj][:, :3], np.tile(np.float32(Lenv), ((~obj).sum(), 1)))
        exact = (img[obj][:, :3] == aL).all(axis=1)
        # silhouette pixels whose view direction is within 1e-6 of the shading plane are rejected
        assert exact.mean() > 0.999, exact.mean()
        # each object pixel traced spp continuation rays; each escaped
        assert segs == obj.sum() * spp
        lum = np.float32(0.299) * aL[0] + np.float32(0.587) * aL[1] + np.float32(0.114) * aL[2]
        assert np.allclose(img[obj][exact][:, 3], lum * lum, rtol=1e-6)


def test_path_length_one_and_emission_on_primary():
    sc = convex_scene()
    img, gb, segs = render(sc, 2, 1)
    obj = gb["instance_id"] != 0xFFFFFFFF
    assert segs == 0
    assert (img[obj][:, :3] == 0).all()
    img, gb, _ = render(sc, 2, 3, emission_on_primary=False)
    assert (img[~obj][:, :3] == 0).all()


def test_emission_front_face_only():
    b = scenes.SceneBuilder("emitter")
    m = scenes.pbrt_diffuse((0.0, 0.0, 0.0))
    m["emission"] = (3.0, 2.0, 1.0)
    mid = b.add_material(m)
    P, N, T = scenes.quad((-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0))  # front = +z
    b.add_mesh(P, N, T, mid, flags=abi.GEOM_TWO_SIDED)
    sc = b.build(**scenes.constant_env((0.0, 0.0, 0.0)))
    for z, expect in ((3.0, (3.0, 2.0, 1.0)), (-3.0, (0.0, 0.0, 0.0))):
        sc.camera = scenes.CameraSpec((0.0, 0.0, z), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.radians(20), 16, 16,
                                      jitter=abi.JITTER_NONE)
        img, gb, _ = render(sc, 1, 1)
        assert (gb["instance_id"] != 0xFFFFFFFF).all()
        assert np.array_equal(img[..., :3].reshape(-1, 3), np.tile(np.float32(expect), (256, 1)))


def test_latlong_rows():
    # env radiance = row index (constant along u): the background pixel for direction d is the
    # bilinear filter at v = acos(d.y) / pi, i.e. row v * H - 0.5
    H, W = 64, 128
    rgba = np.zeros((H, W, 4), np.float32)
    rgba[..., :3] = np.arange(H, dtype=np.float32)[:, None, None]
    b = scenes.SceneBuilder("sky")
    mid = b.add_material(scenes.diffuse_material((0.5, 0.5, 0.5)))
    P, N, T = scenes.quad((100, 100, 100), (100.1, 100, 100), (100.1, 100.1, 100), (100, 100.1, 100))
    b.add_mesh(P, N, T, mid)
    sc = b.build(env_type=abi.ENV_LATLONG, env_wrap=abi.WRAP_REPEAT, env_rgba=rgba)
    sc.camera = scenes.CameraSpec((0.0, 0.0, 0.0), (0.0, 0.3, -1.0), (0.0, 1.0, 0.0), math.radians(90), 32, 32,
                                  jitter=abi.JITTER_NONE)
    img, gb, _ = render(sc, 1, 1)
    rays = O.camera_rays(O.camera_look_at(sc.camera, 0), 32, 32)
    y = rays["direction"][:, 1].astype(np.float64)
    row = np.arccos(y) / np.pi * H - 0.5
    assert np.allclose(img[..., 0].reshape(-1), np.clip(row, 0, H - 1), atol=2e-4)
