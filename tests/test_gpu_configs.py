"""The serving mix and learned guiding at the sizes their benchmarks run.

Serving (mrt_server_*): the eight resident scenes of the serving mix (scenes.serving_scenes(8),
103,682 to 1,999,896 triangles) with jobs of every resolution and spp class of the mix,
including a 1920x1080 64-spp job on the 2M-triangle scene and an NEE job. Each job's image
rendered alone (a one-slot server) equals the same job rendered in the mix (eight slots, the
jobs submitted at once and admitted as slots and budget free up) and, for a subset, the
one-frame entry point mrt_render_frame, bit for bit.

Learned guiding (mrt_guided_render_dev): the 1080p workload's scene and view (full_scene(seed=2):
1,999,896 triangles in the enclosing room, a 2048x1024 sky, 1920x1080, L = 8) with the cached
settings, the sky's sun at 100x its dimmest texel instead of 5e4x (with the strong sun the
guided estimator is heavy-tailed, and a block z-test reads its skew as bias). After online
training, frames with the guide frozen agree with the NEE estimator (unbiased, tested against
the BSDF-only estimator in test_gpu_nee.py) block by block within their standard errors, from
the per-pixel second moment the renderer writes (render_pt.slang:119-127).
"""
import json
import os

import numpy as np
import pytest

from merian_amd import abi, scenes
from merian_amd.runtime import Guide, Server, guide_config, pt_params

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def mix(ctx):
    resident = scenes.serving_scenes(8)
    assert [s.triangle_count for s in resident][0] == 103682 and resident[-1].triangle_count == 1999896
    dss = [ctx.scene(s) for s in resident]
    yield resident, dss
    for d in dss:
        d.destroy()


# (scene, width, height, spp, max path length, frame, mode): every resolution and spp class of the
# mix; the 1080p 64-spp job on the 2M-triangle scene runs in passes (2^24 paths per job)
MIX_JOBS = [
    (7, 1920, 1080, 64, 8, 5, abi.MODE_BSDF),
    (0, 256, 256, 1, 4, 1, abi.MODE_BSDF),
    (3, 1024, 768, 16, 6, 2, abi.MODE_BSDF),
    (5, 512, 512, 4, 8, 3, abi.MODE_NEE),
    (1, 256, 256, 64, 5, 4, abi.MODE_BSDF),
    (4, 512, 512, 16, 7, 6, abi.MODE_BSDF),
    (6, 1024, 768, 1, 8, 7, abi.MODE_BSDF),
    (2, 1920, 1080, 4, 4, 8, abi.MODE_BSDF),
]


def _args(ctx, resident, dss, j):
    k, w, h, spp, L, frame, mode = j
    c = resident[k].camera
    cam = ctx.camera(scenes.CameraSpec(c.eye, c.target, c.up, c.vfov, w, h), frame)
    return dss[k], pt_params(spp=spp, max_path_length=L, mode=mode, record_signature=False), cam, frame, w, h


def _serve(ctx, resident, dss, jobs, **kw):
    with Server(ctx, **kw) as srv:
        subs = [srv.submit(*_args(ctx, resident, dss, j)) for j in jobs]
        sts = [s.wait(120_000) for s in subs]
        stats = srv.stats()
    for j, st in zip(jobs, sts):
        assert st.state == abi.JOB_DONE, (j, abi.JOB_STATE_NAMES[st.state], st.error)
    return [s.irradiance for s in subs], sts, stats


def test_serving_mix_alone_equals_in_mix(ctx, mix):
    resident, dss = mix
    alone = []
    for j in MIX_JOBS:
        print(f"serving mix: job {j} alone", flush=True)
        imgs, _, _ = _serve(ctx, resident, dss, [j], max_slots=1)
        alone.append(imgs[0])
    print("serving mix: all jobs in eight slots", flush=True)
    mixed, sts, stats = _serve(ctx, resident, dss, MIX_JOBS, max_slots=8)
    for j, a, b in zip(MIX_JOBS, alone, mixed):
        assert a.tobytes() == b.tobytes(), j
        assert np.isfinite(b).all(), j
    assert stats.completed == len(MIX_JOBS) and stats.peak_in_flight > 1
    assert stats.segments == sum(s.segments for s in sts)
    for j, st in zip(MIX_JOBS, sts):
        assert st.segments >= j[1] * j[2]
    # the one-frame entry point on a few of them (the 1080p job included)
    for k in (0, 1, 3):
        img, _, _, _ = ctx.render(*_args(ctx, resident, dss, MIX_JOBS[k]), signature=False, gbuffer=False)
        assert img.tobytes() == mixed[k].tobytes(), MIX_JOBS[k]


def _frames(ctx, ds, cam, params, n, base, w, h, guide=None, train=False):
    import torch
    strm = torch.cuda.current_stream().cuda_stream
    gb = torch.empty(w * h * 16, dtype=torch.uint8, device="cuda")
    out = torch.empty(w * h * 4, dtype=torch.float32, device="cuda")
    acc = torch.zeros(w * h * 4, dtype=torch.float64, device="cuda")
    ctx.gbuffer_dev(ds, cam, w, h, gb.data_ptr(), stream=strm)
    for f in range(n):
        if guide is None:
            ctx.pt_render_dev(ds, params, cam, base + f, w, h, gb.data_ptr(), out.data_ptr(), stream=strm)
        else:
            guide.render_dev(ds, params, cam, base + f, w, h, gb.data_ptr(), out.data_ptr(), train=train, stream=strm)
        acc += out
    return (acc / n).cpu().numpy().reshape(h, w, 4)


def test_guided_unbiased_on_the_1080p_view(ctx):
    sc = scenes.full_scene(seed=2, dynamic_range=100.0)
    assert sc.triangle_count == 1999896
    W, H, L = 1920, 1080, 8
    ds = ctx.scene(sc)
    try:
        with open(os.path.join(ROOT, "merian_amd", "tuning", "full_1080p.json")) as f:
            t = ds.tuning()
            for k, v in json.load(f)["tuning"].items():
                setattr(t, k, v)
            ds.set_tuning(t)
        cam = ctx.camera(sc.camera, 0)
        print("guided 1080p: NEE reference", flush=True)
        ref = _frames(ctx, ds, cam, pt_params(spp=64, max_path_length=L, mode=abi.MODE_NEE, record_signature=False),
                      4, 50000, W, H)
        g = Guide(ctx, ds, guide_config(seed=7, train_steps=16))
        try:
            print("guided 1080p: training", flush=True)
            _frames(ctx, ds, cam, pt_params(spp=4, max_path_length=L, mode=abi.MODE_GUIDED, record_signature=False),
                    6, 2000, W, H, guide=g, train=True)
            print("guided 1080p: frozen frames", flush=True)
            gimg = _frames(ctx, ds, cam, pt_params(spp=64, max_path_length=L, mode=abi.MODE_GUIDED,
                                                   record_signature=False), 4, 3000, W, H, guide=g)
        finally:
            g.close()
    finally:
        ds.destroy()
    assert np.isfinite(gimg).all() and np.isfinite(ref).all()

    def blocks(img, n):
        l = img[..., 0] * 0.299 + img[..., 1] * 0.587 + img[..., 2] * 0.114
        var = np.maximum(img[..., 3] - l * l, 0) / n
        b = lambda a: a.reshape(H // 40, 40, W // 40, 40).mean((1, 3))
        return b(l), np.sqrt(b(var) / 1600)
    m_r, s_r = blocks(ref, 256)
    m_g, s_g = blocks(gimg, 256)
    z = (m_g - m_r) / np.sqrt(s_r ** 2 + s_g ** 2 + 1e-12)
    print(f"guided 1080p vs NEE: max |z| {np.abs(z).max():.2f}, mean z {z.mean():.3f}, "
          f"image means {m_g.mean():.5f} / {m_r.mean():.5f}")
    assert np.abs(z).max() < 6.0, (np.abs(z).max(), z.mean())
    assert abs(z.mean()) < 0.3, z.mean()
