"""The 1080p workload at its full size against the CPU oracle.

Scene: full_scene(seed=2), 256 displaced tori (1,999,896 triangles) in an enclosing room, 258
materials (Lambert, GGX conductors, rough dielectrics), a 2048x1024 lat-long sky with a sun lobe.
Frame: 1920x1080, 64 spp, max path length 8 (render_pt.slang:92-130), with the cached autotuned
settings (every sample of the frame in one pass of 2^27+ paths) and again in passes of at most
2^24 paths.

Checks:
- the whole GBuffer bit-exact against the oracle's;
- pt radiance bit-identical on signature-matched pixels of six seeded 32x18 tiles, rendered by the
  oracle over the same gbuffer (render_pt.slang:92-130 per pixel);
- over the whole frame: every pixel finite, the second moment at least the squared luminance of
  the mean (mean of squares >= square of the mean), primary segments = pixels and continuation
  segments <= pixels * spp * (L - 1);
- the single-pass and the multi-pass frame bit-identical (irradiance, signatures, segment counts).
"""
import json
import os

import numpy as np
import pytest

from merian_amd import scenes
from merian_amd.runtime import pt_params
from oracle import oracle as O

from test_gpu_pt import _compare

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TUNING = os.path.join(ROOT, "merian_amd", "tuning", "full_1080p.json")
W, H, SPP, L, FRAME = 1920, 1080, 64, 8, 1


def _set_tuning(ds, **kv):
    t = ds.tuning()
    for k, v in kv.items():
        setattr(t, k, v)
    ds.set_tuning(t)


@pytest.fixture(scope="module")
def full(ctx):
    sc = scenes.full_scene(seed=2)
    assert sc.triangle_count == 1999896 and sc.env_rgba.shape[:2] == (1024, 2048)
    ds = ctx.scene(sc)
    with open(TUNING) as f:
        _set_tuning(ds, **json.load(f)["tuning"])
    yield sc, ds
    ds.destroy()


def test_full_1080p_frame(ctx, full):
    sc, ds = full
    assert ds.tuning().max_paths >= W * H * SPP  # the cached tuning runs the frame as one pass
    cam = ctx.camera(sc.camera, FRAME)
    print("full 1080p: single-pass frame", flush=True)
    img, sig, gb, st = ctx.render(ds, pt_params(spp=SPP, max_path_length=L), cam, FRAME, W, H)
    print(f"  {int(st.continuation_segments)} continuation segments", flush=True)

    # whole-frame properties
    assert np.isfinite(img).all()
    lum = 0.299 * img[..., 0].astype(np.float64) + 0.587 * img[..., 1] + 0.114 * img[..., 2]
    assert np.all(img[..., 3] >= lum * lum * (1 - 1e-4) - 1e-12)
    assert int(st.primary_segments) == W * H
    cont = int(st.continuation_segments)
    assert W * H * SPP < cont <= W * H * SPP * (L - 1)

    # the same frame in passes of at most 2^24 paths
    saved = ds.tuning().max_paths
    _set_tuning(ds, max_paths=1 << 24)
    try:
        print("full 1080p: multi-pass frame", flush=True)
        img2, sig2, gb2, st2 = ctx.render(ds, pt_params(spp=SPP, max_path_length=L), cam, FRAME, W, H)
    finally:
        _set_tuning(ds, max_paths=saved)
    assert gb2.tobytes() == gb.tobytes()
    assert img2.tobytes() == img.tobytes() and sig2.tobytes() == sig.tobytes()
    assert int(st2.continuation_segments) == cont

    # the oracle: the whole GBuffer, then six seeded tiles of pt
    print("full 1080p: oracle", flush=True)
    osc = O.OracleScene(sc)
    ocam = O.camera_look_at(sc.camera, FRAME)
    assert bytes(ocam) == bytes(cam)
    ref_gb = osc.gbuffer(ocam, W, H)
    assert gb.tobytes() == ref_gb.tobytes(), f"{int((gb != ref_gb).sum())} gbuffer texels differ"
    rng = np.random.default_rng(2024)
    gbf = gb.reshape(-1)
    for _ in range(6):
        x0, y0 = int(rng.integers(0, W - 32)), int(rng.integers(0, H - 18))
        ref, ref_sig, _ = osc.render(O.pt_params(SPP, L), cam, FRAME, W, H, gbf, rect=(x0, y0, 32, 18))
        _compare(img[y0:y0 + 18, x0:x0 + 32], sig[y0:y0 + 18, x0:x0 + 32], ref, ref_sig, min_match=0.99)
