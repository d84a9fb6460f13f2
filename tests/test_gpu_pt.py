"""GPU GBuffer + pt vs the CPU oracle on the same seeded scenes.

GBuffer hits (PackedHitInfo, fp16 barycentrics) are compared bit-exactly. pt radiance is compared
per pixel where the path signature (a hash of every path's hit-id sequence) matches. The stated
tolerance is 1e-4 relative; since device and oracle evaluate the same operation sequences
(no contraction, the same elementary-function algorithms), signature-matched pixels are in fact
required to be bit-identical, and signatures must match on (almost) every pixel.
"""
import numpy as np
import pytest

from merian_amd import abi, scenes
from merian_amd.runtime import pt_params
from oracle import oracle as O

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4


def _compare(img, sig, ref, ref_sig, min_match=0.999):
    match = sig == ref_sig
    frac = match.mean()
    assert frac >= min_match, f"only {frac:.4f} of pixels have matching path signatures"
    a, b = img[match], ref[match]
    err = np.abs(a - b) / np.maximum(np.abs(b), 1e-6)
    worst = float(err.max()) if err.size else 0.0
    assert worst <= REL_TOL, f"max relative error {worst:.3e} on signature-matched pixels"
    exact = (a.view(np.uint32) == b.view(np.uint32)).all(axis=-1).mean() if a.size else 1.0
    assert exact == 1.0, f"only {exact:.5f} of signature-matched pixels are bit-identical"
    return frac, worst


def _with_tree(ds, width, leaf, order=0):
    t = ds.tuning()
    t.bvh_width, t.bvh_max_leaf, t.bvh8_order = width, leaf, order
    ds.set_tuning(t)
    return ds


@pytest.fixture(scope="module")
def small(ctx):
    sc = scenes.plain_small(seed=0)
    ds = ctx.scene(sc)
    osc = O.OracleScene(sc)
    yield sc, ds, osc
    ds.destroy()


def test_gbuffer_bit_exact(ctx, small):
    sc, ds, osc = small
    W = H = 256
    cam = ctx.camera(sc.camera, 3)
    ocam = O.camera_look_at(sc.camera, 3)
    assert bytes(cam) == bytes(ocam)
    _, _, gb, _ = ctx.render(ds, pt_params(spp=1, max_path_length=1), cam, 3, W, H)
    ref = osc.gbuffer(ocam, W, H)
    assert gb.tobytes() == ref.tobytes(), f"{int((gb != ref).sum())} gbuffer texels differ"


@pytest.mark.parametrize("spp,L,frame", [(4, 4, 0), (2, 6, 5), (1, 1, 2), (3, 2, 1)])
def test_pt_vs_oracle(ctx, small, spp, L, frame):
    sc, ds, osc = small
    W = H = 128
    cam = ctx.camera(sc.camera, frame)
    img, sig, gb, st = ctx.render(ds, pt_params(spp=spp, max_path_length=L), cam, frame, W, H)
    ref, ref_sig, segs = osc.render(O.pt_params(spp, L), cam, frame, W, H, gb)
    _compare(img, sig, ref, ref_sig)
    assert st.primary_segments == W * H
    assert abs(int(st.continuation_segments) - segs) <= max(10, segs // 1000)


def test_pt_bvh8_matches_bvh4_bitwise(ctx):
    # the same frame through the 4-wide and the 8-wide tree: every pixel bit-identical
    sc = scenes.plain_small(seed=0)
    ds = ctx.scene(sc)
    try:
        W = H = 128
        cam = ctx.camera(sc.camera, 5)
        p = pt_params(spp=2, max_path_length=6)
        a, sa, _, _ = ctx.render(ds, p, cam, 5, W, H)
        _with_tree(ds, 8, 3)
        b, sb, _, _ = ctx.render(ds, p, cam, 5, W, H)
        assert np.array_equal(sa, sb)
        assert a.tobytes() == b.tobytes()
    finally:
        ds.destroy()


@pytest.mark.parametrize("tree", [(4, 4), (4, 2), (8, 2), (8, 4), (8, 2, 1)])
def test_pt_full_scene_crop(ctx, tree):
    # 256 displaced tori (reduced instance count), dielectrics, lat-long sky; 1080p camera.
    # Every BVH shape the autotuner may pick must give the same image.
    sc = scenes.full_scene(seed=2, instances=32, env_size=(256, 128))
    ds = _with_tree(ctx.scene(sc), *tree)
    osc = O.OracleScene(sc)
    W, H = 192, 108
    frame = 1
    cam = ctx.camera(scenes.CameraSpec(sc.camera.eye, sc.camera.target, sc.camera.up, sc.camera.vfov, W, H), frame)
    img, sig, gb, st = ctx.render(ds, pt_params(spp=4, max_path_length=8), cam, frame, W, H)
    ref, ref_sig, _ = osc.render(O.pt_params(4, 8), cam, frame, W, H, gb)
    ds.destroy()
    _compare(img, sig, ref, ref_sig, min_match=0.999)


def test_converged_image_relmse_against_16384_spp(ctx):
    # the converged image (1024 spp, independent RNG streams) against a 16384-spp render of the
    # same frame: relMSE = mean (x - r)^2 / (r^2 + 1e-2) < 1e-3
    sc = scenes.plain_small(seed=0, res=128)
    ds = ctx.scene(sc)
    try:
        cam = ctx.camera(sc.camera, 0)
        ref = np.zeros((128, 128, 4), np.float64)
        for k in range(4):  # 4 x 4096 spp, distinct frame seeds
            img, _, _, _ = ctx.render(ds, pt_params(spp=4096, max_path_length=4, record_signature=False), cam,
                                      1000 + k, 128, 128, signature=False, gbuffer=False)
            ref += img
        ref /= 4
        img, _, _, _ = ctx.render(ds, pt_params(spp=1024, max_path_length=4, record_signature=False), cam, 7,
                                  128, 128, signature=False, gbuffer=False)
        rel = float(np.mean((img[..., :3] - ref[..., :3]) ** 2 / (ref[..., :3] ** 2 + 1e-2)))
        assert rel < 1e-3, rel
        # and far fewer samples are visibly not converged (the bound is not vacuous)
        few, _, _, _ = ctx.render(ds, pt_params(spp=4, max_path_length=4, record_signature=False), cam, 9,
                                  128, 128, signature=False, gbuffer=False)
        assert float(np.mean((few[..., :3] - ref[..., :3]) ** 2 / (ref[..., :3] ** 2 + 1e-2))) > rel
    finally:
        ds.destroy()


def test_zero_spp_writes_zeros(ctx, small):
    # render_pt.slang:113-129: no sample, no division, irradiance and second moment stay 0
    sc, ds, osc = small
    cam = ctx.camera(sc.camera, 0)
    img, sig, gb, st = ctx.render(ds, pt_params(spp=0, max_path_length=4), cam, 0, 32, 24)
    assert not img.any() and not sig.any()
    assert int(st.continuation_segments) == 0
    ref, _, segs = osc.render(O.pt_params(0, 4), cam, 0, 32, 24, gb)
    assert not ref.any() and segs == 0


@pytest.mark.parametrize("W,H", [(1, 1), (1, 37), (53, 1), (7, 13)])
def test_ragged_image_sizes(ctx, small, W, H):
    sc, ds, osc = small
    cam = ctx.camera(sc.camera, 4)
    img, sig, gb, st = ctx.render(ds, pt_params(spp=3, max_path_length=5), cam, 4, W, H)
    assert img.shape == (H, W, 4)
    ref_gb = osc.gbuffer(O.camera_look_at(sc.camera, 4), W, H)
    assert gb.tobytes() == ref_gb.tobytes()
    ref, ref_sig, _ = osc.render(O.pt_params(3, 5), cam, 4, W, H, gb)
    _compare(img, sig, ref, ref_sig, min_match=1.0 if W * H < 1000 else 0.999)


def test_frame_split_into_passes(ctx, small):
    # paths in flight below spp x pixels: the frame runs in passes of whole samples (2, 2, 1
    # here) and must equal the single-pass frame bit for bit, and the oracle
    sc, ds, osc = small
    W, H = 48, 40
    cam = ctx.camera(sc.camera, 6)
    p = pt_params(spp=5, max_path_length=5)
    one, sig1, gb, st1 = ctx.render(ds, p, cam, 6, W, H)
    t = ds.tuning()
    saved = t.max_paths
    t.max_paths = 2 * W * H + 7
    ds.set_tuning(t)
    try:
        split, sig2, gb2, st2 = ctx.render(ds, p, cam, 6, W, H)
    finally:
        t = ds.tuning()
        t.max_paths = saved
        ds.set_tuning(t)
    assert gb2.tobytes() == gb.tobytes()
    assert split.tobytes() == one.tobytes() and sig2.tobytes() == sig1.tobytes()
    assert int(st2.continuation_segments) == int(st1.continuation_segments)
    ref, ref_sig, _ = osc.render(O.pt_params(5, 5), cam, 6, W, H, gb)
    _compare(split, sig2, ref, ref_sig)


@pytest.mark.parametrize("mode", [abi.MODE_BSDF, abi.MODE_NEE])
def test_wide_shading_tiles_match_narrow(ctx, small, mode):
    # passes of 2^23 paths or more shade wider tiles (three paths per thread instead of two);
    # nothing else depends on the pass size, so a frame run as one pass of 8.4 M paths equals
    # the same frame in two passes of 4.2 M bit for bit
    sc, ds, osc = small
    W = H = 1024
    cam = ctx.camera(sc.camera, 9)
    p = pt_params(spp=8, max_path_length=4, mode=mode)
    t = ds.tuning()
    saved = t.max_paths
    try:
        t.max_paths = 8 * W * H
        ds.set_tuning(t)
        wide, sig1, _, st1 = ctx.render(ds, p, cam, 9, W, H)
        t = ds.tuning()
        t.max_paths = 4 * W * H
        ds.set_tuning(t)
        narrow, sig2, _, st2 = ctx.render(ds, p, cam, 9, W, H)
    finally:
        t = ds.tuning()
        t.max_paths = saved
        ds.set_tuning(t)
    assert int(st1.continuation_segments) == int(st2.continuation_segments) > 0
    assert wide.tobytes() == narrow.tobytes() and sig1.tobytes() == sig2.tobytes()


@pytest.mark.parametrize("which", ["default", "side"])
def test_frames_on_torch_streams(ctx, small, which):
    # frames enqueued on a torch stream (the default one, handle 0 = the HIP null stream, or a
    # side stream) and summed there by torch with no synchronisation in between: the sum equals
    # that of the same frames rendered synchronously
    import torch
    sc, ds, _ = small
    W, H, n = 64, 48, 4
    cam = ctx.camera(sc.camera, 2)
    p = pt_params(spp=4, max_path_length=5, record_signature=False)
    ref = np.zeros((H, W, 4))
    for f in range(n):
        ref += ctx.render(ds, p, cam, 10 + f, W, H, signature=False)[0]
    s = torch.cuda.current_stream() if which == "default" else torch.cuda.Stream()
    with torch.cuda.stream(s):
        gb = torch.empty(W * H * 16, dtype=torch.uint8, device="cuda")
        out = torch.empty(W * H * 4, dtype=torch.float32, device="cuda")
        acc = torch.zeros(W * H * 4, dtype=torch.float64, device="cuda")
        ctx.gbuffer_dev(ds, cam, W, H, gb.data_ptr(), stream=s.cuda_stream)
        for f in range(n):
            ctx.pt_render_dev(ds, p, cam, 10 + f, W, H, gb.data_ptr(), out.data_ptr(), stream=s.cuda_stream)
            acc += out
            out.fill_(float("nan"))  # a frame that read stale or half-written output shows up
        got = acc.cpu().numpy().reshape(H, W, 4)
    assert got.tobytes() == ref.tobytes()


def test_zero_sized_image_is_a_no_op(ctx, small):
    sc, ds, _ = small
    cam = ctx.camera(sc.camera, 0)
    for W, H in ((0, 16), (16, 0)):
        img, sig, gb, st = ctx.render(ds, pt_params(spp=2, max_path_length=3), cam, 0, W, H)
        assert img.size == 0 and int(st.continuation_segments) == 0
