"""CPU pins of the oracle's building blocks (no GPU).

RNG: the PCG step (random.slang:108-128) is the published pcg32 RXS-M-XS 32/32 generator and
pcg4d16 (hash.slang:138-146) the first output of Jarzynski & Olano's pcg4d; both are checked
against an arbitrary-precision Python evaluation (catches missing 2^32 wrap-around or shift
slips) and for uniformity. Elementary functions: accuracy against float64 numpy. fp16 packing:
against numpy's float16 conversion (an independent round-to-nearest-even implementation).
Oct encoding and the ray-origin offset: closed-form properties. Camera: closed-form basis.
"""
import math

import numpy as np
import pytest

from merian_amd import scenes
from oracle import oracle as O

M32 = (1 << 32) - 1


def py_pcg4d16(v):
    v = [(x * 1664525 + 1013904223) & M32 for x in v]
    v[0] = (v[0] + v[1] * v[3]) & M32
    v[1] = (v[1] + v[2] * v[0]) & M32
    v[2] = (v[2] + v[0] * v[1]) & M32
    v[3] = (v[3] + v[1] * v[2]) & M32
    v[0] = (v[0] + v[1] * v[3]) & M32
    return v[0]


def py_pcg_stream(state, n):
    out = []
    for _ in range(n):
        s = state
        state = (state * 747796405 + 2891336453) & M32
        word = (((s >> ((s >> 28) + 4)) ^ s) * 277803737) & M32
        out.append((word >> 22) ^ word)
    return out


def test_pcg4d16_matches_bigint_evaluation():
    rng = np.random.default_rng(0)
    for v in rng.integers(0, 1 << 32, size=(300, 4), dtype=np.uint64):
        v = [int(x) for x in v]
        assert O.pcg4d16(*v) == py_pcg4d16(v)
    # pixel/frame/sample style inputs
    for x, y, f, s in [(0, 0, 0, 0), (511, 511, 15, 3), (1919, 1079, 7, 63)]:
        assert O.pcg4d16(x, y, f, s) == py_pcg4d16([x, y, f, s])


def test_pcg_stream_and_unit_floats():
    for seed in (0, 1, 0xDEADBEEF, O.pcg4d16(3, 4, 5, 6)):
        u, f = O.pcg_stream(seed, 64)
        assert u.tolist() == py_pcg_stream(seed, 64)
        assert np.array_equal(f, ((u >> 8).astype(np.float64) * 2.0 ** -24).astype(np.float32))
    u, f = O.pcg_stream(12345, 1 << 20)
    assert f.min() >= 0.0 and f.max() < 1.0
    hist = np.histogram(f, bins=64, range=(0, 1))[0]
    chi2 = ((hist - hist.mean()) ** 2 / hist.mean()).sum()
    assert chi2 < 120  # 63 dof, p ~ 1e-5
    # the largest input maps to 1 - 2^-24 exactly
    assert np.float32((0xFFFFFFFF >> 8) * 2.0 ** -24) == np.float32(1 - 2.0 ** -24)


def test_neighbouring_pixel_streams_are_decorrelated():
    firsts = np.array([py_pcg_stream(O.pcg4d16(x, 7, 0, 0), 1)[0] for x in range(4096)], dtype=np.uint64)
    f = (firsts >> 8).astype(np.float64) * 2.0 ** -24
    r = np.corrcoef(f[:-1], f[1:])[0, 1]
    assert abs(r) < 0.05


def _ulps(a, ref):
    ref32 = ref.astype(np.float32)
    return np.abs(a.astype(np.float64) - ref) / np.maximum(np.abs(np.spacing(np.abs(ref32))), 1e-45)


@pytest.mark.parametrize("op,fn,lo,hi,bound", [
    ("sin", np.sin, -2 * np.pi, 2 * np.pi, 2.0),
    ("cos", np.cos, -2 * np.pi, 2 * np.pi, 2.0),
    ("atan", np.arctan, -60.0, 60.0, 3.0),
    ("acos", np.arccos, -1.0, 1.0, 3.0),
    ("exp2", np.exp2, -140.0, 127.0, 1.5),
    ("exp", np.exp, -80.0, 80.0, 1.5),
])
def test_elementary_function_accuracy(op, fn, lo, hi, bound):
    x = np.random.default_rng(1).uniform(lo, hi, 400_000).astype(np.float32)
    err = _ulps(O.math(op, x), fn(x.astype(np.float64)))
    assert err.max() <= bound, f"{op}: {err.max():.2f} ulp"


def test_atan2_log2_pow_accuracy_and_special_cases():
    rng = np.random.default_rng(2)
    a, b = rng.normal(size=(2, 400_000)).astype(np.float32)
    assert _ulps(O.math("atan2", a, b), np.arctan2(a.astype(np.float64), b.astype(np.float64))).max() <= 3.0
    p = np.exp(rng.uniform(-80, 80, 400_000)).astype(np.float32)
    assert _ulps(O.math("log2", p), np.log2(p.astype(np.float64))).max() <= 3.0
    x = rng.uniform(0, 1, 400_000).astype(np.float32)
    y = rng.uniform(0, 10, 400_000).astype(np.float32)
    ref = np.power(x.astype(np.float64), y.astype(np.float64))
    rel = np.abs(O.math("pow", x, y) - ref) / np.maximum(ref, 1e-30)
    assert np.median(rel) < 1e-6 and rel.max() < 2e-5
    # quadrants and signed zeros of atan2(y, x), C99 semantics
    ys = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 1.0, -1.0], np.float32)
    xs = np.array([0.0, 0.0, -0.0, -0.0, 0.0, 0.0, -1.0, -1.0], np.float32)
    got = O.math("atan2", ys, xs)
    want = np.arctan2(ys.astype(np.float64), xs.astype(np.float64)).astype(np.float32)
    assert np.allclose(got, want, rtol=3e-7, atol=0) and np.array_equal(np.signbit(got), np.signbit(want))
    assert O.math("acos", np.array([1.0, -1.0], np.float32)).tolist() == [0.0, np.float32(np.pi)]
    assert np.isnan(O.math("acos", np.array([1.0000001], np.float32))[0])
    assert O.math("exp", np.array([100.0, -200.0], np.float32)).tolist() == [np.inf, 0.0]


def test_half_conversion_matches_numpy():
    rng = np.random.default_rng(3)
    vals = np.concatenate([rng.uniform(-1, 1, 20000), rng.uniform(0, 1, 2000) * 6e-5,
                           np.exp(rng.uniform(-20, 11.2, 5000)), [65504, 65519.99, 65520, 1e6, 2 ** -25,
                                                                  2 ** -24 * 1.5, 0.0, -0.0]]).astype(np.float32)
    ours = np.array([O.lib().orc_f32_to_f16(float(v)) for v in vals], np.uint16)
    with np.errstate(over="ignore"):
        want = vals.astype(np.float16).view(np.uint16)
    assert np.array_equal(ours, want)
    back = np.array([O.lib().orc_f16_to_f32(int(h)) for h in ours[:5000]], np.float32)
    assert np.array_equal(back, ours[:5000].view(np.float16).astype(np.float32))


def test_oct_normal_encoding():
    rng = np.random.default_rng(4)
    n = rng.normal(size=(3000, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    n = n.astype(np.float32)
    err = []
    for v in n:
        d = O.decode_normal(O.encode_normal(v))
        assert abs(np.linalg.norm(d.astype(np.float64)) - 1) < 1e-6
        a, b = d.astype(np.float64), v.astype(np.float64)
        err.append(math.atan2(np.linalg.norm(np.cross(a, b)), np.dot(a, b)))
    assert max(err) < 1e-4  # 2 x 16-bit octahedral encoding
    for axis in np.eye(3, dtype=np.float32):
        for s in (1, -1):
            assert np.allclose(O.decode_normal(O.encode_normal(s * axis)), s * axis, atol=1e-7)
    # the product's scene generator encodes normals in numpy with the same operations
    enc = scenes.encode_normals(n[:500])
    assert enc.tolist() == [O.encode_normal(v) for v in n[:500]]


def test_offset_ray_origin_closed_form():
    rng = np.random.default_rng(5)
    for _ in range(500):
        p = rng.uniform(-50, 50, 3).astype(np.float32)
        nrm = rng.normal(size=3)
        nrm = (nrm / np.linalg.norm(nrm)).astype(np.float32)
        o = O.offset_ray_origin(p, nrm)
        for k in range(3):
            if abs(p[k]) < 1.0 / 16.0:
                assert o[k] == np.float32(p[k] + nrm[k] * np.float32(1 / 65536))
            else:
                io = int(np.float32(nrm[k] * np.float32(256)))  # truncation toward zero
                step = p.view(np.int32)[k] + (-io if p[k] < 0 else io)
                assert o.view(np.int32)[k] == step
        # the origin moves to the side the normal points to
        assert np.dot(o.astype(np.float64) - p, nrm) >= 0


def test_camera_basis_closed_form():
    spec = scenes.plain_small(seed=0).camera
    for frame in (0, 1, 7, 16):
        cam = O.camera_look_at(spec, frame)
        U, V, W = (np.array(getattr(cam, k), np.float64) for k in ("U", "V", "W"))
        eye, tgt = np.array(spec.eye), np.array(spec.target)
        focus = np.linalg.norm(tgt - eye)
        assert abs(np.linalg.norm(W) - focus) < 1e-5 * focus
        assert abs(np.dot(U, W)) < 1e-4 and abs(np.dot(V, W)) < 1e-4 and abs(np.dot(U, V)) < 1e-4
        assert abs(np.linalg.norm(V) / focus - math.tan(spec.vfov / 2)) < 1e-6
        assert abs(np.linalg.norm(U) / np.linalg.norm(V) - spec.aspect) < 1e-6
        assert np.dot(np.cross(U, V), W) < 0  # right-handed: x right, y up, looking down W
        j = np.array(cam.jitter)
        n = frame % 16
        assert np.allclose(j, [math.fmod(0.5 + 0.7548776662466927 * n, 1) - 0.5,
                               math.fmod(0.5 + 0.5698402909980532 * n, 1) - 0.5], atol=2e-7)
    # rays through pixel centres, float64 restatement within rounding
    cam = O.camera_look_at(spec, 0)
    rays = O.camera_rays(cam, 64, 64)
    px, py = np.meshgrid(np.arange(64), np.arange(64))
    ndc_x = 2 * (px.reshape(-1) + 0.5) / 64 - 1
    ndc_y = -2 * (py.reshape(-1) + 0.5) / 64 + 1
    U, V, W = (np.array(getattr(cam, k), np.float64) for k in ("U", "V", "W"))
    d = ndc_x[:, None] * U + ndc_y[:, None] * V + W
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    assert np.abs(rays["direction"] - d).max() < 1e-6
    cosw = d @ (W / np.linalg.norm(W))
    assert np.allclose(rays["tmin"], spec.near / cosw, rtol=1e-6)
    assert np.allclose(rays["tmax"], spec.far / cosw, rtol=1e-6)
