"""Pins of the oracle's MCPG restatement (oracle/oracle_mcpg.h) by what the mathematics fixes.

The reference holds no test or golden vector for the MCPG renderer, so each piece is checked
against a closed form, an invariant or an independent float64 evaluation:
- log1p / expm1 / x / expm1(x) against numpy in float64 (few-ulp accuracy, exact small-x limits);
- the vMF density (Tokuyoshi's form) against the textbook density in float64, and its integral
  over the sphere; sampling against the closed-form mean cosine coth(k) - 1/k and the
  distribution of cos theta (KS test);
- the chain update (MCState.add_sample): the first sample reproduces (w, w target, w) exactly,
  later ones the running weights alpha = max(1/N, 0.01), N saturating at 1024, and the derived
  lobe (mu = direction to the target, kappa from the mean cosine) against float64;
- hash-grid addressing: slots in range, probe-aligned, distinct cells of one Morton tile map
  to distinct slots of one 8^L block, the normal bin and level change the key;
- the whole estimator: the oracle's MCPG frames average to the oracle's plain pt image within
  their statistical error (the guided direction is reweighted by the stochastic-MIS pdf).
"""
import numpy as np
import pytest

from merian_amd import abi, scenes
from merian_amd.runtime import mcpg_config
from oracle import oracle as O


def test_log1p_expm1_accuracy():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-0.999, 3.0, 4000), 10.0 ** rng.uniform(-12, 0, 2000),
                        -(10.0 ** rng.uniform(-12, -0.01, 2000))]).astype(np.float32)
    l = O.mcpg_scalar("log1p", x).astype(np.float64)
    ref = np.log1p(x.astype(np.float64))
    assert np.max(np.abs(l - ref) / np.maximum(np.abs(ref), 1e-30)) < 2e-6
    e = O.mcpg_scalar("expm1", x).astype(np.float64)
    ref = np.expm1(x.astype(np.float64))
    assert np.max(np.abs(e - ref) / np.maximum(np.abs(ref), 1e-30)) < 2e-6
    y = np.concatenate([x, -(10.0 ** rng.uniform(0, 3, 500)).astype(np.float32), np.float32([0.0, -0.0])])
    q = O.mcpg_scalar("x_over_expm1", y).astype(np.float64)
    yd = y.astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        ref = np.where(yd == 0, 1.0, yd / np.expm1(yd))
    assert np.max(np.abs(q - ref) / ref) < 2e-6
    # tiny arguments take the exact branches
    assert O.mcpg_scalar("log1p", [1e-9])[0] == np.float32(1e-9)
    assert O.mcpg_scalar("expm1", [1e-9])[0] == np.float32(1e-9)


def _fib_sphere(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    phi = np.pi * (1 + 5 ** 0.5) * i
    r = np.sqrt(1 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)


@pytest.mark.parametrize("kappa", [0.0, 1e-5, 0.3, 2.0, 25.0, 400.0])
def test_vmf_density_closed_form_and_normalisation(kappa):
    rng = np.random.default_rng(2)
    mu = rng.normal(size=3)
    mu /= np.linalg.norm(mu)
    w = _fib_sphere(200000)
    p = O.vmf_pdf_tokuyoshi(np.full(len(w), kappa, np.float32), np.tile(mu, (len(w), 1)).astype(np.float32),
                            w.astype(np.float32)).astype(np.float64)
    c = w @ mu
    if kappa < 1e-4:
        ref = np.full(len(w), 1 / (4 * np.pi))
    else:
        ref = kappa / (2 * np.pi * (1 - np.exp(-2 * kappa))) * np.exp(kappa * (c - 1))
    big = ref > 1e-6 * ref.max()
    assert np.max(np.abs(p[big] - ref[big]) / ref[big]) < 5e-4 * max(1.0, kappa / 25)
    if kappa <= 25:  # quadrature resolves the lobe
        assert abs(p.mean() * 4 * np.pi - 1.0) < 2e-3


@pytest.mark.parametrize("kappa", [0.5, 4.0, 60.0, 2000.0])
def test_vmf_sampling_matches_density(kappa):
    rng = np.random.default_rng(3)
    mu = rng.normal(size=3)
    mu /= np.linalg.norm(mu)
    n = 40000
    d = O.vmf_sample_tokuyoshi(np.full(n, kappa, np.float32), np.tile(mu, (n, 1)).astype(np.float32), seed=5)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-5)
    c = np.clip(d.astype(np.float64) @ mu, -1, 1)
    mean_cos = 1 / np.tanh(kappa) - 1 / kappa
    assert abs(c.mean() - mean_cos) < 5 * np.sqrt(max(c.var(), 1e-12) / n) + 1e-6
    # KS against F(c) = (exp(k c) - exp(-k)) / (exp(k) - exp(-k)) = expm1(k (c + 1)) / expm1(2k)
    cs = np.sort(c)
    if kappa <= 50:
        F = np.expm1(kappa * (cs + 1)) / np.expm1(2 * kappa)
    else:  # the same CDF without overflow: exp(-2k) is negligible
        F = np.exp(kappa * (cs - 1))
    emp = np.arange(1, n + 1) / n
    assert np.max(np.abs(F - emp)) < 1.63 / np.sqrt(n) * 1.5
    # the azimuth around mu is uniform
    t = np.cross(mu, [0.0, 0.0, 1.0] if abs(mu[2]) < 0.9 else [1.0, 0.0, 0.0])
    t /= np.linalg.norm(t)
    b = np.cross(mu, t)
    az = np.arctan2(d @ b, d @ t)
    hist, _ = np.histogram(az, bins=8, range=(-np.pi, np.pi))
    assert np.all(np.abs(hist - n / 8) < 6 * np.sqrt(n / 8))


def _state(w_tgt=(0, 0, 0), sum_w=0.0, w_cos=0.0, N=0):
    s = np.zeros((), abi.MCPG_STATE_DTYPE)
    s["w_tgt"], s["sum_w"], s["w_cos"], s["N"] = w_tgt, sum_w, w_cos, N
    return s


def test_chain_update_closed_forms():
    pos = np.float32([0.5, 1.0, -2.0])
    tgt = np.float32([3.0, 5.0, 1.0])
    prior = 0.2
    s, kv = O.mcpg_add_sample(_state(), pos, 0.75, tgt, prior)
    # first sample: alpha = 1 -> the sample itself
    assert s["N"] == 1 and s["sum_w"] == np.float32(0.75)
    assert np.array_equal(s["w_tgt"], np.float32(0.75) * tgt)
    d = (tgt - pos).astype(np.float64)
    dist2 = d @ d
    mu = d / np.sqrt(dist2)
    assert np.allclose(kv[1:], mu, atol=2e-7)
    assert abs(s["w_cos"] - 0.75) <= 2e-7  # w * cos(0) up to the rounding of normalize
    pr = max(1e-4, prior / dist2)
    r = min(max((1 * float(s["w_cos"]) / 0.75) / (1 + pr), 0.0), 0.9999999)
    assert abs(kv[0] - (3 * r - r ** 3) / (1 - r * r)) <= 1e-5 * kv[0]
    # running weights: the n-th sample enters with alpha = 1/n
    s2, _ = O.mcpg_add_sample(s, pos, 0.25, tgt * 2, prior)
    assert s2["N"] == 2
    assert abs(s2["sum_w"] - 0.5) < 1e-7
    assert np.allclose(s2["w_tgt"], 0.5 * 0.75 * tgt + 0.5 * 0.25 * 2 * tgt, rtol=1e-6)
    # N saturates at 1024; alpha floors at 0.01
    s3 = _state(w_tgt=tuple(tgt), sum_w=1.0, w_cos=0.5, N=1024)
    s4, _ = O.mcpg_add_sample(s3, pos, 3.0, tgt, prior)
    assert s4["N"] == 1024
    assert abs(s4["sum_w"] - (0.99 * 1.0 + 0.01 * 3.0)) < 1e-6
    # an empty state yields the degenerate lobe (kappa 0, mu 0)
    _, kv0 = O.mcpg_add_sample(_state(N=0), pos, 0.0, tgt, prior)
    assert kv0.tolist() == [0.0, 0.0, 0.0, 0.0]


def _grid(size, probes, locality):
    g = abi.McpgGridConfig()
    g.buffer_size, g.probe_count, g.locality_bits = size, probes, locality
    g.tan_alpha_half, g.level_bias, g.distribution_dimension = 0.006, 0.0, 2.0
    return g


def test_hash_grid_addressing_invariants():
    rng = np.random.default_rng(4)
    n = 20000
    P = rng.uniform(-50, 50, (n, 3)).astype(np.float32)
    N = rng.normal(size=(n, 3)).astype(np.float32)
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    lv = rng.integers(20, 40, n).astype(np.uint32)
    for size, probes, loc in [(4000037, 4, 3), (32777259, 2, 3), (1531, 3, 0), (1 << 20, 8, 5), (256, 2, 3)]:
        c = O.mcpg_cells(_grid(size, probes, loc), P, N, lv)
        assert (c[:, 0] < size).all()
        if probes & (probes - 1) == 0:
            assert (c[:, 0] % probes == 0).all()
        again = O.mcpg_cells(_grid(size, probes, loc), P, N, lv)
        assert np.array_equal(c, again)
    # one Morton tile: 8^L neighbouring cells at one level/normal -> 8^L distinct consecutive
    # slots of one block, distinct keys
    L, size = 3, 4000037
    g = _grid(size, 1, L)
    level = 30
    inv_w = (0.5 / np.float32(0.006)) * 2.0 ** (32 - level)
    base = np.array([37, -12, 5]) * 8  # tile-aligned lattice index
    idx = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij"), -1).reshape(-1, 3) + base
    pos = ((idx + 0.25) / inv_w).astype(np.float32)  # inside each cell, clear of its edges
    nrm = np.tile(np.float32([0, 0, 1]), (len(pos), 1))
    c = O.mcpg_cells(g, pos, nrm, np.full(len(pos), level, np.uint32))
    slots = np.sort(c[:, 0])
    assert len(np.unique(c[:, 0])) == 512 and slots[-1] - slots[0] == 511 and slots[0] % 512 == 0
    assert len(np.unique(c[:, 1])) == 512
    # the normal bin and the level are part of the key
    c2 = O.mcpg_cells(g, pos, -nrm, np.full(len(pos), level, np.uint32))
    assert (c2[:, 1] != c[:, 1]).all()
    c3 = O.mcpg_cells(g, pos, nrm, np.full(len(pos), level + 2, np.uint32))
    assert (c3[:, 1] != c[:, 1]).mean() > 0.99


def test_mcpg_frames_are_unbiased_against_pt():
    sc = scenes.full_scene(seed=5, instances=8, env_size=(64, 32), res=(24, 14))
    osc = O.OracleScene(sc)
    W, H = 24, 14
    cam = O.camera_look_at(scenes.CameraSpec(sc.camera.eye, sc.camera.target, sc.camera.up, sc.camera.vfov, W, H,
                                             jitter=abi.JITTER_NONE), 0)
    gb = osc.gbuffer(cam, W, H)
    L = 5
    ref = np.mean([osc.render(O.pt_params(256, L), cam, 50 + f, W, H, gb)[0] for f in range(8)], 0)
    m = O.McpgOracle(mcpg_config(mc_buffer_size=1 << 16, lc_buffer_size=1 << 14))
    p = O.pt_params(8, L)
    for f in range(6):  # learning
        m.render(osc, p, cam, f, W, H, gb)
    imgs = [m.render(osc, p, cam, 100 + f, W, H, gb)[0] for f in range(24)]
    img = np.mean(imgs, 0)
    lum = lambda x: x[..., 0] * 0.299 + x[..., 1] * 0.587 + x[..., 2] * 0.114
    # frame-to-frame spread of the image mean gives its error
    means = np.array([lum(x).mean() for x in imgs])
    err = means.std(ddof=1) / np.sqrt(len(means))
    assert abs(lum(img).mean() - lum(ref).mean()) < 5 * err + 1e-3 * lum(ref).mean(), (lum(img).mean(), lum(ref).mean(), err)
    # the grids learned something
    mk, ms, lk, le = m.grids()
    assert (ms["sum_w"] > 0).sum() > 100 and (le["N"] > 0).sum() > 100
