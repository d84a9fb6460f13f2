"""Pins of the guide oracle (oracle/oracle_guide.cpp) by mathematics, not by restating it.

- the SH basis is orthonormal on the sphere (quadrature);
- a vMF lobe and the decoded 8-lobe mixture integrate to 1;
- the exact-mode gradient of the training loss equals central finite differences of the
  exact-mode loss (weights of every layer, grid entries of dense and hashed levels);
- trilinear grid interpolation reproduces a linear function on dense levels exactly;
- the mirror mode (fp16 rounding points of the kernels) stays close to the exact mode.
"""
import math

import numpy as np
import pytest

from merian_amd.runtime import guide_config, guide_layout
from oracle import oracle as O


def sphere_quadrature(n_theta=200, n_phi=400):
    """Gauss-Legendre in cos(theta) x uniform phi: directions [n, 3] and weights summing to 4 pi."""
    x, w = np.polynomial.legendre.leggauss(n_theta)
    phi = (np.arange(n_phi) + 0.5) * 2 * np.pi / n_phi
    ct = np.repeat(x, n_phi)
    st = np.sqrt(1 - ct * ct)
    ph = np.tile(phi, n_theta)
    d = np.stack([st * np.cos(ph), st * np.sin(ph), ct], 1)
    return d, np.repeat(w, n_phi) * (2 * np.pi / n_phi)


def small_spec(log2_table=10, base=2, top=64):
    cfg = guide_config(log2_table=log2_table, base_resolution=base, max_resolution=top)
    res, table = guide_layout(cfg)
    return O.guide_spec(res, table, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)), res, table


def random_params(table, seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    W = np.zeros((5, 64, 64), np.float32)
    for k in range(5):
        n_out = 64 if k < 4 else 40
        W[k, :n_out] = rng.uniform(-1, 1, (n_out, 64)) * math.sqrt(6 / (64 + n_out)) * scale
    G = rng.uniform(-0.5, 0.5, (8, table, 4)).astype(np.float32)
    return W, G


def random_records(n, seed=1):
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.05, 0.95, (n, 3))
    wo = rng.normal(size=(n, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.exponential(1.0, n)
    R = np.zeros((n, 12), np.float32)
    R[:, 0:3], R[:, 3], R[:, 4:7], R[:, 8:11] = p, w, wo, d
    return R


def test_sh_basis_is_orthonormal():
    d, w = sphere_quadrature(64, 128)
    Y = np.stack([O.guide_sh(v) for v in d])
    gram = (Y * w[:, None]).T @ Y
    assert np.abs(gram - np.eye(16)).max() < 2e-5


@pytest.mark.parametrize("kappa", [0.1, 1.0, 10.0, 100.0, 1000.0])
def test_vmf_lobe_integrates_to_one(kappa):
    x, w = np.polynomial.legendre.leggauss(4000)
    # integral over the sphere = 2 pi * integral over cos in [-1, 1]
    total = 2 * np.pi * sum(wi * O.vmf_pdf(kappa, xi) for xi, wi in zip(x, w))
    assert abs(total - 1.0) < 1e-6 * max(1.0, kappa / 100)


def test_mixture_integrates_to_one():
    rng = np.random.default_rng(3)
    out = rng.normal(size=40).astype(np.float32)
    for j in range(8):
        out[5 * j + 1] = rng.uniform(-1, 3)  # log kappa in a range the quadrature resolves
    d, w = sphere_quadrature(400, 800)
    q = np.array([O.guide_mixture(out, v)[0] for v in d[:: 1]])
    assert abs(float((q * w).sum()) - 1.0) < 1e-4
    _, pi, ka, mu = O.guide_mixture(out, d[0])
    assert abs(pi.sum() - 1) < 1e-6 and np.allclose(np.linalg.norm(mu, axis=1), 1, atol=1e-6)
    assert np.all((ka >= 0.1 - 1e-6) & (ka <= 5000 * (1 + 1e-6)))


def test_dense_levels_reproduce_linear_functions():
    spec, res, table = small_spec(log2_table=16, base=2, top=64)
    # feature 0 = a . (vertex index / res) + b on every dense level; interpolation is exact
    a, b = np.array([0.3, -0.2, 0.5]), 0.1
    G = np.zeros((8, table, 4), np.float32)
    dense = [l for l in range(8) if (res[l] + 1) ** 3 <= table]
    assert len(dense) >= 4
    for l in dense:
        s = res[l] + 1
        idx = np.arange(s ** 3)
        xyz = np.stack([idx % s, (idx // s) % s, idx // (s * s)], 1) / res[l]
        G[l, idx, 0] = xyz @ a + b
        G[l, idx, 1] = 1.0
        G[l, idx, 3] = xyz @ a
    W = np.zeros((5, 64, 64), np.float32)
    pts = np.random.default_rng(4).uniform(0, 1, (64, 3)).astype(np.float32)
    _, enc = O.guide_forward(spec, W, G, pts, np.tile([0, 0, 1.0], (64, 1)), exact=True, encoding=True)
    for l in dense:
        assert np.allclose(enc[:, 4 * l], pts @ a + b, atol=1e-6)
        assert np.allclose(enc[:, 4 * l + 1], 1.0, atol=1e-6)  # the weights sum to one
        assert np.allclose(enc[:, 4 * l + 2], 0.0) and np.allclose(enc[:, 4 * l + 3], pts @ a, atol=1e-6)


def test_exact_gradient_matches_finite_differences():
    spec, res, table = small_spec()
    W, G = random_params(table, seed=5, scale=1.5)
    R = random_records(24, seed=6)
    loss, dW, dG = O.guide_gradient(spec, W, G, R, exact=True)
    n = len(R)
    rng = np.random.default_rng(7)
    h = 1e-3

    def loss_at(W2, G2):
        return O.guide_gradient(spec, W2, G2, R, exact=True)[0]

    checked = 0
    for k in range(5):
        n_out = 64 if k < 4 else 40
        for _ in range(6):
            o, i = rng.integers(n_out), rng.integers(64)
            Wp, Wm = W.copy(), W.copy()
            Wp[k, o, i] += h
            Wm[k, o, i] -= h
            fd = (loss_at(Wp, G) - loss_at(Wm, G)) / (2 * h)
            an = dW[k, o, i] / n
            assert abs(fd - an) <= 2e-3 * max(1.0, abs(an)) + 2e-4, (k, o, i, fd, an)
            checked += 1
    # grid entries with a nonzero gradient, on dense and hashed levels
    nz = np.argwhere(np.abs(dG) > 1e-3)
    for l in sorted(set(nz[:, 0].tolist()))[:: 3]:
        e = nz[nz[:, 0] == l][0]
        Gp, Gm = G.copy(), G.copy()
        Gp[tuple(e)] += h
        Gm[tuple(e)] -= h
        fd = (loss_at(W, Gp) - loss_at(W, Gm)) / (2 * h)
        an = dG[tuple(e)] / n
        assert abs(fd - an) <= 2e-3 * max(1.0, abs(an)) + 2e-4, (tuple(e), fd, an)
        checked += 1
    assert checked >= 32 and np.isfinite(loss)


def test_mirror_mode_close_to_exact():
    spec, res, table = small_spec()
    W, G = random_params(table, seed=8)
    R = random_records(64, seed=9)
    ex = O.guide_forward(spec, W, G, R[:, 0:3], R[:, 4:7], exact=True)
    mi = O.guide_forward(spec, W, G, R[:, 0:3], R[:, 4:7], exact=False)
    assert np.abs(mi - ex).max() < 2e-2 * max(1.0, np.abs(ex).max())
    le, dWe, _ = O.guide_gradient(spec, W, G, R, exact=True)
    lm, dWm, _ = O.guide_gradient(spec, W, G, R, exact=False)
    assert abs(le - lm) < 1e-2 * abs(le) + 1e-3
    for k in range(5):
        assert np.linalg.norm(dWm[k] - dWe[k]) <= 1e-1 * np.linalg.norm(dWe[k]) + 1e-4
    # record by record: a ReLU unit whose pre-activation lies within fp16 rounding of zero
    # switches between the modes and moves one row of that layer's gradient (a few percent of
    # records, whatever the grid's hash); every other record agrees within 2 %
    close = 0
    for i in range(len(R)):
        _, a, _ = O.guide_gradient(spec, W, G, R[i:i + 1], exact=True)
        _, b, _ = O.guide_gradient(spec, W, G, R[i:i + 1], exact=False)
        close += all(np.linalg.norm(b[k] - a[k]) <= 2e-2 * np.linalg.norm(a[k]) + 1e-6 for k in range(5))
    assert close >= 0.9 * len(R), close


def test_guide_layout():
    res, table = guide_layout(guide_config())
    assert len(res) == 8 and res[0] == 16 and res[-1] == 2048 and table == 1 << 14
    assert all(b > a for a, b in zip(res, res[1:]))
    # a hashed level keeps each grid row along x whole, so the finest row must fit one table
    assert guide_layout(guide_config(log2_table=10, base_resolution=2, max_resolution=1023))[1] == 1024
    with pytest.raises(Exception, match="out of range"):
        guide_layout(guide_config(log2_table=10, base_resolution=2, max_resolution=1024))
