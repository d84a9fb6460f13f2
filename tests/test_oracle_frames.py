"""CPU pins of the oracle's shading frames and equal-area environment mapping (no GPU).

Each check follows from the mathematics the reference's formulas implement, not from the
formulas themselves, so a dropped term, a wrong sign or a swapped operand fails one of them:

- make_frame (frame.slang:7-13, Duff et al. 2017): rows orthonormal, third row = z exactly,
  right-handed (det +1) for every unit z, including both poles and the z = -1 branch;
- OrthonormalFrame(n, t, hs) (frame.slang:90-96): N = n/|n|, T is the unit vector of the plane
  of (t, N) perpendicular to N on t's side, det [T; B; N] = hs;
- apply_normal_map (frame.slang:123-132): the new normal is the frame's to_world of the local
  normal, the new tangent stays on the old tangent's side, the result is right-handed;
- the Lengyel UV-tangent fallback of Scene.get_surface (scene.slang:357-371): for a triangle
  that is an affine image p = O + u*A + v*B of its texture coordinates the tangent is A/|A|
  exactly (dp/du), and the reference's winding sign makes the bitangent point along -dp/dv
  with and without FlipFacing; FlipFacing negates the face normal; degenerate UVs fall back
  to make_frame;
- EqualAreaOctMap.dir_to_uv (environment-map.slang:44-63, Clarberg 2008): uniform sphere
  directions give a uniform uv histogram (the map is equal-area), the published inverse
  (written here from the paper's construction) recovers the direction, and sign flips of x, y
  and z fold the square as the octahedral layout prescribes.
"""
import numpy as np
import pytest

from merian_amd import abi
from oracle import oracle as O


def unit(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def check_orthonormal(F, tol=2e-6):
    F = F.astype(np.float64)
    G = F @ F.T
    assert np.abs(G - np.eye(3)).max() < tol, G


def special_normals():
    axes = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    near = [(1e-4, 2e-4, -1.0), (-3e-4, 1e-4, 1.0), (0.6, 0.0, -0.8), (0.0, 0.6, 0.8)]
    out = [np.array(a, np.float32) for a in axes]
    out += [np.array(v, np.float32) / np.float32(np.linalg.norm(v)) for v in near]
    return out


def test_make_frame_orthonormal_right_handed():
    rng = np.random.default_rng(1)
    for z in list(unit(rng, 4000)) + special_normals():
        F = O.frame("normal", z)
        assert np.array_equal(F[2], z)
        check_orthonormal(F)
        assert np.linalg.det(F.astype(np.float64)) == pytest.approx(1.0, abs=5e-6)
        v = rng.normal(size=3).astype(np.float32)
        local = F.astype(np.float64) @ v                   # to_local: row dots
        back = local @ F.astype(np.float64)                # to_world: mul(v, basis)
        assert np.allclose(back, v, atol=1e-5)


def test_make_frame_on_poles_is_the_identity_up_to_the_flip():
    # at z = +Z the basis is the identity; at z = -Z a right-handed basis with the same
    # tangent x must have the bitangent -y
    assert np.array_equal(O.frame("normal", [0, 0, 1]), np.eye(3, dtype=np.float32))
    F = O.frame("normal", [0, 0, -1])
    assert np.array_equal(F[0], [1, 0, 0]) and np.array_equal(F[1], [0, -1, 0])


@pytest.mark.parametrize("hs", [1.0, -1.0])
def test_frame_from_normal_and_tangent(hs):
    rng = np.random.default_rng(2)
    for _ in range(2000):
        n = rng.normal(size=3).astype(np.float32) * np.float32(rng.uniform(0.2, 5))
        t = rng.normal(size=3).astype(np.float32)
        F = O.frame("nt", n, t, hs).astype(np.float64)
        T, B, N = F
        n64, t64 = n.astype(np.float64), t.astype(np.float64)
        assert np.allclose(N, n64 / np.linalg.norm(n64), atol=2e-7)
        # float32 Gram-Schmidt loses digits as t approaches N: scale by |t| / |t_perp|
        cond = np.linalg.norm(t64) / np.linalg.norm(t64 - N * np.dot(t64, N))
        check_orthonormal(F.astype(np.float32), tol=2e-6 * max(1.0, cond))
        # T lies in span(t, N), on t's side: its component along cross(N, t) vanishes
        side = np.cross(N, t64)
        if np.linalg.norm(side) > 1e-3 * np.linalg.norm(t64):
            assert abs(np.dot(T, side / np.linalg.norm(side))) < 2e-6
            assert np.dot(T, t64) > 0
        assert np.linalg.det(F) == pytest.approx(hs, abs=4e-6 * max(1.0, cond))


def test_apply_normal_map():
    rng = np.random.default_rng(3)
    for _ in range(1000):
        n, t = rng.normal(size=3), rng.normal(size=3)
        hs = float(rng.choice([-1.0, 1.0]))
        base = O.frame("nt", n, t, hs).astype(np.float64)
        nl = rng.normal(size=3)
        nl[2] = abs(nl[2]) + 0.3
        nl = (nl / np.linalg.norm(nl)).astype(np.float32)
        F = O.frame("normal_map", n, t, hs, nl).astype(np.float64)
        assert np.allclose(F[2], nl.astype(np.float64) @ base, atol=3e-6)   # to_world(n_local)
        cond = 1.0 / max(1e-3, np.linalg.norm(base[0] - F[2] * np.dot(base[0], F[2])))
        # the new normal is not renormalised (to_world only): it inherits the base frame's rounding
        check_orthonormal(F.astype(np.float32), tol=6e-6 * cond)
        assert np.dot(F[0], base[0]) > 0
        assert abs(np.dot(F[0], np.cross(F[2], base[0]))) < 3e-6 * cond     # T in span(T_old, N)
        assert np.linalg.det(F) == pytest.approx(1.0, abs=6e-6 * cond)      # always right-handed
    # the identity map keeps T and N and rebuilds B right-handed
    F0 = O.frame("nt", [0.2, 0.3, 0.9], [1, 0, 0], -1.0)
    F1 = O.frame("normal_map", [0.2, 0.3, 0.9], [1, 0, 0], -1.0, [0, 0, 1])
    assert np.allclose(F1[0], F0[0], atol=1e-6) and np.allclose(F1[2], F0[2], atol=1e-6)
    assert np.allclose(F1[1], -F0[1], atol=1e-6)


def half(x):
    return np.float16(x).view(np.uint16)


def affine_triangle(rng, uv):
    """Triangle whose vertex positions are O + u*A + v*B for the given texture coordinates."""
    Ov, A, B = rng.normal(size=3), rng.normal(size=3), rng.normal(size=3)
    v = np.zeros(3, dtype=abi.VERTEX_DTYPE)
    for i in range(3):
        v["position"][i] = Ov + uv[i][0] * A + uv[i][1] * B
        v["uv"][i] = [half(uv[i][0]), half(uv[i][1])]
    return v, A, B


@pytest.mark.parametrize("flip", [False, True])
def test_uv_tangent_fallback_is_dp_du(flip):
    rng = np.random.default_rng(4 + flip)
    flags = abi.GEOM_FLAT_SHADING | (abi.GEOM_FLIP_FACING if flip else 0)
    for _ in range(500):
        # texture coordinates exactly representable in fp16 (multiples of 1/64)
        uv = rng.integers(-64, 64, size=(3, 2)) / 64.0
        d1, d2 = uv[1] - uv[0], uv[2] - uv[0]
        if abs(d1[0] * d2[1] - d2[0] * d1[1]) < 0.05:
            continue
        v, A, B = affine_triangle(rng, uv)
        if np.linalg.norm(np.cross(A, B)) < 0.2 * np.linalg.norm(A) * np.linalg.norm(B):
            continue
        bary = rng.dirichlet([1, 1, 1])[1:].astype(np.float32)
        r = O.surface_frame(v, flags, bary)
        p = v["position"].astype(np.float64)
        ng = np.cross(p[1] - p[0], p[2] - p[0])
        ng /= np.linalg.norm(ng)
        assert np.allclose(r["face_normal"], -ng if flip else ng, atol=2e-6)
        assert np.allclose(r["N"], r["face_normal"], atol=1e-7)                # FlatShading
        assert np.allclose(r["T"], A / np.linalg.norm(A), atol=3e-5)          # dp/du
        # bitangent along -dp/dv: the component of B perpendicular to A, negated
        Bp = B - A * np.dot(A, B) / np.dot(A, A)
        assert np.allclose(r["B"], -Bp / np.linalg.norm(Bp), atol=3e-5)


def test_uv_degenerate_falls_back_to_make_frame():
    rng = np.random.default_rng(6)
    uv = np.array([[0.25, 0.5], [0.25, 0.5], [0.25, 0.5]])
    v, _, _ = affine_triangle(rng, uv)
    v["position"] += rng.normal(size=(3, 3)) * 0.5                              # any triangle
    r = O.surface_frame(v, abi.GEOM_FLAT_SHADING, np.float32([0.3, 0.3]))
    F = np.stack([r["T"], r["B"], r["N"]])
    assert np.array_equal(F, O.frame("normal", r["N"]))


def test_vertex_tangents_and_handedness():
    rng = np.random.default_rng(7)
    for _ in range(300):
        v = np.zeros(3, dtype=abi.VERTEX_DTYPE)
        v["position"] = rng.normal(size=(3, 3))
        n = unit(rng, 1)[0]
        t = np.cross(n, unit(rng, 1)[0])
        t /= np.linalg.norm(t)
        w = int(rng.integers(0, 2))
        for i in range(3):
            v["encoded_normal"][i] = O.encode_normal(n)
            v["encoded_tangent"][i] = (O.encode_normal(t) & ~1) | w
        r = O.surface_frame(v, abi.GEOM_HAS_TANGENTS, np.float32([0.2, 0.5]))
        assert np.allclose(r["N"], n, atol=1e-4)
        assert np.allclose(r["T"], t, atol=2e-4)
        # handedness bit set -> B = -cross(N, T)
        assert np.allclose(r["B"], (-1 if w else 1) * np.cross(r["N"], r["T"]), atol=1e-6)


def clarberg_inverse(uv):
    """uv in [0,1]^2 -> unit direction, from the construction of the equal-area octahedral map:
    r = |u| + |v| on the folded square, |z| = 1 - r^2, azimuth (pi/2) * |v| / r in the octant."""
    u = 2.0 * uv[:, 0] - 1.0
    v = 2.0 * uv[:, 1] - 1.0
    au, av = np.abs(u), np.abs(v)
    lower = au + av > 1.0
    fu = np.where(lower, 1.0 - av, au)
    fv = np.where(lower, 1.0 - au, av)
    r = fu + fv
    phi = np.where(r > 0, fv / np.where(r > 0, r, 1), 0.0) * (np.pi / 2)
    z = 1.0 - r * r
    s = r * np.sqrt(np.maximum(0.0, 2.0 - r * r))
    x, y = s * np.cos(phi), s * np.sin(phi)
    return np.stack([np.sign(u) * x, np.sign(v) * y, np.where(lower, -z, z)], axis=1)


def test_equal_area_uniform_histogram():
    rng = np.random.default_rng(8)
    d = unit(rng, 1 << 20)
    uv = O.equal_area_uv(d)
    assert uv.min() >= 0.0 and uv.max() <= 1.0
    K = 32
    h = np.histogram2d(uv[:, 0], uv[:, 1], bins=K, range=[[0, 1], [0, 1]])[0]
    chi2 = ((h - h.mean()) ** 2 / h.mean()).sum()
    assert chi2 < K * K + 6 * np.sqrt(2 * K * K), chi2       # 1023 dof, ~6 sigma


def test_equal_area_inverse_round_trip():
    rng = np.random.default_rng(9)
    d = np.concatenate([unit(rng, 200000), np.float32(np.vstack([np.eye(3), -np.eye(3)]))])
    back = clarberg_inverse(O.equal_area_uv(d).astype(np.float64))
    err = np.abs(back - d).max(axis=1)
    assert err.max() < 5e-5 and np.quantile(err, 0.99) < 1e-6
    assert np.allclose(O.equal_area_uv([[0, 0, 1]]), [[0.5, 0.5]])
    c = O.equal_area_uv([[0, 0, -1]])[0]
    assert set(np.round(c, 6)) <= {0.0, 1.0}


def test_equal_area_folds():
    rng = np.random.default_rng(10)
    d = unit(rng, 50000)
    uv = O.equal_area_uv(d).astype(np.float64)
    fx = O.equal_area_uv(d * np.float32([-1, 1, 1])).astype(np.float64)
    fy = O.equal_area_uv(d * np.float32([1, -1, 1])).astype(np.float64)
    assert np.allclose(fx[:, 0], 1 - uv[:, 0], atol=1e-6) and np.allclose(fx[:, 1], uv[:, 1], atol=1e-6)
    assert np.allclose(fy[:, 1], 1 - uv[:, 1], atol=1e-6) and np.allclose(fy[:, 0], uv[:, 0], atol=1e-6)
    # the upper hemisphere maps into the centred diamond, the lower one onto its complement
    l1 = np.abs(2 * uv[:, 0] - 1) + np.abs(2 * uv[:, 1] - 1)
    assert np.all(l1[d[:, 2] > 1e-3] <= 1 + 1e-6) and np.all(l1[d[:, 2] < -1e-3] >= 1 - 1e-6)
    # z -> -z reflects across the diamond's edge: (|u|, |v|) -> (1 - |v|, 1 - |u|)
    fz = O.equal_area_uv(d * np.float32([1, 1, -1])).astype(np.float64)
    au, av = np.abs(2 * uv - 1).T
    bu, bv = np.abs(2 * fz - 1).T
    assert np.allclose(bu, 1 - av, atol=2e-6) and np.allclose(bv, 1 - au, atol=2e-6)
