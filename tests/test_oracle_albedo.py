"""CPU pins of the oracle's directional albedo (BSDF::get_albedo) and of "demodulate albedo".

The GBuffer stores each surface's directional albedo towards the camera (gbuffer_rt.slang:76-77)
and "demodulate albedo" divides every sample by it, floored at 1e-2 (render_pt.slang:104-116,
gbuffer.slang:135-140). The albedo of the GGX-based lobes is a fitted approximation (Ray Tracing
Gems ch. 32 eq. 4, brdf-ggx.slang:283-317) and the sheen's a fitted table (brdf-sheen.slang:92-118),
so they are pinned by what they approximate: the lobe's own hemispherical reflectance
E[sample weight] = integral of f cos, estimated by Monte Carlo with the oracle's sampling, which
the BSDF invariant tests pin separately. A transposed fit matrix or a wrong coefficient moves
the fit far outside the agreement measured here. Closed forms pin the combinators (Lambert is
its base colour, a Fresnel mix with f0 = 0 is its base, a conductor with F0 = 0 is black, the
transmission lobe is eta^2 tint). End to end, a convex Lambert object under a constant
environment demodulates to the environment radiance, and the background to one.
"""
import numpy as np
import pytest

from merian_amd import abi
from oracle import oracle as O

from test_oracle_pt import convex_scene


def dirs(mu, n):
    return np.tile(np.array([[np.sqrt(1 - mu * mu), 0.0, mu]], np.float32), (n, 1))


def mc_albedo(cfg, mu, n=160000, seed=5):
    wi = dirs(mu, n)
    _, _, w, v = O.bsdf(cfg, "sample_eval", wi, seed=seed)
    return (w * v[:, None]).astype(np.float64).mean(axis=0), O.bsdf(cfg, "albedo", wi[:1])[0]


LOBES = {
    # name: (config for alpha, (max, mean) of |fit - Monte Carlo| over the grid)
    "ggx": (lambda a: O.bsdf_cfg(1, alpha=(a, a)), (0.07, 0.03)),
    "conductor": (lambda a: O.bsdf_cfg(3, albedo=(0.9, 0.5, 0.1), alpha=(a, a)), (0.05, 0.025)),
    "sheen": (lambda a: O.bsdf_cfg(5, albedo=(1, 1, 1), alpha=(a, a)), (0.1, 0.035)),
    "rough_dielectric": (lambda a: O.bsdf_cfg(4, albedo=(1, 1, 1), alpha=(a, a), eta=1 / 1.5), (0.15, 0.06)),
    "fresnel_mix": (lambda a: O.bsdf_cfg(2, albedo=(0.5, 0.5, 0.5), alpha=(a, a), f0=0.04), (0.2, 0.03)),
}


@pytest.mark.parametrize("lobe", sorted(LOBES))
def test_albedo_matches_the_lobes_reflectance(lobe):
    make, (tol_max, tol_mean) = LOBES[lobe]
    diffs = []
    for a in (0.05, 0.2, 0.5, 0.8, 1.0):
        for mu in (1.0, 0.7, 0.4, 0.15):
            mc, fit = mc_albedo(make(a), mu)
            assert np.all(np.isfinite(fit)) and np.all(fit >= 0)
            diffs.append(np.abs(mc - fit).max())
    assert max(diffs) < tol_max and np.mean(diffs) < tol_mean, (max(diffs), np.mean(diffs))


def test_albedo_closed_forms():
    wi = np.stack([dirs(mu, 1)[0] for mu in (1.0, 0.6, 0.2)])
    base = (0.3, 0.6, 0.9)
    assert np.array_equal(O.bsdf(O.bsdf_cfg(0, albedo=base), "albedo", wi), np.tile(np.float32(base), (3, 1)))
    # f0 = 0: the specular layer reflects nothing (the fit's bias is faded out) -> the base
    got = O.bsdf(O.bsdf_cfg(2, albedo=base, alpha=(0.3, 0.3), f0=0.0), "albedo", wi)
    assert np.array_equal(got, np.tile(np.float32(base), (3, 1)))
    assert np.array_equal(O.bsdf(O.bsdf_cfg(3, albedo=(0, 0, 0), alpha=(0.3, 0.3)), "albedo", wi), np.zeros((3, 3)))
    eta = np.float32(1 / 1.5)
    got = O.bsdf(O.bsdf_cfg(6, albedo=base, alpha=(0.3, 0.3), eta=float(eta)), "albedo", wi)
    assert np.allclose(got, (eta * eta) * np.float32(base), rtol=1e-6)
    # the conductor is linear in F0 on top of a shared bias for grey F0
    g1 = O.bsdf(O.bsdf_cfg(3, albedo=(1, 1, 1), alpha=(0.4, 0.4)), "albedo", wi)
    g5 = O.bsdf(O.bsdf_cfg(3, albedo=(0.5, 0.5, 0.5), alpha=(0.4, 0.4)), "albedo", wi)
    g25 = O.bsdf(O.bsdf_cfg(3, albedo=(0.25, 0.25, 0.25), alpha=(0.4, 0.4)), "albedo", wi)
    assert np.allclose(g1 - g5, 2 * (g5 - g25), atol=1e-6)


def material_albedo(mat, mu):
    cfg = O.bsdf_cfg(100, material=mat)
    return O.bsdf(cfg, "albedo", dirs(mu, 1))[0]


def test_openpbr_tree_albedo_combines_its_lobes():
    from merian_amd import scenes
    # metalness 1: the conductor with F0 = base colour; metalness 0, no specular: Lambert
    metal = scenes.pbrt_conductor((0.9, 0.6, 0.3), 0.2)
    diffuse = scenes.pbrt_diffuse((0.7, 0.5, 0.3))
    m = material_albedo(metal, 0.8)
    cond = O.bsdf(O.bsdf_cfg(3, albedo=(0.9, 0.6, 0.3), alpha=(0.2, 0.2)), "albedo", dirs(0.8, 1))[0]
    assert np.allclose(m, cond, rtol=1e-6)
    assert np.allclose(material_albedo(diffuse, 0.5), (0.7, 0.5, 0.3), rtol=1e-6)


def render_demod(sc, spp, L, demod=True, frame=0):
    osc = O.OracleScene(sc)
    cam = O.camera_look_at(sc.camera, frame)
    W, H = sc.camera.width, sc.camera.height
    gb = osc.gbuffer(cam, W, H)
    img, _, _ = osc.render(O.pt_params(spp, L, demodulate_albedo=demod), cam, frame, W, H, gb)
    return img, gb.reshape(H, W), osc.gbuffer_albedo(cam, W, H, gb)


def test_demodulated_convex_lambert_is_the_environment():
    albedo, Lenv = (0.6, 0.4, 0.2), (2.0, 1.5, 0.5)
    sc = convex_scene(albedo, Lenv)
    img, gb, alb = render_demod(sc, 2, 3)
    obj = gb["instance_id"] != 0xFFFFFFFF
    assert 0.2 < obj.mean() < 0.8
    # the GBuffer albedo: the Lambert albedo on the object, the environment on a miss
    assert np.array_equal(alb[obj], np.tile(np.float32([*albedo, 1]), (obj.sum(), 1)))
    assert np.array_equal(alb[~obj], np.tile(np.float32([*Lenv, 1]), ((~obj).sum(), 1)))
    # background L / L = 1 exactly; object (a L) / a = L within rounding
    assert np.array_equal(img[~obj][:, :3], np.ones(((~obj).sum(), 3), np.float32))
    exact = np.isclose(img[obj][:, :3], np.float32(Lenv), rtol=3e-7, atol=0).all(axis=1)
    assert exact.mean() > 0.999, exact.mean()
    # the second moment is that of the demodulated samples
    lum = 0.299 * Lenv[0] + 0.587 * Lenv[1] + 0.114 * Lenv[2]
    assert np.allclose(img[obj][exact][:, 3], lum * lum, rtol=1e-5)


def test_demodulation_floor():
    albedo, Lenv = (0.004, 0.4, 0.2), (2.0, 1.5, 0.5)
    sc = convex_scene(albedo, Lenv)
    img, gb, _ = render_demod(sc, 1, 2)
    obj = gb["instance_id"] != 0xFFFFFFFF
    ref, _, _ = render_demod(sc, 1, 2, demod=False)
    # red is divided by the floor 1e-2, not by 0.004
    assert np.allclose(img[obj][:, 0], ref[obj][:, 0] / np.float32(0.01), rtol=1e-6)
    assert np.allclose(img[obj][:, 1], ref[obj][:, 1] / np.float32(0.4), rtol=1e-6)
