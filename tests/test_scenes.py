"""Procedural workloads: sizes as named, determinism, outward winding, pbrt-v4 export."""
import numpy as np

from merian_amd import abi, scenes


def test_plain_small_sizes_and_determinism():
    a, b = scenes.plain_small(seed=0), scenes.plain_small(seed=0)
    assert a.triangle_count == 64 * 1280 + 2 == 81922
    assert a.digest() == b.digest()
    assert scenes.plain_small(seed=1).digest() != a.digest()
    assert len(a.materials) == 65
    assert (a.camera.width, a.camera.height) == (512, 512)


def test_full_scene_sizes():
    s = scenes.full_scene(seed=2, instances=256, env_size=(64, 32))
    assert s.triangle_count == 256 * 7812 + 24 == 1999896  # 12 room quads: 4 faces + 2 with an opening
    assert len(s.geometries) == 256 + 12
    feats = s.materials["features"][:256]
    metal = s.materials["metalness"][:256]
    assert (feats & abi.PBR_TRANSMISSION).sum() > 20 and (metal == 1).sum() > 20
    two = (s.geometries["flags"] & abi.GEOM_TWO_SIDED) != 0
    # dielectrics are TwoSided (pbrt_scene.cpp:440-466), nothing else is
    assert np.array_equal(two[:256], (feats & abi.PBR_TRANSMISSION) != 0)
    assert (s.camera.width, s.camera.height) == (1920, 1080)
    # the open set: the same tori and materials, a floor and three walls, the camera outside
    o = scenes.full_scene(seed=2, instances=256, env_size=(64, 32), enclosed=False)
    assert o.triangle_count == 256 * 7812 + 8 and len(o.geometries) == 256 + 4
    nv = int(s.geometries["vertex_offset"][256])  # the tori's vertices come first
    assert nv == int(o.geometries["vertex_offset"][256]) and o.vertices[:nv].tobytes() == s.vertices[:nv].tobytes()
    assert o.camera.eye[2] > o.bounds()[1][2]


def test_sky_dynamic_range():
    e = scenes.sky_env(256, 128)
    lum = e["env_rgba"][..., :3].mean(-1)
    assert 2e4 < lum.max() / lum.min() < 1e5


def test_icosphere_and_torus_wind_outward():
    for P, T, centre in [(*scenes.icosphere(3), None), (*scenes.torus(20, 12, 1.0, 0.3), "ring")]:
        n = np.cross(P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]])
        c = P[T].mean(1)
        if centre == "ring":
            ang = np.arctan2(c[:, 2], c[:, 0])
            c = c - np.stack([np.cos(ang), np.zeros_like(ang), np.sin(ang)], 1)
        assert ((n * c).sum(1) > 0).all()


def test_pbrt_export(tmp_path):
    s = scenes.plain_small(seed=0)
    p = tmp_path / "scene.pbrt"
    scenes.write_pbrt(s, str(p))
    text = p.read_text()
    assert text.count('Shape "trianglemesh"') == len(s.geometries)
    assert 'LightSource "infinite" "rgb L" [1.0 1.0 1.0]' in text
    assert text.count('Material "conductor"') + text.count('Material "diffuse"') == len(s.geometries)
