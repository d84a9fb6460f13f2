"""The pbrt-v4 loader of the "PBRT Scene" node: shapes, PLY meshes and the default camera (host only).

- The reference's own fixture (tests/test_pbrt_scene.cpp:22-38: a named diffuse quad and an
  emissive sphere, 64x64 camera) loads with the counts that test expects: two materials (the quad's
  and the sphere's emissive copy), a camera, a valid bounding box.
- "sphere" is tessellated into 64 slices x 32 stacks (pbrt_scene.cpp:692-733): every vertex lies on
  the sphere and its normal is radial; "disk" into a 64-slice ring.
- "plymesh" (pbrt_scene.cpp:577-690): the same mesh written as ascii, binary little-endian,
  binary big-endian and gzip-compressed PLY loads to exactly the trianglemesh it holds; quads are
  split into two triangles.
- Without a camera in the file the default camera frames the scene's bounds
  (pbrt_scene.cpp:994-1003, camera.cpp:373-423): every corner of the box projects inside the
  field of view and the box is centred.
"""
import gzip
import math
import struct

import numpy as np

from merian_amd import abi, pbrt_io

FIXTURE = """
Integrator "path"
LookAt 0 1 5  0 1 0  0 1 0
Camera "perspective" "float fov" 40
Film "rgb" "integer xresolution" 64 "integer yresolution" 64
WorldBegin
MakeNamedMaterial "White" "string type" "diffuse" "rgb reflectance" [0.7 0.7 0.7]
NamedMaterial "White"
Shape "trianglemesh" "integer indices" [0 1 2 0 2 3]
    "point3 P" [-1 0 -1  -1 0 1  1 0 1  1 0 -1]
AttributeBegin
    AreaLightSource "diffuse" "rgb L" [10 10 10]
    Translate 0 2 0
    Shape "sphere" "float radius" 0.25
AttributeEnd
"""


def arrays(ls):
    d = ls.desc()
    V = np.ctypeslib.as_array(C_cast(d.vertices, abi.VERTEX_DTYPE, d.vertex_count))
    I = np.ctypeslib.as_array((np_u32 * d.index_count).from_address(d.indices)) if d.index_count else np.zeros(0)
    G = np.ctypeslib.as_array(C_cast(d.geometries, abi.GEOMETRY_DTYPE, d.geometry_count))
    M = np.ctypeslib.as_array(C_cast(d.materials, abi.MATERIAL_DTYPE, d.material_count))
    return V.copy(), np.array(I, np.uint32), G.copy(), M.copy()


import ctypes as C  # noqa: E402

np_u32 = C.c_uint32


def C_cast(addr, dtype, n):
    buf = (C.c_uint8 * (dtype.itemsize * n)).from_address(addr)
    return np.frombuffer(buf, dtype=dtype, count=n)


def decode(enc):
    from oracle import oracle as O
    return np.array([O.decode_normal(int(e)) for e in enc])


def test_reference_fixture(tmp_path):
    p = tmp_path / "merian-test-scene.pbrt"
    p.write_text(FIXTURE)
    info = pbrt_io.inspect(p)
    assert len(info["materials"]) == 2
    assert info["camera"]["present"] and info["camera"]["from_file"]
    assert (info["camera"]["width"], info["camera"]["height"]) == (64, 64)
    assert info["warnings"] == []
    ls = pbrt_io.LoadedScene(p)
    V, I, G, M = arrays(ls)
    assert len(G) == 2 and G[1]["primitive_count"] == 64 * 32 * 2
    assert tuple(M[1]["emission"]) == (10.0, 10.0, 10.0) and tuple(M[0]["emission"]) == (0.0, 0.0, 0.0)
    assert all(G["flags"] & abi.GEOM_TWO_SIDED)   # not a two-sided emitter: visible from both sides
    P = V["position"].astype(np.float64)
    assert np.isfinite(P).all() and (P.max(0) > P.min(0)).all()   # a valid bounding box


def test_sphere_and_disk(tmp_path):
    p = tmp_path / "shapes.pbrt"
    p.write_text("""WorldBegin
AttributeBegin
  Translate 1 2 3
  Shape "sphere" "float radius" 0.5
AttributeEnd
AttributeBegin
  Translate 0 -1 0
  Shape "disk" "float radius" 2 "float innerradius" 0.5 "float height" 0.25
AttributeEnd
""")
    V, I, G, M = arrays(pbrt_io.LoadedScene(p))
    assert len(G) == 2
    vs = slice(int(G[0]["vertex_offset"]), int(G[1]["vertex_offset"]))
    P = V["position"][vs].astype(np.float64) - (1, 2, 3)
    assert len(P) == 65 * 33
    assert np.allclose(np.linalg.norm(P, axis=1), 0.5, atol=1e-6)
    N = decode(V["encoded_normal"][vs])
    assert np.allclose(N, P / 0.5, atol=1e-4)
    D = V["position"][int(G[1]["vertex_offset"]):].astype(np.float64)
    # the disk lies in z = height, translated by -1 in y (the ring is centred at (0, -1))
    r = np.hypot(D[:, 0], D[:, 1] + 1)
    assert len(D) == 2 * 65 and np.allclose(D[:, 2], 0.25)
    outer, inner = np.isclose(r, 2.0, atol=1e-5), np.isclose(r, 0.5, atol=1e-5)
    assert (outer | inner).all() and outer.sum() == inner.sum() == 65
    assert G[1]["primitive_count"] == 128


MESH_P = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 0.5, 1]], np.float32)
MESH_F = [[0, 1, 2, 3], [0, 1, 4], [1, 2, 4], [2, 3, 4], [3, 0, 4]]


def write_ply(path, fmt, gz=False):
    head = [b"ply", f"format {fmt} 1.0".encode(), b"comment test mesh", f"element vertex {len(MESH_P)}".encode(),
            b"property float x", b"property float y", b"property float z",
            f"element face {len(MESH_F)}".encode(), b"property list uchar int vertex_indices", b"end_header"]
    data = b"\n".join(head) + b"\n"
    if fmt == "ascii":
        data += b"".join(f"{x} {y} {z}\n".encode() for x, y, z in MESH_P)
        data += b"".join((" ".join(map(str, [len(f)] + f)) + "\n").encode() for f in MESH_F)
    else:
        e = "<" if fmt == "binary_little_endian" else ">"
        data += b"".join(struct.pack(e + "3f", *v) for v in MESH_P)
        data += b"".join(struct.pack(e + "B" + "i" * len(f), len(f), *f) for f in MESH_F)
    if gz:
        data = gzip.compress(data)
    path.write_bytes(data)


def test_plymesh_formats_equal_the_trianglemesh(tmp_path):
    tris = []
    for f in MESH_F:
        for q in range(1, len(f) - 1):
            tris += [f[0], f[q], f[q + 1]]
    ref = tmp_path / "ref.pbrt"
    ref.write_text('WorldBegin\nShape "trianglemesh" "integer indices" [' + " ".join(map(str, tris)) +
                   '] "point3 P" [' + " ".join(f"{x:.9g}" for x in MESH_P.reshape(-1)) + "]\n")
    rV, rI, rG, _ = arrays(pbrt_io.LoadedScene(ref))
    assert len(rI) == 18
    for name, fmt, gz in [("a.ply", "ascii", False), ("l.ply", "binary_little_endian", False),
                          ("b.ply", "binary_big_endian", False), ("g.ply.gz", "binary_little_endian", True)]:
        write_ply(tmp_path / name, fmt, gz)
        sc = tmp_path / (name + ".pbrt")
        sc.write_text(f'WorldBegin\nShape "plymesh" "string filename" "{name}"\n')
        V, I, G, _ = arrays(pbrt_io.LoadedScene(sc))
        assert V.tobytes() == rV.tobytes() and np.array_equal(I, rI) and G.tobytes() == rG.tobytes(), name
    bad = tmp_path / "bad.pbrt"
    bad.write_text('WorldBegin\nShape "plymesh" "string filename" "missing.ply"\n')
    info = pbrt_io.inspect(bad)
    assert info["triangles"] == 0 and any("missing.ply" in w for w in info["warnings"])


def test_default_camera_frames_the_scene(tmp_path):
    p = tmp_path / "nocam.pbrt"
    p.write_text('WorldBegin\nShape "trianglemesh" "integer indices" [0 1 2 0 2 3] '
                 '"point3 P" [4 0 -2  9 0 -2  9 3 1  4 3 1]\n')
    info = pbrt_io.inspect(p)
    cam = info["camera"]
    assert cam["present"] and not cam["from_file"]
    assert abs(cam["aspect"] - 1920 / 1080) < 1e-6 and abs(cam["vfov"] - math.radians(60)) < 1e-6
    eye, fwd, up = (np.array(cam[k], np.float64) for k in ("eye", "forward", "up"))
    lo, hi = np.array([4, 0, -2.0]), np.array([9, 3, 1.0])
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)
    u = np.cross(right, fwd)
    tan_y = math.tan(math.radians(60) / 2)
    tan_x = tan_y * 1920 / 1080
    px, py = [], []
    for i in range(8):
        c = np.where([(i >> k) & 1 for k in range(3)], hi, lo)
        v = c - eye
        d = v @ fwd
        assert d > 0
        px.append(v @ right / d / tan_x)
        py.append(v @ u / d / tan_y)
    # inside the field of view (fitted to 85 % of it), and centred
    assert max(map(abs, px)) <= 1.0 and max(map(abs, py)) <= 1.0
    assert max(max(map(abs, px)), max(map(abs, py))) > 0.7
    assert abs(min(px) + max(px)) < 0.05 and abs(min(py) + max(py)) < 0.05
