"""pbrt-v4 files as the "PBRT Scene" node loads them (mrt_pbrt_*), and PFM images."""
from __future__ import annotations

import ctypes as C
import json

import numpy as np

from . import abi


def inspect(path: str) -> dict:
    """Geometry/material/env/camera summary of a pbrt file (host only)."""
    lib = abi.load_library()
    out, err = C.c_void_p(), C.c_void_p()
    st = lib.mrt_pbrt_inspect(str(path).encode(), C.byref(out), C.byref(err))
    def take(p):
        if not p.value:
            return ""
        s = C.cast(p, C.c_char_p).value.decode()
        lib.mrt_string_free(p)
        return s
    text, errors = take(out), take(err)
    if st != abi.MRT_OK:
        raise RuntimeError(errors.strip() or "pbrt load failed")
    return json.loads(text)


class LoadedScene:
    """A pbrt file loaded into a host scene description (mrt_pbrt_load_desc), owned by the library.

    desc() returns the mrt_scene_desc, valid while this object lives; `camera` the file's camera
    (or the default camera framing the scene) as a dict."""

    def __init__(self, path: str):
        self.lib = abi.load_library()
        self.p = C.POINTER(abi.PbrtDesc)()
        err = C.c_void_p()
        st = self.lib.mrt_pbrt_load_desc(str(path).encode(), C.byref(self.p), C.byref(err))
        if st != abi.MRT_OK:
            msg = C.cast(err, C.c_char_p).value.decode() if err.value else "pbrt load failed"
            if err.value:
                self.lib.mrt_string_free(err)
            raise RuntimeError(msg.strip())
        d = self.p.contents
        self.camera = {"present": bool(d.has_camera), "from_file": bool(d.camera_from_file),
                       "eye": list(d.eye), "forward": list(d.forward), "up": list(d.up), "vfov": d.vfov,
                       "aspect": d.aspect, "width": d.width, "height": d.height}

    def desc(self) -> abi.SceneDesc:
        return self.p.contents.scene

    def close(self):
        if self.p:
            self.lib.mrt_pbrt_desc_free(self.p)
            self.p = None

    def __del__(self):
        self.close()


def read_pfm(path: str) -> np.ndarray:
    """PFM (PF / Pf) as float32 [h, w, 3], top row first."""
    with open(path, "rb") as f:
        magic = f.readline().strip()
        w, h = map(int, f.readline().split())
        scale = float(f.readline())
        ch = 3 if magic == b"PF" else 1
        data = np.frombuffer(f.read(4 * w * h * ch), dtype="<f4" if scale < 0 else ">f4").reshape(h, w, ch)
    return np.ascontiguousarray(data[::-1]).astype(np.float32)


def run_graph(args: list[str]) -> int:
    """mrt_graph_run_main in this process (argv[0] is supplied)."""
    lib = abi.load_library()
    argv = ["merian-amd-graph-run"] + [str(a) for a in args]
    arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
    return int(lib.mrt_graph_run_main(len(argv), arr))
