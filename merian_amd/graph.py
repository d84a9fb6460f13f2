"""Graph descriptions over the C ABI (mrt_graph_*): merian's graph JSON (v4 stored, v1-v3
read), fragment merge, includes and command-line overrides. Host-only; no device needed.
"""
from __future__ import annotations

import ctypes as C
import json
import os

from . import abi


def _lib():
    return abi.load_library()


def _take(lib, p: C.c_void_p) -> str:
    if not p.value:
        return ""
    s = C.cast(p, C.c_char_p).value.decode()
    lib.mrt_string_free(p)
    return s


def _dirs(search_dirs) -> bytes | None:
    if not search_dirs:
        return None
    return ":".join(os.fspath(d) for d in search_dirs).encode()


class GraphError(RuntimeError):
    pass


def apply_cli(config: dict, args: list[str], search_dirs=None) -> tuple[bool, dict, list[str]]:
    """GraphDescription::apply_cli: (ok, updated config, error messages)."""
    lib = _lib()
    argv = (C.c_char_p * max(1, len(args)))(*[a.encode() for a in args])
    out, err = C.c_void_p(), C.c_void_p()
    st = lib.mrt_graph_apply_cli(json.dumps(config).encode(), len(args), argv, _dirs(search_dirs), C.byref(out),
                                 C.byref(err))
    text, errors = _take(lib, out), _take(lib, err)
    cfg = json.loads(text) if text else config
    return st == abi.MRT_OK, cfg, [e for e in errors.splitlines() if e]


def resolve_includes(config: dict, search_dirs=None) -> tuple[bool, dict, list[str]]:
    lib = _lib()
    out, err = C.c_void_p(), C.c_void_p()
    st = lib.mrt_graph_resolve_includes(json.dumps(config).encode(), _dirs(search_dirs), C.byref(out), C.byref(err))
    text, errors = _take(lib, out), _take(lib, err)
    return st == abi.MRT_OK, (json.loads(text) if text else config), [e for e in errors.splitlines() if e]


def merge(base: dict, overwrite: dict) -> dict:
    lib = _lib()
    out = C.c_void_p()
    if lib.mrt_graph_merge(json.dumps(base).encode(), json.dumps(overwrite).encode(), C.byref(out)) != abi.MRT_OK:
        raise GraphError("merge: invalid JSON")
    return json.loads(_take(lib, out))


def store(graph: dict) -> dict:
    """Loads a graph description of any schema version and stores it as v4."""
    lib = _lib()
    out, err = C.c_void_p(), C.c_void_p()
    st = lib.mrt_graph_store(json.dumps(graph).encode(), C.byref(out), C.byref(err))
    text, errors = _take(lib, out), _take(lib, err)
    if st != abi.MRT_OK:
        raise GraphError(errors.strip())
    return json.loads(text)


def cli_help(config: dict, search_dirs=None) -> str:
    lib = _lib()
    out = C.c_void_p()
    if lib.mrt_graph_cli_help(json.dumps(config).encode(), _dirs(search_dirs), C.byref(out)) != abi.MRT_OK:
        raise GraphError("cli_help: invalid JSON")
    return _take(lib, out)
