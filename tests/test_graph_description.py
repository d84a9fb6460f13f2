"""Graph descriptions and command-line overrides (mrt_graph_*), host only.

The cases restate merian's own boundary tests (tests/test_graph_load_store.cpp:17-345) on this
library's engine, plus include resolution and schema parsing cases of
src/merian-graph/graph/graph_description.cpp.
"""
import json

import pytest

from merian_amd import graph as G


def write(path, obj):
    path.write_text(json.dumps(obj))
    return str(path)


def test_cli_round_trip():  # test_graph_load_store.cpp:17-48
    config = {"version": 4, "nodes": {},
              "cli": {"--": {"pointer": "/nodes/shell/properties/startup commands"},
                      "scene": {"pointer": "/nodes/scene/properties/file", "required": True}}}
    assert G.store(config)["cli"] == config["cli"]
    assert "cli" not in G.store({"version": 4, "nodes": {}})
    meta = {"version": 4, "nodes": {"color": {"type": "Color", "metadata": {"pos": [10, 20]}}}}
    assert G.store(meta)["nodes"]["color"]["metadata"] == {"pos": [10, 20]}


def test_v3_array_back_compat():  # :50-64
    stored = G.store({"version": 3, "nodes": [
        {"id": "color", "type": "Color", "enabled": False, "outputs": ["out->blit.src"]},
        {"id": "blit", "type": "Blit"}]})
    assert stored["version"] == 4
    assert stored["nodes"]["color"]["enabled"] is False
    assert stored["nodes"]["blit"]["type"] == "Blit"
    assert stored["nodes"]["color"]["outputs"] == ["out->blit.src"]


def test_v1_and_object_outputs():
    v1 = G.store({"nodes": {"a": {"type": "A", "disable": True}, "b": {"type": "B"}},
                  "connections": [{"src": "a", "dst": "b", "src_output": "out", "dst_input": "in"}]})
    assert v1["nodes"]["a"]["enabled"] is False and v1["nodes"]["a"]["outputs"] == ["out->b.in"]
    v4 = G.store({"version": 4, "nodes": {"a": {"type": "A", "outputs": {"out": ["b.x", "c.y"]}},
                                          "b": {"type": "B"}, "c": {"type": "C"}}})
    assert v4["nodes"]["a"]["outputs"] == ["out->b.x", "out->c.y"]
    with pytest.raises(G.GraphError):
        G.store({"version": 4, "nodes": {"a": {"properties": {}}}})  # no type
    with pytest.raises(G.GraphError):
        G.store({"version": 4, "nodes": {"a.b": {"type": "A"}}})  # reserved character
    with pytest.raises(G.GraphError):
        G.store({"version": 4, "nodes": {"a": {"type": "A", "outputs": ["bad"]}}})
    with pytest.raises(G.GraphError):
        G.store({"version": 9, "nodes": {}})


def test_value_with_set():  # :66-87
    config = {"nodes": {"render": {"properties": {"samples per pixel": 2}},
                        "scene": {"properties": {"file": "", "env": {"Type": "Empty", "Path": ""}}}},
              "cli": {"--": {"pointer": "/nodes/scene/properties/file"},
                      "env-map": {"pointer": "/nodes/scene/properties/env/Path",
                                  "set": {"/nodes/scene/properties/env/Type": "LatLong"}},
                      "spp": {"pointer": "/nodes/render/properties/samples per pixel"}}}
    ok, c, _ = G.apply_cli(config, ["--env-map", "/tmp/x.hdr", "--spp", "8", "/scene.gltf"])
    assert ok
    assert c["nodes"]["scene"]["properties"]["env"] == {"Type": "LatLong", "Path": "/tmp/x.hdr"}
    assert c["nodes"]["render"]["properties"]["samples per pixel"] == 8
    assert isinstance(c["nodes"]["render"]["properties"]["samples per pixel"], int)
    assert c["nodes"]["scene"]["properties"]["file"] == "/scene.gltf"


VARIANT = {"v": 0, "cli": {"quality": {"type": "variant", "default": "high",
                                       "variants": {"low": {"set": {"/v": 1}}, "high": {"set": {"/v": 8}}}}}}


def test_variant():  # :89-107
    ok, c, _ = G.apply_cli(VARIANT, [])
    assert ok and c["v"] == 8
    ok, c, _ = G.apply_cli(VARIANT, ["--quality", "low"])
    assert ok and c["v"] == 1
    ok, c, errs = G.apply_cli(VARIANT, ["--quality", "ultra"])
    assert not ok and c["v"] == 0 and any("ultra" in e for e in errs)


def test_flag():  # :109-124
    base = {"d": False, "cli": {"denoise": {"type": "flag", "set": {"/d": True}, "off": {"set": {"/d": False}}}}}
    ok, c, _ = G.apply_cli(base, ["--denoise", "on"])
    assert ok and c["d"] is True
    off = dict(base, d=True)
    ok, c, _ = G.apply_cli(off, ["--denoise", "off"])
    assert ok and c["d"] is False
    ok, _, _ = G.apply_cli(base, ["--denoise", "maybe"])
    assert not ok


def test_required_and_unknown():  # :126-145
    config = {"nodes": {"scene": {"properties": {"file": ""}}},
              "cli": {"scene": {"pointer": "/nodes/scene/properties/file", "required": True}}}
    assert not G.apply_cli(config, [])[0]
    ok, given, _ = G.apply_cli(config, ["--scene", "/s.gltf"])
    assert ok and given["nodes"]["scene"]["properties"]["file"] == "/s.gltf"
    assert G.apply_cli(given, [])[0]  # a stored value satisfies the requirement
    assert not G.apply_cli(config, ["--nope", "x"])[0]  # positional without "--"


def test_order_last_wins():  # :147-151
    ok, c, _ = G.apply_cli({"v": 0, "cli": {"spp": {"pointer": "/v"}}}, ["--spp", "1", "--spp", "2"])
    assert ok and c["v"] == 2


def test_missing_value():
    ok, _, errs = G.apply_cli({"v": 0, "cli": {"spp": {"pointer": "/v"}}}, ["--spp"])
    assert not ok and any("missing value" in e for e in errs)


def test_merge_declares_override(tmp_path):  # :153-173
    f = write(tmp_path / "fragment.json", {"nodes": {"render": {"properties": {"samples per pixel": 2}}},
                                           "cli": {"spp": {"pointer": "/nodes/render/properties/samples per pixel"}}})
    for args in (["--merge", f, "--spp", "8"], ["--spp", "8", "--merge", f]):
        ok, c, _ = G.apply_cli({"nodes": {"scene": {}}}, args)
        assert ok and c["nodes"]["render"]["properties"]["samples per pixel"] == 8


def test_default_variant_declares_override(tmp_path):  # :175-197
    f = write(tmp_path / "renderer.json", {"nodes": {"render": {"properties": {"samples per pixel": 2}}},
                                           "cli": {"spp": {"pointer": "/nodes/render/properties/samples per pixel"}}})
    config = {"nodes": {"scene": {}},
              "cli": {"renderer": {"type": "variant", "default": "pt", "variants": {"pt": {"merge": f}}}}}
    ok, c, _ = G.apply_cli(config, ["--spp", "8"])
    assert ok and c["nodes"]["render"]["properties"]["samples per pixel"] == 8
    assert c["cli"]["renderer"]["selected"] == "pt"


def test_variant_sticky():  # :199-225
    ok, c, _ = G.apply_cli(VARIANT, [])
    assert ok and c["v"] == 8 and c["cli"]["quality"]["selected"] == "high"
    c["v"] = 42
    ok, c, _ = G.apply_cli(c, [])
    assert ok and c["v"] == 42  # a persisted selection is not re-applied
    ok, c, _ = G.apply_cli(c, ["--quality", "high"])
    assert ok and c["v"] == 8
    ok, c, _ = G.apply_cli(c, ["--quality", "low"])
    assert ok and c["v"] == 1 and c["cli"]["quality"]["selected"] == "low"


def test_variant_custom_file(tmp_path):  # :227-255
    f = write(tmp_path / "custom-renderer.json", {"nodes": {"render": {"properties": {"samples per pixel": 16}}}})
    config = {"nodes": {"scene": {}},
              "cli": {"renderer": {"type": "variant", "default": "pt",
                                   "variants": {"pt": {"set": {"/nodes/scene/properties/x": 1}}}}}}
    ok, c, _ = G.apply_cli(config, ["--renderer", f])
    assert ok and c["nodes"]["render"]["properties"]["samples per pixel"] == 16
    assert c["cli"]["renderer"]["selected"] == f and "properties" not in c["nodes"]["scene"]
    missing = {"cli": {"renderer": {"type": "variant", "variants": {}}}}
    ok, c, _ = G.apply_cli(missing, ["--renderer", "does/not/exist.json"])
    assert not ok and "selected" not in c["cli"]["renderer"]


def test_patch_merge_then_set(tmp_path):  # :257-276
    f = write(tmp_path / "merge-set.json", {"nodes": {"render": {"properties": {"max path length": 12}}}})
    config = {"nodes": {"scene": {}},
              "cli": {"renderer": {"type": "variant", "variants": {"pt": {
                  "set": {"/nodes/render/properties/max path length": 3}, "merge": f}}}}}
    ok, c, _ = G.apply_cli(config, ["--renderer", "pt"])
    assert ok and c["nodes"]["render"]["properties"]["max path length"] == 3


def test_null_removes_override(tmp_path):  # :278-296
    f = write(tmp_path / "null.json", {"cli": {"max-path-length": None}})
    ok, c, _ = G.apply_cli({"v": 5, "cli": {"max-path-length": {"pointer": "/v"}}}, ["--merge", f])
    assert ok and "max-path-length" not in c["cli"]
    assert not G.apply_cli(c, ["--max-path-length", "3"])[0]


def test_merge_file(tmp_path):  # :298-313
    f = write(tmp_path / "merge.json", {"nodes": {"render": {"properties": {"samples per pixel": 16}}}})
    ok, c, _ = G.apply_cli({"nodes": {"render": {"properties": {"samples per pixel": 2, "max path length": 5}}}},
                           ["--merge", f])
    assert ok and c["nodes"]["render"]["properties"] == {"samples per pixel": 16, "max path length": 5}


def test_merge_rules():  # :315-345
    assert G.merge({"a": {"x": 1, "y": 2}, "b": 1}, {"a": {"y": 9, "z": 3}, "c": 4}) == \
        {"a": {"x": 1, "y": 9, "z": 3}, "b": 1, "c": 4}
    base = {"nodes": {"render": {"properties": {"spp": 2, "len": 3}}, "scene": {"properties": {"file": "a"}}}}
    m = G.merge(base, {"nodes": {"render": {"properties": {"spp": 8}}, "extra": {"properties": {"k": 1}}}})
    assert len(m["nodes"]) == 3 and m["nodes"]["render"]["properties"] == {"spp": 8, "len": 3}
    assert m["nodes"]["extra"]["properties"]["k"] == 1
    assert G.merge({"x": 1}, {"x": [1, 2]}) == {"x": [1, 2]}
    assert G.merge({"a": [1, 2, 3]}, {"a": [9]}) == {"a": [9]}
    # "$+$key" appends missing elements instead of replacing
    assert G.merge({"a": [1, 2]}, {"$+$a": [2, 3]}) == {"a": [1, 2, 3]}


def test_includes(tmp_path):
    sub = tmp_path / "sub"
    sub.mkdir()
    write(sub / "inner.json", {"nodes": {"render": {"properties": {"samples per pixel": 1, "max path length": 5}}}})
    write(tmp_path / "outer.json", {"include": "sub/inner.json",
                                    "nodes": {"render": {"properties": {"samples per pixel": 2}}}})
    ok, c, _ = G.resolve_includes({"include": ["outer.json"], "nodes": {"render": {"type": "Render (Path-traced)"}}},
                                  [tmp_path])
    assert ok
    assert c["nodes"]["render"] == {"type": "Render (Path-traced)",
                                    "properties": {"samples per pixel": 2, "max path length": 5}}
    # cycles and missing files are errors
    write(tmp_path / "a.json", {"include": "b.json"})
    write(tmp_path / "b.json", {"include": "a.json"})
    ok, _, errs = G.resolve_includes({"include": "a.json"}, [tmp_path])
    assert not ok and any("cycle" in e for e in errs)
    ok, _, errs = G.resolve_includes({"include": "nope.json"}, [tmp_path])
    assert not ok and any("not found" in e for e in errs)


def test_string_targets_stay_strings():
    config = {"nodes": {"scene": {"properties": {"file": ""}}},
              "cli": {"--": {"pointer": "/nodes/scene/properties/file"}}}
    ok, c, _ = G.apply_cli(config, ["123"])
    assert ok and c["nodes"]["scene"]["properties"]["file"] == "123"
    ok, c, _ = G.apply_cli(config, ["a b", "c"])  # several positionals are quoted and joined
    assert ok and c["nodes"]["scene"]["properties"]["file"] == '"a b" c'


def test_cli_help():
    text = G.cli_help({"cli": {"spp": {"pointer": "/nodes/render/properties/samples per pixel"},
                               "scene": {"pointer": "/x", "required": True}, **VARIANT["cli"]}})
    assert "--spp VALUE" in text and "(required)" in text and "--quality high|low|FILE.json" in text
    assert "sets /nodes/render/properties/samples per pixel" in text and "variant, now high" in text
    assert G.cli_help({}) == "this graph declares no options\n"
