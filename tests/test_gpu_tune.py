"""Autotuning (mrt_tune, mrt_scene_set_tuning): any BVH / launch settings render the same image
bit for bit (exact traversal, sample-order accumulation), and the tuner returns settings that
are no slower than the ones it started from."""
import numpy as np
import pytest

from merian_amd import abi, scenes
from merian_amd.runtime import MrtError, pt_params

pytestmark = pytest.mark.gpu


def test_tuning_keeps_results_and_does_not_slow_down(ctx):
    sc = scenes.full_scene(seed=11, instances=24, env_size=(128, 64), res=(160, 90))
    ds = ctx.scene(sc)
    try:
        cam = ctx.camera(sc.camera, 0)
        p = pt_params(spp=4, max_path_length=6, record_signature=False)
        ref, _, _, _ = ctx.render(ds, p, cam, 3, 160, 90, signature=False, gbuffer=False)
        cases = [(4, 8, 2.0, 32, 1, 50, 1 << 15, 0), (4, 1, 0.5, 8, 64, 75, 1 << 20, 0), (8, 2, 1.0, 48, 8, 100, 0, 0),
                 (8, 4, 1.0, 16, 32, 100, 1 << 18, 0), (8, 2, 1.0, 16, 16, 100, 0, 1)]
        for width, leaf, cost, refill, leaf_lanes, grid, paths, order in cases:
            t = ds.tuning()
            t.bvh_width, t.bvh_max_leaf, t.bvh_traversal_cost = width, leaf, cost
            t.refill_lanes, t.leaf_lanes, t.grid_percent, t.max_paths = refill, leaf_lanes, grid, paths
            t.bvh8_order = order
            ds.set_tuning(t)
            img, _, _, _ = ctx.render(ds, p, cam, 3, 160, 90, signature=False, gbuffer=False)
            assert img.tobytes() == ref.tobytes(), (width, leaf, cost, refill, leaf_lanes, grid, paths, order)
        for field, bad in (("refill_lanes", 0), ("leaf_lanes", 65), ("bvh_width", 6), ("bvh_max_leaf", 9),
                           ("bvh8_order", 2)):
            t = ds.tuning()
            setattr(t, field, bad)
            with pytest.raises(MrtError):
                ds.set_tuning(t)
        # back to the defaults, then tune
        t = ds.tuning()
        t.bvh_width, t.bvh_max_leaf, t.bvh_traversal_cost = 4, 4, 1.0
        t.refill_lanes, t.leaf_lanes, t.grid_percent, t.max_paths, t.bvh8_order = 16, 16, 100, 0, 0
        ds.set_tuning(t)
        best = ds.tune(cam, 160, 90, p, budget_ms=8000)
        assert 1 <= best.bvh_max_leaf <= 8 and best.bvh_width in (4, 8) and best.ms_per_frame > 0
        assert 1 <= best.leaf_lanes <= 64
        assert ds.tuning().as_dict() == best.as_dict()
        img, _, _, _ = ctx.render(ds, p, cam, 3, 160, 90, signature=False, gbuffer=False)
        assert img.tobytes() == ref.tobytes()
    finally:
        ds.destroy()
