"""mrt_trace_closest_noalloc_dev (the allocation-free ray query) against mrt_trace_closest_dev.

The entry point has no counterpart in the reference; it serves callers that cannot allocate
device memory. Its traversal scratch is a fixed device array shared by every call on the GPU,
which keeps the previous call's contents, so the test runs two calls in a row, the second on a
smaller input, and compares every hit (triangle id, t, u, v) bit for bit with the allocating
entry point on the same rays, for the 4-wide and the 8-wide tree. It creates its own context, so
it runs alone in a fresh process as well as after the other GPU tests.
"""
import numpy as np
import pytest

from merian_amd import abi, scenes
from merian_amd.runtime import Context, MrtError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("width", [4, 8])
def test_noalloc_trace_matches_trace_dev(width):
    import torch
    pos, idx = scenes.ray_query(seed=3, n_tri=200_000)
    rays = scenes.incoherent_rays(1 << 21, seed=4)
    with Context(0) as ctx:
        ds = ctx.triangles(pos, idx)
        t = ds.tuning()
        t.bvh_width, t.bvh_max_leaf = width, 2
        ds.set_tuning(t)
        try:
            d_rays = torch.from_numpy(rays.view(np.uint8)).cuda()
            torch.cuda.synchronize()
            for n in (len(rays), 300_001):
                print(f"noalloc trace: width {width}, {n} rays", flush=True)
                d_ref = torch.zeros(n * 16, dtype=torch.uint8, device="cuda")
                d_got = torch.full((n * 16,), 0xAB, dtype=torch.uint8, device="cuda")
                torch.cuda.synchronize()
                before = ctx.memory_in_use
                ctx.trace_noalloc_dev(ds, d_rays.data_ptr(), n, d_got.data_ptr())
                ctx.synchronize()
                assert ctx.memory_in_use == before       # nothing charged to the context
                ctx.trace_dev(ds, d_rays.data_ptr(), n, d_ref.data_ptr())
                ctx.synchronize()
                assert torch.equal(d_got, d_ref), f"{int((d_got != d_ref).sum())} bytes differ at n={n}"
                hit = (d_ref.view(torch.int32).view(-1, 4)[:, 0] != -1).float().mean().item()
                assert hit > 0.2
        finally:
            ds.destroy()


def test_noalloc_rejects_oversize_requests():
    pos, idx = scenes.ray_query(seed=3, n_tri=1000)
    with Context(0) as ctx:
        ds = ctx.triangles(pos, idx)
        try:
            # host-side checks: nothing is launched
            with pytest.raises(MrtError) as e:
                ctx.trace_noalloc_dev(ds, 16, 1 << 32, 16)
            assert e.value.status == abi.MRT_ERR_INVALID
            with pytest.raises(MrtError):
                ctx.trace_noalloc_dev(ds, 0, 10, 0)
            ctx.trace_noalloc_dev(ds, 0, 0, 0)        # an empty query is a no-op
        finally:
            ds.destroy()
