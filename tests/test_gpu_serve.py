"""Serving loop (mrt_server_*, mrt_job_*): continuous batching of render jobs of uneven size.

The bar: a job's result is the same alone, in a mix and through the one-frame entry point
(mrt_render_frame), bit for bit, because per-job RNG depends on (pixel, frame, sample) only and
accumulation runs in sample order whatever the chunking; jobs overlap on the device; the
admission respects the HBM budget and submission order; graph replays equal eager frames.
"""
import math

import numpy as np
import pytest

from merian_amd import abi, scenes
from merian_amd.runtime import MrtError, Server, pt_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def resident(ctx):
    a = scenes.plain_small(seed=101, subdiv=2, grid=6)
    b = scenes.full_scene(seed=102, instances=12, env_size=(128, 64), name="serve")
    ds = [ctx.scene(a), ctx.scene(b)]
    yield [(a, ds[0]), (b, ds[1])]
    for d in ds:
        d.destroy()


def _spec(sc, w, h):
    c = sc.camera
    return scenes.CameraSpec(c.eye, c.target, c.up, c.vfov, w, h)


# (scene index, width, height, spp, max path length, frame, mode)
JOBS = [
    (0, 64, 64, 4, 4, 3, abi.MODE_BSDF),
    (1, 96, 54, 16, 6, 7, abi.MODE_BSDF),
    (0, 128, 96, 1, 8, 11, abi.MODE_BSDF),
    (1, 80, 60, 8, 5, 2, abi.MODE_NEE),
    (0, 48, 32, 64, 4, 5, abi.MODE_BSDF),
    (1, 160, 90, 2, 8, 9, abi.MODE_BSDF),
    (0, 64, 48, 0, 4, 1, abi.MODE_BSDF),
]


def _job_args(ctx, resident, j):
    si, w, h, spp, L, frame, mode = j
    sc, ds = resident[si]
    cam = ctx.camera(_spec(sc, w, h), frame)
    return ds, pt_params(spp=spp, max_path_length=L, mode=mode, record_signature=False), cam, frame, w, h


def _eager(ctx, resident, j):
    ds, p, cam, frame, w, h = _job_args(ctx, resident, j)
    img, _, _, _ = ctx.render(ds, p, cam, frame, w, h, signature=False, gbuffer=False)
    return img


def _run(ctx, resident, jobs, **kw):
    with Server(ctx, **kw) as srv:
        subs = [srv.submit(*_job_args(ctx, resident, j)) for j in jobs]
        sts = [s.wait(60_000) for s in subs]
        stats = srv.stats()
    for j, st in zip(jobs, sts):
        assert st.state == abi.JOB_DONE, (j, abi.JOB_STATE_NAMES[st.state], st.error)
    return [s.irradiance for s in subs], sts, stats


@pytest.mark.parametrize("graphs", [True, False])
def test_server_equals_one_frame_entry_point(ctx, resident, graphs):
    eager = [_eager(ctx, resident, j) for j in JOBS]
    got, sts, stats = _run(ctx, resident, JOBS, max_slots=4, use_graphs=graphs)
    for j, a, b, st in zip(JOBS, eager, got, sts):
        assert a.tobytes() == b.tobytes(), j
        w, h = j[1], j[2]
        assert st.segments >= w * h
    assert stats.completed == len(JOBS)
    if graphs:
        assert stats.graph_instantiations + stats.graph_updates == len(JOBS)


def test_alone_equals_in_mix_and_jobs_overlap(ctx, resident):
    alone = []
    for j in JOBS:
        imgs, _, st = _run(ctx, resident, [j], max_slots=1)
        alone.append(imgs[0])
    # a small per-job path limit forces chunking in the mix as well
    mixed, sts, stats = _run(ctx, resident, JOBS * 2, max_slots=6, max_paths_per_job=1 << 16)
    for k, img in enumerate(mixed):
        assert img.tobytes() == alone[k % len(JOBS)].tobytes(), JOBS[k % len(JOBS)]
    assert stats.peak_in_flight > 1
    assert stats.p50_ms <= stats.p99_ms <= stats.max_ms
    assert stats.segments == sum(s.segments for s in sts)


def test_admission_by_budget(ctx, resident):
    j = JOBS[1]
    with Server(ctx, hbm_budget=1 << 30, max_slots=4) as probe:
        need = probe.job_bytes(*_job_args(ctx, resident, j))
    # room for one job at a time: all complete, never two in flight, in submission order
    with Server(ctx, hbm_budget=need + need // 2, max_slots=4) as srv:
        subs = [srv.submit(*_job_args(ctx, resident, j)) for _ in range(4)]
        sts = [s.wait(60_000) for s in subs]
        stats = srv.stats()
        assert all(s.state == abi.JOB_DONE for s in sts)
        assert stats.peak_in_flight == 1
        assert stats.peak_bytes <= need + need // 2
        # a job that can never fit is refused at submission
        big = _job_args(ctx, resident, (1, 1920, 1080, 64, 8, 0, abi.MODE_BSDF))
        with pytest.raises(MrtError) as e:
            srv.submit(*big)
        assert e.value.status == abi.MRT_ERR_BUDGET
    ref = _eager(ctx, resident, j)
    for s in subs:
        assert s.irradiance.tobytes() == ref.tobytes()


def test_cancel_and_invalid_jobs(ctx, resident):
    with Server(ctx, max_slots=1) as srv:
        subs = [srv.submit(*_job_args(ctx, resident, JOBS[1])) for _ in range(3)]
        assert subs[2].cancel()  # still queued behind the first two
        sts = [s.wait(60_000) for s in subs]
        assert sts[2].state == abi.JOB_CANCELLED
        assert not subs[2].cancel()  # already finished
        assert sts[0].state == sts[1].state == abi.JOB_DONE
        assert not np.any(subs[2].irradiance)
        ds, p, cam, frame, w, h = _job_args(ctx, resident, JOBS[0])
        p.max_path_length = 0
        with pytest.raises(MrtError) as e:
            srv.submit(ds, p, cam, frame, w, h)
        assert e.value.status == abi.MRT_ERR_INVALID
        st = srv.stats()
        assert st.cancelled == 1 and st.completed == 2


def test_graph_replays_equal_eager(ctx, resident):
    # the same job replayed through one slot's executable graph several times, with eager
    # frames and synchronisations of the program in between
    j = JOBS[3]
    ref = _eager(ctx, resident, j)
    with Server(ctx, max_slots=1, use_graphs=True) as srv:
        for k in range(4):
            s = srv.submit(*_job_args(ctx, resident, j))
            assert s.wait(60_000).state == abi.JOB_DONE
            assert s.irradiance.tobytes() == ref.tobytes(), f"replay {k}"
            other = _eager(ctx, resident, JOBS[k])
            ctx.synchronize()
            assert np.isfinite(other).all()
        st = srv.stats()
        assert st.graph_instantiations >= 1 and st.graph_instantiations + st.graph_updates == 4
