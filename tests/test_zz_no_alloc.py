"""svo_batch_run_static, the batched entry point that allocates nothing, against
svo_batch_run on the same jobs, through the C ABI: every output field equal,
bit for bit, on two calls in a row, the second with a smaller input (fewer
jobs, fewer tiles), which must not see the first call's table or counters.
(Named to run last: a failure here would hide the tests collected after it.)"""
from __future__ import annotations

import pytest

from shocovox_amd import Batch, DeviceTree, Octree, _native as N, alloc_hits, workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ("hit", "pix", "albedo", "voxel", "distance", "impact", "normal", "flags")


def _say(msg: str) -> None:
    print(msg, flush=True)


def _static_fn():
    return N.lib().svo_batch_run_static


def _jobs(specs, dev, rays):
    """(jobs for the static call, jobs for the context, their output buffers)"""
    import torch
    js, jb, outs = [], [], []
    for kind, tree, a, b in specs:
        if kind == "render":
            vp, w, h = a, b[0], b[1]
            bs, bb = alloc_hits(w * h, dev, FIELDS), alloc_hits(w * h, dev, FIELDS)
            for t in (*bs.values(), *bb.values()):
                t.fill_(-7)
            js.append(Batch.render_job(tree, vp, w, h, bs))
            jb.append(Batch.render_job(tree, vp, w, h, bb))
        else:
            n = rays.shape[0]
            bs, bb = alloc_hits(n, dev, FIELDS), alloc_hits(n, dev, FIELDS)
            for t in (*bs.values(), *bb.values()):
                t.fill_(-7)
            js.append(Batch.cast_job(tree, rays, bs))
            jb.append(Batch.cast_job(tree, rays, bb))
        outs.append((bs, bb))
    torch.cuda.synchronize()
    return js, jb, outs


def test_static_batch_equals_batch_run(gpu_device):
    import torch
    plain = W.build_scene(W.PLAIN, seed=1)
    caves = Octree.scene(2, 128, 8, 5)
    dp, dc = DeviceTree(plain), DeviceTree(caves)
    rays = torch.from_numpy(W.incoherent_rays(128, 5000, seed=9)).to(gpu_device)
    vp_p = W.viewports(W.PLAIN)[0]
    vp_c = W.look_at_viewport((200.0, 150.0, -40.0), (64.0, 40.0, 64.0), 160, 96)
    fn = _static_fn()
    stream = torch.cuda.current_stream()
    calls = [
        ("first call: 2 trees, 2 renders + 1 cast", [
            ("render", dp, vp_p, (200, 150)), ("render", dc, vp_c, (160, 96)),
            ("cast", dc, None, None)]),
        ("second call: 1 smaller render", [("render", dc, vp_c, (64, 40))]),
    ]
    for label, specs in calls:
        _say(f"no-alloc batch: {label}")
        js, jb, outs = _jobs(specs, gpu_device, rays)
        arr = (N.Job * len(js))(*js)
        st = fn(arr, len(js), stream.cuda_stream)
        assert st == 0, N.lib().svo_last_error().decode()
        ctx = Batch(len(jb))
        ctx.run(jb, stream)
        torch.cuda.synchronize()
        for (bs, bb), spec in zip(outs, specs):
            for k in FIELDS:
                assert torch.equal(bs[k], bb[k]), f"{label}: {spec[0]} field {k} differs"
        hits = sum(int(bs["hit"].sum()) for bs, _ in outs)
        _say(f"no-alloc batch: {label}: equal, {hits} hits")
        assert hits > 0


def test_static_batch_rejects_more_jobs_than_its_table_holds():
    fn = _static_fn()
    jobs = (N.Job * 257)()
    assert fn(jobs, 257, None) == 5  # SVO_INVALID_ARGUMENT, before any device call
    assert "at most 256 jobs" in N.lib().svo_last_error().decode()
