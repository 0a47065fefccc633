"""Host-side checks of served mode (no GPU): the request plan, edit replay,
configuration errors and admission by a device-memory budget."""
from __future__ import annotations

import dataclasses
import math
from collections import Counter

import numpy as np
import pytest

from shocovox_amd import Octree, OctreeError, serve

SIZES = [64, 128, 256, 512, 1024, 64, 128, 256]


def test_plan_is_deterministic_and_seeded():
    cfg = serve.ServeConfig(qps=500, duration_s=1.0, seed=7)
    a, b = serve.plan(cfg, SIZES), serve.plan(cfg, SIZES)
    assert a == b
    c = serve.plan(dataclasses.replace(cfg, seed=8), SIZES)
    assert [r.arrival_s for r in c] != [r.arrival_s for r in a]


def test_plan_poisson_rate_zipf_shares_and_edit_fraction():
    cfg = serve.ServeConfig(qps=2000, duration_s=10.0, seed=3, edit_fraction=0.05)
    p = serve.plan(cfg, SIZES)
    n = len(p)
    # Poisson count: mean 20000, sd ~141
    assert abs(n - 20000) < 6 * math.sqrt(20000)
    t = np.array([r.arrival_s for r in p])
    assert (np.diff(t) > 0).all() and t[-1] < 10.0
    gaps = np.diff(t)
    assert abs(gaps.mean() - 1 / 2000) < 0.05 / 2000
    # exponential gaps: coefficient of variation 1
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.05
    edits = sum(r.kind for r in p)
    assert abs(edits / n - 0.05) < 0.01
    # Zipf(1.1) over trees: share of rank k ~ k^-1.1 / H
    cnt = Counter(r.tree for r in p)
    h = sum(k ** -1.1 for k in range(1, len(SIZES) + 1))
    for k in range(len(SIZES)):
        assert abs(cnt[k] / n - (k + 1) ** -1.1 / h) < 0.015
    res = Counter((r.width, r.height) for r in p if r.kind == 0)
    order = [res[wh] for wh in serve.SERVED_RESOLUTIONS]
    assert order == sorted(order, reverse=True)
    for r in p:
        if r.kind == 1:
            assert 1 <= r.edit_voxels <= 1000 and r.edit_clear in (0, 1)


def test_plan_viewports_look_at_the_tree_from_outside():
    cfg = serve.ServeConfig(qps=300, duration_s=1.0, seed=5, edit_fraction=0.0)
    for r in serve.plan(cfg, SIZES):
        s = SIZES[r.tree]
        o = np.array(r.viewport.origin, np.float64)
        d = np.array(r.viewport.direction, np.float64)
        assert abs(np.linalg.norm(d) - 1) < 1e-6
        assert np.linalg.norm(o - s / 2) == pytest.approx(1.5 * s, rel=1e-5)
        to_c = (s / 2 - o) / np.linalg.norm(s / 2 - o)
        assert np.dot(to_c, d) > 0.99999
        assert r.viewport.frustum[1] == pytest.approx(2 * math.tan(math.pi / 6), rel=1e-6)
        assert r.viewport.frustum[0] == pytest.approx(r.viewport.frustum[1] * r.width / r.height,
                                                      rel=1e-6)


def test_edit_replay_is_reproducible():
    cfg = serve.ServeConfig(qps=400, duration_s=1.0, seed=11, edit_fraction=0.5,
                            max_edit_voxels=300)
    p = [dataclasses.replace(r, tree=0) for r in serve.plan(cfg, [64]) if r.kind == 1][:6]
    a, b = Octree.scene(2, 64, 4, 3), Octree.scene(2, 64, 4, 3)
    for r in p:
        serve.apply_edit(a, r)
        serve.apply_edit(b, r)
    ma, mb = a.export_model(), b.export_model()
    assert ma.keys() == mb.keys()
    for k in ma:
        assert np.array_equal(np.asarray(ma[k]), np.asarray(mb[k])), k
    # edits changed the tree
    c = Octree.scene(2, 64, 4, 3)
    mc = c.export_model()
    assert any(not np.array_equal(np.asarray(ma[k]), np.asarray(mc[k])) for k in ma)


@pytest.mark.parametrize("bad", [dict(qps=0.0), dict(duration_s=-1.0), dict(mode="x"),
                                 dict(max_inflight=0), dict(max_inflight=17),
                                 dict(max_batch_rays=10), dict(max_batch_jobs=0),
                                 dict(edit_fraction=1.5), dict(resolutions=((0, 5),)),
                                 dict(qps=1e9, duration_s=1.0), dict(deadline_ms=-1.0)])
def test_bad_configs_are_refused(bad):
    cfg = serve.ServeConfig(qps=100, duration_s=1.0)
    if bad.get("mode") == "x":
        with pytest.raises(KeyError):
            cfg.mode = "x"
            cfg.to_c()
        return
    cfg = dataclasses.replace(cfg, **bad)
    with pytest.raises(OctreeError) as e:
        serve.plan(cfg, [64])
    assert e.value.status == 5  # SVO_INVALID_ARGUMENT


def test_admission_refuses_beyond_budget_without_device():
    t = Octree.scene(2, 128, 8, 2)
    with pytest.raises(OctreeError) as e:
        serve.admit(t, budget_bytes=100_000)
    assert e.value.status == 6 and "device bytes" in str(e.value)
    # the reserve counts against the budget as well
    with pytest.raises(OctreeError):
        serve.admit(t, budget_bytes=10 << 20, reserve_bytes=10 << 20)
