"""The real-time loop as a request stream: the reference re-voxelizes every
frame at the current time (voxelizeRealTime, vulkan/src/application.cpp:
68-82).  Requests with FREP_REQUEST_AT_TIME render their scene's tree built
at their own tape time; interleaved with untimed requests (the scene's
registered time) and with another animated scene, all submitted at once so
that batches hold requests at conflicting times, every result equals
frep_voxelize + frep_render at its time, and sampled frames equal the
oracle's."""
import numpy as np
import pytest

from frep import csg, lib, shapes

pytestmark = pytest.mark.gpu

MISS = np.uint64(0xFFFFFFFFFFFFFFFF)


def _scenes():
    balls = csg.compile(csg.OrganicBalls())            # sin(t) moves a ball
    g_balls = shapes.make_grid([8, 8, 4], low=csg.DEMO_LOW, world_size=csg.DEMO_WORLD)
    morph = csg.compile(csg.Morph())                   # blend and rotation follow t
    g_morph = shapes.make_grid([8, 4, 4], low=csg.DEMO_LOW, world_size=csg.DEMO_WORLD,
                               tape_time=0.3)
    return [(balls, g_balls), (morph, g_morph)]


def _cam(k, w=96, h=64):
    az = 0.4 + 0.37 * k
    b = shapes.orbit_basis(az, 0.2, 45.0, w, h)
    return lib.make_camera(b, w, h)


def test_timed_requests_match_the_plain_path(ctx, oracle):
    scenes = _scenes()
    ths = [ctx.upload(t) for t, _ in scenes]
    srv = lib.Server(ctx, max_batch=16, max_inflight=2)
    try:
        sids = [srv.scene(th, g) for th, (_, g) in zip(ths, scenes)]
        plan = []           # (scene index, time or None, camera index)
        for f in range(18):
            plan.append((0, 0.25 * f, f))
            if f % 3 == 0:
                plan.append((0, None, f + 100))
            plan.append((1, 0.3 + 0.4 * (f // 2), f + 200))
            if f % 4 == 1:
                plan.append((0, 0.25 * f, f + 300))    # same time again: no rebuild needed
        tickets = [srv.submit(sids[s], _cam(c), want=("rgba", "hits"), tape_time=t)
                   for s, t, c in plan]
        results = [srv.wait(tk) for tk in tickets]
        stats = srv.stats()
    finally:
        srv.close()
    assert all(r[0].status == 0 for r in results)
    assert stats.mean_batch > 1.0
    # the plain path at each request's time, one tree per (scene, time)
    trees = {}
    checked = 0
    for (s, t, c), (res, rgba, hits, _) in zip(plan, results):
        tape, g = scenes[s]
        tt = g["tape_time"] if t is None else t
        key = (s, np.float32(tt).tobytes())
        if key not in trees:
            gg = dict(g, tape_time=tt)
            trees[key] = (ctx.voxelize(ths[s], gg), gg)
        tree, gg = trees[key]
        prgba, phits, _ = ctx.render(tree, _cam(c), want=("rgba", "hits"))
        assert np.array_equal(hits, phits), (s, t, c)
        assert np.array_equal(rgba, prgba), (s, t, c)
        if checked < 3 and t is not None and t > 1.0:
            ohits, _, orgba = oracle.Tree(tape, gg).render(oracle.make_camera(
                shapes.orbit_basis(0.4 + 0.37 * c, 0.2, 45.0, 96, 64), 96, 64))
            assert np.array_equal(hits, ohits) and np.array_equal(rgba, orgba)
            checked += 1
    assert checked == 3
    # frames at different times differ (the shapes move)
    assert any(not np.array_equal(results[0][2], r[2]) for r in results[1:6])
    # every (scene, time) was built at least once; repeats of a resident
    # time were rendered without a rebuild
    assert len(trees) <= stats.voxelizations < len(plan)
    for tree, _ in trees.values():
        ctx.tree_free(tree)
