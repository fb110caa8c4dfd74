"""The served and heavy-transcendental workloads at their full sizes.

Served: Poisson arrivals of voxelize + render(960x540) requests for
24-primitive random CSG scenes at 1024^3 (branching 16/8/8) under the default
64 GiB device budget, replayed in real time through frep_server; every
result equals frep_voxelize + frep_render of its scene and camera, and the
results of several requests equal the oracle's tree rendered by the
oracle's raytracer (vulkan/src/engine/voxelizer.cpp:429-440,
vulkan/src/engine/raytracer.cpp:243-272)."""
import time

import numpy as np
import pytest

from frep import lib, shapes

pytestmark = pytest.mark.gpu

MISS = np.uint64(0xFFFFFFFFFFFFFFFF)


def _served_camera(az, el, seed, w=960, h=540):
    return shapes.orbit_basis(az, el, 3.0, w, h), dict(seed=seed)


def test_served_960x540_at_1024_default_budget(ctx, oracle):
    grid = shapes.make_grid([16, 8, 8])
    reqs = shapes.served_requests(48, seed=7, rate=500.0)
    scenes = sorted({k for _, k, _, _, _ in reqs})
    tapes = {k: shapes.random_csg(24, 10000 + k) for k in scenes}
    ths = {k: ctx.upload(t) for k, t in tapes.items()}
    srv = lib.Server(ctx)                       # default budget: 64 GiB
    try:
        handles = {k: srv.scene(ths[k], grid) for k in scenes}
        cams = []
        for _, _, az, el, sd in reqs:
            b, _ = _served_camera(az, el, sd)
            cams.append(lib.make_camera(b, 960, 540, seed=sd))
        tickets = []
        t0 = time.perf_counter()
        for (arr, k, _, _, _), cam in zip(reqs, cams):   # arrivals in real time
            d = t0 + arr - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets.append(srv.submit(handles[k], cam, want=("rgba", "hits", "normals")))
        results = [srv.wait(t) for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    assert st.budget_bytes == 64 << 30
    assert st.completed == len(reqs) and st.failed == 0 and st.evictions == 0
    assert st.voxelizations == len(scenes)           # each scene built once, then resident
    trees = {}
    for (_, k, _, _, _), cam, (res, rgba, hits, nrm) in zip(reqs, cams, results):
        assert res.status == lib.FREP_OK
        if k not in trees:
            trees[k] = ctx.voxelize(ths[k], grid)
        prgba, phits, pnrm = ctx.render(trees[k], cam)
        assert np.array_equal(hits, phits), k
        assert np.array_equal(nrm.view(np.uint32), pnrm.view(np.uint32)), k
        assert np.array_equal(rgba, prgba), k
    # the oracle's tree and raytracer for the first requests of three scenes
    checked = set()
    for (_, k, az, el, sd), (res, rgba, hits, _) in zip(reqs, results):
        if k in checked or len(checked) == 3:
            continue
        b, _ = _served_camera(az, el, sd)
        otree = oracle.Tree(tapes[k], grid)
        assert np.array_equal(ctx.export(trees[k]), otree.canonical()), k
        ohits, _, orgba = otree.render(oracle.make_camera(b, 960, 540, seed=sd))
        assert np.array_equal(hits, ohits), k
        assert np.array_equal(rgba.reshape(-1), orgba.reshape(-1)), k
        assert (hits != MISS).sum() > 10000
        checked.add(k)
    assert len(checked) == 3
    for tr in trees.values():
        ctx.tree_free(tr)
    for th in ths.values():
        ctx.tape_free(th)


def test_gyroid_10k_ops_at_4096(ctx, oracle):
    """The heavy transcendental workload: the gyroid lattice smooth-unioned
    with 416 spheres (a 9607-op tape with sin, cos, sqrt, min, max) at
    4096^3 with branching 16/16/16, and its 1920x1080 frame at 4 jittered
    samples per pixel.  The tree is too large for the oracle to build whole:
    level 0 is compared in full (so the node count of level 1 is the
    oracle's), then a sampled descent -- every 16th level-1 node evaluated by
    the oracle, every 16th child of those evaluated at level 2 -- checks
    records found by coords (coords, shortened tape, masks, child entries),
    and every 8th sampled level-2 node's leaf bits against dense point
    sampling of the full 9607-op tape.  The oracle's raytracer renders every
    24th row over the exported tree: hit ids and pixels equal the GPU's."""
    import os
    from concurrent.futures import ThreadPoolExecutor

    tape, grid = shapes.workload("gyroid_4096")
    assert tape.n > 9000
    th = ctx.upload(tape)
    t0 = time.perf_counter()
    tree = ctx.voxelize(th, grid)
    info = ctx.tree_info(tree)
    t1 = time.perf_counter()
    blob = ctx.export(tree)
    t2 = time.perf_counter()
    print(f"gyroid 4096^3: nodes {list(info.node_count[:3])}, build {t1 - t0:.2f} s, "
          f"export {t2 - t1:.2f} s ({blob.nbytes / 2**30:.2f} GiB)", flush=True)
    threads = max(1, min(16, os.cpu_count() or 1))
    bad, rep = oracle.check_blob(tape, grid, blob, n_full=1, sample_every=16, dense_every=8,
                                 threads=threads)
    print(f"oracle check {time.perf_counter() - t2:.2f} s: {rep}", flush=True)
    assert bad == 0, rep
    assert rep["counts"][:2] == [int(v) for v in info.node_count[:2]]
    assert rep["sampled_records"] > 1000 and rep["dense_voxels"] > 1_000_000
    assert int(info.node_count[2]) > 1_000_000

    f = shapes.frame("gyroid_4096")
    b = shapes.camera_basis(f["pos"], f["fwd"], f["up"], 45.0, f["width"], f["height"])
    W, H, spp = f["width"], f["height"], f["spp"]
    cam = lib.make_camera(b, W, H, spp=spp, jitter=True, seed=17)
    rgba, hits, _ = ctx.render(tree, cam, want=("rgba", "hits"))
    ocam = oracle.make_camera(b, W, H, spp=spp, jitter=True, seed=17)
    t3 = time.perf_counter()
    otree = oracle.Tree(tape, grid, blob=blob)
    print(f"oracle tree from blob {time.perf_counter() - t3:.2f} s", flush=True)
    step, first, workers = 24, 5, 8

    def part(i):
        return otree.render_rows(ocam, first + step * i, step * workers)

    t3 = time.perf_counter()
    with ThreadPoolExecutor(workers) as ex:
        parts = list(ex.map(part, range(workers)))
    print(f"oracle tree + rows {time.perf_counter() - t3:.2f} s", flush=True)
    rows = np.arange(first, H, step)
    ohits = np.zeros((H, W, spp), dtype=np.uint64)
    orgba = np.zeros((H, W, 4), dtype=np.uint8)
    for i, (ph, _, pr) in enumerate(parts):
        r = np.arange(first + step * i, H, step * workers)
        ohits[r] = ph.reshape(H, W, spp)[r]
        orgba[r] = pr[r]
    gh = hits.reshape(H, W, spp)
    assert np.array_equal(gh[rows], ohits[rows])
    assert np.array_equal(rgba.reshape(H, W, 4)[rows], orgba[rows])
    assert (gh[rows] != MISS).sum() > 0.2 * rows.size * W * spp
    ctx.tree_free(tree)
    ctx.tape_free(th)
