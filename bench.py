"""Benchmark: voxelization throughput of the 1024^3 random-CSG workload.

One step = one full voxelization (all tree levels, tape shortening included)
of a 24-primitive random CSG tape over a 1024^3 grid (branching 16/8/8) on
one MI355X, through the C ABI, with the tape already resident on the device.
The value is executed voxel-evals per second: one interval evaluation per
cell of every inner-level node plus one point evaluation per voxel of every
last-level node, per step / step time.  The dense-equivalent rate (1024^3
classified voxels per step / step time) is reported beside it as
effective_gvoxel_per_s.

Also reported: per-level kernel times (HIP events on the library's stream),
the roofline fraction of the dominant kernel (algorithmic FP32 flops per
launch / its mean duration / MI355X FP32 vector peak), a 1920x1080 render
rate, and the reference's plain CPU evaluator (dense point sampling,
restated in oracle/) timed on bounded samples on this host, on all threads of
its CPU share and on one thread.
"""
from __future__ import annotations

import argparse
from concurrent.futures import ThreadPoolExecutor
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "f-rep-voxel-rendering_amd"))

FP32_PEAK_TFLOPS = 157.3      # MI355X published FP32 vector peak
HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "voxelize_csg24_1024_hbm.md")
COUNTER_PROFILE = os.path.join(ROOT, "profiles", "voxelize_csg24_1024_valu_counters.md")
WORKLOAD_TEXT = {
    "csg24_1024": "random CSG of 24 primitives (seed 2)",
    "spheres8_256": "union of 8 spheres (seed 1)",
    "gyroid_4096": "gyroid lattice (k = 20) thickened, clipped by a box, smooth-unioned with "
                   "416 spheres (balanced smooth-union tree)",
}


def profiled_traffic(level):
    """HBM bytes per launch of a level kernel of the headline workload, from
    the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the
    copy-kernel calibration); None if there is no such summary."""
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None
    for line in text.splitlines():
        if line.startswith("<!-- traffic:"):
            table = json.loads(line[len("<!-- traffic:"):].rstrip("-> ").strip())
            v = table.get(f"level{level}")
            return None if v is None else int(v)
    return None


def profiled_counters(level):
    """FP32 VALU figures of a level kernel of the headline workload from the
    committed rocprofv3 counter summary: VALU instructions, FP32 add/mul/
    transcendental instructions per dispatch (x 64 lanes = FP32 lane ops),
    and the counter-derived FP32 op rate over the traced mean duration."""
    try:
        text = open(COUNTER_PROFILE).read()
    except OSError:
        return None
    for line in text.splitlines():
        if line.startswith("<!-- counters:"):
            table = json.loads(line[len("<!-- counters:"):].rstrip("-> ").strip())
            v = table.get(f"level{level}")
            if v is not None:
                v = dict(v)
                v["source"] = os.path.relpath(COUNTER_PROFILE, ROOT)
            return v
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(tape, grid, budgets=(("faithful", 0, 10.0), ("tight", 0, 5.0),
                                       ("faithful", 1, 5.0), ("tight", 1, 4.0))):
    """The reference's plain evaluator (SceneBuilder dense sampling of every
    fine voxel with Tape::Eval semantics) on x-slabs of the same grid.
    "faithful" keeps the per-call slot-vector allocation of
    vulkan/src/csg/tape.cpp:276, "tight" reuses one slot array.  Runs on the
    host's full CPU share (at most 16 threads: the share of one GPU on these
    boxes) and on 1 thread; each run times whole x-slabs for about its budget
    in seconds.  The headline figure is the faithful run on all threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    F = 1
    for d in grid["dims"]:
        F *= d
    n_threads = max(1, min(16, os.cpu_count() or 1))
    runs = []
    x = 0
    for mode, single, budget_s in budgets:
        threads = 1 if single else n_threads
        slabs, t_used = 0, 0.0
        while t_used < budget_s and slabs < F:
            x0 = x % F
            x1 = min(F, x0 + threads)
            t0 = time.perf_counter()
            O.dense_occupancy(tape, grid, x0, x1, faithful=(mode == "faithful"),
                              threads=threads)
            t_used += time.perf_counter() - t0
            slabs += x1 - x0
            x = x1
        voxels = slabs * F * F
        runs.append({"mode": mode, "threads": threads,
                     "gvoxel_evals_per_s": round(voxels / t_used / 1e9, 5),
                     "sample": f"{slabs} of {F} x-slabs, {voxels} voxels, {t_used:.1f} s"})
    head = runs[0]
    return {"value": head["gvoxel_evals_per_s"], "unit": "Gvoxel-evals/s",
            "cores": head["threads"], "kind": "port",
            "sample": f"{head['sample']}: dense Tape::Eval of the {tape.n}-op tape at every "
                      f"voxel of the {F}^3 grid, per-call slot allocation",
            "cpu_model": cpu_model(), "host_threads": os.cpu_count(), "runs": runs}


def bench_batched(ctx, shapes, reps=5, warm=8):
    """The 64-shape mix (sphere unions of 4..400 primitives at 128^3..1024^3):
    requests/s of one frep_voxelize per shape (plain) against one
    frep_voxelize_batch over all 64 (batched), same trees, same results.
    The untimed batched calls grow every tree's buffers, then run the
    autotuner's timed trials and capture the tuned batch's graph."""
    mix = shapes.batch_mix(64)
    ths = [ctx.upload(t) for t, _, _ in mix]
    grids = [g for _, g, _ in mix]
    trees = [ctx.tree_create() for _ in mix]
    for _ in range(warm):
        ctx.voxelize_batch(ths, grids, trees)
    for th, g, tr in zip(ths, grids, trees):
        ctx.voxelize(th, g, tr)
    t0 = time.perf_counter()
    for _ in range(reps):
        for th, g, tr in zip(ths, grids, trees):
            ctx.voxelize(th, g, tr)
    plain = (time.perf_counter() - t0) / reps
    ctx.set_timing(True)
    bms = 0.0
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.voxelize_batch(ths, grids, trees)
        bms += ctx.timing().batch_ms
    batched = (time.perf_counter() - t0) / reps
    ctx.set_timing(False)
    voxels = float(sum(r ** 3 for _, _, r in mix))
    out = {
        "workload": "64 sphere unions, 4..400 primitives (log-uniform, seeds 1000+k), "
                    "resolutions {128,256,512,1024}^3",
        "tape_ops_total": sum(t.n for t, _, _ in mix),
        "plain_req_per_s": round(len(mix) / plain, 2),
        "batched_req_per_s": round(len(mix) / batched, 2),
        "speedup": round(plain / batched, 3),
        "batched_ms": round(batched * 1e3, 3),
        "batched_gpu_ms": round(bms / reps, 3),
        "batched_gvoxel_evals_per_s": round(voxels / batched / 1e9, 2),
    }
    for tr in trees:
        ctx.tree_free(tr)
    for th in ths:
        ctx.tape_free(th)
    return out


def bench_served(ctx, lib, shapes, rates=(50, 100, 200, 500, 1000, 2000, 3000, 4000, 6000, 8000),
                 size=(960, 540), burst=512, p99_cap_ms=50.0):
    """Served workload: Poisson arrivals of voxelize+render requests (new
    24-primitive random CSG scenes at 1024^3, or with p = 0.5 a re-render of
    a resident one, 960x540 orbit views), swept from 50 req/s past
    saturation.  Each rate's stream (200..800 requests) is replayed in real
    time through frep_server (continuous batching).  The same stream runs
    through the plain path (frep_voxelize + frep_render per request, one at a
    time); its latencies follow from its measured service times and the same
    arrival times (a FIFO queue).  Both report sustained throughput
    (completed / (last completion - first arrival)) and p50/p99 latency at
    every rate, and the capacity of each path is the highest throughput it
    sustains at p99 <= p99_cap_ms over the swept rates.  The closed loop
    submits `burst` requests at once to each path."""
    grid = shapes.make_grid([16, 8, 8])
    tapes = {}

    def tape(k):
        if k not in tapes:
            tapes[k] = ctx.upload(shapes.random_csg(24, 10000 + k))
        return tapes[k]

    def camera(az, el, seed):
        b = shapes.orbit_basis(az, el, 3.0, size[0], size[1])
        return lib.make_camera(b, size[0], size[1], seed=seed)

    def pct(v, q):
        v = sorted(v)
        return v[min(len(v) - 1, max(0, int(math.ceil(q * len(v))) - 1))]

    def served_run(reqs, realtime):
        srv = lib.Server(ctx, budget_bytes=64 << 30)
        handles = {}
        for _, k, _, _, _ in reqs:
            if k not in handles:
                handles[k] = srv.scene(tape(k), grid)
        cams = [camera(az, el, sd) for _, _, az, el, sd in reqs]
        # clients: this thread submits on the arrival schedule, a pool of
        # threads collects results (each wait copies one image out)
        pool = ThreadPoolExecutor(max_workers=8)
        done = []

        def collect(t):
            r = srv.wait(t)
            done.append(time.perf_counter())
            return r

        futures = []
        t0 = time.perf_counter()
        for (arr, k, _, _, _), cam in zip(reqs, cams):
            if realtime:
                d = t0 + arr - time.perf_counter()
                if d > 0:
                    time.sleep(d)
            futures.append(pool.submit(collect, srv.submit(handles[k], cam)))
        for f in futures:
            f.result()
        span = max(done) - t0 - (reqs[0][0] if realtime else 0.0)
        pool.shutdown()
        st = srv.stats()
        res = {"req_per_s": round(st.completed / span, 2), "p50_ms": round(st.p50_ms, 3),
               "p99_ms": round(st.p99_ms, 3), "mean_batch": round(st.mean_batch, 2),
               "builds": int(st.voxelizations), "failed": int(st.failed)}
        srv.close()
        return res

    def plain_service(reqs):
        trees = {}
        service = []
        for _, k, az, el, sd in reqs:
            cam = camera(az, el, sd)
            t0 = time.perf_counter()
            if k not in trees:
                trees[k] = ctx.voxelize(tape(k), grid)
            ctx.render(trees[k], cam, want=("rgba",))
            service.append(time.perf_counter() - t0)
        for tr in trees.values():
            ctx.tree_free(tr)
        return service

    def plain_queue(reqs, service):
        finish, lat = 0.0, []
        for (arr, _, _, _, _), sv in zip(reqs, service):
            finish = max(arr, finish) + sv
            lat.append((finish - arr) * 1e3)
        span = finish - reqs[0][0]
        return {"req_per_s": round(len(reqs) / span, 2), "p50_ms": round(pct(lat, 0.5), 3),
                "p99_ms": round(pct(lat, 0.99), 3)}

    # warm both paths (kernels, allocator, capacity hints)
    warm = shapes.served_requests(64, seed=3, rate=1e6)
    served_run(warm, False)
    plain_service(warm[:16])
    per_rate = []
    cap = {"served": 0.0, "plain": 0.0}
    for rate in rates:
        n = int(min(800, max(200, rate)))
        reqs = shapes.served_requests(n, seed=7, rate=rate)
        served = served_run(reqs, True)
        plain = plain_queue(reqs, plain_service(reqs))
        per_rate.append({"rate": rate, "requests": n, "served": served, "plain": plain})
        for key, r in (("served", served), ("plain", plain)):
            if r["p99_ms"] <= p99_cap_ms:
                cap[key] = max(cap[key], r["req_per_s"])
    reqs = shapes.served_requests(burst, seed=11, rate=1e6)
    served = served_run(reqs, False)
    plain_sv = plain_service(reqs)
    for th in tapes.values():
        ctx.tape_free(th)
    return {
        "workload": "voxelize (24-primitive random CSG, 1024^3, branching 16/8/8) + render "
                    "960x540 orbit view per request; p = 0.5 re-render of a resident scene",
        "per_rate": per_rate,
        "capacity": {"p99_cap_ms": p99_cap_ms, "served_req_per_s": cap["served"],
                     "plain_req_per_s": cap["plain"],
                     "speedup": round(cap["served"] / cap["plain"], 3) if cap["plain"] else None,
                     "basis": "highest sustained throughput (completed / (last completion - "
                              "first arrival)) over the swept Poisson rates with p99 <= the cap"},
        "closed_loop": {"requests": burst, "served_req_per_s": served["req_per_s"],
                        "plain_req_per_s": round(len(reqs) / sum(plain_sv), 2),
                        "speedup": round(served["req_per_s"] * sum(plain_sv) / len(reqs), 3),
                        "served_p99_ms": served["p99_ms"], "mean_batch": served["mean_batch"]},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="csg24_1024")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-render", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the voxelizer runs on one GPU; --gpus {args.gpus} is not supported "
              f"(trees would be sharded by level-0 cell across GPUs, which is not built)",
              file=sys.stderr)
        sys.exit(2)

    from frep import lib, shapes

    tape, grid = shapes.workload(args.workload)
    ctx = lib.Context(0)
    th = ctx.upload(tape)
    tree = ctx.tree_create()
    for _ in range(max(1, args.warmup)):
        ctx.voxelize(th, grid, tree)
    info = ctx.tree_info(tree)
    nl = len(grid["dims"])

    # timed region: the plain path as a caller runs it (each call returns
    # after its tree is complete; no per-launch events)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.voxelize(th, grid, tree)
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    # per-level kernel times: a second pass with HIP events around each launch
    ctx.set_timing(True)
    level_ms = [0.0] * nl
    for _ in range(args.steps):
        ctx.voxelize(th, grid, tree)
        tm = ctx.timing()
        for L in range(nl):
            level_ms[L] += tm.level_ms[L]
    ctx.set_timing(False)
    level_ms = [v / args.steps for v in level_ms]

    F = 1
    for d in grid["dims"]:
        F *= d
    info = ctx.tree_info(tree)
    # executed voxel-evals: every cell of every node of every level is one
    # interval (inner levels) or point (last level) evaluation
    evals = [int(info.node_count[L]) * int(grid["dims"][L]) ** 3 for L in range(nl)]
    value = sum(evals) / (ms_per_step * 1e-3) / 1e9
    dense_equivalent = float(F) ** 3 / (ms_per_step * 1e-3) / 1e9

    work = [{"flops": int(info.work_flops[L]), "op_evals": int(info.op_evals[L]),
             "kind": "point" if L == nl - 1 else "interval"} for L in range(nl)]
    dom = max(range(nl), key=lambda L: level_ms[L])
    achieved = work[dom]["flops"] / (level_ms[dom] * 1e-3) / 1e12
    out = {
        "metric": "Gvoxel-evals/s (1024^3, 200-op tape); batched vs plain req/s; p50/p99 latency",
        "metric_note": "the metric's name is the configuration's; the headline tape compiles to "
                       f"{tape.n} ops, and value counts executed voxel-evals (value_basis), the "
                       "dense-equivalent rate is effective_gvoxel_per_s",
        "value": round(value, 3),
        "unit": "Gvoxel-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {WORKLOAD_TEXT[args.workload]}, "
                        f"{tape.n}-op tape, {F}^3 grid, branching {grid['dims']}",
            "tape_ops": tape.n, "slots": tape.slot_count,
            "nodes_per_level": [int(info.node_count[L]) for L in range(nl)],
            "level_ms": [round(v, 4) for v in level_ms],
            "executed_op_evals": sum(w["op_evals"] for w in work),
        },
        "value_basis": "executed voxel-evals per second: interval evaluations of every cell "
                       "of every inner-level node plus point evaluations of every voxel of "
                       "every last-level node",
        "executed_voxel_evals_per_step": sum(evals),
        "executed_voxel_evals_per_level": evals,
        "effective_gvoxel_per_s": round(dense_equivalent, 3),
        "effective_basis": f"dense-equivalent: {F}^3 voxels classified per step / step time",
        "roofline": {
            "bound": "valu",
            "kernel": f"level {dom} ({work[dom]['kind']} evaluation)",
            "achieved": round(achieved, 4),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 5),
            "traffic": profiled_traffic(dom) if args.workload == "csg24_1024" else None,
            "traffic_source": "profiles/voxelize_csg24_1024_hbm.md (rocprofv3 FETCH_SIZE, "
                              "WRITE_SIZE per dispatch x copy-kernel calibration)",
            "algorithmic_flops_per_launch": work[dom]["flops"],
            "per_level": [{"level": L, "kind": work[L]["kind"], "ms": round(level_ms[L], 4),
                           "flops": work[L]["flops"],
                           "tflops": round(work[L]["flops"] / (level_ms[L] * 1e-3) / 1e12, 4),
                           "frac": round(work[L]["flops"] / (level_ms[L] * 1e-3) / 1e12
                                         / FP32_PEAK_TFLOPS, 5)} for L in range(nl)],
        },
    }
    counters = profiled_counters(dom) if args.workload == "csg24_1024" else None
    if counters:
        out["roofline"]["counters"] = counters
    if not args.no_render:
        fr = shapes.frame(args.workload)
        w, h = fr["width"], fr["height"]
        basis = shapes.camera_basis(fr["pos"], fr["fwd"], fr["up"], 45.0, w, h,
                                    ortho=fr["ortho"], extent=fr.get("extent", (2.0, 2.0)))
        cam = lib.make_camera(basis, w, h, spp=fr["spp"], ortho=fr["ortho"],
                              jitter=fr["jitter"], seed=1234)
        ctx.render(tree, cam, want=("rgba",))
        ctx.set_timing(True)
        t0 = time.perf_counter()
        reps = 5
        rms = 0.0
        for _ in range(reps):
            ctx.render(tree, cam, want=("rgba",))
            rms += ctx.timing().render_ms
        rwall = (time.perf_counter() - t0) / reps
        ctx.set_timing(False)
        kind = "orthographic" if fr["ortho"] else "perspective"
        jit = ", jittered" if fr["jitter"] else ""
        out["render"] = {"frame": f"{w}x{h} {kind}, {fr['spp']} spp{jit}",
                         "kernel_ms": round(rms / reps, 4),
                         "frames_per_s_incl_copy": round(1.0 / rwall, 2)}
    if not args.no_batched:
        out["batched"] = bench_batched(ctx, shapes)
    if not args.no_served:
        out["served"] = bench_served(ctx, lib, shapes)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(tape, grid)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
