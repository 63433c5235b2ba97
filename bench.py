"""Benchmark: Mrays/s of the MI355X traversal path on the incoherent-ray workload.

Workload (one "step" = one frame): a 1,310,720-triangle displaced icosphere
(frequency 256, radius 32, 4-octave value-noise displacement of 0.15 R),
BVH built on the GPU (binned SAH, 8 bins, leaf size <= 2), 1920x1080 pinhole
primary rays in 8x8-tile order, nearest-hit trace, one cosine-weighted
diffuse bounce per hit (PCG32-seeded, origin offset 1e-4 along the normal;
generated grouped by direction bin with each ray's source pixel, see
bvht_bounce_sorted), nearest-hit trace of the bounce rays.  Inputs are resident in HBM before the
timed region; every step regenerates its rays on the GPU.  The timed steps
replay the frame captured once as a hipGraph (the frame's calls only enqueue
kernels); a second set of eager frames with HIP events between the launches
gives the per-kernel breakdown.

value = (primary + bounce rays) / frame time.  Primary and bounce Mrays/s are
reported separately from their own kernel times.  The roofline object prices
the dominant kernel, k_trace_blas (two launches per frame), at its
algorithmic bytes per launch -- per ray 28 + 20 + 32 + 64 P + 36 T (ray, hit,
root node, P child-pair fetches, T triangle tests; P and T from the
instrumented trace of the same rays), averaged over the primary and bounce
launches -- divided by the mean launch duration from HIP events, against
8 TB/s; `traffic` is the measured HBM bytes per launch from rocprofv3
FETCH_SIZE/WRITE_SIZE (profiles/).  The CPU baseline is the oracle (C
restatement of the reference's traversal) on the same BVH and the same
frame's rays, on up to 16 host threads, repeated to fill --cpu-seconds.

Extras (not part of `value`): 16 instances with per-frame deform/refit/TLAS
rebuild in one hipGraph; 64 uneven scenes traced by one persistent launch vs
64 plain launches; and the differentiable depth/shading loop (forward +
adjoint + SGD + refit, 100 steps on the 81,920-face icosphere); and the
reference's own build/refit benchmarks on its unity.tri and bigben.tri.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--width", type=int, default=1920)
    p.add_argument("--height", type=int, default=1080)
    p.add_argument("--freq", type=int, default=256)
    p.add_argument("--seed", type=int, default=0xB0B)
    p.add_argument("--cpu-seconds", type=float, default=10.0, help="budget for the CPU baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "k_trace_blas_sphere1m_traffic.json"),
                   help="measured HBM bytes per launch of the dominant kernel (rocprofv3 --pmc summary)")
    p.add_argument("--no-autotune", action="store_true", help="use the default trace configuration")
    p.add_argument("--trace-config", default=None, metavar="WGxLDS",
                   help="fix the trace kernel configuration (e.g. 64x24) instead of autotuning")
    p.add_argument("--no-extras", action="store_true", help="skip the batched and instanced workloads")
    p.add_argument("--batch-scenes", type=int, default=64)
    p.add_argument("--instanced-frames", type=int, default=64)
    p.add_argument("--adjoint-steps", type=int, default=100)
    p.add_argument("--only", choices=["adjoint", "instanced", "batched", "meshes", "physics"], default=None,
                   help="run one extra workload alone and print its JSON")
    p.add_argument("--copy-only", type=int, default=0, metavar="BYTES",
                   help="only run the stream-copy kernel 5 times on BYTES (profiler counter calibration)")
    return p.parse_args()


def stream_copy_gbs(ctx, nbytes=1 << 30, reps=10):
    """Measured device copy bandwidth (read + write bytes / time)."""
    from bvhtracer_amd import _lib as L

    a, b = ctx.alloc(nbytes), ctx.alloc(nbytes)
    L.check(L.lib().bvht_stream_copy(ctx.handle, b.ptr, a.ptr, nbytes))
    e0, e1 = ctx.event(), ctx.event()
    ctx.sync()
    ctx.record(e0)
    for _ in range(reps):
        L.check(L.lib().bvht_stream_copy(ctx.handle, b.ptr, a.ptr, nbytes))
    ctx.record(e1)
    ms = ctx.elapsed_ms(e0, e1)
    a.free()
    b.free()
    return 2.0 * nbytes * reps / (ms * 1e-3) / 1e9


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); every GPU box here has one GPU",
              file=sys.stderr)
        sys.exit(2)
    import bvhtracer_amd as B
    from bvhtracer_amd import scenes

    if args.copy_only:
        from bvhtracer_amd import _lib as L

        ctx = B.Context(0)
        a, b = ctx.alloc(args.copy_only), ctx.alloc(args.copy_only)
        for _ in range(5):
            L.check(L.lib().bvht_stream_copy(ctx.handle, b.ptr, a.ptr, args.copy_only))
        ctx.sync()
        print(json.dumps({"copy_bytes": args.copy_only, "launches": 5}))
        return
    if args.only:
        fn = {"adjoint": bench_adjoint, "instanced": bench_instanced, "batched": bench_batched,
              "meshes": bench_meshes, "physics": bench_physics}[args.only]
        print(json.dumps({args.only: fn(args, B.Context(0))}), flush=True)
        return

    W, H = args.width, args.height
    n_primary = W * H
    t0 = time.time()
    mesh = scenes.icosphere(args.freq, radius=32.0)
    log(f"[bench] icosphere {mesh.shape[0]} triangles generated in {time.time() - t0:.1f}s")
    ctx = B.Context(0)
    t0 = time.time()
    m = mesh.copy()
    bvh = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(m)
    build_s = time.time() - t0
    log(f"[bench] GPU build {build_s * 1e3:.1f} ms, nodes_used {bvh.nodes_used()}, depth {bvh.depth()}")
    cam = scenes.sphere_camera(32.0, W, H)

    d_rays = ctx.alloc(n_primary * B.RAY_DTYPE.itemsize)
    d_hits = ctx.alloc(n_primary * B.HIT_DTYPE.itemsize)
    d_brays = ctx.alloc(n_primary * B.RAY_DTYPE.itemsize)
    d_bhits = ctx.alloc(n_primary * B.HIT_DTYPE.itemsize)
    d_bsrc = ctx.alloc(n_primary * 4)
    d_count = ctx.alloc(4)
    import ctypes

    from bvhtracer_amd import _lib as L

    lib = L.lib()
    cam_rec = np.ascontiguousarray(cam)

    def frame(seed, evs=None):
        if evs:
            ctx.record(evs[0])
        L.check(lib.bvht_raygen(ctx.handle, L.ptr(cam_rec), W, H, B.ORDER_TILES, d_rays.ptr))
        if evs:
            ctx.record(evs[1])
        L.check(lib.bvht_trace_blas(ctx.handle, bvh.handle, d_rays.ptr, n_primary, d_hits.ptr, None))
        if evs:
            ctx.record(evs[2])
        L.check(lib.bvht_bounce_sorted(ctx.handle, bvh.handle, d_rays.ptr, d_hits.ptr, n_primary, seed, d_brays.ptr,
                                       d_bsrc.ptr, d_count.ptr))
        if evs:
            ctx.record(evs[3])
        L.check(lib.bvht_trace_blas_count(ctx.handle, bvh.handle, d_brays.ptr, d_count.ptr, n_primary,
                                          d_bhits.ptr, None))
        if evs:
            ctx.record(evs[4])

    for w in range(args.warmup):
        frame(args.seed + w)
    ctx.sync()
    n_bounce = int(d_count.download(np.uint32, 1)[0])
    log(f"[bench] warmup done; bounce rays per frame ~{n_bounce}")
    tuned = None
    if args.trace_config:
        from bvhtracer_amd import autotune

        wg, d = (int(x) for x in args.trace_config.lower().split("x"))
        autotune.set_config(ctx, wg, d)
        tuned = {"workgroup": wg, "lds_stack_depth": d, "fixed": True}
    elif not args.no_autotune:
        from bvhtracer_amd import autotune

        def traces():
            L.check(lib.bvht_trace_blas(ctx.handle, bvh.handle, d_rays.ptr, n_primary, d_hits.ptr, None))
            L.check(lib.bvht_trace_blas(ctx.handle, bvh.handle, d_brays.ptr, n_bounce, d_bhits.ptr, None))

        best, table = autotune.tune(ctx, traces, reps=3, log=log)
        tuned = {"workgroup": best[0], "lds_stack_depth": best[1],
                 "sweep_ms": {f"{k[0]}x{k[1]}": round(v, 4) for k, v in table.items()}}
        log(f"[bench] autotuned trace config: workgroup {best[0]}, LDS depth {best[1]}")

    # per-kernel breakdown: K eager frames with HIP events between the launches
    evs = [[ctx.event() for _ in range(5)] for _ in range(args.steps)]
    start, stop = ctx.event(), ctx.event()
    ctx.sync()
    ctx.record(start)
    for s in range(args.steps):
        frame(args.seed, evs[s])
    ctx.record(stop)
    ctx.sync()
    eager_ms = ctx.elapsed_ms(start, stop) / args.steps
    per = np.array([[ctx.elapsed_ms(e[k], e[k + 1]) for k in range(4)] for e in evs])
    raygen_ms, prim_ms, bounce_gen_ms, bounce_ms = per.mean(axis=0)
    # timed region: the frame captured once as a hipGraph, K replays
    graph = B.Graph.capture(ctx, lambda: frame(args.seed))
    graph.launch()
    ctx.sync()
    start, stop = ctx.event(), ctx.event()
    ctx.record(start)
    for s in range(args.steps):
        graph.launch()
    ctx.record(stop)
    ctx.sync()
    ms_per_step = ctx.elapsed_ms(start, stop) / args.steps
    graph.close()
    n_bounce = int(d_count.download(np.uint32, 1)[0])
    rays_per_step = n_primary + n_bounce
    value = rays_per_step / (ms_per_step * 1e-3) / 1e6
    log(f"[bench] {args.steps} steps: {ms_per_step:.3f} ms/frame as a graph, {eager_ms:.3f} eager (raygen "
        f"{raygen_ms:.3f}, primary {prim_ms:.3f}, bounce-gen {bounce_gen_ms:.3f}, bounce {bounce_ms:.3f})")

    # algorithmic bytes per bounce ray from the instrumented trace of the same rays (untimed)
    d_stats = ctx.alloc(n_primary * B.STATS_DTYPE.itemsize)
    L.check(lib.bvht_trace_blas(ctx.handle, bvh.handle, d_brays.ptr, n_bounce, d_bhits.ptr, d_stats.ptr))
    st = d_stats.download(B.STATS_DTYPE, n_bounce)
    P, T = st["pairs"].astype(np.float64).mean(), st["tests"].astype(np.float64).mean()
    b_ray = 28 + 20 + 32 + 64 * P + 36 * T
    L.check(lib.bvht_trace_blas(ctx.handle, bvh.handle, d_rays.ptr, n_primary, d_hits.ptr, d_stats.ptr))
    stp = d_stats.download(B.STATS_DTYPE, n_primary)
    Pp, Tp = stp["pairs"].astype(np.float64).mean(), stp["tests"].astype(np.float64).mean()
    b_ray_p = 28 + 20 + 32 + 64 * Pp + 36 * Tp
    # dominant kernel: k_trace_blas, two launches per frame (primary, bounce);
    # average algorithmic bytes per launch over the average launch duration
    bytes_per_launch = (b_ray_p * n_primary + b_ray * n_bounce) / 2.0
    avg_launch_ms = (prim_ms + bounce_ms) / 2.0
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    bounce_achieved = b_ray * n_bounce / (bounce_ms * 1e-3) / 1e9
    copy_gbs = stream_copy_gbs(ctx)
    log(f"[bench] stream copy {copy_gbs:.0f} GB/s")
    traffic = None
    if args.traffic and os.path.exists(args.traffic):
        with open(args.traffic) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    result = {
        "metric": "Mrays/s primary+diffuse on 1 MI355X",
        "value": round(value, 2),
        "unit": "Mrays/s",
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
            "workload": f"displaced icosphere {mesh.shape[0]} tris (freq {args.freq}, R 32), {W}x{H} primary + "
                        f"1 cosine diffuse bounce per hit (grouped by direction), nearest hit, SAH BVH (8 bins, "
                        f"leaf<=2) built on GPU",
            "primary_rays": n_primary,
            "bounce_rays": n_bounce,
        },
        "primary_mrays_s": round(n_primary / (prim_ms * 1e-3) / 1e6, 2),
        "bounce_mrays_s": round(n_bounce / (bounce_ms * 1e-3) / 1e6, 2),
        "eager_ms_per_step": round(eager_ms, 4),
        "frame_breakdown_ms": {"raygen": round(raygen_ms, 4), "primary_trace": round(prim_ms, 4),
                               "bounce_gen": round(bounce_gen_ms, 4), "bounce_trace": round(bounce_ms, 4)},
        "blas_build_ms": round(build_s * 1e3, 1),
        "trace_config": tuned,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_trace_blas (primary and bounce launches, averaged per launch)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": round(bytes_per_launch, 0),
            "avg_launch_ms": round(avg_launch_ms, 4),
            "bounce": {"bytes_per_ray": round(b_ray, 1), "pairs_per_ray": round(P, 2),
                       "tests_per_ray": round(T, 3), "achieved": round(bounce_achieved, 1),
                       "frac": round(bounce_achieved / HBM_PEAK_GBS, 4)},
            "primary": {"bytes_per_ray": round(b_ray_p, 1), "pairs_per_ray": round(Pp, 2),
                        "tests_per_ray": round(Tp, 3),
                        "frac": round(b_ray_p * n_primary / (prim_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
            "stream_copy_gbs": round(copy_gbs, 1),
            "frac_of_stream_copy": round(achieved / copy_gbs, 4),
        },
        "cpu_baseline": None,
    }

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(ctx, bvh, d_brays, d_bhits, n_bounce, d_rays, d_hits, n_primary,
                                              args.cpu_seconds)
    if not args.no_extras:
        for name, fn in (("instanced", bench_instanced), ("batched", bench_batched), ("adjoint", bench_adjoint),
                         ("meshes", bench_meshes), ("physics", bench_physics)):
            try:
                result[name] = fn(args, ctx)
            except Exception as e:  # a failing extra must not hide the main measurement
                result[name] = {"error": f"{type(e).__name__}: {e}"}
                log(f"[bench] {name} failed: {e}")
    print(json.dumps(result), flush=True)


def bench_instanced(args, ctx):
    """16 instances of a 200k-triangle displaced icosphere; per frame: deform
    the original vertices (y += 0.2 sin(4x + 0.1 frame)), BLAS refit, object
    bounds, TLAS rebuild, 1280x720 primary + 1 bounce -- one hipGraph per
    frame, replayed for --instanced-frames frames."""
    import bvhtracer_amd as B
    from bvhtracer_amd import _lib as L
    from bvhtracer_amd import scenes

    lib = L.lib()
    W, H = 1280, 720
    n = W * H
    mesh = scenes.icosphere(100, radius=4.0)
    m = mesh.copy()
    bvh = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(m)
    M, Minv = scenes.random_affines(16, seed=0x16, trans_range=20.0)
    objs = [B.SceneObject(0, M[k].reshape(4, 4), Minv[k]) for k in range(16)]
    tlas = B.TlasBuilder(ctx=ctx).build_for(objs, [bvh])
    tlas.rebuild_async()
    batch = B.Batch(ctx, [tlas])
    cam = np.ascontiguousarray(scenes.camera(eye=(0, 0, -60), forward=(0, 0, 1), right=(1, 0, 0), up=(0, 1, 0),
                                             fovy_deg=45, aspect=W / H))
    d_orig = ctx.to_device(m)
    d_frame = ctx.to_device(np.array([0], np.uint32))
    d_rays, d_hits = ctx.alloc(n * 32), ctx.alloc(n * 24)
    d_brays, d_bhits, d_cnt = ctx.alloc(n * 32), ctx.alloc(n * 24), ctx.alloc(4)

    def update():
        B.deform_wave(ctx, bvh, d_orig, 0.2, 4.0, 0.1, d_frame)
        batch.refit()
        batch.rebuild()

    def trace():
        L.check(lib.bvht_raygen(ctx.handle, L.ptr(cam), W, H, B.ORDER_TILES, d_rays.ptr))
        L.check(lib.bvht_trace_tlas(ctx.handle, tlas.handle, d_rays.ptr, n, d_hits.ptr, None))
        L.check(lib.bvht_bounce_tlas(ctx.handle, tlas.handle, d_rays.ptr, d_hits.ptr, n, 0xB0B, d_brays.ptr,
                                     d_cnt.ptr))
        L.check(lib.bvht_trace_tlas_count(ctx.handle, tlas.handle, d_brays.ptr, d_cnt.ptr, n, d_bhits.ptr, None))

    def frame():
        update()
        trace()
        B.counter_add(ctx, d_frame, 1)

    frame()  # allocates scratch / spill before capture
    ctx.sync()
    e = [ctx.event() for _ in range(3)]
    ctx.record(e[0])
    update()
    ctx.record(e[1])
    trace()
    ctx.record(e[2])
    update_ms, trace_ms = ctx.elapsed_ms(e[0], e[1]), ctx.elapsed_ms(e[1], e[2])
    graph = B.Graph.capture(ctx, frame)
    graph.launch()
    ctx.sync()
    t0, t1 = ctx.event(), ctx.event()
    ctx.record(t0)
    for _ in range(args.instanced_frames):
        graph.launch()
    ctx.record(t1)
    ms = ctx.elapsed_ms(t0, t1) / args.instanced_frames
    nb = int(d_cnt.download(np.uint32, 1)[0])
    graph.close()
    return {"workload": "16 instances x 200,000-triangle icosphere, per-frame deform + refit + bounds + TLAS "
                        "rebuild + 1280x720 primary + 1 bounce, one hipGraph per frame",
            "frames": args.instanced_frames, "ms_per_frame": round(ms, 4),
            "mrays_s": round((n + nb) / (ms * 1e-3) / 1e6, 2), "bounce_rays": nb,
            "eager_update_ms": round(update_ms, 4), "eager_trace_ms": round(trace_ms, 4)}


def bench_batched(args, ctx):
    """64 uneven scenes (triangles log-uniform in [1e3, 1e6], 1-8 instances,
    rays log-uniform in [4e3, 2e6]): one frame = deform + refit + bounds +
    TLAS rebuild + trace of every scene.  Batched: one launch per stage (one
    persistent trace fed by an atomic queue), captured in one hipGraph.
    Plain: the same stages launched scene by scene.  Hits are compared."""
    import bvhtracer_amd as B
    from bvhtracer_amd import _lib as L
    from bvhtracer_amd import autotune, scenes

    lib = L.lib()
    t0 = time.time()
    work = scenes.batch_workload(args.batch_scenes)
    log(f"[bench] batched workload generated in {time.time() - t0:.1f}s")
    models, tlases, d_origs = [], [], []
    for w in work:
        m = w["mesh"].copy()
        bvh = B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(m)
        objs = [B.SceneObject(0, w["M"][k].reshape(4, 4), w["Minv"][k]) for k in range(w["M"].shape[0])]
        t = B.TlasBuilder(ctx=ctx).build_for(objs, [bvh])
        t.rebuild_async()
        models.append(bvh)
        tlases.append(t)
        d_origs.append(ctx.to_device(m))
    rays = np.concatenate([w["rays"] for w in work])
    offsets = np.concatenate([[0], np.cumsum([w["rays"].shape[0] for w in work])]).astype(np.uint32)
    n = int(offsets[-1])
    d_rays = ctx.to_device(rays)
    d_hits, d_hits_plain = ctx.alloc(n * 24), ctx.alloc(n * 24)
    d_frame = ctx.to_device(np.array([0], np.uint32))
    batch = B.Batch(ctx, tlases)
    batch.set_rays(offsets, 128)
    best, table = autotune.tune_batch(batch, offsets, lambda: batch.trace(d_rays, d_hits), reps=2)
    log(f"[bench] batched trace config {best}")

    def deform():
        for bvh, d_o in zip(models, d_origs):
            B.deform_wave(ctx, bvh, d_o, 0.2, 4.0, 0.1, d_frame)

    def batched_frame():
        deform()
        batch.refit()
        batch.rebuild()
        batch.trace(d_rays, d_hits)

    def plain_frame():
        deform()
        for s, (bvh, t) in enumerate(zip(models, tlases)):
            L.check(lib.bvht_blas_refit(ctx.handle, bvh.handle))
            L.check(lib.bvht_tlas_refresh_bounds(ctx.handle, t.handle))
            L.check(lib.bvht_tlas_rebuild_async(ctx.handle, t.handle))
            cnt = int(offsets[s + 1] - offsets[s])
            L.check(lib.bvht_trace_tlas(ctx.handle, t.handle, d_rays.offset(32 * int(offsets[s])), cnt,
                                        d_hits_plain.offset(24 * int(offsets[s])), None))

    batched_frame()
    plain_frame()
    ctx.sync()
    graph = B.Graph.capture(ctx, batched_frame)

    def timed(fn, reps):
        fn()
        a, b = ctx.event(), ctx.event()
        ctx.sync()
        ctx.record(a)
        for _ in range(reps):
            fn()
        ctx.record(b)
        return ctx.elapsed_ms(a, b) / reps

    reps = 5
    plain_ms = timed(plain_frame, reps)
    batched_ms = timed(graph.launch, reps)
    trace_only_ms = timed(lambda: batch.trace(d_rays, d_hits), reps)
    ctx.sync()
    same = d_hits.download(np.uint8).tobytes() == d_hits_plain.download(np.uint8).tobytes()
    graph.close()
    return {"workload": f"{len(work)} scenes, {sum(w['mesh'].shape[0] for w in work)} triangles, "
                        f"{sum(w['M'].shape[0] for w in work)} instances, {n} rays (half camera, half random)",
            "batched_ms_per_frame": round(batched_ms, 4), "plain_ms_per_frame": round(plain_ms, 4),
            "batched_mrays_s": round(n / (batched_ms * 1e-3) / 1e6, 2),
            "plain_mrays_s": round(n / (plain_ms * 1e-3) / 1e6, 2),
            "speedup": round(plain_ms / batched_ms, 3), "batched_trace_only_ms": round(trace_only_ms, 4),
            "hits_identical": bool(same),
            "trace_config": {"workgroup": best[0], "lds_stack_depth": best[1], "rays_per_item": best[2]}}


def bench_adjoint(args, ctx):
    """Differentiable depth/shading: the 81,920-face icosphere (frequency 64,
    R 32) against a target with seeded radial vertex perturbation (sigma
    0.05) rendered from 8 seeded cameras at 512x512 (2,097,152 rays).  One
    step = forward (trace + shade), adjoint (dLoss/dvertex, dLoss/dM^-1,
    loss), SGD on the vertices, gather + BVH refit + vertex normals; the step
    is captured once as a hipGraph and replayed for --adjoint-steps steps.
    fwd+bwd Mrays/s = rays / (forward + adjoint time) from eager per-phase
    events; the adjoint's algorithmic bytes per contributing ray are
    32 ray + 4 hit id + 8 target + 36 triangle + 12 vertex ids + 36 normals
    + 36 scattered gradient (9 float atomics), 44 B for the others."""
    import bvhtracer_amd as B
    from bvhtracer_amd import scenes

    w = scenes.diff_workload(64, width=512, n_cameras=8)
    n1 = 512 * 512
    n = 8 * n1
    d_rays = ctx.alloc(n * B.RAY_DTYPE.itemsize)
    for k in range(8):
        B.raygen(ctx, w["cameras"][k:k + 1], 512, 512, B.ORDER_TILES,
                 d_rays=B.DeviceArray.view(ctx, d_rays.value + k * n1 * 32, n1 * 32))

    def make(V):
        soup = np.ascontiguousarray(V[w["F"]].reshape(-1, 9))
        return B.Mesh(B.BvhBuilder(B.BUILD_SAH, 2, ctx=ctx).build_for(soup), V, w["F"])

    target_mesh = make(w["V_target"])
    d_hits, d_img, d_target = ctx.alloc(n * 24), ctx.alloc(n * 8), ctx.alloc(n * 8)
    B.trace_forward(target_mesh, w["Minv"], w["light"], d_rays, n, d_hits, d_target)
    mesh = make(w["V"])
    d_gv, d_gm, d_loss = ctx.alloc(mesh.n_vertices * 12), ctx.alloc(64), ctx.alloc(8)
    lr, clip = scenes.DIFF_LR, scenes.DIFF_MAX_STEP

    def fwd():
        B.trace_forward(mesh, w["Minv"], w["light"], d_rays, n, d_hits, d_img)

    def adj():
        B.trace_adjoint(mesh, w["Minv"], w["light"], d_rays, n, d_hits, d_target, d_gv, d_gm, d_loss)

    def upd():
        mesh.sgd(d_gv, lr, clip)
        mesh.update(normals=True)

    def step():
        fwd()
        adj()
        upd()

    # step 0 eagerly, with per-phase events (also sizes the adjoint's scratch)
    e = [ctx.event() for _ in range(4)]
    ctx.sync()
    ctx.record(e[0])
    fwd()
    ctx.record(e[1])
    adj()
    ctx.record(e[2])
    upd()
    ctx.record(e[3])
    ctx.sync()
    fwd_ms, adj_ms, upd_ms = (ctx.elapsed_ms(e[k], e[k + 1]) for k in range(3))
    loss0 = float(d_loss.download(np.float64, 1)[0])
    hits = d_hits.download(B.HIT_DTYPE, n)
    tgt = d_target.download(np.float32, 2 * n).reshape(-1, 2)
    contributing = int(((hits["prim"] != B.MISS) & (tgt[:, 0] < 3e38)).sum())
    graph = B.Graph.capture(ctx, step)
    steps = max(args.adjoint_steps - 1, 1)
    t0, t1 = ctx.event(), ctx.event()
    ctx.record(t0)
    for _ in range(steps):
        graph.launch()
    ctx.record(t1)
    ctx.sync()
    step_ms = ctx.elapsed_ms(t0, t1) / steps
    loss_last = float(d_loss.download(np.float64, 1)[0])
    # eager phase times again on the evolved mesh (steady state)
    ctx.record(e[0])
    fwd()
    ctx.record(e[1])
    adj()
    ctx.record(e[2])
    ctx.sync()
    fwd2, adj2 = ctx.elapsed_ms(e[0], e[1]), ctx.elapsed_ms(e[1], e[2])
    graph.close()
    adj_bytes = contributing * 200 + (n - contributing) * 44
    return {"workload": f"icosphere {w['F'].shape[0]} faces / {w['V'].shape[0]} vertices, 8 cameras x 512x512 "
                        f"({n} rays), depth + Lambert L2 loss, SGD lr {lr} (vertex step clip {clip}) + refit per "
                        f"step, hipGraph per step",
            "steps": steps + 1, "ms_per_step": round(step_ms, 4),
            "fwd_bwd_mrays_s": round(n / ((fwd2 + adj2) * 1e-3) / 1e6, 2),
            "step_mrays_s": round(n / (step_ms * 1e-3) / 1e6, 2),
            "forward_ms": round(fwd2, 4), "adjoint_ms": round(adj2, 4), "update_ms_eager": round(upd_ms, 4),
            "first_step_ms": {"forward": round(fwd_ms, 4), "adjoint": round(adj_ms, 4)},
            "contributing_rays": contributing,
            "adjoint_roofline": {"bound": "hbm", "algorithmic_bytes": adj_bytes,
                                 "achieved": round(adj_bytes / (adj2 * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                                 "frac": round(adj_bytes / (adj2 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
            "loss_first": loss0, "loss_last": loss_last}


def bench_meshes(args, ctx):
    """The reference's criterion benchmarks on its own assets
    (tests/golden/): bvh_construction = BvhBuilder on unity.tri
    (benches/bench_bvh_build.rs); bvh_rebuild = a fresh build of bigben.tri
    after animate(0.05), and bvh_refit = Bvh::refit of that animated mesh
    (benches/bench_bvh_refit_rebuild.rs).  The reference builder (8 bins,
    zero-initialised) on the GPU; the oracle on one host thread beside it.
    Build times are wall clock around the synchronous build call; refit time
    from HIP events, averaged over 100 refits."""
    import oracle

    import bvhtracer_amd as B
    from bvhtracer_amd import mesh as M
    from bvhtracer_amd import scenes

    golden = os.path.join(ROOT, "tests", "golden")
    unity = M.load_tri(os.path.join(golden, "unity.tri.gz"))
    bigben = M.load_tri(os.path.join(golden, "bigben.tri.gz"))

    def gpu_build(tris, reps=10):
        times = []
        for _ in range(reps):
            t = tris.copy()
            t0 = time.perf_counter()
            bvh = B.BvhBuilder(B.BUILD_REFERENCE, ctx=ctx).build_for(t)
            times.append(time.perf_counter() - t0)
            last = (bvh, t)
        return float(np.median(times)) * 1e3, last

    def cpu_build(tris, reps=3):
        times = []
        for _ in range(reps):
            t = tris.copy()
            t0 = time.perf_counter()
            out = oracle.bvh_build(t, True, 0)
            times.append(time.perf_counter() - t0)
        return float(np.median(times)) * 1e3, out

    build_gpu, (ubvh, ut) = gpu_build(unity)
    build_cpu, (un, uu, utris) = cpu_build(unity)
    same_build = ubvh.download()[2].tobytes() == utris.tobytes()
    # the refit benchmark: build the original mesh, animate(0.05), refit
    bb = bigben.copy()
    bvh = B.BvhBuilder(B.BUILD_REFERENCE, ctx=ctx).build_for(bb)
    animated = scenes.animate(bb, 0.05)
    bvh.refit(animated)
    e0, e1 = ctx.event(), ctx.event()
    ctx.sync()
    ctx.record(e0)
    for _ in range(100):
        bvh.refit()
    ctx.record(e1)
    refit_gpu = ctx.elapsed_ms(e0, e1) / 100
    nodes0, used, _ = oracle.bvh_build(bb.copy(), True, 0)
    t0 = time.perf_counter()
    for _ in range(20):
        ref_nodes = oracle.bvh_refit(nodes0.copy(), used, animated)
    refit_cpu = (time.perf_counter() - t0) / 20 * 1e3
    same_refit = bvh.download()[0][:used].tobytes() == ref_nodes[:used].tobytes()
    rebuild_gpu, _ = gpu_build(scenes.animate(bigben, 0.05))
    rebuild_cpu, _ = cpu_build(scenes.animate(bigben, 0.05))
    return {"workload": f"unity.tri ({unity.shape[0]} tris) build; bigben.tri ({bigben.shape[0]} tris) "
                        f"animate(0.05) rebuild and refit; reference builder",
            "bvh_construction_ms": {"gpu": round(build_gpu, 3), "cpu_oracle_1_thread": round(build_cpu, 3)},
            "bvh_rebuild_ms": {"gpu": round(rebuild_gpu, 3), "cpu_oracle_1_thread": round(rebuild_cpu, 3)},
            "bvh_refit_ms": {"gpu": round(refit_gpu, 4), "cpu_oracle_1_thread": round(refit_cpu, 4)},
            "build_identical_to_oracle": bool(same_build), "refit_identical_to_oracle": bool(same_refit)}


def bench_physics(args, ctx, n=1 << 20, steps=50):
    """World::run_physics (physics/world.rs:57-60) on n rigid bodies: every
    body under Gravity and a Spring to its neighbour, awake, half of them
    allowed to sleep; dt = 1/60.  GPU time from HIP events over `steps`
    steps; the oracle (one host thread) on the first 65,536 bodies and their
    forces beside it.  Bytes per body per step: 272 B record read twice
    (forces, integrate) and written once, 2 x 48 B force entries, the
    neighbour's 64 B transform, 28 B of accumulators written."""
    import ctypes

    import oracle

    from bvhtracer_amd import _lib as L

    rng = np.random.default_rng(0xF00D)
    bodies = np.zeros(n, L.BODY_DTYPE)
    bodies["inverse_mass"] = (1.0 / rng.uniform(1, 5, n)).astype(np.float32)
    eye3 = np.eye(3, dtype=np.float32).reshape(9)
    bodies["inverse_inertia_tensor"] = eye3
    bodies["inverse_inertia_tensor_world"] = eye3
    bodies["linear_drag"] = rng.uniform(0.95, 1.0, n).astype(np.float32)
    bodies["angular_drag"] = 1.0
    bodies["position"] = rng.uniform(-100, 100, (n, 3)).astype(np.float32)
    q = rng.normal(size=(n, 4)).astype(np.float32)
    bodies["orientation"] = q / np.linalg.norm(q, axis=1, keepdims=True).astype(np.float32)
    bodies["rotation"] = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    bodies["is_awake"] = 1
    bodies["motion"] = np.float32(0.6)
    bodies["can_sleep"] = np.arange(n) % 2
    bodies["transform"] = np.eye(4, dtype=np.float32).reshape(16)
    forces = np.zeros(2 * n, L.FORCE_DTYPE)
    forces["kind"][0::2], forces["body"][0::2] = L.FORCE_GRAVITY, np.arange(n)
    forces["v0"][0::2] = (0, -9.81, 0)
    forces["kind"][1::2], forces["body"][1::2], forces["other"][1::2] = L.FORCE_SPRING, np.arange(n), (np.arange(n) + 1) % n
    forces["v0"][1::2], forces["v1"][1::2] = (0.1, 0, 0), (0, 0.1, 0)
    forces["spring_constant"][1::2], forces["rest_length"][1::2] = 0.2, 1.0
    h = ctypes.c_void_p()
    L.check(L.lib().bvht_world_create(ctx.handle, L.ptr(bodies), n, L.ptr(forces), 2 * n, ctypes.byref(h)))
    try:
        L.check(L.lib().bvht_world_start_frame(ctx.handle, h))
        dt = np.float32(1 / 60)
        for _ in range(3):
            L.check(L.lib().bvht_world_run_physics(ctx.handle, h, dt))
        ctx.sync()
        e0, e1 = ctx.event(), ctx.event()
        ctx.record(e0)
        for _ in range(steps):
            L.check(L.lib().bvht_world_run_physics(ctx.handle, h, dt))
        ctx.record(e1)
        ms = ctx.elapsed_ms(e0, e1) / steps
    finally:
        L.lib().bvht_world_destroy(ctx.handle, h)
    m = 1 << 16
    sub = bodies[:m].view(oracle.BODY_DTYPE).copy()
    fsub = forces[: 2 * m].copy()
    fsub["other"] = np.where(fsub["kind"] == L.FORCE_SPRING, fsub["other"] % m, fsub["other"])
    oracle.world_start_frame(sub)
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        oracle.world_run_physics(sub, fsub, dt)
    cpu_s = (time.perf_counter() - t0) / reps
    bytes_per_body = 3 * 272 + 2 * 48 + 64 + 28
    return {"workload": f"{n} rigid bodies, Gravity + Spring each, dt 1/60 (World::run_physics)",
            "ms_per_step": round(ms, 4), "bodies_per_s": round(n / (ms * 1e-3), 1),
            "achieved_gbs": round(n * bytes_per_body / (ms * 1e-3) / 1e9, 1), "bytes_per_body": bytes_per_body,
            "cpu_oracle_1_thread_bodies_per_s": round(m / cpu_s, 1)}


def cpu_baseline(ctx, bvh, d_brays, d_bhits, n_bounce, d_rays, d_hits, n_primary, budget_s):
    """The oracle (C restatement of the reference traversal, -O3) on the same
    BVH and a bounded, interleaved sample of the same primary and bounce rays."""
    import oracle

    import bvhtracer_amd as B

    nodes, used, tris = bvh.download()
    cores = max(1, min(16, os.cpu_count() or 1))
    rays = d_rays.download(B.RAY_DTYPE, n_primary)
    brays = d_brays.download(B.RAY_DTYPE, n_bounce)
    hits = d_hits.download(B.HIT_DTYPE, n_primary)
    bhits = d_bhits.download(B.HIT_DTYPE, n_bounce)
    # calibrate the sample on 20k rays, then size it to the budget
    stride = 101
    probe = np.concatenate([rays[::stride][:10000], brays[::stride][:10000]]).view(oracle.RAY_DTYPE)
    t0 = time.perf_counter()
    oracle.bvh_trace(nodes, tris, probe, nthreads=cores)
    rate = probe.shape[0] / max(time.perf_counter() - t0, 1e-6)
    total = n_primary + n_bounce
    want = int(rate * budget_s)
    frac = min(want, total) / total
    kp = max(1, int(n_primary * frac))
    kb = max(1, int(n_bounce * frac))
    # a whole frame takes less than the budget on many cores: trace it again
    reps = max(1, min(64, round(want / total))) if want > total else 1
    ip = np.linspace(0, n_primary - 1, kp).astype(np.int64)
    ib = np.linspace(0, n_bounce - 1, kb).astype(np.int64)
    sample = np.concatenate([rays[ip], brays[ib]]).view(oracle.RAY_DTYPE)
    t0 = time.perf_counter()
    ref = oracle.bvh_trace(nodes, tris, sample, nthreads=cores)
    for _ in range(reps - 1):
        oracle.bvh_trace(nodes, tris, sample, nthreads=cores)
    dt = (time.perf_counter() - t0) / reps
    t1 = time.perf_counter()
    single = np.concatenate([rays[ip[:20000]], brays[ib[:20000]]]).view(oracle.RAY_DTYPE)
    oracle.bvh_trace(nodes, tris, single, nthreads=1)
    dt1 = time.perf_counter() - t1
    gpu = np.concatenate([hits[ip], bhits[ib]])
    parity = bool(gpu.tobytes() == ref.tobytes())
    return {
        "value": round(sample.shape[0] / dt / 1e6, 3),
        "unit": "Mrays/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{kp} primary + {kb} bounce rays of one frame (every {max(1, n_primary // kp)}th), same BVH, "
                  f"traced {reps}x; {dt * reps:.1f} s on {cores} threads",
        "single_thread_mrays_s": round(single.shape[0] / dt1 / 1e6, 3),
        "sample_hits_bit_identical": parity,
    }


if __name__ == "__main__":
    main()
