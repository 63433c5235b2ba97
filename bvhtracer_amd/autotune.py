"""Autotuner for the trace kernels.

Sweeps the trace kernel configurations compiled into libbvht_hip.so
(workgroup size 64/128/256 lanes x LDS stack depth 8/16/24/32 entries per
lane; entries beyond the LDS depth spill to HBM) on the caller's own
workload, and leaves the fastest configuration set on the context.  Results
are identical for every configuration (the traversal order is fixed), so the
tuner only measures time.
"""
import ctypes

from . import _lib as L

WORKGROUPS = (64, 128, 256)
LDS_DEPTHS = (8, 16, 24, 32)


def configurations():
    return [(wg, d) for wg in WORKGROUPS for d in LDS_DEPTHS]


def current(ctx):
    wg, d = ctx_i32(), ctx_i32()
    L.check(L.lib().bvht_get_trace_config(ctx.handle, ctypes.byref(wg), ctypes.byref(d)))
    return wg.value, d.value


def ctx_i32():
    return ctypes.c_int32()


def set_config(ctx, wg, lds_depth):
    L.check(L.lib().bvht_set_trace_config(ctx.handle, wg, lds_depth))


def time_ms(ctx, run, reps):
    e0, e1 = ctx.event(), ctx.event()
    ctx.sync()
    ctx.record(e0)
    for _ in range(reps):
        run()
    ctx.record(e1)
    return ctx.elapsed_ms(e0, e1) / reps


def tune(ctx, run, reps=3, configs=None, log=None):
    """Time ``run()`` (which enqueues trace work on ctx's stream) under each
    configuration; set the fastest on ctx.  Returns (best, {config: ms})."""
    table = {}
    for wg, d in configs or configurations():
        set_config(ctx, wg, d)
        run()  # warm up this variant
        table[(wg, d)] = time_ms(ctx, run, reps)
        if log:
            log(f"[autotune] wg {wg:3d} lds {d:2d}: {table[(wg, d)]:.4f} ms")
    best = min(table, key=table.get)
    set_config(ctx, *best)
    return best, table


BATCH_CONFIGS = [(wg, d, r) for wg in (64, 256) for d in (16, 24) for r in (64, 128, 256)]


def tune_batch(batch, offsets, run, reps=3, log=None):
    """Sweep the persistent batched trace over workgroup size x LDS depth x
    rays per work item; leave the fastest on the batch.  ``run()`` enqueues
    the batched trace.  Returns (best, {config: ms})."""
    ctx = batch.ctx
    table = {}
    for wg, d, r in BATCH_CONFIGS:
        batch.set_rays(offsets, r)
        batch.config(wg, d, r)
        run()
        table[(wg, d, r)] = time_ms(ctx, run, reps)
        if log:
            log(f"[autotune] batch wg {wg:3d} lds {d:2d} rays/item {r:3d}: {table[(wg, d, r)]:.4f} ms")
    best = min(table, key=table.get)
    batch.set_rays(offsets, best[2])
    batch.config(*best)
    return best, table
