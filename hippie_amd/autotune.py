"""Autotuner for libhx's launch shapes and build parameters.

No reference counterpart: the reference hard-codes its warp counts, stack
depth (HipTracerBinKernels.cu:5-15), radix digit width (RadixSortKernels.h
:13-22) and treelet size (ATRBuilder.h:13).  Here each is a knob chosen by
measurement on the caller's own workload:

  sort_digit_bits      radix digit width of every sort       8, 10
  sort_items           keys per thread in a sort pass        8, 16, 32
  sort_lookback        tiles read per look-back round trip   4, 8, 16
  trace_lds_stack      traversal-stack entries kept in LDS   0, 16, 24, 32, 64
  trace_fetch          wave refill threshold (busy lanes)    8 .. 64
  trace_blocks_per_cu  persistent trace workgroups per CU    0 (occupancy), 1, 2, 4
  trace_postpone       leaves held back per lane             1 .. 4
  trace_woop_ahead     Woop rows of the next triangle read   1, 2
                       during a triangle test
  atr_treelet_size     ATRBVH treelet (quality vs time)      5 .. 16

The search is coordinate descent from the current values: one knob at a time,
every candidate timed with HIP events on the context's stream (median of
`repeats` runs after a warm-up), the best kept before the next knob.  Every
candidate gives the same results (the tests check each value), so tuning
changes speed only -- except the treelet size, which changes the tree and is
scored by build time plus the trace time it buys.  Results can be cached in a
JSON file keyed by device and workload size.
"""
import json
import os
import statistics

import torch

from . import core

SORT_DIGITS = (8, 10)
SORT_ITEMS = (8, 16, 32)
SORT_LOOKBACK = (4, 8, 16)
LDS_STACK = (0, 16, 24, 32, 64)
FETCH = (8, 16, 24, 32, 40, 48, 56, 64)
BLOCKS_PER_CU = (0, 1, 2, 4)
POSTPONE = (1, 2, 3, 4)
WOOP_AHEAD = (1, 2)
TREELETS = (5, 7, 9, 12, 16)


def _time_ms(fn, repeats):
    fn()
    torch.cuda.synchronize()
    out = []
    for _ in range(repeats):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        out.append(a.elapsed_time(b))
    return statistics.median(out)


def _bucket(n):
    """Workload size class for the cache key (powers of 4)."""
    k = 0
    while (4 ** (k + 1)) <= max(int(n), 1):
        k += 1
    return f"4^{k}"


class Autotuner:
    def __init__(self, ctx=None, cache_path=None, repeats=3, log=None):
        self.ctx = ctx or core.Context.default()
        self.cache_path = cache_path
        self.repeats = repeats
        self.log = log or (lambda msg: None)
        self.history = []
        self.cache = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                self.cache = json.load(f)

    # -- helpers -----------------------------------------------------------
    def _key(self, kind, n):
        name = torch.cuda.get_device_name(torch.cuda.current_device())
        return f"{name}|{kind}|{_bucket(n)}"

    def _save(self):
        if self.cache_path:
            with open(self.cache_path, "w") as f:
                json.dump(self.cache, f, indent=1, sort_keys=True)

    def _descend(self, knobs, run):
        """Coordinate descent over libhx tunables; run() is the timed workload."""
        best = {k: self.ctx.get_tuning(k) for k, _ in knobs}
        for knob, values in knobs:
            timings = {}
            for v in values:
                self.ctx.set_tuning(knob, v)
                timings[v] = _time_ms(run, self.repeats)
                self.history.append((knob, v, timings[v]))
                self.log(f"autotune {knob}={v}: {timings[v]:.3f} ms")
            best[knob] = min(timings, key=timings.get)
            self.ctx.set_tuning(knob, best[knob])
        return best

    def apply(self, config):
        for k, v in config.items():
            if k != "atr_treelet_size":
                self.ctx.set_tuning(k, int(v))

    # -- tuners ------------------------------------------------------------
    def tune_sort(self, n, bits=60):
        """Digit width for sorting n keys of `bits` bits (the build's Morton sort)."""
        key = self._key(f"sort{bits}", n)
        if key in self.cache:
            self.apply(self.cache[key])
            return self.cache[key]
        g = torch.Generator(device="cuda")
        g.manual_seed(1)
        src = torch.randint(0, 1 << min(bits, 62), (n,), generator=g, device="cuda", dtype=torch.int64)
        keys = torch.empty_like(src)
        vals = torch.empty(n, dtype=torch.int32, device="cuda")

        def run():
            keys.copy_(src)
            core.sort_pairs(self.ctx, keys, vals, bits)
        best = self._descend([("sort_digit_bits", SORT_DIGITS), ("sort_items", SORT_ITEMS),
                              ("sort_lookback", SORT_LOOKBACK)], run)
        self.cache[key] = best
        self._save()
        return best

    def tune_trace(self, bvh, rays, sort_method=None, label=""):
        """Trace launch shape for this BVH and ray set (RayBuffer), kept in
        rays.launch so that HipTracer applies it to this ray set's traces only
        (coherent primaries and incoherent secondaries differ); the context's
        own values are left as they were.  label: tells apart ray sets of one
        size in the cache."""
        key = self._key(f"trace-{sort_method or 'unsorted'}{'-' + label if label else ''}", rays.size)
        if key in self.cache:
            rays.launch = dict(self.cache[key])
            return self.cache[key]
        tracer = core.HipTracer(self.ctx, bvh)
        knobs = [("trace_lds_stack", LDS_STACK), ("trace_fetch", FETCH), ("trace_blocks_per_cu", BLOCKS_PER_CU),
                 ("trace_postpone", POSTPONE), ("trace_woop_ahead", WOOP_AHEAD)]
        saved = {k: self.ctx.get_tuning(k) for k, _ in knobs}
        launch, rays.launch = rays.launch, {}

        def run():
            if sort_method:
                tracer.trace_sort(rays, method=sort_method, bits=60)
            else:
                tracer.trace(rays)
        try:
            best = self._descend(knobs, run)
        finally:
            self.apply(saved)
            rays.launch = launch
        rays.launch = dict(best)
        self.cache[key] = best
        self._save()
        return best

    def tune_batch_trace(self, batch, rays, ray_mesh, label=""):
        """Trace launch shape of a batched trace (HipBatch) over these rays,
        kept in batch.launch; the context's own values are left as they were."""
        n = int(rays.shape[0])
        key = self._key(f"batch-trace{'-' + label if label else ''}", n)
        if key in self.cache:
            batch.launch = dict(self.cache[key])
            return self.cache[key]
        knobs = [("trace_lds_stack", LDS_STACK), ("trace_fetch", FETCH), ("trace_blocks_per_cu", BLOCKS_PER_CU),
                 ("trace_postpone", POSTPONE), ("trace_woop_ahead", WOOP_AHEAD)]
        saved = {k: self.ctx.get_tuning(k) for k, _ in knobs}
        launch, batch.launch = batch.launch, {}
        results = torch.zeros((n, 4), dtype=torch.int32, device=rays.device)
        try:
            best = self._descend(knobs, lambda: batch.trace(rays, ray_mesh, results=results))
        finally:
            self.apply(saved)
            batch.launch = launch
        batch.launch = dict(best)
        self.cache[key] = best
        self._save()
        return best

    def tune_treelet(self, scene, rays, sizes=TREELETS, iterations=2):
        """ATRBVH treelet size scored by build time + closest-hit trace time."""
        key = self._key(f"atr-i{iterations}", scene.n_triangles)
        if key in self.cache:
            return self.cache[key]
        scores = {}
        for t in sizes:
            b = core.ATRBuilder(self.ctx, atr_treelet_size=t, atr_iterations=iterations)
            bvh = b.build(scene)
            build_ms = _time_ms(lambda: b.rebuild(bvh, scene), self.repeats)
            tracer = core.HipTracer(self.ctx, bvh)
            trace_ms = _time_ms(lambda: tracer.trace(rays), self.repeats)
            scores[t] = build_ms + trace_ms
            self.history.append(("atr_treelet_size", t, scores[t]))
            self.log(f"autotune atr_treelet_size={t}: build {build_ms:.3f} + trace {trace_ms:.3f} ms")
        best = {"atr_treelet_size": min(scores, key=scores.get)}
        self.cache[key] = best
        self._save()
        return best
