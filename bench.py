"""Benchmark: LBVH build (Mtris/s) and sorted closest-hit trace (Mrays/s) on one GPU.

Workload (one step): rebuild a 60-bit LBVH over 16,777,216 triangles (an
8,388,608-triangle 64-component Gaussian-mixture soup plus a 2048x2048-cell
ground grid), then trace 16,777,216 primary rays (4096x4096 pinhole) and
16,777,216 cosine-hemisphere secondaries from the primary hits, each set
sorted by the 60-bit Aila ray key before tracing.  Geometry and rays are
seeded synthetic data resident in HBM before timing starts.

value = build throughput in Mtris/s; trace throughput is reported beside it.

Also measured in the same run (reported beside `value`):
  builders  PLOC (radius 25), ATRBVH (treelet 9, 2 iterations), HLBVH (15 SAH
            bits), insertion (mod 1) and presplit (beta 0.3) LBVH rebuilds of
            the same scene, with their SAH cost against the LBVH's;
  batched   4096 meshes of round(exp(U[ln 64, ln 262144])) triangles (half
            displaced icospheres, half soups, each in its own unit box): one
            batched build against a loop of per-mesh rebuilds, and one
            differentiable step (refit -> trace of 256 rays per 1k triangles ->
            loss -> adjoint scatter) batched and graph-captured against the
            per-mesh loop.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "lbvh60_16M_hbm_traffic.md")


# The kernels behind each timed build stage, as the PMC summary names them.
PMC_KERNELS = {"scene_box": ("k_scene_box_partial", "k_scene_box_final"), "morton": ("k_morton",),
               "sort_hist": ("k_sort_hist",), "climb": ("k_lbvh_local", "k_lbvh_top"),
               "sort_local": ("k_sort_local",), "collapse_emit": ("k_emit",)}


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` (calibrated FETCH_SIZE + WRITE_SIZE
    read + written MB) from the committed rocprofv3 PMC summary, or None."""
    try:
        with open(TRAFFIC_PROFILE) as fh:
            for line in fh:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) == 4 and cells[0] == kernel:
                    return int(round((float(cells[2]) + float(cells[3])) * 1e6))
    except (OSError, ValueError):
        return None
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mixture", type=int, default=8388608, help="Gaussian-mixture triangles")
    ap.add_argument("--grid", type=int, default=2048, help="ground grid cells per side")
    ap.add_argument("--res", type=int, default=4096, help="primary ray resolution (res x res)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=0, help="triangles in the CPU baseline sample (0 = all)")
    ap.add_argument("--cpu-rays", type=int, default=4194304, help="rays in the CPU trace sample")
    ap.add_argument("--builders", default="ploc,atr,tr,hlbvh,insertion,presplit",
                    help="other builders to time on the same scene ('' = none)")
    ap.add_argument("--batch-meshes", type=int, default=4096, help="meshes in the batched workload (0 = skip)")
    ap.add_argument("--batch-lo", type=int, default=64)
    ap.add_argument("--batch-hi", type=int, default=262144)
    ap.add_argument("--batch-rays-per-1k", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--no-autotune", action="store_true", help="keep libhx's default launch shapes")
    ap.add_argument("--no-diff-single", action="store_true", help="skip the one-mesh differentiable step")
    ap.add_argument("--c5-meshes", type=int, default=512, help="meshes of the presplit PLOC differentiable cycle (0 = skip)")
    ap.add_argument("--c5-rays", type=int, default=64 * 1024 * 1024, help="rays per step of that cycle")
    ap.add_argument("--tune", action="append", default=[], metavar="NAME=VALUE",
                    help="libhx launch tunable (hx_context_set_tuning), e.g. trace_lds_stack=16")
    ap.add_argument("--layouts", default="quad,oct", help="wide layouts to time on the same scene and rays ('' = none)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run one plain device copy of this many MiB first (calibrates PMC byte counters)")
    return ap.parse_args()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(verts, tris, rays, args):
    """Oracle (C++ restatement, all host threads) on a bounded sample of the same workload:
    one warm-up build, then the median of 5 timed builds."""
    import numpy as np
    from oracle import hx_oracle as O
    rng = np.random.default_rng(11)
    n = tris.shape[0]
    m = min(args.cpu_sample, n) if args.cpu_sample > 0 else n
    sel = np.sort(rng.choice(n, m, replace=False)) if m < n else np.arange(n)
    t_s = tris[sel]
    O.lib()
    threads = O.get_threads()
    b = O.build_lbvh(verts, t_s, bits=60)
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        b = O.build_lbvh(verts, t_s, bits=60)
        times.append(time.perf_counter() - t0)
    build_s = float(np.median(times))
    what = "all" if m == n else f"{m} of the"
    r = rays[: min(args.cpu_rays, rays.shape[0])]
    t0 = time.perf_counter()
    O.trace(b["nodes"], b["woop"], b["tri_idx"], r)
    trace_s = time.perf_counter() - t0
    return dict(value=round(m / build_s / 1e6, 3), unit="Mtris/s", cores=threads, kind="port", cpu=cpu_model(),
                sample=f"LBVH-60 build of {what} {n} triangles (median of 5 after a warm-up): "
                       f"{build_s:.2f} s; closest-hit trace of {r.shape[0]} primary rays on that BVH: "
                       f"{r.shape[0] / trace_s / 1e6:.3f} Mrays/s",
                trace_mrays_s=round(r.shape[0] / trace_s / 1e6, 3))


# FP32 VALU issue rate of one MI355X without packed math: 256 CUs x 64 lanes x
# 2.4 GHz lane-operations per second.
VALU_LANE_OPS = 256 * 64 * 2.4e9


def traversal_model(ctx, bvh, rb, hx, valu=True):
    """Per-ray nodes visited and triangles tested (the stats kernel) and, for
    the Bin layout, the VALU model of a ray's cost, 40 ops per node + 26 per
    triangle test, with the Mrays/s that VALU issue would allow at that cost."""
    import numpy as np
    r = hx.RayBuffer(rb.rays)
    hx.HipTracer(ctx, bvh).trace_stats(r)
    st = r.stats.cpu().numpy().astype(np.float64)
    nodes, tris = float(st[:, 0].mean()), float(st[:, 1].mean())
    if not valu:
        return {"nodes_per_ray": round(nodes, 2), "tris_per_ray": round(tris, 2)}
    ops = 40.0 * nodes + 26.0 * tris
    return {"nodes_per_ray": round(nodes, 2), "tris_per_ray": round(tris, 2), "valu_ops_per_ray": round(ops, 1),
            "valu_bound_mrays_s": round(VALU_LANE_OPS / ops / 1e6, 1)}


def sort_and_trace_ms(tracer, rbs, reps=3):
    """(ray keys + sort ms, trace ms) per step over the ray sets, from
    HipTracer::traceSort(rays, sortTime, traceTime)."""
    s = t = 0.0
    for _ in range(reps):
        for rb in rbs:
            a, b = tracer.trace_sort_timed(rb, method="aila", bits=60)
            s += a
            t += b
    return 1e3 * s / reps, 1e3 * t / reps


def bench_layouts(ctx, scene, n, rbs, n_rays, args):
    """Quad and Oct layouts of the same LBVH-60 (Bvh.layout quad / oct): build,
    Aila-sorted trace of the same rays, trace-only rate, ray-sort time, SAH."""
    import hippie_amd as hx
    out = {}
    for name in [x for x in args.layouts.split(",") if x]:
        log(f"layout {name}")
        b = hx.LBVHBuilder(ctx, morton_bits=60, layout=name)
        bv = b.build(scene)
        ms = timed(lambda: b.rebuild(bv, scene), 3, 1)
        tr = hx.HipTracer(ctx, bv)
        for rb in rbs:
            tr.trace_sort(rb, method="aila", bits=60)
        full = timed(lambda: [tr.trace_sort(rb, method="aila", bits=60) for rb in rbs], 3, 0)
        sort_ms, trace_ms = sort_and_trace_ms(tr, rbs)
        out[name] = {"build_ms": round(ms, 3), "build_mtris_s": round(n / (ms * 1e-3) / 1e6, 2),
                     "nodes": bv.n_interior + bv.n_leaves, "sah": round(bv.get_cost(), 4),
                     "sorted_trace_ms": round(full, 3), "sorted_trace_mrays_s": round(n_rays / (full * 1e-3) / 1e6, 2),
                     "ray_sort_ms": round(sort_ms, 3), "trace_only_ms": round(trace_ms, 3),
                     "trace_only_mrays_s": round(n_rays / (trace_ms * 1e-3) / 1e6, 2),
                     "work_primary": traversal_model(ctx, bv, rbs[0], hx, valu=False),
                     "work_diffuse": traversal_model(ctx, bv, rbs[1], hx, valu=False)}
        del tr, bv
    return out


def timed(fn, steps, warmup):
    """Average ms of fn() over `steps` calls after `warmup`, by HIP events on torch's stream."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def bench_builders(ctx, scene, n, args):
    """PLOC r=25, ATRBVH t=9 x2, TR x2, HLBVH, insertion and presplit LBVH rebuilds of the benchmark scene,
    with SAH costs."""
    import hippie_amd as hx
    out = {}
    kinds = [k for k in args.builders.split(",") if k]
    specs = {"ploc": ("ploc_r25", hx.PLOCBuilder, dict(ploc_radius=25)),
             "atr": ("atr_t9_i2", hx.ATRBuilder, dict(atr_treelet_size=9, atr_iterations=2)),
             "tr": ("tr_i2", hx.TRBuilder, dict(tr_iterations=2)),
             "hlbvh": ("hlbvh_s15", hx.HLBVHBuilder, dict(hlbvh_sah_bits=15)),
             "insertion": ("insertion_m1", hx.InsertionBuilder, dict(insertion_mod=1)),
             "presplit": ("presplit_b03_lbvh", hx.LBVHBuilder, dict(presplitting=1, presplitter_beta=0.3))}
    for k in kinds:
        name, cls, kw = specs[k]
        b = cls(ctx, morton_bits=60, **kw)
        bvh = b.build(scene)
        reps = 1 if k == "insertion" else max(1, args.steps // 2)
        ms = timed(lambda: b.rebuild(bvh, scene), reps, 0 if k == "insertion" else 1)
        out[name] = {"ms": round(ms, 3), "mtris_s": round(n / (ms * 1e-3) / 1e6, 2), "sah": round(bvh.get_cost(), 4)}
        if k in ("atr", "tr"):
            out[name]["treelets_kept"] = bvh.info("atr_accepted")
        if k == "ploc":
            out[name]["rounds"] = bvh.info("ploc_rounds")
        if k == "hlbvh":
            out[name]["clusters"] = bvh.info("hlbvh_clusters")
        if k == "presplit":
            out[name]["references"] = bvh.info("references")
        if k == "insertion":
            out[name]["steps"] = bvh.info("insertion_steps")
            out[name]["reinserted"] = bvh.info("insertion_reinserted")
        log(f"{name}: {ms:.3f} ms ({out[name]['mtris_s']} Mtris/s), SAH {out[name]['sah']}")
        del bvh
    return out


def batch_workload(args, device):
    """Config 3 geometry and rays, generated on the device (hippie_amd.workloads)."""
    from hippie_amd import workloads
    return workloads.batched_build_trace(device, args.batch_meshes, args.batch_lo, args.batch_hi,
                                         args.batch_rays_per_1k)


def bench_diff_single(ctx, args):
    """Differentiable refit -> trace -> loss -> adjoint step on one mesh: a
    displaced level-8 icosphere (1,310,720 triangles) and 16 orbit cameras of
    512x512 rays (4,194,304 rays), loss sum (t - t_target)^2 over rays that hit
    in both, t_target from a second displacement seed; beside it the plain
    refit + trace of the same step."""
    import numpy as np
    import torch
    import hippie_amd as hx
    from hippie_amd import scenes
    from hippie_amd.diff import DifferentiableTracer
    dev = torch.device("cuda", 0)
    t0 = time.perf_counter()
    v0, f = scenes.displaced_icosphere(8, seed=1)
    v1, _ = scenes.displaced_icosphere(8, seed=2)
    rays = torch.as_tensor(scenes.orbit_rays(16, 512), device=dev)
    tris = torch.as_tensor(f, device=dev)
    verts0 = torch.as_tensor(v0, device=dev)
    verts1 = torch.as_tensor(v1, device=dev)
    log(f"diff workload: {f.shape[0]} triangles, {rays.shape[0]} rays ({time.perf_counter() - t0:.1f} s)")
    with torch.no_grad():
        t_tgt, _, _, ids_t = DifferentiableTracer(verts1, tris, ctx)(verts1, rays)
    dt = DifferentiableTracer(verts0, tris, ctx)
    vp = verts0.clone().requires_grad_()
    zero = torch.zeros((), device=dev)

    def step():
        t, _, _, ids = dt(vp, rays)
        loss = torch.where((ids >= 0) & (ids_t >= 0), (t - t_tgt) ** 2, zero).sum()
        vp.grad = None
        loss.backward()

    step_ms = timed(step, args.batch_steps, 1)
    tracer = dt.tracer()
    rb = hx.RayBuffer(rays)
    scene = hx.Scene(verts0, tris)

    def plain():
        dt.bvh.update(scene)
        tracer.trace(rb)

    plain_ms = timed(plain, args.batch_steps, 1)
    n_rays = int(rays.shape[0])
    log(f"diff step (one mesh) {step_ms:.3f} ms ({n_rays / (step_ms * 1e-3) / 1e6:.1f} Mrays/s), "
        f"plain refit + trace {plain_ms:.3f} ms")
    return {"workload": f"displaced icosphere level 8 ({f.shape[0]} triangles), 16 cameras x 512^2 rays",
            "triangles": int(f.shape[0]), "rays": n_rays, "step_ms": round(step_ms, 3),
            "step_mrays_s": round(n_rays / (step_ms * 1e-3) / 1e6, 1), "plain_refit_trace_ms": round(plain_ms, 3)}


def bench_config5(ctx, args):
    """Batched differentiable cycle at scale: 512 meshes of log-uniform 1k..200k
    triangles, 30% of them long thin slivers (aspect up to 1:200), presplit
    (beta 0.3) PLOC (r = 25) rebuild once per 10 steps, and per step refit ->
    trace -> loss -> adjoint (hipGraph-captured) of 64M rays spread over the
    meshes in proportion to their triangles; against the per-mesh loop of the
    same cycle.  The vertices are refit in place each step (same work as a
    deformation)."""
    import torch
    import hippie_amd as hx
    from hippie_amd import workloads
    from hippie_amd.diff import backward as single_backward
    dev = torch.device("cuda", 0)
    t0 = time.perf_counter()
    w = workloads.batched_presplit_diff(dev, args.c5_meshes, args.c5_rays)
    M, meshes, counts, rays, owner = args.c5_meshes, w["meshes"], w["counts"], w["rays"], w["owner"]
    mb = hx.MeshBatch.from_meshes(meshes)
    N = int(counts.sum())
    R = int(rays.shape[0])
    torch.cuda.synchronize()
    log(f"config 5 workload: {M} meshes, {N} triangles, {R} rays ({time.perf_counter() - t0:.1f} s)")
    params = dict(method="ploc", ploc_radius=25, presplitting=1, presplitter_beta=0.3)
    hb = hx.HipBatch(ctx, mb, **params)
    steps = args.batch_steps
    rebuild_ms = timed(hb.rebuild, steps, 1)
    tuned = None
    if not args.no_autotune:
        from hippie_amd.autotune import Autotuner
        tuned = Autotuner(ctx, repeats=2, log=log).tune_batch_trace(hb, rays, owner, label="config5")
    res = hb.trace(rays, owner)
    t_ref = res[:, 1].view(torch.float32).clone() * 0.999
    grad = torch.zeros_like(mb.scene.vertices)
    dl_dt = torch.zeros(R, dtype=torch.float32, device=dev)

    def batched_step():
        hb.refit()
        hb.trace(rays, owner, results=res)
        hit = res[:, 0] >= 0
        t = res[:, 1].view(torch.float32)
        torch.where(hit, 2.0 * (t - t_ref), torch.zeros_like(t), out=dl_dt)
        grad.zero_()
        hb.backward(rays, owner, res, dl_dt=dl_dt, out=grad)

    e_ms = timed(batched_step, steps, 1)
    graph_ms = None
    try:
        st = torch.cuda.Stream()
        st.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(st):
            ctx.set_stream(st)
            batched_step()
        torch.cuda.current_stream().wait_stream(st)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            ctx.set_stream(torch.cuda.current_stream())
            batched_step()
        ctx.set_stream(torch.cuda.current_stream())
        graph_ms = timed(graph.replay, steps, 1)
    except Exception as e:  # capture is optional; the eager step is the measured fallback
        log(f"graph capture failed: {e}")
        ctx.set_stream(torch.cuda.current_stream())
    step_ms = graph_ms if graph_ms is not None else e_ms
    cycle_ms = rebuild_ms + 10 * step_ms

    # Per-mesh loop of the same cycle.
    builder = hx.PLOCBuilder(ctx, **{k: v for k, v in params.items() if k != "method"})
    scenes_m = [hx.Scene(v, t) for v, t in meshes]
    bvhs = [builder.build(sc) for sc in scenes_m]
    roffs = w["roffs"]
    res_m = [res[int(roffs[m]):int(roffs[m + 1])] for m in range(M)]
    rays_m = [rays[int(roffs[m]):int(roffs[m + 1])] for m in range(M)]
    dl_m = [dl_dt[int(roffs[m]):int(roffs[m + 1])] for m in range(M)]
    grads_m = [torch.zeros_like(sc.vertices) for sc in scenes_m]
    tracer = hx.HipTracer(ctx)
    # the loop's traces get the batched trace's launch shape too
    rbs_m = [hx.RayBuffer.wrap(rays_m[m], res_m[m]) for m in range(M)]
    for rb in rbs_m:
        rb.launch = dict(hb.launch)

    def loop_rebuild():
        for sc, bv in zip(scenes_m, bvhs):
            builder.rebuild(bv, sc)

    def loop_step():
        for m in range(M):
            bv, sc = bvhs[m], scenes_m[m]
            bv.update(sc)
            tracer.set_bvh(bv)
            tracer.trace(rbs_m[m])
            single_backward(ctx, sc.vertices, sc.triangles, rays_m[m], res_m[m], dl_dt=dl_m[m], out=grads_m[m])

    l_rebuild_ms = timed(loop_rebuild, 1, 0)
    l_step_ms = timed(loop_step, 1, 1)
    l_cycle_ms = l_rebuild_ms + 10 * l_step_ms

    # Where a step's time goes: each stage of the batched step on its own, the
    # loop's traces on their own, and the traversal work per ray (stats trace)
    # on sliver and on other meshes.
    b_refit_ms = timed(hb.refit, steps, 1)
    b_trace_ms = timed(lambda: hb.trace(rays, owner, results=res), steps, 1)
    b_back_ms = timed(lambda: hb.backward(rays, owner, res, dl_dt=dl_dt, out=grad), steps, 1)

    def loop_trace():
        for m in range(M):
            tracer.set_bvh(bvhs[m])
            tracer.trace(rbs_m[m])
    l_trace_ms = timed(loop_trace, 1, 1)
    stats = torch.zeros((R, 2), dtype=torch.int32, device=dev)
    hb.trace(rays, owner, results=torch.zeros_like(res), stats=stats)
    sl = torch.as_tensor(w["sliver"], device=dev)[owner.long()]
    st = stats.double()

    def per_ray(mask):
        k = max(int(mask.sum()), 1)
        return {"rays": int(mask.sum()), "nodes_per_ray": round(float(st[mask, 0].sum()) / k, 2),
                "tris_per_ray": round(float(st[mask, 1].sum()) / k, 2)}
    breakdown = {"refit_ms": round(b_refit_ms, 3), "trace_ms": round(b_trace_ms, 3),
                 "backward_ms": round(b_back_ms, 3), "loop_trace_ms": round(l_trace_ms, 3),
                 "sliver_meshes": per_ray(sl), "other_meshes": per_ray(~sl),
                 # the cycle with free refit, loss and adjoint: a bound on the speedup
                 "speedup_bound_trace_only": round(l_cycle_ms / (rebuild_ms + 10 * b_trace_ms), 2)}
    log(f"config 5 breakdown: {breakdown}")
    log(f"config 5: batched rebuild {rebuild_ms:.2f} ms, step {e_ms:.2f} ms (graph {graph_ms}), cycle "
        f"{cycle_ms:.1f} ms; per-mesh loop rebuild {l_rebuild_ms:.1f} ms, step {l_step_ms:.1f} ms, cycle "
        f"{l_cycle_ms:.1f} ms ({l_cycle_ms / cycle_ms:.1f}x)")
    return {"workload": f"{M} meshes, {N} triangles (log-uniform 1000..200000, 30% sliver meshes), {R} rays; "
                        "presplit (0.3) PLOC r=25 rebuild + 10 refit/trace/adjoint steps",
            "meshes": M, "triangles": N, "references": int(hb.info("references")),
            "rays": R, "rebuild_ms": round(rebuild_ms, 3), "step_ms": round(e_ms, 3),
            "step_graph_ms": None if graph_ms is None else round(graph_ms, 3), "cycle_ms": round(cycle_ms, 3),
            "loop_rebuild_ms": round(l_rebuild_ms, 3), "loop_step_ms": round(l_step_ms, 3),
            "loop_cycle_ms": round(l_cycle_ms, 3), "cycle_speedup": round(l_cycle_ms / cycle_ms, 2),
            "breakdown": breakdown, "trace_launch": tuned}


def bench_batch(ctx, args):
    """Batched build and differentiable step against the per-mesh loop."""
    import numpy as np
    import torch
    import hippie_amd as hx
    from hippie_amd.diff import backward as single_backward
    dev = torch.device("cuda", 0)
    t0 = time.perf_counter()
    w = batch_workload(args, dev)
    M, N, R = args.batch_meshes, int(w["offs"][-1]), int(w["rays"].shape[0])
    torch.cuda.synchronize()
    log(f"batch workload: {M} meshes, {N} triangles, {R} rays ({time.perf_counter() - t0:.1f} s)")
    mb = hx.MeshBatch(w["V"], w["T"], w["offs"])
    hb = hx.HipBatch(ctx, mb, morton_bits=60)
    steps = args.batch_steps
    b_ms = timed(hb.rebuild, steps, 1)

    # Plain path: one BVH per mesh (built once), then a timed loop of rebuilds.
    scenes_m, bvhs = [], []
    builder = hx.LBVHBuilder(ctx, morton_bits=60)
    for m in range(M):
        vb0, vb1 = int(w["vbase"][m]), int(w["vbase"][m + 1])
        sc = hx.Scene.__new__(hx.Scene)
        sc.vertices = w["V"][vb0:vb1]
        sc.triangles = w["tris_local"][m]
        scenes_m.append(sc)
        bvhs.append(builder.build(sc))
    torch.cuda.synchronize()

    def loop_build():
        for sc, bv in zip(scenes_m, bvhs):
            builder.rebuild(bv, sc)
    l_ms = timed(loop_build, steps, 1)
    log(f"batched build {b_ms:.3f} ms vs per-mesh loop {l_ms:.3f} ms ({l_ms / b_ms:.1f}x)")

    # Differentiable step: refit -> trace -> loss (t - t_ref)^2 over hits -> adjoint.
    rays, owner = w["rays"], w["owner"]
    tuned = None
    if not args.no_autotune:
        from hippie_amd.autotune import Autotuner
        tuned = Autotuner(ctx, repeats=2, log=log).tune_batch_trace(hb, rays, owner, label="config3")
    res = hb.trace(rays, owner)
    t_ref = res[:, 1].view(torch.float32).clone() * 0.999
    grad = torch.zeros_like(mb.scene.vertices)
    dl_dt = torch.zeros(R, dtype=torch.float32, device=dev)

    def batched_step():
        hb.refit()
        hb.trace(rays, owner, results=res)
        hit = res[:, 0] >= 0
        t = res[:, 1].view(torch.float32)
        torch.where(hit, 2.0 * (t - t_ref), torch.zeros_like(t), out=dl_dt)
        grad.zero_()
        hb.backward(rays, owner, res, dl_dt=dl_dt, out=grad)

    e_ms = timed(batched_step, steps, 1)
    graph_ms = None
    try:
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            ctx.set_stream(s)
            batched_step()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            ctx.set_stream(torch.cuda.current_stream())
            batched_step()
        ctx.set_stream(torch.cuda.current_stream())
        graph_ms = timed(graph.replay, steps, 1)
    except Exception as e:  # capture is optional; the eager step is the measured fallback
        log(f"graph capture failed: {e}")
        ctx.set_stream(torch.cuda.current_stream())

    roffs = w["roffs"]
    res_m = [res[int(roffs[m]):int(roffs[m + 1])] for m in range(M)]
    rays_m = [rays[int(roffs[m]):int(roffs[m + 1])] for m in range(M)]
    dl_m = [dl_dt[int(roffs[m]):int(roffs[m + 1])] for m in range(M)]
    gm = [grad[int(w["vbase"][m]):int(w["vbase"][m + 1])] for m in range(M)]
    tracer = hx.HipTracer(ctx)

    # the loop's traces get the batched trace's launch shape too
    rbs_m = [hx.RayBuffer.wrap(rays_m[m], res_m[m]) for m in range(M)]
    for rb in rbs_m:
        rb.launch = dict(hb.launch)

    def loop_step():
        for m in range(M):
            bv, sc = bvhs[m], scenes_m[m]
            bv.update(sc)
            tracer.set_bvh(bv)
            tracer.trace(rbs_m[m])
            single_backward(ctx, sc.vertices, sc.triangles, rays_m[m], res_m[m], dl_dt=dl_m[m], out=gm[m])
    ld_ms = timed(loop_step, steps, 1)
    best = graph_ms if graph_ms is not None else e_ms
    log(f"diff step: batched {e_ms:.3f} ms, graph {graph_ms} ms, per-mesh loop {ld_ms:.3f} ms")
    out = {
        "workload": f"{M} meshes, {N} triangles (log-uniform {args.batch_lo}..{args.batch_hi}, half displaced "
                    f"icospheres), {R} rays",
        "meshes": M, "triangles": N, "rays": R,
        "build_ms": round(b_ms, 3), "build_mtris_s": round(N / (b_ms * 1e-3) / 1e6, 2),
        "loop_build_ms": round(l_ms, 3), "build_speedup": round(l_ms / b_ms, 2),
        "diff_step_ms": round(e_ms, 3),
        "diff_step_graph_ms": None if graph_ms is None else round(graph_ms, 3),
        "loop_diff_step_ms": round(ld_ms, 3), "diff_step_speedup": round(ld_ms / best, 2),
        "trace_launch": tuned,
    }
    for bv in bvhs:
        del bv
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (got --gpus {args.gpus}); multi-GPU sharding is design-only.",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import hippie_amd as hx
    from hippie_amd import scenes

    if args.copy_probe_mib > 0:
        nelem = args.copy_probe_mib * (1 << 20) // 4
        src = torch.ones(nelem, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {nelem * 4} bytes read, {nelem * 4} bytes written")
        del src, dst

    t_setup = time.perf_counter()
    log("generating scene")
    verts, tris = scenes.config2_scene(args.mixture, args.grid, seed=2)
    n = tris.shape[0]
    log(f"{n} triangles; generating {args.res}x{args.res} primary rays")
    prim = scenes.config2_primary_rays(args.res, args.res)
    ctx = hx.Context.default(0)
    for kv in args.tune:
        name, value = kv.split("=")
        ctx.set_tuning(name, int(value))
    scene = hx.Scene(verts, tris)
    builder = hx.LBVHBuilder(ctx, morton_bits=60)
    bvh = builder.build(scene)
    prim_rb = hx.RayBuffer(prim)
    tracer = hx.HipTracer(ctx, bvh)
    tracer.trace(prim_rb)
    res = prim_rb.results.cpu().numpy()
    log(f"primary hit rate {float((res[:, 0] >= 0).mean()):.3f}; generating secondaries")
    sec = scenes.secondary_rays(verts, tris, prim, res, seed=3)
    sec_rb = hx.RayBuffer(sec)
    n_rays = prim.shape[0] + sec.shape[0]
    tuning = None
    if not args.no_autotune:
        from hippie_amd.autotune import Autotuner
        at = Autotuner(ctx, repeats=2, log=log)
        tuning = dict(at.tune_sort(n, bits=60))
        # each ray set keeps its own trace launch shape (RayBuffer.launch)
        tuning["trace_primary"] = at.tune_trace(bvh, prim_rb, sort_method="aila", label="primary")
        tuning["trace_diffuse"] = at.tune_trace(bvh, sec_rb, sort_method="aila", label="diffuse")
        # ATRBVH treelet size scored by build + trace time of the diffuse rays (reported; the
        # builders table keeps the configured t = 9)
        tuning["atr_treelet"] = at.tune_treelet(scene, sec_rb, iterations=2)["atr_treelet_size"]
        log(f"autotuned: {tuning}")
    torch.cuda.synchronize()
    log(f"setup {time.perf_counter() - t_setup:.1f} s")

    def step(ev=None):
        if ev is not None:
            ev[0].record()
        builder.rebuild(bvh, scene)
        if ev is not None:
            ev[1].record()
        tracer.trace_sort(prim_rb, method="aila", bits=60)
        tracer.trace_sort(sec_rb, method="aila", bits=60)
        if ev is not None:
            ev[2].record()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ctx.profile_reset()
    ctx.set_profiling(True)
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ctx.set_profiling(False)
    prof = ctx.profile()
    build_ms = sum(e[0].elapsed_time(e[1]) for e in evs) / args.steps
    trace_ms = sum(e[1].elapsed_time(e[2]) for e in evs) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    log(f"build {build_ms:.3f} ms, sorted trace of {n_rays} rays {trace_ms:.3f} ms, step {ms_per_step:.3f} ms")
    for name, (cnt, tot) in sorted(prof.items(), key=lambda kv: -kv[1][1]):
        log(f"  {name:18s} {cnt:6d} launches  {tot / max(cnt, 1):9.4f} ms avg  {tot / args.steps:9.4f} ms/step")

    # Roofline of the dominant build kernel: one onesweep radix pass moves
    # 24 B per key (8 B key + 4 B value read and written); the first pass
    # generates the values (20 B).  Averaged over the passes of a 60-bit sort
    # (8 at 8-bit digits, 6 at 10-bit) of the n triangle keys plus the 32M-ray
    # key sorts in the same step.
    sort_cnt, sort_ms = prof.get("sort_onesweep", (0, 0.0))
    digit_bits = ctx.get_tuning("sort_digit_bits")
    passes = (60 + digit_bits - 1) // digit_bits
    # The build's Morton sort runs four passes over the high bits and the
    # local pass when the split sort is on (it is for n >= 2^18 keys); the ray
    # sorts run every pass.
    build_passes = 4 if ctx.get_tuning("sort_local_run") > 0 and n >= 1 << 18 else passes
    bytes_build = n * (20 + 24 * (build_passes - 1))
    bytes_rays = prim.shape[0] * (20 + 24 * (passes - 1)) + sec.shape[0] * (20 + 24 * (passes - 1))
    sort_bytes_per_step = bytes_build + bytes_rays
    launches_per_step = sort_cnt / args.steps
    avg_launch_ms = sort_ms / max(sort_cnt, 1)
    bytes_per_launch = sort_bytes_per_step / max(launches_per_step, 1)
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9 if sort_cnt else 0.0

    # Every LBVH build kernel against the same peak, from its algorithmic
    # bytes per launch (n triangles; SAH-collapsed node count for the emit):
    #   scene_box  12 B index + 36 B vertices per triangle           48n
    #   morton     the same reads + an 8 B key + the 48 B packed
    #              triangle the climb reads                             104n
    #   sort_hist  one read of the 8 B keys                              8 per key
    #   sort_local the split sort's local pass: 12 B key + value read
    #              and written                                          24n
    #   climb      hierarchy built during the climb: 8 B key + 4 B id +
    #              48 B packed triangle read; Woop (48 B), first/last/
    #              left/right (16 B), parents (8 B), state, cover (8 B)
    #              and the 48 B Bin rows of the state-1 nodes (about half
    #              the internal nodes: 24 B per triangle) written       164n
    #   emit       int_slot of every internal node, then per final
    #              interior its 48 B Bin rows, range, children, parent
    #              and their slots (~100 B); 64 B per Bin node          100n + 64 nodes
    nodes_out = bvh.n_interior + bvh.n_leaves
    alg = {"scene_box": 48 * n, "morton": 104 * n, "sort_hist": 8 * n, "sort_local": 24 * n, "climb": 164 * n,
           "collapse_emit": 100 * n + 64 * nodes_out}
    build_kernels = {}
    for k, b in alg.items():
        cnt, tot = prof.get(k, (0, 0.0))
        if not cnt:
            continue
        per_launch = b
        if k == "sort_hist":
            # one histogram per sort: the build's keys and the two ray sets
            per_launch = (8 * n + 8 * (prim.shape[0] + sec.shape[0])) / (cnt / args.steps)
        gbs = per_launch / (tot / cnt * 1e-3) / 1e9
        traffic = [measured_traffic(name) for name in PMC_KERNELS.get(k, ())]
        build_kernels[k] = {"avg_ms": round(tot / cnt, 5), "bytes_per_launch": int(per_launch),
                            "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                            "traffic": sum(traffic) if traffic and None not in traffic else None}

    ray_sort_ms, trace_only_ms = sort_and_trace_ms(tracer, [prim_rb, sec_rb])
    model = {"primary": traversal_model(ctx, bvh, prim_rb, hx), "diffuse": traversal_model(ctx, bvh, sec_rb, hx)}
    layouts = bench_layouts(ctx, scene, n, [prim_rb, sec_rb], n_rays, args) if args.layouts else {}
    builders = bench_builders(ctx, scene, n, args) if args.builders else {}
    if builders:
        builders["lbvh"] = {"ms": round(build_ms, 3), "sah": round(bvh.get_cost(), 4)}
    batched = bench_batch(ctx, args) if args.batch_meshes > 0 else None
    diff_single = bench_diff_single(ctx, args) if not args.no_diff_single else None
    config5 = bench_config5(ctx, args) if args.c5_meshes > 0 else None

    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle)")
        cpu = cpu_baseline(verts, tris, prim, args)

    out = {
        "metric": "BVH build Mtris/s + trace Mrays/s; batched & diff step vs plain, 1 GPU",
        "value": round(n / (build_ms * 1e-3) / 1e6, 2),
        "unit": "Mtris/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"LBVH-60 build of {n} triangles ({args.mixture} Gaussian-mixture soup + "
                        f"{args.grid}x{args.grid} ground grid) + Aila-sorted closest-hit trace of "
                        f"{prim.shape[0]} primary + {sec.shape[0]} diffuse secondary rays",
            "triangles": n,
            "rays": n_rays,
            "builder": "lbvh",
            "morton_bits": 60,
        },
        "build_ms": round(build_ms, 3),
        "build_mtris_s": round(n / (build_ms * 1e-3) / 1e6, 2),
        "trace_ms": round(trace_ms, 3),
        "trace_mrays_s": round(n_rays / (trace_ms * 1e-3) / 1e6, 2),
        "ray_sort_ms": round(ray_sort_ms, 3),
        "trace_only_ms": round(trace_only_ms, 3),
        "trace_only_mrays_s": round(n_rays / (trace_only_ms * 1e-3) / 1e6, 2),
        "traversal_model": model,
        "layouts": layouts,
        "kernels_ms_per_step": {k: round(v[1] / args.steps, 4) for k, v in prof.items()},
        "roofline": {
            "bound": "hbm",
            "kernel": "sort_onesweep",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            # PMC-measured for the 16,777,216-key passes this benchmark runs
            # (the committed profile's command); None if it is missing.
            "traffic": measured_traffic("k_onesweep"),
            "bytes_per_launch": int(bytes_per_launch),
            "avg_launch_ms": round(avg_launch_ms, 5),
        },
        "roofline_build_kernels": build_kernels,
        "cpu_baseline": cpu,
        "tuning": tuning,
        "builders": builders,
        "batched": batched,
        "diff_single": diff_single,
        "batched_diff_presplit_ploc": config5,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
