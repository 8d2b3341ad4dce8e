"""Benchmark: LBVH build (Mtris/s) and sorted closest-hit trace (Mrays/s) on one GPU.

Workload (one step): rebuild a 60-bit LBVH over 16,777,216 triangles (an
8,388,608-triangle 64-component Gaussian-mixture soup plus a 2048x2048-cell
ground grid), then trace 16,777,216 primary rays (4096x4096 pinhole) and
16,777,216 cosine-hemisphere secondaries from the primary hits, each set
sorted by the 60-bit Aila ray key before tracing.  Geometry and rays are
seeded synthetic data resident in HBM before timing starts.

value = build throughput in Mtris/s; trace throughput is reported beside it.
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
    ap.add_argument("--cpu-sample", type=int, default=2097152, help="triangles in the CPU baseline sample")
    ap.add_argument("--cpu-rays", type=int, default=262144, help="rays in the CPU trace sample")
    ap.add_argument("--tune", action="append", default=[], metavar="NAME=VALUE",
                    help="libhx launch tunable (hx_context_set_tuning), e.g. trace_lds_stack=16")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run one plain device copy of this many MiB first (calibrates PMC byte counters)")
    return ap.parse_args()


def cpu_baseline(verts, tris, rays, args):
    """Oracle (C++ restatement, all host threads) on a bounded sample of the same workload."""
    import numpy as np
    from oracle import hx_oracle as O
    rng = np.random.default_rng(11)
    n = tris.shape[0]
    m = min(args.cpu_sample, n)
    sel = np.sort(rng.choice(n, m, replace=False)) if m < n else np.arange(n)
    t_s = tris[sel]
    O.lib()
    threads = O.get_threads()
    times = []
    for _ in range(2):
        t0 = time.perf_counter()
        b = O.build_lbvh(verts, t_s, bits=60)
        times.append(time.perf_counter() - t0)
    build_s = min(times)
    r = rays[: min(args.cpu_rays, rays.shape[0])]
    t0 = time.perf_counter()
    O.trace(b["nodes"], b["woop"], b["tri_idx"], r)
    trace_s = time.perf_counter() - t0
    return dict(value=round(m / build_s / 1e6, 3), unit="Mtris/s", cores=threads, kind="port",
                sample=f"LBVH-60 build of {m} of the {n} triangles (random subset, best of 2): "
                       f"{build_s:.2f} s; closest-hit trace of {r.shape[0]} primary rays on that BVH: "
                       f"{r.shape[0] / trace_s / 1e6:.3f} Mrays/s",
                trace_mrays_s=round(r.shape[0] / trace_s / 1e6, 3))


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
    # generates the values (20 B).  Averaged over the 8 passes of a 60-bit sort
    # of the n triangle keys plus the 32M-ray key sorts in the same step.
    sort_cnt, sort_ms = prof.get("sort_onesweep", (0, 0.0))
    passes = 8
    bytes_build = n * (20 + 24 * (passes - 1))
    bytes_rays = prim.shape[0] * (20 + 24 * (passes - 1)) + sec.shape[0] * (20 + 24 * (passes - 1))
    sort_bytes_per_step = bytes_build + bytes_rays
    launches_per_step = sort_cnt / args.steps
    avg_launch_ms = sort_ms / max(sort_cnt, 1)
    bytes_per_launch = sort_bytes_per_step / max(launches_per_step, 1)
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9 if sort_cnt else 0.0

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
        "kernels_ms_per_step": {k: round(v[1] / args.steps, 4) for k, v in prof.items()},
        "roofline": {
            "bound": "hbm",
            "kernel": "sort_onesweep",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "bytes_per_launch": int(bytes_per_launch),
            "avg_launch_ms": round(avg_launch_ms, 5),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
