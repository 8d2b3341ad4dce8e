"""hx_trace_sorted_noalloc (the sorted trace over a fixed per-device scratch)
against hx_trace_sorted, through the C ABI: results and ray order bit for bit,
on two calls in a row (the second smaller, so it runs over the first call's
leftovers), with no device memory taken by the library."""
import numpy as np
import pytest
import torch

import hippie_amd as hx
from hippie_amd import _lib, scenes
from hippie_amd.core import RAY_KEY_METHODS, _ptr

pytestmark = pytest.mark.gpu

MAX_RAYS = 1 << 20


def _sorted_trace(entry, ctx, bvh, rays, method, bits, results, order):
    _lib.call(entry, ctx.handle, bvh.handle, _ptr(rays), int(rays.shape[0]), 0, RAY_KEY_METHODS[method], bits,
              0.25, _ptr(results), None, _ptr(order))


def _buffers(n):
    return (torch.full((n, 4), -7, dtype=torch.int32, device="cuda"),
            torch.full((n,), -7, dtype=torch.int32, device="cuda"))


def test_noalloc_sorted_trace_matches_the_allocating_entry():
    ctx = hx.Context(0)                       # a fresh context: none of its buffers exist yet
    v, t = scenes.uniform_soup(65536, seed=1)
    bvh = hx.LBVHBuilder(ctx, morton_bits=30).build(hx.Scene(v, t))
    big = torch.as_tensor(scenes.config1_rays(1024, 1024), device="cuda")          # the maximum, 2^20
    small_h = scenes.config1_rays(700, 500)
    small_h[:, 0:3] += np.float32(0.01)                                             # other origins
    small = torch.as_tensor(small_h, device="cuda")
    assert big.shape[0] == MAX_RAYS and small.shape[0] < MAX_RAYS
    cases = [(big, "aila", 60), (small, "twopoint", 48)]
    got = [_buffers(r.shape[0]) for r, _, _ in cases]
    want = [_buffers(r.shape[0]) for r, _, _ in cases]
    warm = _buffers(1000)
    print("noalloc: warm-up call (loads the kernels)", flush=True)
    _sorted_trace("hx_trace_sorted_noalloc", ctx, bvh, small[:1000], "aila", 60, *warm)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for (rays, method, bits), (res, order) in zip(cases, got):
        print(f"noalloc: {rays.shape[0]} rays, {method} {bits}-bit keys", flush=True)
        _sorted_trace("hx_trace_sorted_noalloc", ctx, bvh, rays, method, bits, res, order)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    # The allocating entry would reserve 24 B per ray (keys and order, twice) for 2^20 rays.
    assert free0 - free1 < (4 << 20), f"device memory fell by {free0 - free1} bytes"
    assert ctx.overflow_count() == 0          # the context's own counter buffer was never created
    for (rays, method, bits), (res, order) in zip(cases, want):
        print(f"allocating entry: {rays.shape[0]} rays, {method}", flush=True)
        _sorted_trace("hx_trace_sorted", ctx, bvh, rays, method, bits, res, order)
    torch.cuda.synchronize()
    for (rays, method, _), (rg, og), (rw, ow) in zip(cases, got, want):
        assert torch.equal(og, ow), f"{method}: sorted ray order"
        assert torch.equal(rg, rw), f"{method}: results"
        assert int((rw[:, 0] >= 0).sum()) > 0
    # Again after the allocating calls: the same bits.
    res, order = _buffers(small.shape[0])
    _sorted_trace("hx_trace_sorted_noalloc", ctx, bvh, small, "twopoint", 48, res, order)
    torch.cuda.synchronize()
    assert torch.equal(order, want[1][1]) and torch.equal(res, want[1][0])


def test_noalloc_rejects_more_rays_than_its_scratch():
    ctx = hx.Context.default(0)
    v, t = scenes.uniform_soup(256, seed=2)
    bvh = hx.LBVHBuilder(ctx, morton_bits=30).build(hx.Scene(v, t))
    n = MAX_RAYS + 1
    rays = torch.zeros((n, 8), dtype=torch.float32, device="cuda")
    res, order = _buffers(n)
    with pytest.raises(_lib.HxError, match="exceed"):
        _sorted_trace("hx_trace_sorted_noalloc", ctx, bvh, rays, "aila", 60, res, order)
    assert int((res[:, 0] != -7).sum()) == 0   # nothing ran
