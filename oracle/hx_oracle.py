"""ctypes binding of oracle/liboracle.so -- the CPU restatement of hippie's LBVH path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg.  The product (hippie_amd/) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int
_f32 = ctypes.c_float


def build():
    """Compile liboracle.so with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "hxo_set_threads": (None, [_i32]),
            "hxo_get_threads": (_i32, []),
            "hxo_scene_box": (None, [_p, _p, _i64, _p]),
            "hxo_normalize": (None, [_p, _i64, _p]),
            "hxo_setup_boxes": (None, [_p, _p, _i64, _p, _p]),
            "hxo_morton_codes": (None, [_p, _p, _i64, _p, _i32, _p]),
            "hxo_sort_pairs": (None, [_p, _p, _i64, _i32]),
            "hxo_construct": (None, [_p, _i64, _p, _p, _p]),
            "hxo_refit": (None, [_p, _p, _p, _i64, _p, _p, _p, _p, _p]),
            "hxo_collapse": (_i64, [_i64, _p, _p, _p, _p, _p, _p, _i32, _i32, _f32, _f32,
                                    _p, _p, _p, _p]),
            "hxo_woopify": (None, [_p, _p, _p, _i64, _p]),
            "hxo_refit_final": (None, [_p, _i64, _i64, _p, _p, _p]),
            "hxo_cost": (ctypes.c_double, [_p, _i64, _f32, _f32, _f32]),
            "hxo_trace": (_i64, [_p, _p, _p, _p, _i64, _i32, _p, _p]),
            "hxo_trace_brute": (None, [_p, _p, _i64, _p, _i64, _p]),
            "hxo_bin_canonical": (_i64, [_p, _i64, _i64, _p, _p, _i64]),
            "hxo_validate_bin": (_i32, [_p, _i64, _i64, _p, _i64, _p, _p]),
            "hxo_validate_bin_refs": (_i32, [_p, _i64, _i64, _p, _i64, _i64, _p, _p]),
            "hxo_presplit": (_i64, [_p, _p, _i64, _p, _f32, _p, _p, _p, _i64, _p, _p]),
            "hxo_ps_pow": (_f32, [_f32, _f32]),
            "hxo_ps_split": (None, [_p, _p, _i32, _p, _p, _p, _p, _p]),
            "hxo_build_lbvh": (_i64, [_p, _p, _i64, _p, _i32, _i32, _i32, _f32, _f32,
                                      _p, _p, _p, _p, _p, _p, _f32]),
            "hxo_ploc": (_i32, [_i64, _i32, _i32, _i32, _p, _p, _p, _p, _p]),
            "hxo_atr": (_i64, [_i64, _i32, _i32, _f32, _f32, _p, _p, _p, _p, _p]),
            "hxo_build_atr": (_i64, [_p, _p, _i64, _p, _i32, _i32, _i32, _i32, _i32, _f32, _f32,
                                     _p, _p, _p, _p, _p, _p, _f32]),
            "hxo_build_sbvh": (_i64, [_p, _p, _i64, _f32, _i32, _f32, _f32, _p, _i64, _p, _i64, _p, _p, _p, _p]),
            "hxo_tr": (_i64, [_i64, _i32, _f32, _f32, _p, _p, _p, _p, _p]),
            "hxo_build_tr": (_i64, [_p, _p, _i64, _p, _i32, _i32, _i32, _i32, _f32, _f32,
                                    _p, _p, _p, _p, _p, _p]),
            "hxo_hlbvh": (_i64, [_i64, _i32, _i32, _p, _p, _p, _p, _p, _p]),
            "hxo_build_hlbvh": (_i64, [_p, _p, _i64, _p, _i32, _i32, _i32, _i32, _f32, _f32,
                                       _p, _p, _p, _p, _p, _p, _f32]),
            "hxo_build_ploc": (_i64, [_p, _p, _i64, _p, _i32, _i32, _i32, _i32, _i32, _i32, _f32, _f32,
                                      _p, _p, _p, _p, _p, _p, _f32]),
            "hxo_ins_find": (None, [_i64, _p, _p, _p, _p, _p, _p, _p]),
            "hxo_insertion": (_i64, [_i64, _i32, _f32, _f32, _p, _p, _p, _p, _p, _p]),
            "hxo_build_insertion": (_i64, [_p, _p, _i64, _p, _i32, _i32, _i32, _i32, _f32, _f32,
                                           _p, _p, _p, _p, _p, _p, _f32]),
            "hxo_ray_keys_aila64": (None, [_p, _i64, _i32, _p]),
            "hxo_ray_keys_twopoint64": (None, [_p, _i64, _i32, _f32, _p]),
            "hxo_tea16": (None, [_p, _p, _i64, _p]),
            "hxo_jenkins_mix": (None, [_p, _i64]),
            "hxo_halton": (None, [_p, _i64, _i32, _p]),
            "hxo_sobol2d": (None, [_p, _i64, _p]),
            "hxo_init_seeds": (None, [_i64, _i32, _p]),
            "hxo_primary_rays": (None, [_i32, _p, _p, _i32, _i32, _f32, _p, _p]),
            "hxo_shadow_rays": (None, [_i64, _i64, _i32, _p, _f32, _p, _p, _p, _p]),
            "hxo_ao_rays": (None, [_i64, _i64, _i32, _p, _f32, _p, _p, _p, _p, _p]),
            "hxo_path_rays": (_i64, [_i32, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
            "hxo_reconstruct_occlusion": (None, [_i64, _i64, _i32, _i32, _p, _p, _p, _p, _p]),
            "hxo_interpolate": (None, [_i64, _i64, _f32, _p, _p, _p, _i32]),
            "hxo_collapse_wide": (_i64, [_i32, _i64, _i32, _p, _p, _p, _p, _f32, _f32, _i32, _p, _p]),
            "hxo_cost_wide": (ctypes.c_double, [_i32, _p, _i64, _f32, _f32, _f32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def set_threads(n):
    lib().hxo_set_threads(int(n))


def get_threads():
    return lib().hxo_get_threads()


def normalize(verts):
    """Load-time normalisation: (normalised copy of verts, vertex box)."""
    v = np.array(verts, dtype=np.float32, order="C", copy=True)
    box = np.zeros(6, np.float32)
    lib().hxo_normalize(_ptr(v), v.shape[0], _ptr(box))
    return v, box


def scene_box(verts, tris):
    v, t = _c(verts, np.float32), _c(tris, np.int32)
    out = np.zeros(6, np.float32)
    lib().hxo_scene_box(_ptr(v), _ptr(t), t.shape[0], _ptr(out))
    return out


def setup_boxes(verts, tris):
    v, t = _c(verts, np.float32), _c(tris, np.int32)
    n = t.shape[0]
    bmin = np.zeros((n, 4), np.float32)
    bmax = np.zeros((n, 4), np.float32)
    lib().hxo_setup_boxes(_ptr(v), _ptr(t), n, _ptr(bmin), _ptr(bmax))
    return bmin, bmax


def morton_codes(bmin, bmax, box, bits):
    bmin, bmax = _c(bmin, np.float32), _c(bmax, np.float32)
    box = _c(box, np.float32)
    keys = np.zeros(bmin.shape[0], np.uint64)
    lib().hxo_morton_codes(_ptr(bmin), _ptr(bmax), bmin.shape[0], _ptr(box), int(bits), _ptr(keys))
    return keys


def sort_pairs(keys, vals, end_bit):
    k, v = _c(keys, np.uint64).copy(), _c(vals, np.int32).copy()
    lib().hxo_sort_pairs(_ptr(k), _ptr(v), k.shape[0], int(end_bit))
    return k, v


def construct(sorted_keys):
    k = _c(sorted_keys, np.uint64)
    n = k.shape[0]
    N = 2 * n - 1
    left, right, parent = (np.zeros(N, np.int32) for _ in range(3))
    lib().hxo_construct(_ptr(k), n, _ptr(left), _ptr(right), _ptr(parent))
    return left, right, parent


def refit(sorted_vals, ref_min, ref_max, left, right, parent):
    sv = _c(sorted_vals, np.int32)
    n = sv.shape[0]
    N = 2 * n - 1
    nmin = np.zeros((N, 4), np.float32)
    nmax = np.zeros((N, 4), np.float32)
    lib().hxo_refit(_ptr(sv), _ptr(_c(ref_min, np.float32)), _ptr(_c(ref_max, np.float32)), n,
                    _ptr(_c(left, np.int32)), _ptr(_c(right, np.int32)), _ptr(_c(parent, np.int32)),
                    _ptr(nmin), _ptr(nmax))
    return nmin, nmax


def collapse(sorted_vals, left, right, parent, nmin, nmax, adaptive=True, max_leaf_size=8,
             ct=3.0, ci=2.0):
    sv = _c(sorted_vals, np.int32)
    n = sv.shape[0]
    N = max(2 * n - 1, 2)
    nodes = np.zeros((N, 16), np.int32)
    tri = np.zeros(n, np.int32)
    ni, nl = ctypes.c_int64(0), ctypes.c_int64(0)
    nn = lib().hxo_collapse(n, _ptr(sv), _ptr(_c(left, np.int32)), _ptr(_c(right, np.int32)),
                            _ptr(_c(parent, np.int32)), _ptr(_c(nmin, np.float32)),
                            _ptr(_c(nmax, np.float32)), int(bool(adaptive)), int(max_leaf_size),
                            float(ct), float(ci), _ptr(nodes), _ptr(tri), ctypes.byref(ni),
                            ctypes.byref(nl))
    return nodes[:nn].copy(), tri, ni.value, nl.value


def woopify(verts, tris, tri_idx):
    v, t, ti = _c(verts, np.float32), _c(tris, np.int32), _c(tri_idx, np.int32)
    woop = np.zeros((ti.shape[0], 12), np.float32)
    lib().hxo_woopify(_ptr(v), _ptr(t), _ptr(ti), ti.shape[0], _ptr(woop))
    return woop


def refit_final(nodes, n_interior, n_leaves, tri_idx, verts, tris):
    """Refit a copy of the Bin nodes to new vertex positions (HipBVH::refit)."""
    nd = _c(nodes, np.int32).copy()
    lib().hxo_refit_final(_ptr(nd), int(n_interior), int(n_leaves), _ptr(_c(tri_idx, np.int32)),
                          _ptr(_c(verts, np.float32)), _ptr(_c(tris, np.int32)))
    return nd


def cost(nodes, scene_area, ct=3.0, ci=2.0):
    nd = _c(nodes, np.int32)
    return lib().hxo_cost(_ptr(nd), nd.shape[0], float(scene_area), float(ct), float(ci))


def trace(nodes, woop, tri_idx, rays, any_hit=False, want_stats=False):
    nd, w, ti = _c(nodes, np.int32), _c(woop, np.float32), _c(tri_idx, np.int32)
    r = _c(rays, np.float32).reshape(-1, 8)
    R = r.shape[0]
    res = np.zeros((R, 4), np.int32)
    stats = np.zeros((R, 2), np.int32) if want_stats else None
    ov = lib().hxo_trace(_ptr(nd), _ptr(w), _ptr(ti), _ptr(r), R, int(bool(any_hit)), _ptr(res),
                         _ptr(stats))
    return (res, stats, ov) if want_stats else (res, ov)


def trace_brute(verts, tris, rays):
    v, t = _c(verts, np.float32), _c(tris, np.int32)
    r = _c(rays, np.float32).reshape(-1, 8)
    res = np.zeros((r.shape[0], 4), np.int32)
    lib().hxo_trace_brute(_ptr(v), _ptr(t), t.shape[0], _ptr(r), r.shape[0], _ptr(res))
    return res


def bin_canonical(nodes, n_interior, n_leaves, tri_idx):
    nd, ti = _c(nodes, np.int32), _c(tri_idx, np.int32)
    m = lib().hxo_bin_canonical(_ptr(nd), int(n_interior), int(n_leaves), _ptr(ti), None, 0)
    out = np.zeros(m, np.uint32)
    lib().hxo_bin_canonical(_ptr(nd), int(n_interior), int(n_leaves), _ptr(ti), _ptr(out), m)
    return out


def validate_bin(nodes, n_interior, n_leaves, tri_idx, verts, tris):
    nd, ti = _c(nodes, np.int32), _c(tri_idx, np.int32)
    v, t = _c(verts, np.float32), _c(tris, np.int32)
    return lib().hxo_validate_bin(_ptr(nd), int(n_interior), int(n_leaves), _ptr(ti), ti.shape[0],
                                  _ptr(v), _ptr(t))


def build_lbvh(verts, tris, bits=60, scene_box=None, adaptive=True, max_leaf_size=8, ct=3.0, ci=2.0, presplit_beta=None):
    """Whole LBVH build; returns dict with nodes, tri_idx, woop, counts and stage seconds."""
    v, t = _c(verts, np.float32), _c(tris, np.int32)
    sb = None if scene_box is None else _c(scene_box, np.float32)
    beta = -1.0 if presplit_beta is None else float(presplit_beta)
    n = t.shape[0] if presplit_beta is None else presplit_count(v, t, beta, sb)
    N = max(2 * n - 1, 2)
    nodes = np.zeros((N, 16), np.int32)
    tri = np.zeros(n, np.int32)
    woop = np.zeros((n, 12), np.float32)
    ni, nl = ctypes.c_int64(0), ctypes.c_int64(0)
    sec = np.zeros(8, np.float64)
    nn = lib().hxo_build_lbvh(_ptr(v), _ptr(t), t.shape[0], _ptr(sb), int(bits), int(bool(adaptive)),
                              int(max_leaf_size), float(ct), float(ci), _ptr(nodes), _ptr(tri),
                              _ptr(woop), ctypes.byref(ni), ctypes.byref(nl), _ptr(sec), beta)
    return dict(nodes=nodes[:nn].copy(), tri_idx=tri, woop=woop, n_interior=ni.value,
                n_leaves=nl.value, seconds=sec)


def presplit_count(verts, tris, beta, scene_box=None):
    v, t = _c(verts, np.float32), _c(tris, np.int32)
    sb = None if scene_box is None else _c(scene_box, np.float32)
    return int(lib().hxo_presplit(_ptr(v), _ptr(t), t.shape[0], _ptr(sb), float(beta), None, None, None, 0, None, None))


def presplit(verts, tris, beta, scene_box=None):
    """Presplitter references: dict(ref_min, ref_max (float4 rows), ref_tri, D, priorities)."""
    v, t = _c(verts, np.float32), _c(tris, np.int32)
    sb = None if scene_box is None else _c(scene_box, np.float32)
    m = presplit_count(v, t, beta, sb)
    rmin = np.zeros((m, 4), np.float32)
    rmax = np.zeros((m, 4), np.float32)
    rtri = np.zeros(m, np.int32)
    D = ctypes.c_float(0.0)
    prio = np.zeros(t.shape[0], np.float32)
    lib().hxo_presplit(_ptr(v), _ptr(t), t.shape[0], _ptr(sb), float(beta), _ptr(rmin), _ptr(rmax), _ptr(rtri), m,
                       ctypes.byref(D), _ptr(prio))
    return dict(ref_min=rmin, ref_max=rmax, ref_tri=rtri, D=D.value, priorities=prio)


def ps_pow(x, y):
    return float(lib().hxo_ps_pow(float(x), float(y)))


def ps_split(tri_verts, box, count, scene_box):
    """One presplitter split; returns (left box[6], left count, right box[6], right count)."""
    v = _c(np.asarray(tri_verts, np.float32).reshape(9), np.float32)
    b = _c(np.asarray(box, np.float32).reshape(6), np.float32)
    sb = _c(np.asarray(scene_box, np.float32).reshape(6), np.float32)
    lb, rb = np.zeros(6, np.float32), np.zeros(6, np.float32)
    lc, rc = ctypes.c_int32(0), ctypes.c_int32(0)
    lib().hxo_ps_split(_ptr(v), _ptr(b), int(count), _ptr(sb), _ptr(lb), _ptr(rb), ctypes.byref(lc), ctypes.byref(rc))
    return lb, lc.value, rb, rc.value


def validate_bin_refs(nodes, n_interior, n_leaves, tri_idx, verts, tris):
    nd, ti = _c(nodes, np.int32), _c(tri_idx, np.int32)
    v, t = _c(verts, np.float32), _c(tris, np.int32)
    return lib().hxo_validate_bin_refs(_ptr(nd), int(n_interior), int(n_leaves), _ptr(ti), ti.shape[0], t.shape[0],
                                       _ptr(v), _ptr(t))


def ploc(leaf_min, leaf_max, radius, max_radius=512, adaptive=False):
    """PLOC over sorted leaf boxes (n x 4 each); returns left, right, parent, nmin, nmax (2n-1), rounds."""
    lmin, lmax = _c(leaf_min, np.float32), _c(leaf_max, np.float32)
    n = lmin.shape[0]
    N = 2 * n - 1
    left, right, parent = (np.zeros(N, np.int32) for _ in range(3))
    nmin = np.zeros((N, 4), np.float32)
    nmax = np.zeros((N, 4), np.float32)
    nmin[n - 1:] = lmin
    nmax[n - 1:] = lmax
    rounds = lib().hxo_ploc(n, int(radius), int(max_radius), int(bool(adaptive)), _ptr(left), _ptr(right), _ptr(parent), _ptr(nmin), _ptr(nmax))
    return left, right, parent, nmin, nmax, rounds


def build_ploc(verts, tris, bits=60, radius=25, scene_box=None, adaptive=True, max_leaf_size=8, ct=3.0, ci=2.0,
               max_radius=512, ploc_adaptive=False, presplit_beta=None):
    v, t = _c(verts, np.float32), _c(tris, np.int32)
    sb = None if scene_box is None else _c(scene_box, np.float32)
    beta = -1.0 if presplit_beta is None else float(presplit_beta)
    n = t.shape[0] if presplit_beta is None else presplit_count(v, t, beta, sb)
    N = max(2 * n - 1, 2)
    nodes = np.zeros((N, 16), np.int32)
    tri = np.zeros(n, np.int32)
    woop = np.zeros((n, 12), np.float32)
    ni, nl = ctypes.c_int64(0), ctypes.c_int64(0)
    sec = np.zeros(8, np.float64)
    nn = lib().hxo_build_ploc(_ptr(v), _ptr(t), t.shape[0], _ptr(sb), int(bits), int(radius), int(max_radius),
                              int(bool(ploc_adaptive)), int(bool(adaptive)),
                              int(max_leaf_size), float(ct), float(ci), _ptr(nodes), _ptr(tri), _ptr(woop),
                              ctypes.byref(ni), ctypes.byref(nl), _ptr(sec), beta)
    return dict(nodes=nodes[:nn].copy(), tri_idx=tri, woop=woop, n_interior=ni.value, n_leaves=nl.value,
                seconds=sec)


def atr(left, right, parent, nmin, nmax, treelet_size=9, iterations=2, ct=3.0, ci=2.0):
    """ATR optimisation of a binary tree (copies); returns left, right, parent, nmin, nmax, accepted."""
    l, r, p = (_c(a, np.int32).copy() for a in (left, right, parent))
    mn, mx = _c(nmin, np.float32).copy(), _c(nmax, np.float32).copy()
    n = (l.shape[0] + 1) // 2
    acc = lib().hxo_atr(n, int(treelet_size), int(iterations), float(ct), float(ci), _ptr(l), _ptr(r), _ptr(p),
                        _ptr(mn), _ptr(mx))
    return l, r, p, mn, mx, acc


def build_atr(verts, tris, bits=60, treelet_size=9, iterations=2, scene_box=None, adaptive=True, max_leaf_size=8,
              ct=3.0, ci=2.0, presplit_beta=None):
    v, t = _c(verts, np.float32), _c(tris, np.int32)
    sb = None if scene_box is None else _c(scene_box, np.float32)
    beta = -1.0 if presplit_beta is None else float(presplit_beta)
    n = t.shape[0] if presplit_beta is None else presplit_count(v, t, beta, sb)
    N = max(2 * n - 1, 2)
    nodes = np.zeros((N, 16), np.int32)
    tri = np.zeros(n, np.int32)
    woop = np.zeros((n, 12), np.float32)
    ni, nl, acc = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
    nn = lib().hxo_build_atr(_ptr(v), _ptr(t), t.shape[0], _ptr(sb), int(bits), int(treelet_size), int(iterations),
                             int(bool(adaptive)), int(max_leaf_size), float(ct), float(ci), _ptr(nodes), _ptr(tri),
                             _ptr(woop), ctypes.byref(ni), ctypes.byref(nl), ctypes.byref(acc), beta)
    return dict(nodes=nodes[:nn].copy(), tri_idx=tri, woop=woop, n_interior=ni.value, n_leaves=nl.value,
                accepted=acc.value)


def tr(left, right, parent, nmin, nmax, iterations=2, ct=3.0, ci=2.0):
    """TR optimisation of a binary tree (copies); returns left, right, parent, nmin, nmax, accepted."""
    l, r, p = (_c(a, np.int32).copy() for a in (left, right, parent))
    mn, mx = _c(nmin, np.float32).copy(), _c(nmax, np.float32).copy()
    n = (l.shape[0] + 1) // 2
    acc = lib().hxo_tr(n, int(iterations), float(ct), float(ci), _ptr(l), _ptr(r), _ptr(p), _ptr(mn), _ptr(mx))
    return l, r, p, mn, mx, acc


def build_tr(verts, tris, bits=60, iterations=2, scene_box=None, adaptive=True, max_leaf_size=8, ct=3.0, ci=2.0):
    v, t = _c(verts, np.float32), _c(tris, np.int32)
    sb = None if scene_box is None else _c(scene_box, np.float32)
    n = t.shape[0]
    N = max(2 * n - 1, 2)
    nodes = np.zeros((N, 16), np.int32)
    tri = np.zeros(n, np.int32)
    woop = np.zeros((n, 12), np.float32)
    ni, nl, acc = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
    nn = lib().hxo_build_tr(_ptr(v), _ptr(t), n, _ptr(sb), int(bits), int(iterations), int(bool(adaptive)),
                            int(max_leaf_size), float(ct), float(ci), _ptr(nodes), _ptr(tri), _ptr(woop),
                            ctypes.byref(ni), ctypes.byref(nl), ctypes.byref(acc))
    return dict(nodes=nodes[:nn].copy(), tri_idx=tri, woop=woop, n_interior=ni.value, n_leaves=nl.value,
                accepted=acc.value)


def build_sbvh(verts, tris, alpha=1e-5, max_leaf_size=8, ct=3.0, ci=2.0):
    """CPU SBVH (the GPU builders' quality reference) as Bin nodes; returns dict
    with nodes, tri_idx (reference slots), woop and counts."""
    v, t = _c(verts, np.float32), _c(tris, np.int32)
    n = t.shape[0]
    cap = 4 * n + 64
    while True:
        nodes = np.zeros((2 * cap, 16), np.int32)
        tri = np.zeros(cap, np.int32)
        woop = np.zeros((cap, 12), np.float32)
        ni, nl, nr = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
        nn = lib().hxo_build_sbvh(_ptr(v), _ptr(t), n, float(alpha), int(max_leaf_size), float(ct), float(ci),
                                  _ptr(nodes), 2 * cap, _ptr(tri), cap, _ptr(woop), ctypes.byref(ni), ctypes.byref(nl),
                                  ctypes.byref(nr))
        if nn != -1:
            break
        cap *= 2
    if nn < 0:
        raise ValueError("the SBVH needs at least two triangles")
    r = nr.value
    return dict(nodes=nodes[:nn].copy(), tri_idx=tri[:r].copy(), woop=woop[:r].copy(), n_interior=ni.value,
                n_leaves=nl.value, n_refs=r)


def hlbvh(sorted_keys, left, right, parent, nmin, nmax, bits=60, sah_bits=15):
    """HLBVH upper levels over a Karras tree (copies); returns left, right, parent, nmin, nmax, clusters."""
    k = _c(sorted_keys, np.uint64)
    l, r, p = (_c(a, np.int32).copy() for a in (left, right, parent))
    mn, mx = _c(nmin, np.float32).copy(), _c(nmax, np.float32).copy()
    C = lib().hxo_hlbvh(k.shape[0], int(bits), int(sah_bits), _ptr(k), _ptr(l), _ptr(r), _ptr(p), _ptr(mn),
                        _ptr(mx))
    return l, r, p, mn, mx, C


def build_hlbvh(verts, tris, bits=60, sah_bits=15, scene_box=None, adaptive=True, max_leaf_size=8, ct=3.0,
                ci=2.0, presplit_beta=None):
    v, t = _c(verts, np.float32), _c(tris, np.int32)
    sb = None if scene_box is None else _c(scene_box, np.float32)
    beta = -1.0 if presplit_beta is None else float(presplit_beta)
    n = t.shape[0] if presplit_beta is None else presplit_count(v, t, beta, sb)
    N = max(2 * n - 1, 2)
    nodes = np.zeros((N, 16), np.int32)
    tri = np.zeros(n, np.int32)
    woop = np.zeros((n, 12), np.float32)
    ni, nl, cl = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
    nn = lib().hxo_build_hlbvh(_ptr(v), _ptr(t), t.shape[0], _ptr(sb), int(bits), int(sah_bits), int(bool(adaptive)),
                               int(max_leaf_size), float(ct), float(ci), _ptr(nodes), _ptr(tri), _ptr(woop),
                               ctypes.byref(ni), ctypes.byref(nl), ctypes.byref(cl), beta)
    return dict(nodes=nodes[:nn].copy(), tri_idx=tri, woop=woop, n_interior=ni.value, n_leaves=nl.value,
                clusters=cl.value)


def ins_find(left, right, parent, nmin, nmax):
    """findBestNode for every node of a binary tree; returns out (int32), reduction (float32)."""
    l, r, p = (_c(a, np.int32) for a in (left, right, parent))
    mn, mx = _c(nmin, np.float32), _c(nmax, np.float32)
    N = l.shape[0]
    n = (N + 1) // 2
    out = np.zeros(N, np.int32)
    red = np.zeros(N, np.float32)
    lib().hxo_ins_find(n, _ptr(l), _ptr(r), _ptr(p), _ptr(mn), _ptr(mx), _ptr(out), _ptr(red))
    return out, red


def insertion(left, right, parent, nmin, nmax, mod=1, ct=3.0, ci=2.0):
    """Insertion optimisation of a binary tree (copies); returns left, right, parent, nmin, nmax, reinserted, steps."""
    l, r, p = (_c(a, np.int32).copy() for a in (left, right, parent))
    mn, mx = _c(nmin, np.float32).copy(), _c(nmax, np.float32).copy()
    n = (l.shape[0] + 1) // 2
    steps = ctypes.c_int64(0)
    k = lib().hxo_insertion(n, int(mod), float(ct), float(ci), _ptr(l), _ptr(r), _ptr(p), _ptr(mn), _ptr(mx),
                            ctypes.byref(steps))
    return l, r, p, mn, mx, k, steps.value


def build_insertion(verts, tris, bits=60, mod=1, scene_box=None, adaptive=True, max_leaf_size=8, ct=3.0, ci=2.0, presplit_beta=None):
    v, t = _c(verts, np.float32), _c(tris, np.int32)
    sb = None if scene_box is None else _c(scene_box, np.float32)
    beta = -1.0 if presplit_beta is None else float(presplit_beta)
    n = t.shape[0] if presplit_beta is None else presplit_count(v, t, beta, sb)
    N = max(2 * n - 1, 2)
    nodes = np.zeros((N, 16), np.int32)
    tri = np.zeros(n, np.int32)
    woop = np.zeros((n, 12), np.float32)
    ni, nl, steps = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
    nn = lib().hxo_build_insertion(_ptr(v), _ptr(t), t.shape[0], _ptr(sb), int(bits), int(mod), int(bool(adaptive)),
                                   int(max_leaf_size), float(ct), float(ci), _ptr(nodes), _ptr(tri), _ptr(woop),
                                   ctypes.byref(ni), ctypes.byref(nl), ctypes.byref(steps), beta)
    return dict(nodes=nodes[:nn].copy(), tri_idx=tri, woop=woop, n_interior=ni.value, n_leaves=nl.value,
                steps=steps.value)


def ray_keys_aila64(rays, bits=60):
    r = _c(rays, np.float32).reshape(-1, 8)
    keys = np.zeros(r.shape[0], np.uint64)
    lib().hxo_ray_keys_aila64(_ptr(r), r.shape[0], int(bits), _ptr(keys))
    return keys


def ray_keys_twopoint64(rays, bits=60, ray_length=0.25):
    r = _c(rays, np.float32).reshape(-1, 8)
    keys = np.zeros(r.shape[0], np.uint64)
    lib().hxo_ray_keys_twopoint64(_ptr(r), r.shape[0], int(bits), float(ray_length), _ptr(keys))
    return keys


# --- ray generation, reconstruction, interpolation (hx_oracle_raygen.cpp) -----

MATERIAL_DTYPE = np.dtype([("tex_index", "<i4"), ("shininess", "<f4"), ("diffuse", "<f4", (3,)),
                           ("specular", "<f4", (3,))])


def tea16(v0, v1):
    """(v0, v1) after 16 rounds, shape (n, 2); the seed is column 0."""
    v0, v1 = _c(v0, np.uint32), _c(v1, np.uint32)
    out = np.empty((v0.size, 2), np.uint32)
    lib().hxo_tea16(_ptr(v0), _ptr(v1), v0.size, _ptr(out))
    return out


def jenkins_mix(abc):
    a = _c(abc, np.uint32).copy()
    lib().hxo_jenkins_mix(_ptr(a), a.shape[0])
    return a


def halton(idx, base):
    i = _c(idx, np.int32)
    out = np.empty(i.shape, np.float32)
    lib().hxo_halton(_ptr(i), i.size, int(base), _ptr(out))
    return out


def sobol2d(idx):
    i = _c(idx, np.int32)
    out = np.empty((i.size, 2), np.float32)
    lib().hxo_sobol2d(_ptr(i), i.size, _ptr(out))
    return out


def init_seeds(n, frame):
    out = np.empty(n, np.uint32)
    lib().hxo_init_seeds(int(n), int(frame), _ptr(out))
    return out


def primary_rays(sample_index, origin, m, width, height, max_dist, index_to_pixel=None):
    o = _c(origin, np.float32)
    mm = _c(m, np.float32).reshape(16)
    itp = None if index_to_pixel is None else _c(index_to_pixel, np.int32)
    out = np.empty((width * height, 8), np.float32)
    lib().hxo_primary_rays(int(sample_index), _ptr(o), _ptr(mm), int(width), int(height), float(max_dist),
                           _ptr(itp), _ptr(out))
    return out


def shadow_rays(begin, size, n_samples, seeds, light_radius, light, in_rays, in_results):
    s, L = _c(seeds, np.uint32), _c(light, np.float32)
    r, res = _c(in_rays, np.float32), _c(in_results, np.int32)
    out = np.empty((size * n_samples, 8), np.float32)
    lib().hxo_shadow_rays(int(begin), int(size), int(n_samples), _ptr(s), float(light_radius), _ptr(L), _ptr(r),
                          _ptr(res), _ptr(out))
    return out


def ao_rays(begin, size, n_samples, seeds, max_dist, in_rays, in_results, verts, tris):
    s, r, res = _c(seeds, np.uint32), _c(in_rays, np.float32), _c(in_results, np.int32)
    t, v = _c(tris, np.int32), _c(verts, np.float32)
    out = np.empty((size * n_samples, 8), np.float32)
    lib().hxo_ao_rays(int(begin), int(size), int(n_samples), _ptr(s), float(max_dist), _ptr(r), _ptr(res), _ptr(t),
                      _ptr(v), _ptr(out))
    return out


def path_rays(russian_roulette, seeds, in_pixel, mat_indices, tris, normals, materials, in_rays, in_results,
              decreases):
    """Returns (out_rays, out_pixel, seeds after, decreases after)."""
    seeds = _c(seeds, np.uint32).copy()
    dec = _c(decreases, np.float32).copy()
    n = seeds.shape[0]
    mats = np.ascontiguousarray(materials, MATERIAL_DTYPE)
    ip, mi, t, nn = _c(in_pixel, np.int32), _c(mat_indices, np.int32), _c(tris, np.int32), _c(normals, np.float32)
    r, res = _c(in_rays, np.float32), _c(in_results, np.int32)
    out = np.empty((n, 8), np.float32)
    op = np.empty(n, np.int32)
    k = lib().hxo_path_rays(int(bool(russian_roulette)), n, _ptr(seeds), _ptr(ip), _ptr(mi), _ptr(t), _ptr(nn),
                            mats.ctypes.data_as(ctypes.c_void_p), _ptr(r), _ptr(res), _ptr(out), _ptr(op), _ptr(dec))
    return out[:k], op[:k], seeds, dec


def reconstruct_occlusion(begin, size, n_samples, replace, results, index_to_pixel, in_pixels, out_pixels,
                          index_to_slot=None):
    res, itp = _c(results, np.int32), _c(index_to_pixel, np.int32)
    its = None if index_to_slot is None else _c(index_to_slot, np.int32)
    inp = _c(in_pixels, np.float32)
    out = _c(out_pixels, np.float32).copy()
    lib().hxo_reconstruct_occlusion(int(begin), int(size), int(n_samples), int(bool(replace)), _ptr(res), _ptr(itp),
                                    _ptr(its), _ptr(inp), _ptr(out))
    return out


def interpolate(n_static, frame0, frame1, t, out, normalize=False):
    a, b = _c(frame0, np.float32), _c(frame1, np.float32)
    o = _c(out, np.float32).copy()
    lib().hxo_interpolate(int(n_static), a.shape[0], float(t), _ptr(a), _ptr(b), _ptr(o), int(bool(normalize)))
    return o


# --- wide (Quad / Oct) layouts (hx_oracle_wide.cpp) ---------------------------

def collapse_wide(N, n, root, left, right, nmin, nmax, ct=3.0, ci=2.0, max_leaf_size=0):
    """Wide nodes (int32 (nodes, 8N)) and the interior count of a binary tree
    (internal 0..n-2, leaves n-1+p, float4 boxes); max_leaf_size <= 0: adaptive."""
    l, r = _c(left, np.int32), _c(right, np.int32)
    mn, mx = _c(nmin, np.float32), _c(nmax, np.float32)
    out = np.zeros((2 * n, 8 * N), np.int32)
    ni = ctypes.c_int64()
    k = lib().hxo_collapse_wide(int(N), int(n), int(root), _ptr(l), _ptr(r), _ptr(mn), _ptr(mx), float(ct),
                                float(ci), int(max_leaf_size), _ptr(out), ctypes.byref(ni))
    return out[:k].copy(), int(ni.value)


def cost_wide(N, nodes, scene_area, ct=3.0, ci=2.0):
    w = _c(nodes, np.int32)
    return lib().hxo_cost_wide(int(N), _ptr(w), w.shape[0], float(scene_area), float(ct), float(ci))
