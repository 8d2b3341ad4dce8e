"""FP64 reference of the closest-hit adjoint's scatter, for tests.

For each hit ray with upstream dL/dt, the adjoint adds u*lambda, v*lambda and
(1-u-v)*lambda to the gradients of the triangle's three vertices, where
lambda = -(e1 x e2) * dL/dt / det and det = e1 . (e2 x -d) (the solve of
o + t d = v2 + u e1 + v e2).  This module computes each term with the adjoint
kernel's FP64 operations in the kernel's order (so a term equals the kernel's
before its rounding to float32, even where the solve is ill-conditioned, as on
grazing rays over slivers), sums them in FP64 with torch, and also returns,
per vertex coordinate, the sum of the terms' magnitudes S and their count K.  A float32 scatter-add of K terms (each rounded to float32 once)
in any order lies within K * 2^-24 * S of the exact sum, which is the bound
the GPU's atomically accumulated gradient is held to.
"""
import torch


def _cross(a, b):
    return torch.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1], a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                        a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], 1)


def _dot(a, b):
    return (a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]) + a[:, 2] * b[:, 2]


def adjoint_terms_fp64(vertices, triangles, rays, results, dl_dt, owner=None, tri_offsets=None, chunk=1 << 22):
    """Returns (g, S, K): FP64 exact-order gradient, sum of |terms| and term
    counts, each flattened to 3V.  `results` hold mesh-local triangle ids when
    `owner` (mesh of each ray) and `tri_offsets` (first triangle of each mesh)
    are given."""
    dev = vertices.device
    V = vertices.shape[0]
    g = torch.zeros(3 * V, dtype=torch.float64, device=dev)
    S = torch.zeros_like(g)
    K = torch.zeros(3 * V, dtype=torch.int64, device=dev)
    toff = None if tri_offsets is None else torch.as_tensor(tri_offsets, dtype=torch.int64, device=dev)
    R = rays.shape[0]
    for r0 in range(0, R, chunk):
        r1 = min(R, r0 + chunk)
        ids = results[r0:r1, 0].long()
        gt = dl_dt[r0:r1].double()
        hit = (ids >= 0) & (gt != 0)
        if not bool(hit.any()):
            continue
        idx = torch.nonzero(hit).squeeze(1)
        tri = ids[idx]
        if toff is not None:
            tri = tri + toff[owner[r0:r1][idx].long()]
        vid = triangles[tri].long()
        v0, v1, v2 = (vertices[vid[:, k]].double() for k in range(3))
        o = rays[r0:r1][idx, 0:3].double()
        d = rays[r0:r1][idx, 4:7].double()
        # The kernel's FP64 operations in its order (left-to-right dot
        # products, no fused multiply-adds), so each term equals its term.
        e1, e2, c, b = v0 - v2, v1 - v2, -d, o - v2
        bc, ca, ab = _cross(e2, c), _cross(c, e1), _cross(e1, e2)
        det = _dot(e1, bc)
        ok = det != 0
        inv = 1.0 / torch.where(ok, det, torch.ones_like(det))
        u = _dot(bc, b) * inv
        v = _dot(ca, b) * inv
        w = (1.0 - u) - v
        gtd = gt[idx]
        zero = torch.zeros_like(gtd)
        lam = torch.stack([-((bc[:, j] * zero + ca[:, j] * zero) + ab[:, j] * gtd) * inv for j in range(3)], 1)
        for k, beta in enumerate((u, v, w)):
            term = beta[:, None] * lam
            term = torch.where(ok[:, None], term, torch.zeros_like(term))
            for j in range(3):
                slot = 3 * vid[:, k] + j
                g.index_add_(0, slot, term[:, j])
                S.index_add_(0, slot, term[:, j].abs())
                K.index_add_(0, slot, ok.long())
    return g, S, K


def assert_scatter_within_bound(grad, g, S, K, what="gradient"):
    """|grad - g| <= (K + 1) * 2^-24 * S for every coordinate (float32 sums of
    K rounded terms in any order), plus a 1e-30 floor for exact zeros."""
    got = grad.reshape(-1).double()
    err = (got - g).abs()
    bound = (K.double() + 1.0) * 2.0 ** -24 * S + 1e-30
    bad = err > bound
    if bool(bad.any()):
        i = int(torch.argmax((err - bound) / bound.clamp_min(1e-30)))
        raise AssertionError(f"{what}: {int(bad.sum())} coordinates outside the float32 summation bound; worst at "
                             f"{i}: got {float(got[i])}, fp64 {float(g[i])}, bound {float(bound[i])} "
                             f"({int(K[i])} terms, sum of magnitudes {float(S[i])})")
