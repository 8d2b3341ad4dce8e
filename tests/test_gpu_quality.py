"""Tree quality of the GPU builders against the CPU SBVH quality reference.

On a 60,000-triangle clustered soup and a 40,000-triangle sliver soup, every
GPU builder's SAH (FP64 sum over the Bin tree) is printed as a ratio to the
SBVH's (oracle/hx_oracle_sbvh.cpp, SBVHBuilder with alpha = 1e-5 and the
adaptive collapse).  The optimising builders (PLOC, ATRBVH, TR, insertion)
must lie below the plain LBVH, and the SBVH below the LBVH.  The presplit LBVH
is printed only: presplitting adds references, which can raise the SAH of a
soup of small triangles, and at beta = 0.3 it splits no sliver at all (their
priorities are nearly equal, so the reference's six bisection rounds of D
leave every floor(D p) at zero).
"""
import numpy as np
import pytest

import hippie_amd as hx
from hippie_amd import scenes

pytestmark = pytest.mark.gpu

BUILDERS = [("lbvh", hx.LBVHBuilder, {}), ("ploc_r25", hx.PLOCBuilder, dict(ploc_radius=25)),
            ("atr_t9", hx.ATRBuilder, dict(atr_treelet_size=9)), ("tr", hx.TRBuilder, {}),
            ("hlbvh", hx.HLBVHBuilder, {}), ("insertion", hx.InsertionBuilder, {}),
            ("presplit_lbvh", hx.LBVHBuilder, dict(presplitting=1, presplitter_beta=0.3))]


@pytest.mark.parametrize("kind", ["clustered", "sliver"])
def test_builders_against_sbvh(oracle, ctx, kind):
    if kind == "clustered":
        v, t = scenes.gaussian_mixture_soup(60000, seed=71)
    else:
        v, t = scenes.sliver_soup(40000, np.random.default_rng(72))
    scene = hx.Scene(v, t)
    costs = {}
    for name, cls, kw in BUILDERS:
        bvh = cls(ctx, morton_bits=60, **kw).build(scene)
        costs[name] = bvh.get_cost()
        box = bvh.scene_box()
    d = box[3:6] - box[0:3]
    area = float(np.float32(2.0) * np.float32(d[0] * d[1] + d[1] * d[2] + d[2] * d[0]))
    sbvh = oracle.cost(oracle.build_sbvh(v, t)["nodes"], area)
    print(f"{kind}: SBVH SAH {sbvh:.3f}; " + ", ".join(f"{k} {c / sbvh:.3f}x" for k, c in costs.items()))
    assert sbvh < costs["lbvh"]
    for name in ("ploc_r25", "atr_t9", "tr", "insertion"):
        assert costs[name] < costs["lbvh"], name
