/* hx_oracle_sbvh.cpp -- CPU SBVH (spatial-split BVH) restatement: the quality
 * reference of the GPU builders.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's report of the
 * builders' SAH against it; the product never links it.
 *
 * Restates src/hippie/rt/bvh/SBVHBuilder.cpp:14-401 (Stich, Friedrich and
 * Dietrich 2009 as the reference implements it): object splits by a full
 * sort of the references along each axis (key = box min + max on the axis,
 * ties by triangle index; lowest SAH, ties by the more balanced split),
 * spatial splits over 128 bins per axis when the object split's children
 * overlap by at least alpha times the scene box area, reference clipping
 * against the split plane, the three-way choice (unsplit left, unsplit right,
 * duplicate) for straddling references, degenerate references removed at
 * every node, a leaf at one reference or when the leaf SAH is the lowest and
 * the node holds at most max_leaf_size references, spatial splits only above
 * depth 48, the right child built first (references are a stack).  Then the
 * adaptive collapse of BVHConverter::convertAdaptive (a subtree becomes a leaf
 * when ci A size < ct A + C_l + C_r; the root stays an interior) and the Bin
 * emission: interiors in depth-first order, then leaves in depth-first order,
 * leaf reference ranges in depth-first order.
 *
 * Float semantics follow util/AABB.h and util/Math.h: MIN(a, b) = b < a ? b :
 * a, MAX(a, b) = b > a ? b : a; area 2 (dx dy + dy dz + dz dx), 0 for an
 * invalid box; mix(lo, hi, t) = lo (1 - t) + hi t; the bin index is the float
 * truncated to int and clamped.  -ffp-contract=off (oracle/Makefile). */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "hx_oracle.h"

namespace {

constexpr float kMaxF = 3.402823466e+38f;
constexpr int kBins = 128;      // NUMBER_OF_BINS, SBVHBuilder.h:16
constexpr int kMaxDepth = 48;   // MAX_DEPTH, SBVHBuilder.h:15

inline float fmin_r(float a, float b) { return (b < a) ? b : a; }
inline float fmax_r(float a, float b) { return (b > a) ? b : a; }

struct Aabb {
    float mn[3] = {kMaxF, kMaxF, kMaxF};
    float mx[3] = {-kMaxF, -kMaxF, -kMaxF};
    void grow_pt(const float* p) {
        for (int k = 0; k < 3; ++k) {
            mn[k] = fmin_r(mn[k], p[k]);
            mx[k] = fmax_r(mx[k], p[k]);
        }
    }
    void grow(const Aabb& b) {
        grow_pt(b.mn);
        grow_pt(b.mx);
    }
    void intersect(const Aabb& b) {
        for (int k = 0; k < 3; ++k) {
            mn[k] = fmax_r(mn[k], b.mn[k]);
            mx[k] = fmin_r(mx[k], b.mx[k]);
        }
    }
    bool valid() const { return mn[0] <= mx[0] && mn[1] <= mx[1] && mn[2] <= mx[2]; }
    float area() const {
        if (!valid()) return 0.0f;
        const float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct Ref {
    int index = -1;
    Aabb box;
};

struct Spec {
    int size = 0;
    Aabb box;
};

struct ObjSplit {
    float sah = kMaxF;
    int axis = 0;
    int left_count = 0;
    Aabb left_box, right_box;
};

struct SpaSplit {
    float sah = kMaxF;
    int axis = 0;
    float position = 0.0f;
};

struct Bin {
    Aabb box;
    int enter = 0, exit = 0;
};

struct TreeNode {
    Aabb box;
    int child[2] = {-1, -1};
    int begin = 0, end = 0;  // leaf: range in tri (build order)
    bool leaf = false;
};

struct Sbvh {
    const float* V;
    const int32_t* T;
    float alpha, ct, ci;
    int max_leaf;
    float scene_area = 0.0f;
    std::vector<Ref> refs;
    std::vector<Aabb> right_boxes;
    Bin bins[3][kBins];
    std::vector<int> tri;
    std::vector<TreeNode> nodes;

    const float* vert(int i) const { return V + 3 * (int64_t)i; }

    /* findObjectSplit (:14-52). */
    ObjSplit find_object_split(const Spec& spec, float node_sah) {
        ObjSplit split;
        const int first = (int)refs.size() - spec.size;
        float best_tie = kMaxF;
        for (int axis = 0; axis < 3; ++axis) {
            std::sort(refs.begin() + first, refs.end(), [axis](const Ref& l, const Ref& r) {
                const float lv = l.box.mn[axis] + l.box.mx[axis];
                const float rv = r.box.mn[axis] + r.box.mx[axis];
                return lv < rv || (lv == rv && l.index < r.index);
            });
            const Ref* rp = refs.data() + first;
            Aabb right;
            for (int i = spec.size - 1; i > 0; --i) {
                right.grow(rp[i].box);
                right_boxes[i - 1] = right;
            }
            Aabb left;
            for (int i = 1; i < spec.size; ++i) {
                left.grow(rp[i - 1].box);
                const float sah = node_sah + left.area() * ci * (float)i +
                                  right_boxes[i - 1].area() * ci * (float)(spec.size - i);
                const float tie = (float)(i * i + (spec.size - i) * (spec.size - i));
                if (sah < split.sah || (sah == split.sah && tie < best_tie)) {
                    split.sah = sah;
                    split.axis = axis;
                    split.left_count = i;
                    split.left_box = left;
                    split.right_box = right_boxes[i - 1];
                    best_tie = tie;
                }
            }
        }
        return split;
    }

    /* performObjectSplit (:54-61). */
    void perform_object_split(Spec& l, Spec& r, const Spec& spec, const ObjSplit& split) {
        const int first = (int)refs.size() - spec.size;
        const int axis = split.axis;
        std::sort(refs.begin() + first, refs.end(), [axis](const Ref& a, const Ref& b) {
            const float av = a.box.mn[axis] + a.box.mx[axis];
            const float bv = b.box.mn[axis] + b.box.mx[axis];
            return av < bv || (av == bv && a.index < b.index);
        });
        l.size = split.left_count;
        l.box = split.left_box;
        r.size = spec.size - split.left_count;
        r.box = split.right_box;
    }

    /* splitReference (:212-248). */
    void split_reference(Ref& l, Ref& r, const Ref& ref, int axis, float pos) {
        l.index = r.index = ref.index;
        l.box = r.box = Aabb();
        const int32_t* ind = T + 3 * (int64_t)ref.index;
        const float* v1 = vert(ind[2]);
        for (int i = 0; i < 3; ++i) {
            const float* v0 = v1;
            v1 = vert(ind[i]);
            const float v0p = v0[axis], v1p = v1[axis];
            if (v0p <= pos) l.box.grow_pt(v0);
            if (v0p >= pos) r.box.grow_pt(v0);
            if ((v0p < pos && v1p > pos) || (v0p > pos && v1p < pos)) {
                float t = (pos - v0p) / (v1p - v0p);
                t = fmax_r(fmin_r(t, 1.0f), 0.0f);  // scalar clamp: MAX(MIN(v, hi), lo)
                float p[3];
                for (int k = 0; k < 3; ++k) p[k] = v0[k] * (1.0f - t) + v1[k] * t;
                l.box.grow_pt(p);
                r.box.grow_pt(p);
            }
        }
        l.box.mx[axis] = pos;
        r.box.mn[axis] = pos;
        l.box.intersect(ref.box);
        r.box.intersect(ref.box);
    }

    static int to_bin(float x) {
        /* Vec3f -> Vec3i truncates (x86 cvttss2si: NaN and out-of-range give INT_MIN). */
        if (!(x > -2147483648.0f && x < 2147483648.0f)) return INT32_MIN;
        return (int)x;
    }

    /* findSpatialSplit (:63-133). */
    SpaSplit find_spatial_split(const Spec& spec, float node_sah) {
        float origin[3], bin_size[3], inv[3];
        for (int k = 0; k < 3; ++k) {
            origin[k] = spec.box.mn[k];
            bin_size[k] = (spec.box.mx[k] - origin[k]) * (1.0f / (float)kBins);
            inv[k] = 1.0f / bin_size[k];
        }
        for (int axis = 0; axis < 3; ++axis)
            for (int i = 0; i < kBins; ++i) bins[axis][i] = Bin();
        for (int ri = (int)refs.size() - spec.size; ri < (int)refs.size(); ++ri) {
            const Ref ref = refs[ri];
            int fb[3], lb[3];
            for (int k = 0; k < 3; ++k) {
                /* vector clamp: MIN(MAX(v, lo), hi) */
                fb[k] = std::min(std::max(to_bin((ref.box.mn[k] - origin[k]) * inv[k]), 0), kBins - 1);
                lb[k] = std::min(std::max(to_bin((ref.box.mx[k] - origin[k]) * inv[k]), fb[k]), kBins - 1);
            }
            for (int axis = 0; axis < 3; ++axis) {
                Ref cur = ref;
                for (int i = fb[axis]; i < lb[axis]; ++i) {
                    Ref lr, rr;
                    split_reference(lr, rr, cur, axis, origin[axis] + bin_size[axis] * (float)(i + 1));
                    bins[axis][i].box.grow(lr.box);
                    cur = rr;
                }
                bins[axis][lb[axis]].box.grow(cur.box);
                bins[axis][fb[axis]].enter++;
                bins[axis][lb[axis]].exit++;
            }
        }
        SpaSplit split;
        for (int axis = 0; axis < 3; ++axis) {
            Aabb right;
            for (int i = kBins - 1; i > 0; --i) {
                right.grow(bins[axis][i].box);
                right_boxes[i - 1] = right;
            }
            Aabb left;
            int left_num = 0, right_num = spec.size;
            for (int i = 1; i < kBins; ++i) {
                left.grow(bins[axis][i - 1].box);
                left_num += bins[axis][i - 1].enter;
                right_num -= bins[axis][i - 1].exit;
                const float sah = node_sah + left.area() * ci * (float)left_num +
                                  right_boxes[i - 1].area() * ci * (float)right_num;
                if (sah < split.sah) {
                    split.sah = sah;
                    split.axis = axis;
                    split.position = origin[axis] + bin_size[axis] * (float)i;
                }
            }
        }
        return split;
    }

    /* performSpatialSplit (:135-210). */
    void perform_spatial_split(Spec& l, Spec& r, const Spec& spec, const SpaSplit& split) {
        const int axis = split.axis;
        const float pos = split.position;
        const int left_start = (int)refs.size() - spec.size;
        int left_end = left_start;
        int right_start = (int)refs.size();
        l.box = r.box = Aabb();
        for (int i = left_end; i < right_start; ++i) {
            if (refs[i].box.mx[axis] <= pos) {
                l.box.grow(refs[i].box);
                std::swap(refs[i], refs[left_end++]);
            } else if (refs[i].box.mn[axis] >= pos) {
                r.box.grow(refs[i].box);
                std::swap(refs[i--], refs[--right_start]);
            }
        }
        while (left_end < right_start) {
            Ref lref, rref;
            split_reference(lref, rref, refs[left_end], axis, pos);
            Aabb lub = l.box, rub = r.box, ldb = l.box, rdb = r.box;
            lub.grow(refs[left_end].box);
            rub.grow(refs[left_end].box);
            ldb.grow(lref.box);
            rdb.grow(rref.box);
            const float lac = ci * (float)(left_end - left_start);
            const float rac = ci * (float)((int)refs.size() - right_start);
            const float lbc = ci * (float)(left_end - left_start + 1);
            const float rbc = ci * (float)((int)refs.size() - right_start + 1);
            const float unsplit_left = lub.area() * lbc + r.box.area() * rac;
            const float unsplit_right = l.box.area() * lac + rub.area() * rbc;
            const float duplicate = ldb.area() * lbc + rdb.area() * rbc;
            const float min_sah = fmin_r(fmin_r(unsplit_left, unsplit_right), duplicate);
            if (min_sah == unsplit_left) {
                l.box = lub;
                left_end++;
            } else if (min_sah == unsplit_right) {
                r.box = rub;
                std::swap(refs[left_end], refs[--right_start]);
            } else {
                l.box = ldb;
                r.box = rdb;
                refs[left_end++] = lref;
                refs.push_back(rref);
            }
        }
        l.size = left_end - left_start;
        r.size = (int)refs.size() - right_start;
    }

    /* buildLeaf (:250-261). */
    int build_leaf(const Spec& spec) {
        for (int i = 0; i < spec.size; ++i) {
            tri.push_back(refs.back().index);
            refs.pop_back();
        }
        TreeNode nd;
        nd.leaf = true;
        nd.box = spec.box;
        nd.begin = (int)tri.size() - spec.size;
        nd.end = (int)tri.size();
        nodes.push_back(nd);
        return (int)nodes.size() - 1;
    }

    /* buildNode (:263-326). */
    int build_node(Spec spec, int depth) {
        {
            const int first = (int)refs.size() - spec.size;
            for (int i = (int)refs.size() - 1; i >= first; --i) {
                const Aabb& b = refs[i].box;
                const float sx = b.mx[0] - b.mn[0], sy = b.mx[1] - b.mn[1], sz = b.mx[2] - b.mn[2];
                const float mn3 = fmin_r(fmin_r(sx, sy), sz), mx3 = fmax_r(fmax_r(sx, sy), sz);
                if (mn3 < 0.0f || (sx + sy + sz) == mx3) {
                    refs[i] = refs.back();
                    refs.pop_back();
                }
            }
            spec.size = (int)refs.size() - first;
        }
        if (spec.size == 1) return build_leaf(spec);
        const float area = spec.box.area();
        const float leaf_sah = area * ci * (float)spec.size;
        const float node_sah = area * ct;
        const ObjSplit object = find_object_split(spec, node_sah);
        SpaSplit spatial;
        if (depth < kMaxDepth) {
            Aabb overlap = object.left_box;
            overlap.intersect(object.right_box);
            if (overlap.area() >= scene_area * alpha) spatial = find_spatial_split(spec, node_sah);
        }
        const float min_sah = fmin_r(fmin_r(leaf_sah, object.sah), spatial.sah);
        if (min_sah == leaf_sah && spec.size <= max_leaf) return build_leaf(spec);
        Spec l, r;
        if (min_sah == spatial.sah) perform_spatial_split(l, r, spec, spatial);
        if (!l.size || !r.size) perform_object_split(l, r, spec, object);
        const int rn = build_node(r, depth + 1);
        const int ln = build_node(l, depth + 1);
        TreeNode nd;
        nd.box = spec.box;
        nd.child[0] = ln;
        nd.child[1] = rn;
        nodes.push_back(nd);
        return (int)nodes.size() - 1;
    }
};

void put_box(int32_t* w, int slot, const Aabb& b) {
    auto f = [](float x) {
        int32_t i;
        std::memcpy(&i, &x, 4);
        return i;
    };
    w[4 * slot + 0] = f(b.mn[0]);
    w[4 * slot + 1] = f(b.mx[0]);
    w[4 * slot + 2] = f(b.mn[1]);
    w[4 * slot + 3] = f(b.mx[1]);
    w[8 + 2 * slot] = f(b.mn[2]);
    w[9 + 2 * slot] = f(b.mx[2]);
}

}  // namespace

extern "C" {

int64_t hxo_build_sbvh(const float* verts, const int32_t* tris, int64_t n_tri, float alpha, int max_leaf_size, float ct,
                       float ci, int32_t* nodes_out, int64_t node_cap, int32_t* tri_out, int64_t ref_cap, float* woop,
                       int64_t* n_interior, int64_t* n_leaves, int64_t* n_refs) {
    if (n_tri < 2) return -2;
    Sbvh s;
    s.V = verts;
    s.T = tris;
    s.alpha = alpha;
    s.ct = ct;
    s.ci = ci;
    s.max_leaf = max_leaf_size;
    Spec root;
    root.size = (int)n_tri;
    s.refs.resize(n_tri);
    for (int64_t i = 0; i < n_tri; ++i) {
        s.refs[i].index = (int)i;
        for (int j = 0; j < 3; ++j) s.refs[i].box.grow_pt(s.vert(tris[3 * i + j]));
        root.box.grow(s.refs[i].box);
    }
    s.scene_area = root.box.area();
    s.right_boxes.resize(std::max<int64_t>(n_tri, kBins) - 1);
    const int top = s.build_node(root, 0);

    /* BVHConverter::convertAdaptive: subtree sizes, the adaptive collapse and
     * the depth-first emission. */
    const int nn = (int)s.nodes.size();
    std::vector<float> cost(nn);
    std::vector<int> size(nn), state(nn);
    std::vector<int> order;  // post-order over the tree
    {
        std::vector<std::pair<int, int>> st{{top, 0}};
        while (!st.empty()) {
            auto& [v, k] = st.back();
            if (!s.nodes[v].leaf && k < 2) {
                const int c = s.nodes[v].child[k++];
                st.push_back({c, 0});
                continue;
            }
            order.push_back(v);
            st.pop_back();
        }
    }
    for (int v : order) {
        const TreeNode& nd = s.nodes[v];
        const float A = nd.box.area();
        if (nd.leaf) {
            size[v] = nd.end - nd.begin;
            cost[v] = ci * A * (float)size[v];
            state[v] = 0;
        } else {
            size[v] = size[nd.child[0]] + size[nd.child[1]];
            const float as_node = ct * A + cost[nd.child[0]] + cost[nd.child[1]];
            const float as_leaf = ci * A * (float)size[v];
            if (as_leaf < as_node) {
                cost[v] = as_leaf;
                state[v] = 0;
            } else {
                cost[v] = as_node;
                state[v] = 1;
            }
        }
    }
    if (!s.nodes[top].leaf) state[top] = 1;
    else return -2;
    /* Interiors in depth-first (pre-)order, then leaves in depth-first order. */
    std::vector<int> interiors, leaves;
    {
        std::vector<int> st{top};
        while (!st.empty()) {
            const int v = st.back();
            st.pop_back();
            if (state[v] == 1) {
                interiors.push_back(v);
                st.push_back(s.nodes[v].child[1]);
                st.push_back(s.nodes[v].child[0]);
            } else {
                leaves.push_back(v);
            }
        }
    }
    const int64_t ni = (int64_t)interiors.size(), nl = (int64_t)leaves.size();
    if (ni + nl > node_cap) return -1;
    int64_t total_refs = 0;
    for (int v : leaves) total_refs += size[v];
    if (total_refs > ref_cap) return -1;
    std::vector<int> slot(nn, -1);
    for (int64_t i = 0; i < ni; ++i) slot[interiors[i]] = (int)i;
    for (int64_t i = 0; i < nl; ++i) slot[leaves[i]] = (int)(ni + i);
    /* Each leaf's references: its subtree's leaves in depth-first order. */
    std::vector<int> parent_slot(nn, -1);
    for (int v : interiors) {
        for (int k = 0; k < 2; ++k) parent_slot[s.nodes[v].child[k]] = slot[v];
    }
    int64_t ref_pos = 0;
    for (int64_t i = 0; i < nl; ++i) {
        const int v = leaves[i];
        int32_t* w = nodes_out + 16 * (ni + i);
        std::memset(w, 0, 64);
        put_box(w, 0, s.nodes[v].box);
        put_box(w, 1, s.nodes[v].box);
        const int64_t b = ref_pos;
        std::vector<int> st{v};
        while (!st.empty()) {
            const int u = st.back();
            st.pop_back();
            if (s.nodes[u].leaf) {
                for (int k = s.nodes[u].begin; k < s.nodes[u].end; ++k) tri_out[ref_pos++] = s.tri[k];
            } else {
                st.push_back(s.nodes[u].child[1]);
                st.push_back(s.nodes[u].child[0]);
            }
        }
        w[12] = (int32_t)b;
        w[13] = (int32_t)ref_pos;
        w[14] = ~(int32_t)(ref_pos - b);
        w[15] = parent_slot[v];
    }
    for (int64_t i = 0; i < ni; ++i) {
        const int v = interiors[i];
        int32_t* w = nodes_out + 16 * i;
        std::memset(w, 0, 64);
        for (int k = 0; k < 2; ++k) {
            const int c = s.nodes[v].child[k];
            put_box(w, k, s.nodes[c].box);
            w[12 + k] = state[c] == 1 ? slot[c] : ~slot[c];
        }
        w[14] = size[v];
        w[15] = parent_slot[v];
    }
    hxo_woopify(verts, tris, tri_out, ref_pos, woop);
    *n_interior = ni;
    *n_leaves = nl;
    *n_refs = ref_pos;
    return ni + nl;
}

}  // extern "C"
