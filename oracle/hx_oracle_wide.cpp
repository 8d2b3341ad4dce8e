/*
 * hx_oracle_wide.cpp -- CPU restatement of hippie's wide (Quad / Oct) BVH
 * layouts (TEST INFRASTRUCTURE; part of liboracle.so, loaded only by tests/).
 *
 * Reference (src/hippie/): rt/bvh/BVH.h:74-120 (recomputeCostsWide), :228-291
 * (collapseWide), :401-423 (collapseAdaptiveWide / collapseWide);
 * rt/bvh/BVHConverter.cpp:12-78 (breadth-first numbering and node records);
 * rt/bvh/HipBVHNode.h:13-114 (HipBVHNodeN word layout); HipBVHKernels.cu
 * :184-238 (wide SAH cost).  Built with -ffp-contract=off.
 *
 * The root is always an interior node: where the reference's cost would make
 * the root a leaf it indexes count[-2] (undefined); here the root keeps its
 * best split.  Words the reference leaves uninitialised (the unused child-index
 * and link slots) are 0.
 *
 * How it is pinned (tests/test_oracle_wide.py): the cost recursion against an
 * exhaustive search over every way of cutting the binary tree into wide nodes
 * on small trees; the layout by structural invariants (every leaf range once,
 * boxes are the binary boxes, parents, breadth-first order); traversal of the
 * wide tree against a BVH-free brute force.
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <vector>

namespace {

constexpr float kMaxFloat = 3.402823466e+38f;

struct Cnt {
    int8_t x, y;   // slots of the left and right child; (-1, -1): a leaf
};

inline float box_area(const float* mn, const float* mx) {
    if (!(mn[0] <= mx[0] && mn[1] <= mx[1] && mn[2] <= mx[2])) return 0.0f;
    const float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

inline int div_ceil(int a, int b) { return (a + b - 1) / b; }

int32_t fbits(float f) {
    int32_t i;
    std::memcpy(&i, &f, 4);
    return i;
}

}  // namespace

extern "C" {

/* Wide layout of a binary tree: internal nodes 0..n-2 (children left/right),
 * leaves n-1+p (sorted reference p), boxes nmin/nmax as float4 (2n-1), root id.
 * max_leaf_size <= 0: adaptive (BVH::collapseAdaptiveWide), else
 * BVH::collapseWide(N, max_leaf_size).  Writes the nodes (8N int32 each:
 * interiors breadth-first, then leaves breadth-first) and returns the node count;
 * *n_interior receives the interior count. */
int64_t hxo_collapse_wide(int32_t N, int64_t n, int32_t root, const int32_t* left, const int32_t* right,
                          const float* nmin, const float* nmax, float ct, float ci, int32_t max_leaf_size,
                          int32_t* nodes_out, int64_t* n_interior) {
    const int64_t total = 2 * n - 1;
    const int S = N - 1;
    std::vector<float> cost((size_t)total * S);
    std::vector<Cnt> cnt((size_t)total * S);
    std::vector<Cnt> best0((size_t)total);
    std::vector<int> begin((size_t)total), end((size_t)total);
    const int64_t mls = max_leaf_size <= 0 ? INT32_MAX : max_leaf_size;
    const float* rmn = nmin + 4 * (int64_t)root;
    const float* rmx = nmax + 4 * (int64_t)root;
    const float root_area = box_area(rmn, rmx);
    auto is_leaf = [&](int64_t b) { return b >= n - 1; };
    // Post-order over the binary tree (recomputeCostsWide, BVH.h:74-120).
    std::vector<std::pair<int64_t, int>> st;
    st.push_back({root, 0});
    while (!st.empty()) {
        auto& top = st.back();
        const int64_t b = top.first;
        if (!is_leaf(b) && top.second == 0) {
            top.second = 1;
            st.push_back({right[b], 0});
            st.push_back({left[b], 0});
            continue;
        }
        st.pop_back();
        const float A = box_area(nmin + 4 * b, nmax + 4 * b);
        if (is_leaf(b)) {
            begin[b] = (int)(b - (n - 1));
            end[b] = begin[b] + 1;
            const float c = ci * A / root_area * (float)div_ceil(1, N) * (float)N;
            for (int j = 1; j < N; ++j) {
                cost[b * S + j - 1] = c;
                cnt[b * S + j - 1] = Cnt{-1, -1};
            }
            best0[b] = Cnt{-1, -1};
            continue;
        }
        const int64_t l = left[b], r = right[b];
        begin[b] = begin[l] < begin[r] ? begin[l] : begin[r];
        end[b] = end[l] > end[r] ? end[l] : end[r];
        const int size = end[b] - begin[b];
        const float leaf_cost = size > mls ? kMaxFloat : ci * A / root_area * (float)div_ceil(size, N) * (float)N;
        float subtree = ct * A / root_area;
        float cd = kMaxFloat;
        Cnt bc{-1, -1};
        for (int k = 1; k < N; ++k) {
            const float c = cost[l * S + k - 1] + cost[r * S + N - k - 1];
            if (cd > c) {
                cd = c;
                bc = Cnt{(int8_t)k, (int8_t)(N - k)};
            }
        }
        subtree += cd;
        best0[b] = bc;
        if (leaf_cost > subtree) {
            cost[b * S] = subtree;
            cnt[b * S] = bc;
        } else {
            cost[b * S] = leaf_cost;
            cnt[b * S] = Cnt{-1, -1};
        }
        for (int j = 2; j < N; ++j) {
            cd = kMaxFloat;
            for (int k = 1; k < j; ++k) {
                const float c = cost[l * S + k - 1] + cost[r * S + j - k - 1];
                if (cd > c) {
                    cd = c;
                    bc = Cnt{(int8_t)k, (int8_t)(j - k)};
                }
            }
            if (cost[b * S + j - 2] <= cd) {
                cost[b * S + j - 1] = cost[b * S + j - 2];
                cnt[b * S + j - 1] = cnt[b * S + j - 2];
            } else {
                cost[b * S + j - 1] = cd;
                cnt[b * S + j - 1] = bc;
            }
        }
    }

    // Breadth-first conversion (BVHConverter.cpp:12-78) with each interior's
    // expansion (collapseWide, BVH.h:228-291) done when it is dequeued.
    struct Entry {
        int64_t b;     // binary node
        int budget;    // slots it was given
        bool leaf;
        int64_t index; // wide node index
        int64_t parent;
    };
    int64_t next_int = 0, next_leaf = 0;
    std::vector<Entry> interiors, leaves;
    std::deque<Entry> q;
    const bool root_leaf = is_leaf(root);
    q.push_back(Entry{root, 1, false, next_int++, -1});
    std::vector<std::vector<Entry>> kids;
    std::vector<Entry> order;   // dequeue order
    while (!q.empty()) {
        Entry e = q.front();
        q.pop_front();
        order.push_back(e);
        kids.emplace_back();
        if (e.leaf) continue;
        struct Ch {
            int64_t b;
            int budget;
        };
        std::vector<Ch> ch;
        if (root_leaf && e.b == root) {
            ch.push_back(Ch{root, 1});
        } else {
            Cnt c0 = cnt[e.b * S + e.budget - 1];
            if (c0.x < 0) c0 = best0[e.b];   // the root stays interior
            ch.push_back(Ch{left[e.b], c0.x});
            ch.push_back(Ch{right[e.b], c0.y});
            while ((int)ch.size() < N) {
                int s = -1;
                for (int i = 0; i < (int)ch.size(); ++i) {
                    const Cnt cc = cnt[ch[i].b * S + ch[i].budget - 1];
                    if (cc.x > 0 && cc.y > 0 && cc.x + cc.y < N) {
                        s = i;
                        break;
                    }
                }
                if (s < 0) break;
                const Cnt cc = cnt[ch[s].b * S + ch[s].budget - 1];
                const int64_t sb = ch[s].b;
                ch[s] = Ch{left[sb], cc.x};
                ch.insert(ch.begin() + s + 1, Ch{right[sb], cc.y});
            }
        }
        for (const Ch& c : ch) {
            const Cnt cc = cnt[c.b * S + c.budget - 1];
            const bool leaf = cc.x == -1 && cc.y == -1;
            Entry k{c.b, c.budget, leaf, leaf ? next_leaf++ : next_int++, e.index};
            kids.back().push_back(k);
            q.push_back(k);
        }
    }
    const int64_t ni = next_int, nl = next_leaf;
    *n_interior = ni;
    const int W = 8 * N;
    std::memset(nodes_out, 0, sizeof(int32_t) * (size_t)(ni + nl) * W);
    auto slot = [&](const Entry& e) { return e.leaf ? ni + e.index : e.index; };
    auto enc = [&](const Entry& e) { return e.leaf ? ~(int32_t)(ni + e.index) : (int32_t)e.index; };
    auto set_box = [&](int32_t* w, int i, int64_t b) {
        const float* mn = nmin + 4 * b;
        const float* mx = nmax + 4 * b;
        w[4 * i + 0] = fbits(mn[0]);
        w[4 * i + 2] = fbits(mn[1]);
        w[4 * (N + i) + 0] = fbits(mn[2]);
        w[4 * i + 1] = fbits(mx[0]);
        w[4 * i + 3] = fbits(mx[1]);
        w[4 * (N + i) + 1] = fbits(mx[2]);
    };
    for (size_t t = 0; t < order.size(); ++t) {
        const Entry& e = order[t];
        int32_t* w = nodes_out + (size_t)slot(e) * W;
        int size = end[e.b] - begin[e.b];
        int ccnt = 0;
        if (e.leaf) {
            for (int i = 0; i < N; ++i) set_box(w, i, e.b);
            w[4 * N + 2] = begin[e.b];
            w[4 * N + 3] = end[e.b];
            size = ~size;
        } else {
            const auto& ks = kids[t];
            ccnt = (int)ks.size();
            for (int i = 0; i < ccnt; ++i) {
                set_box(w, i, ks[i].b);
                w[4 * (N + i) + 2] = enc(ks[i]);
            }
            const float* mn0 = nmin + 4 * ks[0].b;
            for (int i = ccnt; i < N; ++i) {
                w[4 * i + 0] = fbits(mn0[0]);
                w[4 * i + 1] = fbits(mn0[0]);
                w[4 * i + 2] = fbits(mn0[1]);
                w[4 * i + 3] = fbits(mn0[1]);
                w[4 * (N + i) + 0] = fbits(mn0[2]);
                w[4 * (N + i) + 1] = fbits(mn0[2]);
                w[4 * (N + i) + 2] = enc(ks[0]);
            }
        }
        w[8 * N - 9] = ccnt;
        w[8 * N - 5] = size;
        w[8 * N - 1] = e.parent < 0 ? -1 : (int32_t)e.parent;
    }
    return ni + nl;
}

/* SAH cost of a wide layout (HipBVHKernels.cu:184-238), summed in FP64. */
double hxo_cost_wide(int32_t N, const int32_t* nodes, int64_t n_nodes, float scene_area, float ct, float ci) {
    double sum = 0.0;
    const int W = 8 * N;
    for (int64_t i = 0; i < n_nodes; ++i) {
        const int32_t* w = nodes + i * W;
        float mn[3] = {kMaxFloat, kMaxFloat, kMaxFloat}, mx[3] = {-kMaxFloat, -kMaxFloat, -kMaxFloat};
        for (int c = 0; c < N; ++c) {
            float b[6];
            std::memcpy(&b[0], &w[4 * c + 0], 4);
            std::memcpy(&b[1], &w[4 * c + 2], 4);
            std::memcpy(&b[2], &w[4 * (N + c) + 0], 4);
            std::memcpy(&b[3], &w[4 * c + 1], 4);
            std::memcpy(&b[4], &w[4 * c + 3], 4);
            std::memcpy(&b[5], &w[4 * (N + c) + 1], 4);
            for (int k = 0; k < 3; ++k) {
                mn[k] = b[k] < mn[k] ? b[k] : mn[k];
                mx[k] = b[3 + k] > mx[k] ? b[3 + k] : mx[k];
            }
        }
        const float P = box_area(mn, mx) / scene_area;
        const int sz = w[8 * N - 5];
        if (sz < 0)
            sum += (double)(ci * P * (float)(div_ceil(~sz, N) * N));
        else
            sum += (double)(ct * P);
    }
    return sum;
}

}  // extern "C"
