// hx_api.cpp -- the extern "C" boundary of libhx (declared in include/hx.h).
//
// Host orchestration of the build and trace stages.  Everything is enqueued
// on the context's stream; nothing synchronises except the calls that return
// host data (node counts, cost, profile timings, overflow count).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>

#include "hx_internal.h"

namespace hx {

namespace {
thread_local char g_error[512] = "";
}

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    return e == hipErrorOutOfMemory ? HX_ERR_OUT_OF_MEMORY : HX_ERR_HIP;
}

int check_launch(const char* name) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("kernel %s: launch failed: %s", name, hipGetErrorString(e));
        return HX_ERR_HIP;
    }
    return HX_OK;
}

int DevBuf::reserve(size_t want) {
    if (want <= bytes && ptr) return HX_OK;
    release();
    size_t alloc = std::max<size_t>(want, 256);
    hipError_t e = hipMalloc(&ptr, alloc);
    if (e != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        (void)hipGetLastError();
        set_error("hipMalloc(%zu bytes) failed: %s", alloc, hipGetErrorString(e));
        return HX_ERR_OUT_OF_MEMORY;
    }
    bytes = alloc;
    return HX_OK;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

int Profiler::begin(hipStream_t s, const char* name, hipEvent_t* a, hipEvent_t* b) {
    int idx = -1;
    for (size_t i = 0; i < stats.size(); ++i)
        if (stats[i].name == name) { idx = (int)i; break; }
    if (idx < 0) {
        stats.push_back(KernelStat{name, 0, 0.0});
        idx = (int)stats.size() - 1;
    }
    for (hipEvent_t* e : {a, b}) {
        if (!pool.empty()) {
            *e = pool.back();
            pool.pop_back();
        } else {
            HX_HIP(hipEventCreate(e));
        }
    }
    HX_HIP(hipEventRecord(*a, s));
    pending.push_back(Pending{idx, *a, *b});
    return HX_OK;
}

void Profiler::end(hipStream_t s, hipEvent_t b) { (void)hipEventRecord(b, s); }

int Profiler::flush() {
    for (const Pending& p : pending) {
        HX_HIP(hipEventSynchronize(p.b));
        float ms = 0.0f;
        HX_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        stats[p.stat].count += 1;
        stats[p.stat].total_ms += ms;
        pool.push_back(p.a);
        pool.push_back(p.b);
    }
    pending.clear();
    return HX_OK;
}

void Profiler::reset() {
    (void)flush();
    stats.clear();
}

Profiler::~Profiler() {
    for (const Pending& p : pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

}  // namespace hx

using hx::Context;
using hx::DevBuf;

struct hx_context {
    Context c;
};

struct hx_batch;

struct hx_bvh {
    hx_context* ctx = nullptr;
    hx_build_params params{};
    int64_t n_refs = 0;   // references (leaf slots): the triangles, or the presplitter's pieces
    int64_t n_tri = 0;    // triangles
    const float* verts = nullptr;
    const int32_t* tris = nullptr;
    int64_t n_verts = 0;
    bool keys64 = true;
    int keys_index = 0;
    // Output (owned).
    DevBuf nodes, woop, tri_idx;
    // Build workspace.
    DevBuf keys[2], vals_tmp, left, right, parent, first, last, counters, nmin, nmax, cost, state, cover, leaf_num,
        int_slot, counts, scene_box, cost_out, frontier;
    // PLOC: split position per internal node (kept for refits) and the
    // clustering workspace (cluster ids, boxes and neighbours, ping-pong).
    DevBuf split, ploc_work, atr_work, hlbvh_work, ins_work;
    DevBuf node_at;  // general trees: internal node at each split position
    DevBuf link;     // LBVH builds: arrival words of the climb's top level
    DevBuf proto;    // LBVH builds: Bin rows 0-2 of each internal node from the climb
    DevBuf leaf_stats;  // hx_bvh_leaf_stats: size counts and their sum
    int ploc_rounds = 0;
    int64_t atr_accepted = 0;
    int hlbvh_clusters = 0;
    int64_t ins_steps = 0, ins_reinserted = 0;
    // Presplitting: priorities and scratch, the references (boxes, triangle
    // ids) and the split queues.  refs_active: during a build the leaf slots
    // hold reference ids (mapped to triangle ids at the end).
    DevBuf ps_prio, ps_work, ref_min, ref_max, ref_tri, ps_queue;
    bool refs_active = false;
    float ps_D = 0.0f;
    bool general = false;  // not a Karras tree: the climb uses split[]
    // Per-mesh tables of the emitted layout (one entry unless batched).
    DevBuf mesh_tab, root_slot;
    // Batched builds: meshes occupy triangle ranges [off[m], off[m+1]).
    bool batched = false;
    int n_meshes = 1;
    std::vector<int32_t> h_off;
    std::vector<int4> h_tab;
    bool tab_known = false;
    DevBuf mesh_off, mesh_of, mesh_box, mesh_box_bits, keys_orig, tri_local;
    // With presplitting a mesh's references and triangles have different
    // offsets: mesh_off / mesh_of / h_roff are over reference slots, tri_off /
    // tri_mesh_of over triangles.
    DevBuf tri_off, tri_mesh_of;
    // Packed vertices per triangle (3 float4), written by the Morton pass of
    // a build and gathered by its climbs; stale after a refit.
    DevBuf tri_pack;
    bool pack_valid = false;
    std::vector<int32_t> h_roff;
    // Bin layout, Karras trees: the Bin slot of every binary node, recorded by
    // the first refit after a build (a build leaves it stale); later refits
    // write the Bin rows from the climb.
    DevBuf bin_of;
    bool bin_of_valid = false;
    bool built = false;
    bool counts_known = false;
    int64_t n_interior = 0, n_leaves = 0;
    // Quad / Oct layouts: children per node (0: Bin), the breadth-first entries
    // of the wide nodes (kept for refits) and the collapse workspace.
    int wide = 0;
    bool wide_lbvh_root = false;
    int64_t wide_entries = 0;
    DevBuf wide_cost, wide_cnt, e_b, e_budget, e_leaf, e_parent, e_rank, e_first, e_nchild, wide_scan, wide_partial;
    float host_box[6] = {0, 0, 0, 0, 0, 0};
    bool box_known = false;
};

struct hx_batch {
    hx_bvh bvh;
};

namespace {

#define HX_CHECK_ARG(cond, ...)                   \
    do {                                          \
        if (!(cond)) {                            \
            hx::set_error(__VA_ARGS__);           \
            return HX_ERR_INVALID_ARG;            \
        }                                         \
    } while (0)

int validate_params(const hx_build_params* p) {
    HX_CHECK_ARG(p != nullptr, "params is NULL");
    HX_CHECK_ARG(p->morton_bits >= 6 && p->morton_bits <= 60, "morton_bits must be in [6,60] (got %d)",
                 p->morton_bits);
    HX_CHECK_ARG(p->max_leaf_size > 0 && p->max_leaf_size <= 1024, "max_leaf_size must be in (0,1024] (got %d)",
                 p->max_leaf_size);
    HX_CHECK_ARG(p->ct > 0.0f && p->ci > 0.0f, "ct and ci must be positive");
    HX_CHECK_ARG(p->layout == HX_LAYOUT_BIN || p->layout == HX_LAYOUT_QUAD || p->layout == HX_LAYOUT_OCT,
                 "layout must be HX_LAYOUT_BIN, HX_LAYOUT_QUAD or HX_LAYOUT_OCT (got %d)", p->layout);
    if (p->method != HX_METHOD_LBVH && p->method != HX_METHOD_PLOC && p->method != HX_METHOD_ATRBVH &&
        p->method != HX_METHOD_HLBVH && p->method != HX_METHOD_INSERTION && p->method != HX_METHOD_TR) {
        hx::set_error("builder method %d is not available in this build", p->method);
        return HX_ERR_UNSUPPORTED;
    }
    if (p->method == HX_METHOD_PLOC) {
        // PLOCBuilder::setMaxRadius / setRadius (PLOCBuilder.cpp:261-273).
        HX_CHECK_ARG(p->ploc_max_radius >= 1, "ploc_max_radius must be positive (got %d)", p->ploc_max_radius);
        HX_CHECK_ARG(p->ploc_radius >= 1 && p->ploc_radius <= p->ploc_max_radius,
                     "ploc_radius must be in [1, %d] (got %d)", p->ploc_max_radius, p->ploc_radius);
    }
    if (p->method == HX_METHOD_HLBVH) {
        // HLBVHBuilder::setMortonCodeSAHBits (HLBVHBuilder.cpp:422-426).
        HX_CHECK_ARG(p->hlbvh_sah_bits >= 6 && p->hlbvh_sah_bits <= 15, "hlbvh_sah_bits must be in [6,15] (got %d)",
                     p->hlbvh_sah_bits);
        HX_CHECK_ARG(p->hlbvh_sah_bits <= p->morton_bits, "hlbvh_sah_bits must not exceed morton_bits");
    }
    if (p->method == HX_METHOD_INSERTION) {
        // InsertionBuilder::setMod (InsertionBuilder.cpp:491-494).
        HX_CHECK_ARG(p->insertion_mod >= 1 && p->insertion_mod <= 64, "insertion_mod must be in [1,64] (got %d)",
                     p->insertion_mod);
    }
    if (p->method == HX_METHOD_ATRBVH) {
        // ATRBuilder::setTreeletSize / setIterations (ATRBuilder.cpp:273-290).
        HX_CHECK_ARG(p->atr_treelet_size >= 4 && p->atr_treelet_size <= 30,
                     "atr_treelet_size must be in [4,30] (got %d)", p->atr_treelet_size);
        HX_CHECK_ARG(p->atr_iterations >= 0 && p->atr_iterations <= 25, "atr_iterations must be in [0,25] (got %d)",
                     p->atr_iterations);
    }
    if (p->method == HX_METHOD_TR) {
        // TRBuilder::setIterations (TRBuilder.cpp:329-332).
        HX_CHECK_ARG(p->tr_iterations >= 0 && p->tr_iterations <= 25, "tr_iterations must be in [0,25] (got %d)",
                     p->tr_iterations);
    }
    if (p->presplitting) {
        // Presplitter::setBeta (Presplitter.cpp:328-331).
        HX_CHECK_ARG(p->presplitter_beta >= 0.0f, "presplitter_beta must be non-negative (got %g)",
                     (double)p->presplitter_beta);
    }
    return HX_OK;
}

// PLOC workspace: cl[2] (int), ca[2] (float4), cb[2] (float2), nb (int).
// PLOC workspace: cl[2] (int), ca[2] (float4), cb[2] (float2), nb (int), and
// in batched builds cm[2] (int).
size_t ploc_work_bytes(int64_t n) { return (size_t)n * (2 * 4 + 2 * 16 + 2 * 8 + 4 + 2 * 4) + 2048; }

// ATR workspace: area, cost (2n-1 floats), refs, height, list (n ints),
// hist, cursor (levels + 1 ints), accepted.
size_t atr_work_bytes(int64_t n) {
    return (size_t)(2 * n) * 8 + (size_t)n * 12 + 2 * sizeof(int32_t) * (2 * hx::kAtrMaxLevels + 1) + 4096;
}

// HLBVH workspace: two n-int arrays, cluster arrays of up to 2^15 entries.
size_t hlbvh_work_bytes(int64_t n) {
    const size_t K = (size_t)hx::kHlbvhMaxClusters;
    return (size_t)n * 8 + K * (4 * 8 + 16 * 4) + 8192;
}

// Insertion workspace: out, red, locks, the best tree and the packed walk
// records (2n-1 entries each), refs, height and the height list (n each),
// bucket counts, counters.
size_t ins_work_bytes(int64_t n) {
    const size_t N = (size_t)(2 * n);
    return N * (4 + 4 + 8 + 12 + 64) + (size_t)n * 12 + 2 * sizeof(int32_t) * (2 * hx::kAtrMaxLevels + 1) + 8192;
}

int reserve_bvh(hx_bvh* b, int64_t n) {
    const int64_t N = std::max<int64_t>(2 * n - 1, 2);
    const size_t ks = b->keys64 ? 8 : 4;
    const int wide = b->params.layout == HX_LAYOUT_BIN ? 0 : b->params.layout;
    HX_TRY(b->nodes.reserve((wide ? 32 * (size_t)wide : 64) * (size_t)std::max<int64_t>(2 * n, 2)));
    HX_TRY(b->woop.reserve(48 * (size_t)n));
    HX_TRY(b->tri_idx.reserve(4 * (size_t)n));
    if (!wide) HX_TRY(b->bin_of.reserve(sizeof(int32_t) * (size_t)std::max<int64_t>(2 * n, 2)));
    HX_TRY(b->keys[0].reserve(ks * (size_t)n));
    HX_TRY(b->keys[1].reserve(ks * (size_t)n));
    HX_TRY(b->vals_tmp.reserve(4 * (size_t)n));
    HX_TRY(b->left.reserve(4 * (size_t)N));
    HX_TRY(b->right.reserve(4 * (size_t)N));
    HX_TRY(b->parent.reserve(4 * (size_t)N));
    HX_TRY(b->first.reserve(4 * (size_t)N));
    HX_TRY(b->last.reserve(4 * (size_t)N));
    HX_TRY(b->counters.reserve(4 * (size_t)n));
    HX_TRY(b->nmin.reserve(16 * (size_t)N));
    HX_TRY(b->nmax.reserve(16 * (size_t)N));
    HX_TRY(b->cost.reserve(4 * (size_t)N));
    HX_TRY(b->state.reserve(4 * (size_t)n));
    HX_TRY(b->cover.reserve(4 * (size_t)n));
    HX_TRY(b->leaf_num.reserve(4 * (size_t)n));
    HX_TRY(b->int_slot.reserve(4 * (size_t)n));
    HX_TRY(b->frontier.reserve(4 * (size_t)n));
    HX_TRY(b->counts.reserve(16));
    HX_TRY(b->scene_box.reserve(32));
    HX_TRY(b->cost_out.reserve(16));
    HX_TRY(b->mesh_tab.reserve(16 * (size_t)b->n_meshes));
    HX_TRY(b->root_slot.reserve(4 * (size_t)b->n_meshes));
    if (b->batched) {
        const size_t M = (size_t)b->n_meshes;
        HX_TRY(b->mesh_off.reserve(4 * (M + 1)));
        HX_TRY(b->mesh_of.reserve(4 * (size_t)n));
        HX_TRY(b->mesh_box.reserve(24 * M));
        HX_TRY(b->mesh_box_bits.reserve(24 * M));
        HX_TRY(b->keys_orig.reserve(8 * (size_t)n));
        HX_TRY(b->tri_local.reserve(4 * (size_t)n));
    }
    if (b->params.method == HX_METHOD_LBVH) {
        HX_TRY(b->link.reserve(8 * (size_t)n));
        HX_TRY(b->proto.reserve(48 * (size_t)n));
    }
    if (b->params.method == HX_METHOD_PLOC) {
        HX_TRY(b->split.reserve(4 * (size_t)n));
        HX_TRY(b->node_at.reserve(4 * (size_t)n));
        HX_TRY(b->ploc_work.reserve(ploc_work_bytes(n)));
    }
    if (b->params.method == HX_METHOD_ATRBVH || b->params.method == HX_METHOD_TR) {
        HX_TRY(b->split.reserve(4 * (size_t)n));
        HX_TRY(b->node_at.reserve(4 * (size_t)n));
        HX_TRY(b->atr_work.reserve(atr_work_bytes(n)));
    }
    if (b->params.method == HX_METHOD_HLBVH) {
        HX_TRY(b->split.reserve(4 * (size_t)n));
        HX_TRY(b->node_at.reserve(4 * (size_t)n));
        HX_TRY(b->hlbvh_work.reserve(hlbvh_work_bytes(n)));
    }
    if (b->params.method == HX_METHOD_INSERTION) {
        HX_TRY(b->split.reserve(4 * (size_t)n));
        HX_TRY(b->node_at.reserve(4 * (size_t)n));
        HX_TRY(b->ins_work.reserve(ins_work_bytes(n)));
    }
    return HX_OK;
}

hx::ClimbArgs climb_args(hx_bvh* b, bool refit_only) {
    const hx_build_params& p = b->params;
    hx::ClimbArgs ca{};
    ca.n = b->n_refs;
    ca.sorted_tri = b->tri_idx.as<int32_t>();
    ca.verts = b->verts;
    ca.tris = b->tris;
    ca.left = b->left.as<int32_t>();
    ca.right = b->right.as<int32_t>();
    ca.parent = b->parent.as<int32_t>();
    ca.first = b->first.as<int32_t>();
    ca.last = b->last.as<int32_t>();
    ca.split = b->general ? b->split.as<int32_t>() : nullptr;
    ca.counters = b->counters.as<int32_t>();
    ca.nmin = b->nmin.as<float4>();
    ca.nmax = b->nmax.as<float4>();
    ca.cost = b->cost.as<float>();
    ca.state = b->state.as<int32_t>();
    ca.end_at = b->cover.as<int32_t>();
    ca.woop = b->woop.as<float4>();
    ca.frontier = b->frontier.as<int32_t>();
    ca.frontier_count = b->counts.as<int32_t>() + 2;
    ca.mesh_of = b->batched ? b->mesh_of.as<int32_t>() : nullptr;
    ca.mesh_off = b->batched ? b->mesh_off.as<int32_t>() : nullptr;
    if (b->refs_active) {
        ca.ref_min = b->ref_min.as<float4>();
        ca.ref_max = b->ref_max.as<float4>();
        ca.ref_tri = b->ref_tri.as<int32_t>();
    }
    ca.tri_pack = b->pack_valid ? b->tri_pack.as<float4>() : nullptr;
    ca.refit_only = refit_only ? 1 : 0;
    ca.adaptive = (p.adaptive_leaf_size && !refit_only) ? 1 : 0;
    ca.max_leaf_size = p.max_leaf_size;
    ca.ct = p.ct;
    ca.ci = p.ci;
    return ca;
}

// The LBVH build's climb, which also builds the hierarchy from the keys.
hx::ClimbArgs lbvh_climb_args(hx_bvh* b, const void* keys, bool keys64) {
    hx::ClimbArgs ca = climb_args(b, false);
    ca.keys = keys;
    ca.keys64 = keys64 ? 1 : 0;
    ca.link = b->link.as<uint64_t>();
    ca.root_id = b->counts.as<int32_t>() + 3;
    ca.all_boxes = b->params.keep_debug || b->params.layout != HX_LAYOUT_BIN ? 1 : 0;
    ca.proto = b->proto.as<float4>();
    return ca;
}

hx::EmitArgs emit_args(hx_bvh* b) {
    hx::EmitArgs ea{};
    ea.n = b->n_refs;
    ea.left = b->left.as<int32_t>();
    ea.right = b->right.as<int32_t>();
    ea.parent = b->parent.as<int32_t>();
    ea.first = b->first.as<int32_t>();
    ea.last = b->last.as<int32_t>();
    ea.state = b->state.as<int32_t>();
    ea.nmin = b->nmin.as<float4>();
    ea.nmax = b->nmax.as<float4>();
    ea.cover = b->cover.as<int32_t>();
    ea.leaf_num = b->leaf_num.as<int32_t>();
    ea.int_slot = b->int_slot.as<int32_t>();
    ea.counts = b->counts.as<int32_t>();
    ea.nodes = b->nodes.as<int4>();
    ea.mesh_of = b->batched ? b->mesh_of.as<int32_t>() : nullptr;
    ea.mesh_off = b->batched ? b->mesh_off.as<int32_t>() : nullptr;
    ea.n_meshes = b->n_meshes;
    ea.mesh_tab = b->mesh_tab.as<int4>();
    ea.root_slot = b->root_slot.as<int32_t>();
    if (b->general) {
        ea.split = b->split.as<int32_t>();
        ea.node_at = b->node_at.as<int32_t>();
    }
    return ea;
}

// Presplitter::presplit (Presplitter.cpp:304-324) in place of setupBoxes:
// priorities and D, then the references; the BVH's buffers are sized for
// the reference count.  In a batch every mesh is presplit alone (its own box
// and budget) and h_roff receives the per-mesh reference offsets.
int run_presplit(hx_bvh* b) {
    Context* c = &b->ctx->c;
    const int64_t n = b->n_tri;
    const int M = b->batched ? b->n_meshes : 1;
    const int64_t ntiles = (n + 4095) / 4096;
    HX_TRY(b->ps_prio.reserve(4 * (size_t)n));
    HX_TRY(b->ps_work.reserve(16 * (size_t)M + 4 * (size_t)M + 4 * (size_t)ntiles + 2048));
    hx::PsArgs a{};
    a.n = n;
    a.verts = b->verts;
    a.tris = b->tris;
    a.box = b->scene_box.as<float>();
    if (b->batched) {
        a.mesh_of = b->tri_mesh_of.as<int32_t>();
        a.mesh_box = b->mesh_box.as<float>();
    }
    a.prio = b->ps_prio.as<float>();
    char* w = b->ps_work.as<char>();
    auto carve_w = [&](size_t bytes) {
        char* r = w;
        w += (bytes + 255) & ~(size_t)255;
        return r;
    };
    a.macc = reinterpret_cast<unsigned long long*>(carve_w(16 * (size_t)M));
    a.Dm = reinterpret_cast<float*>(carve_w(4 * (size_t)M));
    a.qcount = reinterpret_cast<int32_t*>(carve_w(8));
    a.partial = reinterpret_cast<int32_t*>(carve_w(4 * (size_t)ntiles));
    std::vector<int64_t> tri_counts(M), ref_counts(M);
    std::vector<float> D(M);
    for (int m = 0; m < M; ++m) tri_counts[m] = b->batched ? (int64_t)b->h_off[m + 1] - b->h_off[m] : n;
    HX_TRY(hx::launch_presplit_count(c, a, b->params.presplitter_beta, M, tri_counts.data(), ref_counts.data(),
                                     D.data()));
    int64_t n_refs = 0;
    for (int m = 0; m < M; ++m) n_refs += ref_counts[m];
    if (n_refs >= ((int64_t)1 << 30)) {
        hx::set_error("presplitting gives %lld references (limit 2^30)", (long long)n_refs);
        return HX_ERR_UNSUPPORTED;
    }
    b->n_refs = n_refs;
    b->ps_D = D[0];
    if (b->batched) {
        b->h_roff.assign(M + 1, 0);
        for (int m = 0; m < M; ++m) b->h_roff[m + 1] = b->h_roff[m] + (int32_t)ref_counts[m];
    }
    HX_TRY(reserve_bvh(b, n_refs));
    HX_TRY(b->ref_min.reserve(16 * (size_t)n_refs));
    HX_TRY(b->ref_max.reserve(16 * (size_t)n_refs));
    HX_TRY(b->ref_tri.reserve(4 * (size_t)n_refs));
    a.queue_cap = std::max<int64_t>(1, n_refs - n);
    const size_t qb = (size_t)a.queue_cap * 16;
    HX_TRY(b->ps_queue.reserve(6 * ((qb + 255) & ~(size_t)255) + 256));
    char* q = b->ps_queue.as<char>();
    auto carve = [&](size_t bytes) {
        char* r = q;
        q += (bytes + 255) & ~(size_t)255;
        return r;
    };
    for (int k = 0; k < 2; ++k) {
        a.task[k] = reinterpret_cast<int4*>(carve(qb));
        a.tmn[k] = reinterpret_cast<float4*>(carve(qb));
        a.tmx[k] = reinterpret_cast<float4*>(carve(qb));
    }
    a.ref_min = b->ref_min.as<float4>();
    a.ref_max = b->ref_max.as<float4>();
    a.ref_tri = b->ref_tri.as<int32_t>();
    HX_TRY(hx::launch_presplit_refs(c, a, n_refs));
    b->refs_active = true;
    return HX_OK;
}

// LBVHBuilder::build front half (LBVHBuilder.cpp:187-213), shared by PLOC
// (PLOCBuilder.cpp:161-187): scene box, Morton codes, sort; the sorted
// triangle ids end in tri_idx.
int run_front(hx_bvh* b) {
    Context* c = &b->ctx->c;
    const hx_build_params& p = b->params;
    HX_HIP(hipSetDevice(c->device));
    float* box = b->scene_box.as<float>();
    if (p.use_scene_box) {
        HX_HIP(hipMemcpyAsync(box, p.scene_box, 6 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    } else {
        HX_TRY(hx::launch_scene_box(c, b->verts, b->tris, b->n_tri, box));
    }
    b->box_known = false;
    b->refs_active = false;
    if (p.presplitting) HX_TRY(run_presplit(b));
    const int64_t n = b->n_refs;
    const float4* rmin = b->refs_active ? b->ref_min.as<float4>() : nullptr;
    const float4* rmax = b->refs_active ? b->ref_max.as<float4>() : nullptr;
    float4* pack = nullptr;
    if (!b->refs_active) {
        HX_TRY(b->tri_pack.reserve(48 * (size_t)n));
        pack = b->tri_pack.as<float4>();
    }
    HX_TRY(hx::launch_morton(c, b->verts, b->tris, n, box, p.morton_bits, b->keys[0].ptr, b->keys64, nullptr, rmin,
                             rmax, pack));
    b->pack_valid = pack != nullptr;
    // The sorted triangle ids must end in tri_idx: pass p reads vals[p&1].
    const int passes = hx::sort_pass_count(c, p.morton_bits, n, b->keys64);
    int32_t* vals[2];
    vals[passes & 1] = b->tri_idx.as<int32_t>();
    vals[(passes & 1) ^ 1] = b->vals_tmp.as<int32_t>();
    void* keys[2] = {b->keys[0].ptr, b->keys[1].ptr};
    int ri = 0;
    HX_TRY(hx::launch_sort_pairs(c, keys, vals, n, p.morton_bits, b->keys64, true, &ri));
    if (vals[ri] != b->tri_idx.as<int32_t>()) {
        HX_HIP(hipMemcpyAsync(b->tri_idx.ptr, vals[ri], 4 * (size_t)n, hipMemcpyDeviceToDevice, c->stream));
    }
    b->keys_index = ri;
    return HX_OK;
}

int run_ploc(hx_bvh* b) {
    Context* c = &b->ctx->c;
    const hx_build_params& p = b->params;
    const int64_t n = b->n_refs;
    hx::PlocArgs a{};
    a.n = n;
    a.verts = b->verts;
    a.tris = b->tris;
    a.sorted_tri = b->tri_idx.as<int32_t>();
    a.tri_out = b->vals_tmp.as<int32_t>();
    if (b->refs_active) {
        a.ref_min = b->ref_min.as<float4>();
        a.ref_max = b->ref_max.as<float4>();
    }
    char* w = b->ploc_work.as<char>();
    auto carve = [&](size_t bytes) {
        char* q = w;
        w += (bytes + 255) & ~(size_t)255;
        return q;
    };
    for (int k = 0; k < 2; ++k) a.ca[k] = reinterpret_cast<float4*>(carve(16 * (size_t)n));
    for (int k = 0; k < 2; ++k) a.cb[k] = reinterpret_cast<float2*>(carve(8 * (size_t)n));
    for (int k = 0; k < 2; ++k) a.cl[k] = reinterpret_cast<int32_t*>(carve(4 * (size_t)n));
    a.nb = reinterpret_cast<int32_t*>(carve(4 * (size_t)n));
    if (b->batched) {
        for (int k = 0; k < 2; ++k) a.cm[k] = reinterpret_cast<int32_t*>(carve(4 * (size_t)n));
        a.mesh_of = b->mesh_of.as<int32_t>();
        a.n_meshes = b->n_meshes;
    }
    a.left = b->left.as<int32_t>();
    a.right = b->right.as<int32_t>();
    a.parent = b->parent.as<int32_t>();
    a.first = b->first.as<int32_t>();
    a.last = b->last.as<int32_t>();
    a.split = b->split.as<int32_t>();
    a.counters = b->counters.as<int32_t>();
    a.end_at = b->cover.as<int32_t>();
    a.frontier_count = b->counts.as<int32_t>() + 2;
    a.radius = p.ploc_radius;
    a.max_radius = p.ploc_max_radius;
    a.adaptive = p.ploc_adaptive ? 1 : 0;
    HX_TRY(hx::launch_ploc(c, a, &b->ploc_rounds));
    b->general = true;
    // The triangle ids in depth-first leaf order become the output array.
    std::swap(b->tri_idx.ptr, b->vals_tmp.ptr);
    std::swap(b->tri_idx.bytes, b->vals_tmp.bytes);
    HX_TRY(hx::launch_climb(c, climb_args(b, false)));
    return hx::launch_collapse_emit(c, emit_args(b));
}

hx::PlocArgs relabel_args(hx_bvh* b) {
    hx::PlocArgs a{};
    a.n = b->n_refs;
    a.verts = b->verts;
    a.tris = b->tris;
    a.sorted_tri = b->tri_idx.as<int32_t>();
    a.tri_out = b->vals_tmp.as<int32_t>();
    a.left = b->left.as<int32_t>();
    a.right = b->right.as<int32_t>();
    a.parent = b->parent.as<int32_t>();
    a.first = b->first.as<int32_t>();
    a.last = b->last.as<int32_t>();
    a.split = b->split.as<int32_t>();
    a.counters = b->counters.as<int32_t>();
    a.end_at = b->cover.as<int32_t>();
    a.frontier_count = b->counts.as<int32_t>() + 2;
    return a;
}

// ATRBuilder::build (ATRBuilder.cpp:179-236): Karras hierarchy, refit with
// the ATR inputs, treelet optimisation, then (as for PLOC) depth-first
// relabelling, the general climb and the collapse.
int run_atr(hx_bvh* b) {
    Context* c = &b->ctx->c;
    const bool tr = b->params.method == HX_METHOD_TR;
    const hx_build_params& p = b->params;
    const int64_t n = b->n_refs;
    HX_TRY(hx::launch_construct(c, b->keys[b->keys_index].ptr, b->keys64, n, b->left.as<int32_t>(),
                                b->right.as<int32_t>(), b->parent.as<int32_t>(), b->first.as<int32_t>(),
                                b->last.as<int32_t>(), b->counters.as<int32_t>(), b->cover.as<int32_t>(),
                                b->counts.as<int32_t>() + 2));
    char* w = b->atr_work.as<char>();
    auto carve = [&](size_t bytes) {
        char* q = w;
        w += (bytes + 255) & ~(size_t)255;
        return q;
    };
    hx::AtrArgs a{};
    a.n = n;
    // TRBuilder: treelets of 7 leaves, Bvh.trIterations (TRBuilder.cpp:281-332).
    a.treelet_size = tr ? 7 : p.atr_treelet_size;
    a.iterations = tr ? p.tr_iterations : p.atr_iterations;
    a.ct = p.ct;
    a.ci = p.ci;
    a.left = b->left.as<int32_t>();
    a.right = b->right.as<int32_t>();
    a.parent = b->parent.as<int32_t>();
    a.nmin = b->nmin.as<float4>();
    a.nmax = b->nmax.as<float4>();
    a.area = reinterpret_cast<float*>(carve(4 * (size_t)(2 * n)));
    a.cost = reinterpret_cast<float*>(carve(4 * (size_t)(2 * n)));
    a.refs = reinterpret_cast<int32_t*>(carve(4 * (size_t)n));
    a.height = reinterpret_cast<int32_t*>(carve(4 * (size_t)n));
    a.list = reinterpret_cast<int32_t*>(carve(4 * (size_t)n));
    a.hist = reinterpret_cast<int32_t*>(carve(4 * (size_t)(2 * hx::kAtrMaxLevels + 1)));
    a.cursor = reinterpret_cast<int32_t*>(carve(4 * (size_t)(2 * hx::kAtrMaxLevels + 1)));
    a.accepted = reinterpret_cast<int32_t*>(carve(16));
    hx::ClimbArgs ca = climb_args(b, false);
    ca.split = nullptr;  // still a Karras tree
    ca.atr_area = a.area;
    ca.atr_cost = a.cost;
    ca.atr_refs = a.refs;
    ca.atr_height = a.height;
    HX_TRY(hx::launch_climb(c, ca));
    const hx::PlocArgs rl = relabel_args(b);
    if (tr) HX_TRY(hx::launch_tr(c, a, rl, &b->atr_accepted));
    else HX_TRY(hx::launch_atr(c, a, rl, &b->atr_accepted));
    b->general = true;
    std::swap(b->tri_idx.ptr, b->vals_tmp.ptr);
    std::swap(b->tri_idx.bytes, b->vals_tmp.bytes);
    // Fresh climb state for the general tree.
    HX_TRY(hx::launch_refit_reset(c, b->counters.as<int32_t>(), n - 1, b->counts.as<int32_t>() + 2));
    HX_TRY(hx::launch_fill_i32(c, b->cover.as<int32_t>(), -1, n));
    HX_TRY(hx::launch_climb(c, climb_args(b, false)));
    return hx::launch_collapse_emit(c, emit_args(b));
}

// InsertionBuilder::build with the LBVH start (InsertionBuilder.cpp:379-440):
// Karras hierarchy and its boxes, the insertion loop, then the depth-first
// relabelling (by height), the general climb and the collapse.
int run_insertion(hx_bvh* b) {
    Context* c = &b->ctx->c;
    const hx_build_params& p = b->params;
    const int64_t n = b->n_refs;
    HX_TRY(hx::launch_construct(c, b->keys[b->keys_index].ptr, b->keys64, n, b->left.as<int32_t>(),
                                b->right.as<int32_t>(), b->parent.as<int32_t>(), b->first.as<int32_t>(),
                                b->last.as<int32_t>(), b->counters.as<int32_t>(), b->cover.as<int32_t>(),
                                b->counts.as<int32_t>() + 2));
    hx::ClimbArgs ca = climb_args(b, true);
    ca.split = nullptr;  // still a Karras tree
    HX_TRY(hx::launch_climb(c, ca));
    char* w = b->ins_work.as<char>();
    auto carve = [&](size_t bytes) {
        char* q = w;
        w += (bytes + 255) & ~(size_t)255;
        return q;
    };
    const size_t N = (size_t)(2 * n - 1);
    hx::InsArgs a{};
    a.n = n;
    a.mod = p.insertion_mod;
    a.ct = p.ct;
    a.ci = p.ci;
    a.left = b->left.as<int32_t>();
    a.right = b->right.as<int32_t>();
    a.parent = b->parent.as<int32_t>();
    a.nmin = b->nmin.as<float4>();
    a.nmax = b->nmax.as<float4>();
    a.counters = b->counters.as<int32_t>();
    a.out = reinterpret_cast<int32_t*>(carve(4 * N));
    a.red = reinterpret_cast<float*>(carve(4 * N));
    a.locks = reinterpret_cast<unsigned long long*>(carve(8 * N));
    a.best_left = reinterpret_cast<int32_t*>(carve(4 * N));
    a.best_right = reinterpret_cast<int32_t*>(carve(4 * N));
    a.best_parent = reinterpret_cast<int32_t*>(carve(4 * N));
    a.acc = reinterpret_cast<unsigned long long*>(carve(32));
    a.refs = reinterpret_cast<int32_t*>(carve(4 * (size_t)n));
    a.height = reinterpret_cast<int32_t*>(carve(4 * (size_t)n));
    a.rec = reinterpret_cast<int4*>(carve(64 * N));
    HX_TRY(hx::launch_insertion(c, a, &b->ins_steps, &b->ins_reinserted));
    hx::AtrArgs rel{};
    rel.n = n;
    rel.refs = a.refs;
    rel.height = a.height;
    rel.list = reinterpret_cast<int32_t*>(carve(4 * (size_t)n));
    rel.hist = reinterpret_cast<int32_t*>(carve(4 * (size_t)(2 * hx::kAtrMaxLevels + 1)));
    rel.cursor = reinterpret_cast<int32_t*>(carve(4 * (size_t)(2 * hx::kAtrMaxLevels + 1)));
    HX_TRY(hx::launch_relabel_by_height(c, rel, relabel_args(b)));
    b->general = true;
    std::swap(b->tri_idx.ptr, b->vals_tmp.ptr);
    std::swap(b->tri_idx.bytes, b->vals_tmp.bytes);
    HX_TRY(hx::launch_refit_reset(c, b->counters.as<int32_t>(), n - 1, b->counts.as<int32_t>() + 2));
    HX_TRY(hx::launch_fill_i32(c, b->cover.as<int32_t>(), -1, n));
    HX_TRY(hx::launch_climb(c, climb_args(b, false)));
    return hx::launch_collapse_emit(c, emit_args(b));
}

// Batched LBVH: per-mesh boxes and Morton codes, (mesh, Morton) order from
// two stable sorts, one Karras construction over the composite key, then the
// shared climb and collapse with per-mesh tables.
int run_batch(hx_bvh* b) {
    Context* c = &b->ctx->c;
    const hx_build_params& p = b->params;
    const int M = b->n_meshes;
    const int64_t n_tri = b->n_tri;
    HX_HIP(hipSetDevice(c->device));
    HX_TRY(b->tri_off.reserve(4 * (size_t)(M + 1)));
    int32_t* toff = b->tri_off.as<int32_t>();
    int32_t* tmesh;
    if (p.presplitting) {
        HX_TRY(b->tri_mesh_of.reserve(4 * (size_t)n_tri));
        tmesh = b->tri_mesh_of.as<int32_t>();
    } else {
        // Without presplitting the references are the triangles: one mesh map
        // serves the boxes and the build.
        HX_TRY(b->mesh_of.reserve(4 * (size_t)n_tri));
        tmesh = b->mesh_of.as<int32_t>();
    }
    HX_HIP(hipMemcpyAsync(toff, b->h_off.data(), 4 * (size_t)(M + 1), hipMemcpyHostToDevice, c->stream));
    HX_TRY(hx::launch_mesh_of(c, toff, M, n_tri, tmesh));
    HX_TRY(hx::launch_mesh_boxes(c, b->verts, b->tris, n_tri, tmesh, M, b->mesh_box_bits.as<uint32_t>(),
                                 b->mesh_box.as<float>()));
    b->refs_active = false;
    if (p.presplitting) {
        HX_TRY(run_presplit(b));
    } else {
        b->n_refs = n_tri;
        b->h_roff = b->h_off;
    }
    const int64_t n = b->n_refs;
    int32_t* off = b->mesh_off.as<int32_t>();
    int32_t* mesh_of = b->mesh_of.as<int32_t>();
    HX_HIP(hipMemcpyAsync(off, b->h_roff.data(), 4 * (size_t)(M + 1), hipMemcpyHostToDevice, c->stream));
    if (p.presplitting) HX_TRY(hx::launch_mesh_of(c, off, M, n, mesh_of));
    const float4* rmin = b->refs_active ? b->ref_min.as<float4>() : nullptr;
    const float4* rmax = b->refs_active ? b->ref_max.as<float4>() : nullptr;
    uint64_t* korig = b->keys_orig.as<uint64_t>();
    float4* pack = nullptr;
    if (!b->refs_active) {
        HX_TRY(b->tri_pack.reserve(48 * (size_t)n));
        pack = b->tri_pack.as<float4>();
    }
    HX_TRY(hx::launch_morton(c, b->verts, b->tris, n, b->mesh_box.as<float>(), p.morton_bits, korig, true, mesh_of,
                             rmin, rmax, pack));
    b->pack_valid = pack != nullptr;
    HX_HIP(hipMemcpyAsync(b->keys[0].ptr, korig, 8 * (size_t)n, hipMemcpyDeviceToDevice, c->stream));
    void* keys[2] = {b->keys[0].ptr, b->keys[1].ptr};
    int32_t* vals[2] = {b->tri_idx.as<int32_t>(), b->vals_tmp.as<int32_t>()};
    int ri = 0;
    HX_TRY(hx::launch_sort_pairs(c, keys, vals, n, p.morton_bits, true, true, &ri));
    int32_t* sorted = vals[ri];
    if (M > 1) {
        int mbits = 0;
        while ((1 << mbits) < M) ++mbits;
        HX_TRY(hx::launch_gather_mesh_keys(c, sorted, off, M, n, static_cast<uint32_t*>(keys[0])));
        int32_t* v2[2] = {sorted, vals[ri ^ 1]};
        int r2 = 0;
        HX_TRY(hx::launch_sort_pairs(c, keys, v2, n, mbits, false, false, &r2));
        sorted = v2[r2];
    }
    if (sorted != b->tri_idx.as<int32_t>())
        HX_HIP(hipMemcpyAsync(b->tri_idx.ptr, sorted, 4 * (size_t)n, hipMemcpyDeviceToDevice, c->stream));
    b->general = false;
    if (p.method == HX_METHOD_PLOC) {
        // Clusters merge inside their own mesh, so each mesh gets the tree a
        // PLOC build of it alone gives.
        HX_TRY(run_ploc(b));
    } else {
        HX_TRY(hx::launch_gather_keys(c, b->tri_idx.as<int32_t>(), korig, n, static_cast<uint64_t*>(keys[0])));
        b->keys_index = 0;
        hx::ClimbArgs ca = lbvh_climb_args(b, keys[0], true);
        HX_TRY(hx::launch_lbvh_climb(c, ca));
        hx::EmitArgs ea = emit_args(b);
        ea.proto = ca.proto;
        HX_TRY(hx::launch_collapse_emit(c, ea));
    }
    if (b->refs_active) {
        HX_TRY(hx::launch_refs_to_tris(c, b->tri_idx.as<int32_t>(), b->ref_tri.as<int32_t>(), n));
        b->refs_active = false;
    }
    HX_TRY(hx::launch_tri_local(c, b->tri_idx.as<int32_t>(), mesh_of, toff, n, b->tri_local.as<int32_t>()));
    b->counts_known = false;
    b->tab_known = false;
    b->bin_of_valid = false;
    b->built = true;
    return HX_OK;
}

int read_tables(hx_bvh* b) {
    if (b->tab_known) return HX_OK;
    Context* c = &b->ctx->c;
    b->h_tab.resize(b->n_meshes);
    HX_HIP(hipMemcpyAsync(b->h_tab.data(), b->mesh_tab.ptr, sizeof(int4) * (size_t)b->n_meshes, hipMemcpyDeviceToHost,
                          c->stream));
    HX_HIP(hipStreamSynchronize(c->stream));
    b->tab_known = true;
    return HX_OK;
}

// HLBVHBuilder::build (HLBVHBuilder.cpp:332-410): Karras hierarchy and its
// boxes, clusters and the binned-SAH upper levels, then (unless everything is
// one cluster) the general climb and the collapse.
int run_hlbvh(hx_bvh* b) {
    Context* c = &b->ctx->c;
    const hx_build_params& p = b->params;
    const int64_t n = b->n_refs;
    HX_TRY(hx::launch_construct(c, b->keys[b->keys_index].ptr, b->keys64, n, b->left.as<int32_t>(),
                                b->right.as<int32_t>(), b->parent.as<int32_t>(), b->first.as<int32_t>(),
                                b->last.as<int32_t>(), b->counters.as<int32_t>(), b->cover.as<int32_t>(),
                                b->counts.as<int32_t>() + 2));
    HX_TRY(hx::launch_climb(c, climb_args(b, false)));
    hx::HlbvhArgs a{};
    a.n = n;
    a.bits = p.morton_bits;
    a.sah_bits = p.hlbvh_sah_bits;
    a.nbins = 1 << ((p.hlbvh_sah_bits + 2) / 3);
    a.keys = b->keys[b->keys_index].ptr;
    a.sorted_tri = b->tri_idx.as<int32_t>();
    a.tri_out = b->vals_tmp.as<int32_t>();
    a.left = b->left.as<int32_t>();
    a.right = b->right.as<int32_t>();
    a.parent = b->parent.as<int32_t>();
    a.first = b->first.as<int32_t>();
    a.last = b->last.as<int32_t>();
    a.split = b->split.as<int32_t>();
    a.nmin = b->nmin.as<float4>();
    a.nmax = b->nmax.as<float4>();
    char* w = b->hlbvh_work.as<char>();
    auto carve = [&](size_t bytes) {
        char* q = w;
        w += (bytes + 255) & ~(size_t)255;
        return q;
    };
    const size_t K = (size_t)hx::kHlbvhMaxClusters;
    a.cl_of = reinterpret_cast<int32_t*>(carve(4 * (size_t)n));
    a.tmp = reinterpret_cast<int32_t*>(carve(4 * (size_t)n));
    int32_t** ints[] = {&a.ca, &a.cb, &a.croot, &a.cl_off, &a.cl_par, &a.upper, &a.task_size, &a.child_counts};
    for (int32_t** q : ints) *q = reinterpret_cast<int32_t*>(carve(4 * K));
    a.members[0] = reinterpret_cast<int32_t*>(carve(4 * K));
    a.members[1] = reinterpret_cast<int32_t*>(carve(4 * K));
    a.coord = reinterpret_cast<int4*>(carve(16 * K));
    a.tasks[0] = reinterpret_cast<int4*>(carve(16 * K));
    a.tasks[1] = reinterpret_cast<int4*>(carve(16 * K));
    a.decision = reinterpret_cast<int4*>(carve(16 * K));
    int C = 0;
    HX_TRY(hx::launch_hlbvh(c, a, b->keys64, &C));
    b->hlbvh_clusters = C;
    if (C <= 1) {
        b->general = false;
        return hx::launch_collapse_emit(c, emit_args(b));
    }
    b->general = true;
    std::swap(b->tri_idx.ptr, b->vals_tmp.ptr);
    std::swap(b->tri_idx.bytes, b->vals_tmp.bytes);
    HX_TRY(hx::launch_refit_reset(c, b->counters.as<int32_t>(), n - 1, b->counts.as<int32_t>() + 2));
    HX_TRY(hx::launch_fill_i32(c, b->cover.as<int32_t>(), -1, n));
    HX_TRY(hx::launch_climb(c, climb_args(b, false)));
    return hx::launch_collapse_emit(c, emit_args(b));
}

// The Quad / Oct layout from the build's binary tree (BVHCollapser::
// collapseAdaptive / collapse for layouts other than Bin, BVHCollapser.cpp
// :573-595), on the device (hx_wide.hip).  lbvh_root: the root id is the split
// position the LBVH climb recorded (else node 0).
hx::WideArgs wide_args(hx_bvh* b, bool lbvh_root) {
    const hx_build_params& p = b->params;
    hx::WideArgs a{};
    a.n = b->n_refs;
    a.N = p.layout;
    a.left = b->left.as<int32_t>();
    a.right = b->right.as<int32_t>();
    a.parent = b->parent.as<int32_t>();
    a.first = b->first.as<int32_t>();
    a.last = b->last.as<int32_t>();
    a.nmin = b->nmin.as<float4>();
    a.nmax = b->nmax.as<float4>();
    a.root_id = lbvh_root ? b->counts.as<int32_t>() + 3 : nullptr;
    a.ct = p.ct;
    a.ci = p.ci;
    a.max_leaf = p.adaptive_leaf_size ? INT32_MAX : p.max_leaf_size;
    a.cost = b->wide_cost.as<float>();
    a.cnt = b->wide_cnt.as<uint16_t>();
    a.counters = b->counters.as<int32_t>();
    a.e_b = b->e_b.as<int32_t>();
    a.e_budget = b->e_budget.as<int8_t>();
    a.e_leaf = b->e_leaf.as<uint8_t>();
    a.e_parent = b->e_parent.as<int32_t>();
    a.e_rank = b->e_rank.as<int32_t>();
    a.e_first = b->e_first.as<int32_t>();
    a.e_nchild = b->e_nchild.as<int32_t>();
    a.scan = b->wide_scan.as<int2>();
    a.scan_part = b->wide_scan.as<int2>() + (2 * b->n_refs + 1);
    a.nodes = b->nodes.as<int4>();
    return a;
}

int run_wide(hx_bvh* b, bool lbvh_root) {
    Context* c = &b->ctx->c;
    const int64_t n = b->n_refs, T = std::max<int64_t>(2 * n - 1, 1);
    const int N = b->params.layout;
    HX_TRY(b->wide_cost.reserve(4 * (size_t)(N - 1) * (size_t)T));
    HX_TRY(b->wide_cnt.reserve(2 * (size_t)(N - 1) * (size_t)T));
    for (DevBuf* d : {&b->e_b, &b->e_parent, &b->e_rank, &b->e_first, &b->e_nchild})
        HX_TRY(d->reserve(4 * (size_t)(2 * n)));
    HX_TRY(b->e_budget.reserve((size_t)(2 * n)));
    HX_TRY(b->e_leaf.reserve((size_t)(2 * n)));
    HX_TRY(b->wide_scan.reserve(8 * (size_t)(2 * n + 1 + (2 * n) / 2048 + 2)));
    int64_t ni = 0, nl = 0, ne = 0;
    HX_TRY(hx::launch_wide_build(c, wide_args(b, lbvh_root), &ni, &nl, &ne));
    b->wide = N;
    b->wide_lbvh_root = lbvh_root;
    b->wide_entries = ne;
    b->n_interior = ni;
    b->n_leaves = nl;
    b->counts_known = true;
    return HX_OK;
}

int run_build(hx_bvh* b) {
    Context* c = &b->ctx->c;
    const hx_build_params& p = b->params;
    // Every build starts from a Karras tree; only the general stages set this.
    b->general = false;
    HX_TRY(run_front(b));
    const int64_t n = b->n_refs;
    const int32_t* sorted_tri = b->tri_idx.as<int32_t>();
    if (n == 1) {
        HX_TRY(hx::launch_single_ref(c, b->verts, b->tris, sorted_tri, b->nmin.as<float4>(), b->nmax.as<float4>(),
                                     b->woop.as<float4>(), b->nodes.as<int4>(), b->counts.as<int32_t>()));
    } else if (p.method == HX_METHOD_PLOC) {
        HX_TRY(run_ploc(b));
    } else if (p.method == HX_METHOD_ATRBVH || p.method == HX_METHOD_TR) {
        HX_TRY(run_atr(b));
    } else if (p.method == HX_METHOD_HLBVH) {
        HX_TRY(run_hlbvh(b));
    } else if (p.method == HX_METHOD_INSERTION) {
        HX_TRY(run_insertion(b));
    } else {
        const hx::ClimbArgs ca = lbvh_climb_args(b, b->keys[b->keys_index].ptr, b->keys64);
        HX_TRY(hx::launch_lbvh_climb(c, ca));
        hx::EmitArgs ea = emit_args(b);
        ea.root_id = ca.root_id;
        ea.proto = ca.proto;
        HX_TRY(hx::launch_collapse_emit(c, ea));
    }
    if (b->refs_active) {
        HX_TRY(hx::launch_refs_to_tris(c, b->tri_idx.as<int32_t>(), b->ref_tri.as<int32_t>(), n));
        b->refs_active = false;
    }
    b->counts_known = false;
    b->bin_of_valid = false;
    b->wide = 0;
    if (p.layout != HX_LAYOUT_BIN) HX_TRY(run_wide(b, p.method == HX_METHOD_LBVH && n > 1));
    b->built = true;
    return HX_OK;
}

// HipBVH::update (HipBVH.cpp:472-477): same topology, boxes and Woop rows
// from the new vertices.  The binary climb reruns in boxes-only mode and the
// emit kernel rewrites every node with the build's slots and collapse states.
int run_refit(hx_bvh* b) {
    Context* c = &b->ctx->c;
    const int64_t n = b->n_refs;
    b->pack_valid = false;  // the packed vertices are the build's
    HX_HIP(hipSetDevice(c->device));
    const int32_t* sorted_tri = b->tri_idx.as<int32_t>();
    if (n == 1) {
        HX_TRY(hx::launch_single_ref(c, b->verts, b->tris, sorted_tri, b->nmin.as<float4>(), b->nmax.as<float4>(),
                                     b->woop.as<float4>(), b->nodes.as<int4>(), b->counts.as<int32_t>()));
        if (b->wide) return hx::launch_wide_emit(c, wide_args(b, false), b->wide_entries, b->n_interior);
        return HX_OK;
    }
    HX_TRY(hx::launch_refit_reset(c, b->counters.as<int32_t>(), n - 1, b->counts.as<int32_t>() + 2));
    // Karras trees (LBVH builds) keep their Bin slots, links and sizes in a
    // refit: once the first refit after a build has recorded every binary
    // node's Bin slot, the climb writes the Bin rows itself and the emit is
    // skipped.  General trees rerun the emit.
    hx::ClimbArgs ca = climb_args(b, true);
    const bool fused = !b->wide && !b->general && b->bin_of_valid;
    if (fused) {
        ca.bin_of = b->bin_of.as<int32_t>();
        ca.bin_nodes = b->nodes.as<int4>();
    }
    HX_TRY(hx::launch_climb(c, ca));
    if (b->wide) return hx::launch_wide_emit(c, wide_args(b, b->wide_lbvh_root), b->wide_entries, b->n_interior);
    if (fused) return HX_OK;
    hx::EmitArgs ea = emit_args(b);
    if (!b->general) {
        const int64_t slots = std::max<int64_t>(2 * n, 2);
        HX_TRY(hx::launch_fill_i32(c, b->bin_of.as<int32_t>(), -1, slots));
        ea.bin_of = b->bin_of.as<int32_t>();
    }
    HX_TRY(hx::launch_emit_only(c, ea));
    // Under stream capture the recording emit only runs when the graph is
    // replayed: the slots stay unrecorded until an uncaptured refit.
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    HX_HIP(hipStreamIsCapturing(c->stream, &capture));
    b->bin_of_valid = !b->general && capture == hipStreamCaptureStatusNone;
    return HX_OK;
}

int read_counts(hx_bvh* b) {
    if (b->counts_known) return HX_OK;
    Context* c = &b->ctx->c;
    int32_t cnt[2] = {0, 0};
    HX_HIP(hipMemcpyAsync(cnt, b->counts.ptr, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    HX_HIP(hipStreamSynchronize(c->stream));
    b->n_interior = cnt[0];
    b->n_leaves = cnt[1];
    b->counts_known = true;
    return HX_OK;
}

int read_box(hx_bvh* b) {
    if (b->box_known) return HX_OK;
    Context* c = &b->ctx->c;
    HX_HIP(hipMemcpyAsync(b->host_box, b->scene_box.ptr, sizeof(b->host_box), hipMemcpyDeviceToHost, c->stream));
    HX_HIP(hipStreamSynchronize(c->stream));
    b->box_known = true;
    return HX_OK;
}

// One persistent trace over Bin BVH arrays (TRACE_FUNC's arguments); the ray
// counter and overflow count live in `control` (2 int32).
int trace_arrays(Context* c, const void* nodes, const void* woop, const int32_t* tri_idx, const float* d_rays,
                 int64_t n_rays, int32_t any_hit, int32_t* d_results, int32_t* d_stats, const int32_t* d_order,
                 int32_t* control, int wide = 0) {
    hx::TraceArgs a{};
    a.n_rays = n_rays;
    a.any_hit = any_hit ? 1 : 0;
    a.rays = reinterpret_cast<const float4*>(d_rays);
    a.results = reinterpret_cast<int4*>(d_results);
    a.nodes = static_cast<const int4*>(nodes);
    a.woop = static_cast<const float4*>(woop);
    a.tri_idx = tri_idx;
    a.stats = reinterpret_cast<int2*>(d_stats);
    a.order = d_order;
    a.control = control;
    a.fetch_threshold = c->tune.trace_fetch;
    if (wide) return hx::launch_trace_wide(c, wide, a);
    return hx::launch_trace(c, a);
}

int trace_common(hx_context* ctx, hx_bvh* bvh, const float* d_rays, int64_t n_rays, int32_t any_hit,
                 int32_t* d_results, int32_t* d_stats, const int32_t* d_order) {
    HX_CHECK_ARG(ctx && bvh, "context or bvh is NULL");
    HX_CHECK_ARG(n_rays >= 0 && n_rays < ((int64_t)1 << 30), "n_rays must be in [0, 2^30)");
    if (n_rays == 0) return HX_OK;
    HX_CHECK_ARG(d_rays && d_results, "rays and results must be device pointers");
    Context* c = &ctx->c;
    HX_TRY(c->control.reserve(64));
    return trace_arrays(c, bvh->nodes.ptr, bvh->woop.ptr, bvh->tri_idx.as<int32_t>(), d_rays, n_rays, any_hit,
                        d_results, d_stats, d_order, c->control.as<int32_t>(), bvh->wide);
}

// Ray keys and their sort into `order` (the sorted ray indices), from the
// context's buffers or, with ws, from caller memory (keys[2], order[2]).
int ray_sort(Context* c, const float* d_rays, int64_t n_rays, int32_t key_method, int32_t key_bits, float ray_length,
             void* keys[2], int32_t* order[2], const hx::SortWorkspace* ws, const int32_t** sorted) {
    HX_TRY(hx::launch_ray_keys(c, reinterpret_cast<const float4*>(d_rays), n_rays, key_method, key_bits, ray_length,
                               static_cast<uint64_t*>(keys[0])));
    int ri = 0;
    HX_TRY(hx::launch_sort_pairs(c, keys, order, n_rays, key_bits, true, true, &ri, false, ws));
    *sorted = order[ri];
    return HX_OK;
}

int check_ray_keys(int32_t key_method, int32_t key_bits) {
    HX_CHECK_ARG(key_method == HX_RAYKEY_AILA || key_method == HX_RAYKEY_TWOPOINT, "unknown ray key method %d",
                 key_method);
    HX_CHECK_ARG(key_bits >= 1 && key_bits <= 64, "key_bits must be in [1,64]");
    return HX_OK;
}

// hx_trace_sorted with optional event timing of its two halves.  With scr,
// the ray keys, order, sort status and trace counters come from that fixed
// scratch and nothing is allocated; else from the context's buffers.
int trace_sorted_common(hx_context* ctx, hx_bvh* bvh, const float* d_rays, int64_t n_rays, int32_t any_hit,
                        int32_t key_method, int32_t key_bits, float ray_length, int32_t* d_results, int32_t* d_stats,
                        int32_t* d_order, float* sort_seconds, float* trace_seconds,
                        const hx::TraceScratch* scr = nullptr) {
    HX_CHECK_ARG(ctx && bvh, "context or bvh is NULL");
    HX_CHECK_ARG(n_rays >= 0 && n_rays < ((int64_t)1 << 30), "n_rays must be in [0, 2^30)");
    HX_TRY(check_ray_keys(key_method, key_bits));
    if (sort_seconds) *sort_seconds = 0.0f;
    if (trace_seconds) *trace_seconds = 0.0f;
    if (n_rays == 0) return HX_OK;
    HX_CHECK_ARG(d_rays && d_results, "rays and results must be device pointers");
    Context* c = &ctx->c;
    void* keys[2];
    int32_t* order[2];
    int32_t* control = nullptr;
    const hx::SortWorkspace* ws = nullptr;
    if (scr) {
        for (int k = 0; k < 2; ++k) {
            keys[k] = scr->keys[k];
            order[k] = scr->order[k];
        }
        control = scr->control;
        ws = &scr->sort;
    } else {
        for (int k = 0; k < 2; ++k) {
            HX_TRY(c->ray_keys[k].reserve(8 * (size_t)n_rays));
            HX_TRY(c->ray_order[k].reserve(4 * (size_t)n_rays));
            keys[k] = c->ray_keys[k].ptr;
            order[k] = c->ray_order[k].as<int32_t>();
        }
        HX_TRY(c->control.reserve(64));
        control = c->control.as<int32_t>();
    }
    const bool timed = sort_seconds || trace_seconds;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    if (timed) {
        for (hipEvent_t& e : ev) HX_HIP(hipEventCreate(&e));
        HX_HIP(hipEventRecord(ev[0], c->stream));
    }
    const int32_t* sorted = nullptr;
    int st = ray_sort(c, d_rays, n_rays, key_method, key_bits, ray_length, keys, order, ws, &sorted);
    if (st == HX_OK && timed) st = hipEventRecord(ev[1], c->stream) == hipSuccess ? HX_OK : HX_ERR_HIP;
    if (st == HX_OK)
        st = trace_arrays(c, bvh->nodes.ptr, bvh->woop.ptr, bvh->tri_idx.as<int32_t>(), d_rays, n_rays, any_hit,
                          d_results, d_stats, sorted, control, bvh->wide);
    if (st == HX_OK && timed) st = hipEventRecord(ev[2], c->stream) == hipSuccess ? HX_OK : HX_ERR_HIP;
    if (st == HX_OK && d_order &&
        hipMemcpyAsync(d_order, sorted, 4 * (size_t)n_rays, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
        st = HX_ERR_HIP;
    if (st == HX_OK && timed) {
        float ms_sort = 0.0f, ms_trace = 0.0f;
        if (hipEventSynchronize(ev[2]) != hipSuccess || hipEventElapsedTime(&ms_sort, ev[0], ev[1]) != hipSuccess ||
            hipEventElapsedTime(&ms_trace, ev[1], ev[2]) != hipSuccess) {
            hx::set_error("trace timing: event query failed");
            st = HX_ERR_HIP;
        }
        if (sort_seconds) *sort_seconds = 1e-3f * ms_sort;
        if (trace_seconds) *trace_seconds = 1e-3f * ms_trace;
    }
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    return st;
}

}  // namespace

extern "C" {

const char* hx_status_string(int status) {
    switch (status) {
        case HX_OK: return "ok";
        case HX_ERR_INVALID_ARG: return "invalid argument";
        case HX_ERR_HIP: return "HIP runtime error";
        case HX_ERR_OUT_OF_MEMORY: return "out of device memory";
        case HX_ERR_UNSUPPORTED: return "unsupported";
        case HX_ERR_STACK_OVERFLOW: return "traversal stack overflow";
        default: return "unknown status";
    }
}

const char* hx_last_error(void) { return hx::g_error; }

int hx_context_create(int device, hx_context** out) {
    HX_CHECK_ARG(out != nullptr, "out is NULL");
    *out = nullptr;
    int count = 0;
    HX_HIP(hipGetDeviceCount(&count));
    HX_CHECK_ARG(device >= 0 && device < count, "device %d out of range (%d devices)", device, count);
    HX_HIP(hipSetDevice(device));
    hx_context* ctx = new (std::nothrow) hx_context();
    if (!ctx) return HX_ERR_OUT_OF_MEMORY;
    ctx->c.device = device;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        delete ctx;
        return hx::hip_fail(e, "hipGetDeviceProperties", __FILE__, __LINE__);
    }
    ctx->c.cu_count = prop.multiProcessorCount;
    const int st = hx::noalloc_setup(device);
    if (st != HX_OK) {
        delete ctx;
        return st;
    }
    *out = ctx;
    return HX_OK;
}

int hx_context_destroy(hx_context* ctx) {
    if (!ctx) return HX_OK;
    (void)hipSetDevice(ctx->c.device);
    (void)hipStreamSynchronize(ctx->c.stream);
    delete ctx;
    return HX_OK;
}

int hx_context_set_stream(hx_context* ctx, void* stream) {
    HX_CHECK_ARG(ctx != nullptr, "context is NULL");
    ctx->c.stream = static_cast<hipStream_t>(stream);
    return HX_OK;
}

int hx_context_synchronize(hx_context* ctx) {
    HX_CHECK_ARG(ctx != nullptr, "context is NULL");
    HX_HIP(hipStreamSynchronize(ctx->c.stream));
    return HX_OK;
}

int hx_context_set_profiling(hx_context* ctx, int enabled) {
    HX_CHECK_ARG(ctx != nullptr, "context is NULL");
    if (!enabled) HX_TRY(ctx->c.prof.flush());
    ctx->c.prof.enabled = enabled != 0;
    return HX_OK;
}

int hx_context_profile_count(hx_context* ctx, int32_t* count) {
    HX_CHECK_ARG(ctx && count, "NULL argument");
    HX_TRY(ctx->c.prof.flush());
    *count = (int32_t)ctx->c.prof.stats.size();
    return HX_OK;
}

int hx_context_profile_entry(hx_context* ctx, int32_t index, char* name, int32_t name_len, int64_t* launches,
                             double* total_ms) {
    HX_CHECK_ARG(ctx != nullptr, "context is NULL");
    HX_TRY(ctx->c.prof.flush());
    HX_CHECK_ARG(index >= 0 && index < (int32_t)ctx->c.prof.stats.size(), "profile index out of range");
    const hx::KernelStat& s = ctx->c.prof.stats[index];
    if (name && name_len > 0) {
        std::strncpy(name, s.name.c_str(), (size_t)name_len - 1);
        name[name_len - 1] = '\0';
    }
    if (launches) *launches = s.count;
    if (total_ms) *total_ms = s.total_ms;
    return HX_OK;
}

int hx_context_set_tuning(hx_context* ctx, const char* name, int32_t value) {
    HX_CHECK_ARG(ctx && name, "NULL argument");
    hx::Tuning& t = ctx->c.tune;
    if (!std::strcmp(name, "trace_lds_stack")) {
        HX_CHECK_ARG(value == 0 || value == 16 || value == 24 || value == 32 || value == 64,
                     "trace_lds_stack must be 0, 16, 24, 32 or 64");
        t.trace_lds_stack = value;
    } else if (!std::strcmp(name, "trace_fetch")) {
        HX_CHECK_ARG(value >= 1 && value <= 65, "trace_fetch must be in [1, 65]");
        t.trace_fetch = value;
    } else if (!std::strcmp(name, "trace_blocks_per_cu")) {
        HX_CHECK_ARG(value >= 0 && value <= 64, "trace_blocks_per_cu must be in [0, 64]");
        t.trace_blocks_per_cu = value;
    } else if (!std::strcmp(name, "trace_postpone")) {
        HX_CHECK_ARG(value >= 1 && value <= 4, "trace_postpone must be in [1, 4]");
        t.trace_postpone = value;
    } else if (!std::strcmp(name, "trace_woop_ahead")) {
        HX_CHECK_ARG(value == 1 || value == 2, "trace_woop_ahead must be 1 or 2");
        t.trace_woop_ahead = value;
    } else if (!std::strcmp(name, "sort_digit_bits")) {
        HX_CHECK_ARG(value == 8 || value == 10, "sort_digit_bits must be 8 or 10");
        t.sort_digit_bits = value;
    } else if (!std::strcmp(name, "sort_lookback")) {
        HX_CHECK_ARG(value == 4 || value == 8 || value == 16, "sort_lookback must be 4, 8 or 16");
        t.sort_lookback = value;
    } else if (!std::strcmp(name, "sort_items")) {
        HX_CHECK_ARG(value == 8 || value == 16 || value == 32, "sort_items must be 8, 16 or 32");
        t.sort_items = value;
    } else if (!std::strcmp(name, "sort_local_run")) {
        HX_CHECK_ARG(value >= 0 && value <= 2048, "sort_local_run must be in [0, 2048]");
        t.sort_local_run = value;
    } else {
        hx::set_error("unknown tuning parameter '%s'", name);
        return HX_ERR_INVALID_ARG;
    }
    return HX_OK;
}

int hx_context_get_tuning(hx_context* ctx, const char* name, int32_t* value) {
    HX_CHECK_ARG(ctx && name && value, "NULL argument");
    const hx::Tuning& t = ctx->c.tune;
    if (!std::strcmp(name, "trace_lds_stack")) *value = t.trace_lds_stack;
    else if (!std::strcmp(name, "trace_fetch")) *value = t.trace_fetch;
    else if (!std::strcmp(name, "trace_blocks_per_cu")) *value = t.trace_blocks_per_cu;
    else if (!std::strcmp(name, "trace_postpone")) *value = t.trace_postpone;
    else if (!std::strcmp(name, "trace_woop_ahead")) *value = t.trace_woop_ahead;
    else if (!std::strcmp(name, "sort_digit_bits")) *value = t.sort_digit_bits;
    else if (!std::strcmp(name, "sort_items")) *value = t.sort_items;
    else if (!std::strcmp(name, "sort_lookback")) *value = t.sort_lookback;
    else if (!std::strcmp(name, "sort_local_run")) *value = t.sort_local_run;
    else {
        hx::set_error("unknown tuning parameter '%s'", name);
        return HX_ERR_INVALID_ARG;
    }
    return HX_OK;
}

int hx_context_profile_reset(hx_context* ctx) {
    HX_CHECK_ARG(ctx != nullptr, "context is NULL");
    ctx->c.prof.reset();
    return HX_OK;
}

int hx_build_params_default(hx_build_params* p) {
    HX_CHECK_ARG(p != nullptr, "params is NULL");
    std::memset(p, 0, sizeof(*p));
    p->method = HX_METHOD_LBVH;
    p->morton_bits = 60;
    p->max_leaf_size = 8;
    p->adaptive_leaf_size = 1;
    p->ct = 3.0f;
    p->ci = 2.0f;
    p->presplitting = 0;
    p->presplitter_beta = 0.1f;
    p->ploc_radius = 1;
    p->ploc_max_radius = 512;
    p->ploc_adaptive = 0;
    p->atr_iterations = 2;
    p->atr_treelet_size = 20;
    p->tr_iterations = 2;
    p->hlbvh_sah_bits = 15;
    p->insertion_mod = 1;
    p->layout = HX_LAYOUT_BIN;
    return HX_OK;
}

int hx_build(hx_context* ctx, const hx_build_params* params, const float* d_vertices, int64_t n_vertices,
             const int32_t* d_triangles, int64_t n_triangles, hx_bvh** out) {
    HX_CHECK_ARG(ctx && out, "context or out is NULL");
    *out = nullptr;
    HX_TRY(validate_params(params));
    HX_CHECK_ARG(n_triangles > 0, "number of triangles must be positive (got %lld)", (long long)n_triangles);
    HX_CHECK_ARG(n_triangles < ((int64_t)1 << 30), "number of triangles must be < 2^30");
    HX_CHECK_ARG(d_vertices && d_triangles && n_vertices > 0, "geometry pointers/counts invalid");
    hx_bvh* b = new (std::nothrow) hx_bvh();
    if (!b) return HX_ERR_OUT_OF_MEMORY;
    b->ctx = ctx;
    b->params = *params;
    b->n_refs = n_triangles;
    b->n_tri = n_triangles;
    b->verts = d_vertices;
    b->tris = d_triangles;
    b->n_verts = n_vertices;
    b->keys64 = params->morton_bits > 32;
    int st = reserve_bvh(b, n_triangles);
    if (st == HX_OK) st = run_build(b);
    if (st != HX_OK) {
        delete b;
        return st;
    }
    *out = b;
    return HX_OK;
}

// The BVH's buffers are sized for the geometry it was built from: new
// geometry must have the same triangle and vertex counts.
int check_same_geometry(const hx_bvh* b, int64_t n_vertices, int64_t n_triangles) {
    HX_CHECK_ARG(n_triangles == b->n_tri, "geometry has %lld triangles, the BVH was built for %lld",
                 (long long)n_triangles, (long long)b->n_tri);
    HX_CHECK_ARG(n_vertices == b->n_verts, "geometry has %lld vertices, the BVH was built for %lld",
                 (long long)n_vertices, (long long)b->n_verts);
    return HX_OK;
}

int hx_rebuild(hx_context* ctx, hx_bvh* bvh, const float* d_vertices, int64_t n_vertices, const int32_t* d_triangles,
               int64_t n_triangles) {
    HX_CHECK_ARG(ctx && bvh, "context or bvh is NULL");
    HX_CHECK_ARG(d_vertices && d_triangles, "geometry pointers invalid");
    HX_TRY(check_same_geometry(bvh, n_vertices, n_triangles));
    bvh->ctx = ctx;
    bvh->verts = d_vertices;
    bvh->tris = d_triangles;
    return run_build(bvh);
}

int hx_refit(hx_context* ctx, hx_bvh* bvh, const float* d_vertices, int64_t n_vertices, const int32_t* d_triangles,
             int64_t n_triangles) {
    HX_CHECK_ARG(ctx && bvh && d_vertices && d_triangles, "NULL argument");
    HX_CHECK_ARG(bvh->built, "the BVH has not been built");
    HX_TRY(check_same_geometry(bvh, n_vertices, n_triangles));
    bvh->ctx = ctx;
    bvh->verts = d_vertices;
    bvh->tris = d_triangles;
    return run_refit(bvh);
}

int hx_backward(hx_context* ctx, const float* d_vertices, const int32_t* d_triangles, const float* d_rays,
                const int32_t* d_results, int64_t n_rays, const float* d_dl_dt, const float* d_dl_du,
                const float* d_dl_dv, float* d_dl_dvertices) {
    HX_CHECK_ARG(ctx && d_vertices && d_triangles && d_dl_dvertices, "NULL argument");
    HX_CHECK_ARG(n_rays >= 0, "n_rays must be >= 0");
    if (n_rays == 0) return HX_OK;
    HX_CHECK_ARG(d_rays && d_results, "rays and results must be device pointers");
    hx::BackwardArgs a{};
    a.n_rays = n_rays;
    a.verts = d_vertices;
    a.tris = d_triangles;
    a.rays = reinterpret_cast<const float4*>(d_rays);
    a.results = reinterpret_cast<const int4*>(d_results);
    a.dl_dt = d_dl_dt;
    a.dl_du = d_dl_du;
    a.dl_dv = d_dl_dv;
    a.dl_dverts = d_dl_dvertices;
    return hx::launch_backward(&ctx->c, a);
}

int hx_recompute_hits(hx_context* ctx, const float* d_vertices, const int32_t* d_triangles, const float* d_rays,
                      const int32_t* d_results, int64_t n_rays, float* d_t, float* d_u, float* d_v) {
    HX_CHECK_ARG(ctx && d_vertices && d_triangles && d_t && d_u && d_v, "NULL argument");
    HX_CHECK_ARG(n_rays >= 0, "n_rays must be >= 0");
    if (n_rays == 0) return HX_OK;
    HX_CHECK_ARG(d_rays && d_results, "rays and results must be device pointers");
    hx::BackwardArgs a{};
    a.n_rays = n_rays;
    a.verts = d_vertices;
    a.tris = d_triangles;
    a.rays = reinterpret_cast<const float4*>(d_rays);
    a.results = reinterpret_cast<const int4*>(d_results);
    return hx::launch_recompute_hits(&ctx->c, a, d_t, d_u, d_v);
}

int hx_build_batch(hx_context* ctx, const hx_build_params* params, const float* d_vertices, int64_t n_vertices,
                   const int32_t* d_triangles, const int64_t* h_mesh_offsets, int32_t n_meshes, hx_batch** out) {
    HX_CHECK_ARG(ctx && out && h_mesh_offsets, "NULL argument");
    *out = nullptr;
    HX_TRY(validate_params(params));
    HX_CHECK_ARG(params->method == HX_METHOD_LBVH || params->method == HX_METHOD_PLOC,
                 "batched builds use the LBVH or PLOC method (got %d)", params->method);
    HX_CHECK_ARG(params->layout == HX_LAYOUT_BIN, "batched builds use the Bin layout (got %d)", params->layout);

    HX_CHECK_ARG(n_meshes >= 1, "n_meshes must be positive (got %d)", n_meshes);
    HX_CHECK_ARG(h_mesh_offsets[0] == 0, "mesh offsets must start at 0");
    const int64_t n = h_mesh_offsets[n_meshes];
    HX_CHECK_ARG(n < ((int64_t)1 << 30), "total number of triangles must be < 2^30");
    for (int32_t m = 0; m < n_meshes; ++m)
        HX_CHECK_ARG(h_mesh_offsets[m + 1] - h_mesh_offsets[m] >= 2, "mesh %d has fewer than 2 triangles", m);
    HX_CHECK_ARG(d_vertices && d_triangles && n_vertices > 0, "geometry pointers/counts invalid");
    hx_batch* bt = new (std::nothrow) hx_batch();
    if (!bt) return HX_ERR_OUT_OF_MEMORY;
    hx_bvh* b = &bt->bvh;
    b->ctx = ctx;
    b->params = *params;
    b->n_refs = n;
    b->n_tri = n;
    b->verts = d_vertices;
    b->tris = d_triangles;
    b->n_verts = n_vertices;
    b->keys64 = true;
    b->batched = true;
    b->n_meshes = n_meshes;
    b->h_off.assign(h_mesh_offsets, h_mesh_offsets + n_meshes + 1);
    int st = reserve_bvh(b, n);
    if (st == HX_OK) st = run_batch(b);
    if (st != HX_OK) {
        delete bt;
        return st;
    }
    *out = bt;
    return HX_OK;
}

int hx_batch_refit(hx_context* ctx, hx_batch* batch, const float* d_vertices, int64_t n_vertices,
                   const int32_t* d_triangles, int64_t n_triangles) {
    HX_CHECK_ARG(ctx && batch && d_vertices && d_triangles, "NULL argument");
    hx_bvh* b = &batch->bvh;
    HX_TRY(check_same_geometry(b, n_vertices, n_triangles));
    b->ctx = ctx;
    b->verts = d_vertices;
    b->tris = d_triangles;
    return run_refit(b);
}

int hx_batch_rebuild(hx_context* ctx, hx_batch* batch, const float* d_vertices, int64_t n_vertices,
                     const int32_t* d_triangles, int64_t n_triangles) {
    HX_CHECK_ARG(ctx && batch && d_vertices && d_triangles, "NULL argument");
    hx_bvh* b = &batch->bvh;
    HX_TRY(check_same_geometry(b, n_vertices, n_triangles));
    b->ctx = ctx;
    b->verts = d_vertices;
    b->tris = d_triangles;
    return run_batch(b);
}

int hx_batch_destroy(hx_batch* batch) {
    if (!batch) return HX_OK;
    if (batch->bvh.ctx) {
        (void)hipSetDevice(batch->bvh.ctx->c.device);
        (void)hipStreamSynchronize(batch->bvh.ctx->c.stream);
    }
    delete batch;
    return HX_OK;
}

int hx_batch_mesh(hx_batch* batch, int32_t mesh, hx_batch_mesh_view* out) {
    HX_CHECK_ARG(batch && out, "NULL argument");
    hx_bvh* b = &batch->bvh;
    HX_CHECK_ARG(mesh >= 0 && mesh < b->n_meshes, "mesh %d out of range (%d meshes)", mesh, b->n_meshes);
    HX_TRY(read_counts(b));
    HX_TRY(read_tables(b));
    const int4 t = b->h_tab[mesh];
    const int next_leaf = mesh + 1 < b->n_meshes ? b->h_tab[mesh + 1].z : (int)b->n_leaves;
    const int64_t off = b->h_roff[mesh];
    out->nodes = b->nodes.as<int32_t>() + 16 * (int64_t)t.x;
    out->woop = b->woop.as<float>() + 12 * off;
    out->tri_indices = b->tri_local.as<int32_t>() + off;
    out->n_interior = t.w;
    out->n_leaves = next_leaf - t.z;
    out->n_references = b->h_roff[mesh + 1] - off;
    out->tri_offset = off;
    return HX_OK;
}

int hx_batch_trace(hx_context* ctx, hx_batch* batch, const float* d_rays, const int32_t* d_ray_mesh,
                   int64_t n_rays, int32_t any_hit, int32_t* d_results, int32_t* d_stats) {
    HX_CHECK_ARG(ctx && batch && d_ray_mesh, "NULL argument");
    HX_CHECK_ARG(n_rays >= 0 && n_rays < ((int64_t)1 << 30), "n_rays must be in [0, 2^30)");
    if (n_rays == 0) return HX_OK;
    HX_CHECK_ARG(d_rays && d_results, "rays and results must be device pointers");
    hx_bvh* b = &batch->bvh;
    Context* c = &ctx->c;
    HX_TRY(c->control.reserve(64));
    hx::TraceArgs a{};
    a.n_rays = n_rays;
    a.any_hit = any_hit ? 1 : 0;
    a.rays = reinterpret_cast<const float4*>(d_rays);
    a.results = reinterpret_cast<int4*>(d_results);
    a.nodes = b->nodes.as<int4>();
    a.woop = b->woop.as<float4>();
    a.tri_idx = b->tri_local.as<int32_t>();
    a.stats = reinterpret_cast<int2*>(d_stats);
    a.control = c->control.as<int32_t>();
    a.fetch_threshold = c->tune.trace_fetch;
    a.ray_mesh = d_ray_mesh;
    a.mesh_tab = b->mesh_tab.as<int4>();
    a.mesh_off = b->mesh_off.as<int32_t>();
    a.n_meshes = b->n_meshes;
    return hx::launch_trace(c, a);
}

int hx_batch_backward(hx_context* ctx, hx_batch* batch, const float* d_vertices, const int32_t* d_triangles,
                      const float* d_rays, const int32_t* d_ray_mesh, const int32_t* d_results, int64_t n_rays,
                      const float* d_dl_dt, const float* d_dl_du, const float* d_dl_dv, float* d_dl_dvertices) {
    HX_CHECK_ARG(ctx && batch && d_vertices && d_triangles && d_ray_mesh && d_dl_dvertices, "NULL argument");
    HX_CHECK_ARG(n_rays >= 0, "n_rays must be >= 0");
    if (n_rays == 0) return HX_OK;
    HX_CHECK_ARG(d_rays && d_results, "rays and results must be device pointers");
    hx::BackwardArgs a{};
    a.n_rays = n_rays;
    a.verts = d_vertices;
    a.tris = d_triangles;
    a.rays = reinterpret_cast<const float4*>(d_rays);
    a.results = reinterpret_cast<const int4*>(d_results);
    a.dl_dt = d_dl_dt;
    a.dl_du = d_dl_du;
    a.dl_dv = d_dl_dv;
    a.dl_dverts = d_dl_dvertices;
    a.ray_mesh = d_ray_mesh;
    a.mesh_off = batch->bvh.tri_off.as<int32_t>();
    return hx::launch_backward(&ctx->c, a);
}

int hx_batch_recompute_hits(hx_context* ctx, hx_batch* batch, const float* d_vertices, const int32_t* d_triangles,
                            const float* d_rays, const int32_t* d_ray_mesh, const int32_t* d_results, int64_t n_rays,
                            float* d_t, float* d_u, float* d_v) {
    HX_CHECK_ARG(ctx && batch && d_vertices && d_triangles && d_ray_mesh && d_t && d_u && d_v, "NULL argument");
    HX_CHECK_ARG(n_rays >= 0, "n_rays must be >= 0");
    if (n_rays == 0) return HX_OK;
    HX_CHECK_ARG(d_rays && d_results, "rays and results must be device pointers");
    hx::BackwardArgs a{};
    a.n_rays = n_rays;
    a.verts = d_vertices;
    a.tris = d_triangles;
    a.rays = reinterpret_cast<const float4*>(d_rays);
    a.results = reinterpret_cast<const int4*>(d_results);
    a.ray_mesh = d_ray_mesh;
    a.mesh_off = batch->bvh.tri_off.as<int32_t>();
    return hx::launch_recompute_hits(&ctx->c, a, d_t, d_u, d_v);
}

int hx_bvh_destroy(hx_bvh* bvh) {
    if (!bvh) return HX_OK;
    if (bvh->ctx) {
        (void)hipSetDevice(bvh->ctx->c.device);
        (void)hipStreamSynchronize(bvh->ctx->c.stream);
    }
    delete bvh;
    return HX_OK;
}

int hx_bvh_get_view(hx_bvh* bvh, hx_bvh_view* out) {
    HX_CHECK_ARG(bvh && out, "NULL argument");
    HX_TRY(read_counts(bvh));
    HX_TRY(read_box(bvh));
    out->nodes = bvh->nodes.ptr;
    out->woop = bvh->woop.ptr;
    out->tri_indices = bvh->tri_idx.as<int32_t>();
    out->n_interior = bvh->n_interior;
    out->n_leaves = bvh->n_leaves;
    out->n_references = bvh->n_refs;
    std::memcpy(out->scene_box, bvh->host_box, sizeof(out->scene_box));
    out->layout = bvh->wide ? bvh->wide : HX_LAYOUT_BIN;
    return HX_OK;
}

int hx_bvh_debug(hx_bvh* bvh, hx_bvh_debug_view* out) {
    HX_CHECK_ARG(bvh && out, "NULL argument");
    out->sorted_keys = bvh->keys[bvh->keys_index].ptr;
    out->keys_are_64bit = bvh->keys64 ? 1 : 0;
    out->sorted_tri = bvh->tri_idx.as<int32_t>();
    out->left = bvh->left.as<int32_t>();
    out->right = bvh->right.as<int32_t>();
    out->parent = bvh->parent.as<int32_t>();
    out->node_min = bvh->nmin.as<float>();
    out->node_max = bvh->nmax.as<float>();
    out->state = bvh->state.as<int32_t>();
    return HX_OK;
}

int hx_bvh_info(hx_bvh* bvh, const char* key, int64_t* value) {
    HX_CHECK_ARG(bvh && key && value, "NULL argument");
    if (!std::strcmp(key, "ploc_rounds")) {
        *value = bvh->ploc_rounds;
    } else if (!std::strcmp(key, "atr_accepted")) {
        *value = bvh->atr_accepted;
    } else if (!std::strcmp(key, "hlbvh_clusters")) {
        *value = bvh->hlbvh_clusters;
    } else if (!std::strcmp(key, "references")) {
        *value = bvh->n_refs;
    } else if (!std::strcmp(key, "presplit_d_bits")) {
        uint32_t u;
        std::memcpy(&u, &bvh->ps_D, 4);
        *value = u;
    } else if (!std::strcmp(key, "insertion_steps")) {
        *value = bvh->ins_steps;
    } else if (!std::strcmp(key, "insertion_reinserted")) {
        *value = bvh->ins_reinserted;
    } else if (!std::strcmp(key, "method")) {
        *value = bvh->params.method;
    } else {
        hx::set_error("unknown build statistic '%s'", key);
        return HX_ERR_INVALID_ARG;
    }
    return HX_OK;
}

int hx_bvh_cost(hx_bvh* bvh, double* out) {
    HX_CHECK_ARG(bvh && out, "NULL argument");
    Context* c = &bvh->ctx->c;
    if (bvh->wide) {
        // computeCost of HipBVHNodeQuad / Oct (HipBVHKernels.cu:184-238) over
        // the scene box's area (HipBVH::getCost, HipBVH.cpp:346-370).
        HX_TRY(read_box(bvh));
        const float* bx = bvh->host_box;
        const float dx = bx[3] - bx[0], dy = bx[4] - bx[1], dz = bx[5] - bx[2];
        const bool valid = bx[0] <= bx[3] && bx[1] <= bx[4] && bx[2] <= bx[5];
        const float area = valid ? 2.0f * (dx * dy + dy * dz + dz * dx) : 0.0f;
        const int64_t nn = bvh->n_interior + bvh->n_leaves;
        HX_TRY(bvh->wide_partial.reserve(8 * (size_t)hx::wide_cost_partials(nn)));
        HX_TRY(hx::launch_wide_cost(c, bvh->wide, bvh->nodes.as<int4>(), nn, area, bvh->params.ct, bvh->params.ci,
                                    bvh->wide_partial.as<double>(), bvh->cost_out.as<double>()));
        HX_HIP(hipMemcpyAsync(out, bvh->cost_out.ptr, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HX_HIP(hipStreamSynchronize(c->stream));
        return HX_OK;
    }
    const int64_t max_nodes = std::max<int64_t>(2 * bvh->n_refs - 1, 2);
    HX_TRY(hx::launch_cost(c, bvh->nodes.as<int4>(), bvh->counts.as<int32_t>(), max_nodes,
                           bvh->scene_box.as<float>(), bvh->params.ct, bvh->params.ci, bvh->cost_out.as<double>()));
    HX_HIP(hipMemcpyAsync(out, bvh->cost_out.ptr, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HX_HIP(hipStreamSynchronize(c->stream));
    return HX_OK;
}

int hx_bvh_validate(hx_bvh* bvh, int32_t* failed) {
    HX_CHECK_ARG(bvh && failed, "NULL argument");
    HX_CHECK_ARG(!bvh->batched, "validation is per BVH; a batch has one BVH per mesh");
    HX_CHECK_ARG(!bvh->wide, "device validation covers the Bin layout");
    HX_CHECK_ARG(bvh->built, "the BVH has not been built");
    Context* c = &bvh->ctx->c;
    const int64_t max_nodes = std::max<int64_t>(2 * bvh->n_refs - 1, 2);
    HX_TRY(bvh->leaf_stats.reserve(sizeof(int32_t) * (size_t)(max_nodes + bvh->n_tri + 64)));
    int32_t* work = bvh->leaf_stats.as<int32_t>() + 64;
    int32_t* flag = bvh->leaf_stats.as<int32_t>();
    HX_TRY(hx::launch_validate(c, bvh->nodes.as<int4>(), bvh->counts.as<int32_t>(), max_nodes,
                               bvh->tri_idx.as<int32_t>(), bvh->n_refs, bvh->verts, bvh->tris, bvh->n_tri, work, flag));
    HX_HIP(hipMemcpyAsync(failed, flag, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HX_HIP(hipStreamSynchronize(c->stream));
    return HX_OK;
}

int hx_bvh_leaf_stats(hx_bvh* bvh, int32_t* histogram33, int64_t* sum_of_leaf_sizes) {
    HX_CHECK_ARG(bvh && histogram33 && sum_of_leaf_sizes, "NULL argument");
    HX_CHECK_ARG(!bvh->batched, "leaf statistics are per BVH; a batch has one BVH per mesh");
    HX_CHECK_ARG(!bvh->wide, "leaf statistics cover the Bin layout");
    Context* c = &bvh->ctx->c;
    HX_TRY(bvh->leaf_stats.reserve(sizeof(int32_t) * (hx::kMaxLeafSizeStat + 2)));
    HX_TRY(hx::launch_leaf_stats(c, bvh->nodes.as<int4>(), bvh->counts.as<int32_t>(),
                                 std::max<int64_t>(bvh->n_refs, 1), bvh->leaf_stats.as<int32_t>()));
    int32_t h[hx::kMaxLeafSizeStat + 2];
    HX_HIP(hipMemcpyAsync(h, bvh->leaf_stats.ptr, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HX_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k <= hx::kMaxLeafSizeStat; ++k) histogram33[k] = h[k];
    *sum_of_leaf_sizes = (uint32_t)h[hx::kMaxLeafSizeStat + 1];
    return HX_OK;
}

int hx_normalize_vertices(hx_context* ctx, float* d_vertices, int64_t n_vertices, float* d_box6) {
    HX_CHECK_ARG(ctx && d_vertices && n_vertices > 0, "invalid argument");
    Context* c = &ctx->c;
    float* box = d_box6;
    if (!box) {
        HX_TRY(c->control.reserve(64));
        box = c->control.as<float>() + 8;
    }
    return hx::launch_normalize(c, d_vertices, n_vertices, box);
}

int hx_init_seeds(hx_context* ctx, int64_t n, int32_t frame_index, uint32_t* d_seeds) {
    HX_CHECK_ARG(ctx && n >= 0 && n < ((int64_t)1 << 31), "n must be in [0, 2^31)");
    if (n == 0) return HX_OK;
    HX_CHECK_ARG(d_seeds, "seeds must be a device pointer");
    return hx::launch_init_seeds(&ctx->c, n, (uint32_t)frame_index, d_seeds);
}

int hx_primary_rays(hx_context* ctx, int32_t sample_index, const float* origin3, const float* screen_to_world16,
                    int32_t width, int32_t height, float max_dist, const int32_t* d_index_to_pixel, float* d_rays) {
    HX_CHECK_ARG(ctx && origin3 && screen_to_world16, "context, origin and matrix must not be NULL");
    HX_CHECK_ARG(width > 0 && height > 0 && (int64_t)width * height < ((int64_t)1 << 31),
                 "image size %d x %d is empty or too large", width, height);
    HX_CHECK_ARG(sample_index >= 0, "sample_index must be >= 0");
    HX_CHECK_ARG(d_rays, "rays must be a device pointer");
    return hx::launch_primary_rays(&ctx->c, sample_index, origin3, screen_to_world16, width, height, max_dist,
                                   d_index_to_pixel, d_rays);
}

int hx_shadow_rays(hx_context* ctx, int64_t batch_begin, int64_t batch_size, int32_t n_samples,
                   const uint32_t* d_seeds, float light_radius, const float* light3, const float* d_in_rays,
                   const int32_t* d_in_results, float* d_out_rays) {
    HX_CHECK_ARG(ctx && light3, "context and light must not be NULL");
    HX_CHECK_ARG(batch_begin >= 0 && batch_size >= 0 && n_samples >= 1, "invalid batch or sample count");
    HX_CHECK_ARG(batch_size * (int64_t)n_samples < ((int64_t)1 << 31), "too many output rays");
    if (batch_size == 0) return HX_OK;
    HX_CHECK_ARG(d_seeds && d_in_rays && d_in_results && d_out_rays, "buffers must be device pointers");
    return hx::launch_secondary_rays(&ctx->c, false, batch_begin, batch_size, n_samples, d_seeds, light_radius, light3,
                                     0.0f, d_in_rays, d_in_results, nullptr, nullptr, d_out_rays);
}

int hx_ao_rays(hx_context* ctx, int64_t batch_begin, int64_t batch_size, int32_t n_samples, const uint32_t* d_seeds,
               float max_dist, const float* d_in_rays, const int32_t* d_in_results, const int32_t* d_triangles,
               const float* d_vertices, float* d_out_rays) {
    HX_CHECK_ARG(ctx, "context is NULL");
    HX_CHECK_ARG(batch_begin >= 0 && batch_size >= 0 && n_samples >= 1, "invalid batch or sample count");
    HX_CHECK_ARG(batch_size * (int64_t)n_samples < ((int64_t)1 << 31), "too many output rays");
    if (batch_size == 0) return HX_OK;
    HX_CHECK_ARG(d_seeds && d_in_rays && d_in_results && d_triangles && d_vertices && d_out_rays,
                 "buffers must be device pointers");
    return hx::launch_secondary_rays(&ctx->c, true, batch_begin, batch_size, n_samples, d_seeds, 0.0f, nullptr,
                                     max_dist, d_in_rays, d_in_results, d_triangles, d_vertices, d_out_rays);
}

int hx_path_rays(hx_context* ctx, int32_t russian_roulette, int64_t n_in, uint32_t* d_seeds,
                 const int32_t* d_in_index_to_pixel, const int32_t* d_mat_indices, const int32_t* d_triangles,
                 const float* d_normals, const void* d_materials, const float* d_in_rays, const int32_t* d_in_results,
                 float* d_out_rays, int32_t* d_out_index_to_pixel, float* d_decreases, int32_t* d_n_out) {
    HX_CHECK_ARG(ctx && n_in >= 0 && n_in < ((int64_t)1 << 31), "n_in must be in [0, 2^31)");
    HX_CHECK_ARG(d_n_out, "n_out must be a device pointer");
    if (n_in > 0)
        HX_CHECK_ARG(d_seeds && d_in_index_to_pixel && d_mat_indices && d_triangles && d_normals && d_materials &&
                         d_in_rays && d_in_results && d_out_rays && d_out_index_to_pixel && d_decreases,
                     "buffers must be device pointers");
    hx::PathRayArgs a{};
    a.n = n_in;
    a.russian_roulette = russian_roulette ? 1 : 0;
    a.seeds = d_seeds;
    a.in_pixel = d_in_index_to_pixel;
    a.mat_indices = d_mat_indices;
    a.tris = d_triangles;
    a.normals = d_normals;
    a.materials = d_materials;
    a.in_rays = d_in_rays;
    a.in_results = d_in_results;
    a.out_rays = d_out_rays;
    a.out_pixel = d_out_index_to_pixel;
    a.decreases = d_decreases;
    a.n_out = d_n_out;
    return hx::launch_path_rays(&ctx->c, a);
}

int hx_reconstruct_occlusion(hx_context* ctx, int64_t batch_begin, int64_t batch_size, int32_t n_samples,
                             int32_t replace, const int32_t* d_results, const int32_t* d_index_to_pixel,
                             const int32_t* d_index_to_slot, const float* d_in_pixels, float* d_out_pixels) {
    HX_CHECK_ARG(ctx && batch_begin >= 0 && batch_size >= 0 && n_samples >= 1, "invalid batch or sample count");
    if (batch_size == 0) return HX_OK;
    HX_CHECK_ARG(d_results && d_index_to_pixel && d_in_pixels && d_out_pixels, "buffers must be device pointers");
    return hx::launch_reconstruct_occlusion(&ctx->c, batch_begin, batch_size, n_samples, replace != 0, d_results,
                                            d_index_to_pixel, d_index_to_slot, d_in_pixels, d_out_pixels);
}

int hx_interpolate_vertices(hx_context* ctx, int64_t n_static, int64_t n_dynamic, float t, const float* d_frame0,
                            const float* d_frame1, float* d_out, int32_t normalize) {
    HX_CHECK_ARG(ctx && n_static >= 0 && n_dynamic >= 0, "invalid argument");
    if (n_dynamic == 0) return HX_OK;
    HX_CHECK_ARG(d_frame0 && d_frame1 && d_out, "buffers must be device pointers");
    return hx::launch_interpolate(&ctx->c, n_static, n_dynamic, t, d_frame0, d_frame1, d_out, normalize != 0);
}

int hx_scene_box(hx_context* ctx, const float* d_vertices, const int32_t* d_triangles, int64_t n, float* d_box6) {
    HX_CHECK_ARG(ctx && d_vertices && d_triangles && d_box6 && n > 0, "invalid argument");
    return hx::launch_scene_box(&ctx->c, d_vertices, d_triangles, n, d_box6);
}

int hx_morton_codes(hx_context* ctx, const float* d_vertices, const int32_t* d_triangles, int64_t n,
                    const float* d_box6, int32_t bits, uint64_t* d_keys) {
    HX_CHECK_ARG(ctx && d_vertices && d_triangles && d_box6 && d_keys && n > 0, "invalid argument");
    HX_CHECK_ARG(bits >= 6 && bits <= 60, "bits must be in [6,60]");
    return hx::launch_morton(&ctx->c, d_vertices, d_triangles, n, d_box6, bits, d_keys, true);
}

int hx_sort_pairs_u64(hx_context* ctx, uint64_t* d_keys, int32_t* d_values, int64_t n, int32_t end_bit) {
    HX_CHECK_ARG(ctx && d_keys && d_values && n >= 0, "invalid argument");
    HX_CHECK_ARG(end_bit >= 0 && end_bit <= 64, "end_bit must be in [0,64]");
    if (n <= 1) return HX_OK;
    Context* c = &ctx->c;
    HX_TRY(c->tmp_keys.reserve(8 * (size_t)n));
    HX_TRY(c->tmp_vals.reserve(4 * (size_t)n));
    void* keys[2] = {d_keys, c->tmp_keys.ptr};
    int32_t* vals[2] = {d_values, c->tmp_vals.as<int32_t>()};
    int ri = 0;
    HX_TRY(hx::launch_sort_pairs(c, keys, vals, n, end_bit, true, false, &ri));
    if (ri == 1) {
        HX_HIP(hipMemcpyAsync(d_keys, keys[1], 8 * (size_t)n, hipMemcpyDeviceToDevice, c->stream));
        HX_HIP(hipMemcpyAsync(d_values, vals[1], 4 * (size_t)n, hipMemcpyDeviceToDevice, c->stream));
    }
    return HX_OK;
}

int hx_trace(hx_context* ctx, hx_bvh* bvh, const float* d_rays, int64_t n_rays, int32_t any_hit, int32_t* d_results,
             int32_t* d_stats) {
    return trace_common(ctx, bvh, d_rays, n_rays, any_hit, d_results, d_stats, nullptr);
}

int hx_ray_keys(hx_context* ctx, const float* d_rays, int64_t n_rays, int32_t key_method, int32_t key_bits,
                float ray_length, uint64_t* d_keys) {
    HX_CHECK_ARG(ctx && d_rays && d_keys && n_rays >= 0, "invalid argument");
    HX_TRY(check_ray_keys(key_method, key_bits));
    return hx::launch_ray_keys(&ctx->c, reinterpret_cast<const float4*>(d_rays), n_rays, key_method, key_bits,
                               ray_length, d_keys);
}

int hx_trace_sorted(hx_context* ctx, hx_bvh* bvh, const float* d_rays, int64_t n_rays, int32_t any_hit,
                    int32_t key_method, int32_t key_bits, float ray_length, int32_t* d_results, int32_t* d_stats,
                    int32_t* d_order) {
    return trace_sorted_common(ctx, bvh, d_rays, n_rays, any_hit, key_method, key_bits, ray_length, d_results,
                               d_stats, d_order, nullptr, nullptr);
}

int hx_trace_sorted_timed(hx_context* ctx, hx_bvh* bvh, const float* d_rays, int64_t n_rays, int32_t any_hit,
                          int32_t key_method, int32_t key_bits, float ray_length, int32_t* d_results,
                          int32_t* d_stats, int32_t* d_order, float* sort_seconds, float* trace_seconds) {
    HX_CHECK_ARG(sort_seconds && trace_seconds, "sort_seconds and trace_seconds must be host pointers");
    return trace_sorted_common(ctx, bvh, d_rays, n_rays, any_hit, key_method, key_bits, ray_length, d_results,
                               d_stats, d_order, sort_seconds, trace_seconds);
}

int hx_trace_sorted_noalloc(hx_context* ctx, hx_bvh* bvh, const float* d_rays, int64_t n_rays, int32_t any_hit,
                            int32_t key_method, int32_t key_bits, float ray_length, int32_t* d_results,
                            int32_t* d_stats, int32_t* d_order) {
    HX_CHECK_ARG(ctx && bvh, "context or bvh is NULL");
    HX_CHECK_ARG(n_rays >= 0, "n_rays must be >= 0");
    hx::TraceScratch scr{};
    HX_TRY(hx::noalloc_trace_scratch(ctx->c.device, n_rays, &scr));
    return trace_sorted_common(ctx, bvh, d_rays, n_rays, any_hit, key_method, key_bits, ray_length, d_results,
                               d_stats, d_order, nullptr, nullptr, &scr);
}

int hx_trace_view(hx_context* ctx, const void* d_nodes, int64_t n_nodes, const void* d_woop,
                  const int32_t* d_tri_indices, int64_t n_references, const float* d_rays, int64_t n_rays,
                  int32_t any_hit, int32_t* d_results, int32_t* d_stats, const int32_t* d_indices) {
    HX_CHECK_ARG(ctx, "context is NULL");
    HX_CHECK_ARG(n_rays >= 0 && n_rays < ((int64_t)1 << 30), "n_rays must be in [0, 2^30)");
    HX_CHECK_ARG(d_nodes && d_woop && d_tri_indices, "nodes, woop triangles and triangle indices must be device "
                                                     "pointers");
    HX_CHECK_ARG(n_nodes >= 2 && n_references >= 1, "a Bin BVH has at least 2 nodes and 1 reference (got %lld, %lld)",
                 (long long)n_nodes, (long long)n_references);
    if (n_rays == 0) return HX_OK;
    HX_CHECK_ARG(d_rays && d_results, "rays and results must be device pointers");
    Context* c = &ctx->c;
    HX_TRY(c->control.reserve(64));
    return trace_arrays(c, d_nodes, d_woop, d_tri_indices, d_rays, n_rays, any_hit, d_results, d_stats, d_indices,
                        c->control.as<int32_t>());
}

int hx_memcpy(hx_context* ctx, void* dst, const void* src, int64_t bytes) {
    HX_CHECK_ARG(ctx && bytes >= 0, "invalid argument");
    if (bytes == 0) return HX_OK;
    HX_CHECK_ARG(dst && src, "NULL pointer");
    HX_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDefault, ctx->c.stream));
    HX_HIP(hipStreamSynchronize(ctx->c.stream));
    return HX_OK;
}

int hx_trace_overflow_count(hx_context* ctx, int64_t* out) {
    HX_CHECK_ARG(ctx && out, "NULL argument");
    Context* c = &ctx->c;
    if (!c->control.ptr) {
        *out = 0;
        return HX_OK;
    }
    int32_t v[2] = {0, 0};
    HX_HIP(hipMemcpyAsync(v, c->control.ptr, sizeof(v), hipMemcpyDeviceToHost, c->stream));
    HX_HIP(hipStreamSynchronize(c->stream));
    *out = v[1];
    return HX_OK;
}

}  // extern "C"
