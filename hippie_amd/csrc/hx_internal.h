// hx_internal.h -- host-side internals shared by the libhx translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/hx.h"

namespace hx {

// Status propagation: every internal call returns an hx status code.
#define HX_TRY(expr)                      \
    do {                                  \
        int _st = (expr);                 \
        if (_st != HX_OK) return _st;     \
    } while (0)

#define HX_HIP(expr)                                   \
    do {                                               \
        hipError_t _e = (expr);                        \
        if (_e != hipSuccess) return ::hx::hip_fail(_e, #expr, __FILE__, __LINE__); \
    } while (0)

int hip_fail(hipError_t e, const char* what, const char* file, int line);
void set_error(const char* fmt, ...);

// Device allocation owned by the library (grow-only).
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    int reserve(size_t want);
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(ptr); }
};

// Per-launch event timing, enabled by hx_context_set_profiling.
struct KernelStat {
    std::string name;
    int64_t count = 0;
    double total_ms = 0.0;
};

struct Profiler {
    bool enabled = false;
    struct Pending {
        int stat;
        hipEvent_t a, b;
    };
    std::vector<KernelStat> stats;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    int begin(hipStream_t s, const char* name, hipEvent_t* a, hipEvent_t* b);
    void end(hipStream_t s, hipEvent_t b);
    int flush();
    void reset();
    ~Profiler();
};

// Launch-shape choices the autotuner may change (hx_context_set_tuning).
struct Tuning {
    int trace_lds_stack = 16;    // traversal-stack entries kept in LDS: 0, 16, 24, 32, 64
    int trace_fetch = 40;        // refill a wave when fewer lanes than this are still busy
    int trace_blocks_per_cu = 0; // persistent trace workgroups per CU (0 = occupancy limit)
    int trace_postpone = 3;      // leaves a lane holds back while its wave traverses (1 to 4)
    int trace_woop_ahead = 1;    // Woop rows of the next triangle read during a test (1 or 2)
    int sort_digit_bits = 8;     // radix digit width: 8 or 10
    int sort_items = 16;         // keys per thread in a sort pass: 8, 16 or 32
    int sort_lookback = 4;       // predecessor tiles read per look-back round trip: 4, 8 or 16
    int sort_local_run = 64;     // longest run of equal high key bits the local pass ranks (0: full LSD)
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    Tuning tune;
    Profiler prof;
    // Scratch reused across calls (grow-only).
    DevBuf sort_status, sort_hist, scan_partials, control;
    DevBuf sort_flag;                  // the split sort's overflow flag (gates its fallback passes)
    DevBuf ray_keys[2], ray_order[2];  // sorted trace
    DevBuf tmp_keys, tmp_vals;         // hx_sort_pairs_u64 ping-pong
    DevBuf segments;                   // relabelling: node ranges for the single-workgroup pass
    std::vector<int2> segments_host;   // their host copy, alive until the next relabelling
};

// RAII-free launch scope: records events around one launch when profiling.
struct LaunchScope {
    Context* ctx;
    hipEvent_t a = nullptr, b = nullptr;
    bool on = false;
    LaunchScope(Context* c, const char* name) : ctx(c) {
        if (ctx->prof.enabled) on = ctx->prof.begin(ctx->stream, name, &a, &b) == HX_OK;
    }
    ~LaunchScope() {
        if (on) ctx->prof.end(ctx->stream, b);
    }
};

int check_launch(const char* name);

// ---------------------------------------------------------------------------
// Kernel launchers (one per .hip translation unit).
// ---------------------------------------------------------------------------

// hx_build.hip
int launch_scene_box(Context* c, const float* verts, const int32_t* tris, int64_t n, float* d_box6);
// Scene normalisation in place: the vertex box (tris = NULL in
// launch_scene_box) to d_box6, then v = (v - mn) / largest extent.
int launch_normalize(Context* c, float* verts, int64_t nv, float* d_box6);
// mesh_of / off (batched builds, else NULL): mesh of each triangle (sorted
// position for construct) and the M+1 triangle offsets; d_box6 then holds
// one box per mesh.
// Keys of n triangles, or (ref_min != NULL) of n presplit reference boxes.
// With pack_out (triangles only) each triangle's vertices are also written as
// three float4 in input order (v0.xyz v1.x | v1.yz v2.xy | v2.z), the record
// the climb gathers instead of the index and three vertex reads.
int launch_morton(Context* c, const float* verts, const int32_t* tris, int64_t n, const float* d_box6,
                  int bits, void* d_keys, bool keys64, const int32_t* mesh_of = nullptr,
                  const float4* ref_min = nullptr, const float4* ref_max = nullptr, float4* pack_out = nullptr);
int launch_construct(Context* c, const void* d_keys, bool keys64, int64_t n, int32_t* left, int32_t* right,
                     int32_t* parent, int32_t* first, int32_t* last, int32_t* counters, int32_t* end_at,
                     int32_t* zero_word, const int32_t* mesh_of = nullptr, const int32_t* off = nullptr);
struct ClimbArgs {
    int64_t n;
    const int32_t* sorted_tri;
    const float* verts;
    const int32_t* tris;
    // Written by launch_lbvh_climb, read by launch_climb.
    int32_t* left;
    int32_t* right;
    int32_t* parent;
    int32_t* first;           // 2n-1: range of sorted (or DFS-ordered) leaves under each node
    int32_t* last;
    const int32_t* split;     // n-1, or NULL for a Karras tree: last leaf of each node's left child
    const void* keys;         // launch_lbvh_climb: the sorted Morton keys,
    int keys64;
    uint64_t* link;           // n-1 arrival words of the top level,
    int32_t* root_id;         // and the root's node id (out)
    int all_boxes;            // 1: store every node's box (else only those a later stage reads)
    float4* proto;            // n-1 x 3: Bin rows 0-2 (both child boxes) of every state-1 or root node
    int32_t* counters;
    float4* nmin;
    float4* nmax;
    float* cost;
    int32_t* state;
    int32_t* end_at;
    float4* woop;
    int32_t* frontier;        // n entries: nodes whose parent straddles a climb block
    int32_t* frontier_count;  // zeroed by launch_construct (or launch_refit_reset)
    // ATRBVH preparation outputs (Karras trees only; NULL otherwise): 2n-1
    // areas and binary SAH costs, n-1 subtree sizes and heights.
    float* atr_area;
    float* atr_cost;
    int32_t* atr_refs;
    int32_t* atr_height;
    const int32_t* mesh_of;   // batched builds: mesh of each sorted position (else NULL)
    const int32_t* mesh_off;  // M+1 triangle offsets
    // Presplit builds: sorted_tri holds reference ids; a leaf's box is its
    // reference's box and its Woop row comes from ref_tri[id] (else NULL).
    const float4* ref_min;
    const float4* ref_max;
    const int32_t* ref_tri;
    const float4* tri_pack;   // packed vertices of each triangle (3 float4), or NULL
    int refit_only;           // boxes and Woop only: keep states, costs and cover
    // Refit of a Karras-tree Bin BVH (else NULL): bin_of[node] is the Bin slot
    // of binary node `node` (>= 0 interior, -2 - slot leaf, -1 none); the
    // climb writes rows 0-2 of those Bin nodes as it finishes each node.
    const int32_t* bin_of;
    int4* bin_nodes;
    int adaptive;
    int max_leaf_size;
    float ct, ci;
};
int launch_climb(Context* c, const ClimbArgs& a);
// Binary radix tree of the sorted keys built bottom-up together with the
// climb (boxes, Woop rows, costs, collapse states); internal nodes are
// numbered by split position.  frontier_count must be zero on entry.
int launch_lbvh_climb(Context* c, const ClimbArgs& a);
struct EmitArgs {
    int64_t n;
    const int32_t* left;
    const int32_t* right;
    const int32_t* parent;
    const int32_t* first;
    const int32_t* last;
    const int32_t* state;
    const float4* nmin;
    const float4* nmax;
    int32_t* cover;      // in: end_at, out: prefix max (in place)
    int32_t* leaf_num;   // out
    int32_t* int_slot;   // out
    int32_t* counts;     // out: [n_interior, n_leaves]
    int4* nodes;         // out: Bin nodes
    const int32_t* mesh_of;   // batched builds (else NULL)
    const int32_t* mesh_off;
    int n_meshes;             // 1 unless batched
    int4* mesh_tab;           // out: per mesh {node base, first interior slot, first leaf number, interiors}
    int32_t* root_slot;       // out: per mesh interior slot of the root
    // General trees (PLOC, ATRBVH, HLBVH, insertion): interior slots follow
    // the split positions, so every mesh's interiors stay contiguous;
    // node_at (n-1) receives the node at each split position.  NULL split:
    // a Karras tree, whose node ids already lie in their ranges.
    const int32_t* split;
    int32_t* node_at;
    const int32_t* root_id;   // single-mesh root's node id (NULL: node 0)
    const float4* proto;      // Bin rows 0-2 per internal node from the climb (NULL: from nmin/nmax)
    int32_t* bin_of;          // optional out: per binary node its Bin slot (see ClimbArgs::bin_of)
};
int launch_collapse_emit(Context* c, const EmitArgs& a);
// Refit: re-zero the climb counters and frontier count, then (after the
// climb) rewrite the Bin nodes from the binary boxes with the build's slots.
int launch_refit_reset(Context* c, int32_t* counters, int64_t n_internal, int32_t* frontier_count);
int launch_emit_only(Context* c, const EmitArgs& a);
// Adjoint of closest-hit t/u/v with respect to the triangle vertices.
struct BackwardArgs {
    int64_t n_rays;
    const float* verts;
    const int32_t* tris;
    const float4* rays;
    const int4* results;
    const float* dl_dt;   // per ray, may be NULL
    const float* dl_du;   // per ray, may be NULL
    const float* dl_dv;   // per ray, may be NULL
    float* dl_dverts;     // float[3V], accumulated
    const int32_t* ray_mesh;  // batched: mesh of each ray (else NULL)
    const int32_t* mesh_off;  // batched: M+1 triangle offsets
};
int launch_backward(Context* c, const BackwardArgs& a);
int launch_recompute_hits(Context* c, const BackwardArgs& a, float* t_out, float* u_out, float* v_out);
int launch_single_ref(Context* c, const float* verts, const int32_t* tris, const int32_t* sorted_tri,
                      float4* nmin, float4* nmax, float4* woop, int4* nodes, int32_t* counts);
int launch_cost(Context* c, const int4* nodes, const int32_t* d_counts, int64_t max_nodes,
                const float* d_box6, float ct, float ci, double* d_out);
// Leaf-size histogram (sizes 0..kMaxLeafSizeStat) and sum of leaf sizes into
// d_out[0..kMaxLeafSizeStat+1] (HipBVHKernels.cu:124-182).
constexpr int kMaxLeafSizeStat = 32;  // MAX_LEAF_SIZE, HipBVHKernels.h:13
int launch_leaf_stats(Context* c, const int4* nodes, const int32_t* d_counts, int64_t max_leaves, int32_t* d_out);
// HipBVH::validate checks on the device; work holds max_nodes + n_tri ints;
// *d_flags receives the OR of the failed checks (0 = valid).
int launch_validate(Context* c, const int4* nodes, const int32_t* d_counts, int64_t max_nodes, const int32_t* tri_idx,
                    int64_t n_refs, const float* verts, const int32_t* tris, int64_t n_tri, int32_t* work,
                    int32_t* d_flags);
int launch_fill_i32(Context* c, int32_t* p, int32_t v, int64_t n);

// hx_ploc.hip: PLOC clustering over the sorted leaves, then depth-first
// relabelling (leaf ranks, first/last/split, tri_out in leaf order) and the
// clears the climb expects.  n >= 2.
struct PlocArgs {
    int64_t n;
    const float* verts;
    const int32_t* tris;
    const int32_t* sorted_tri;   // Morton order
    int32_t* tri_out;            // depth-first leaf order (n)
    const float4* ref_min;       // presplit builds: leaf boxes of the reference ids (else NULL)
    const float4* ref_max;
    int32_t* cl[2];              // cluster node ids, ping-pong (n each)
    int32_t* cm[2];              // batched builds: cluster mesh ids, ping-pong (else NULL)
    const int32_t* mesh_of;      // batched builds: mesh of each sorted position
    int n_meshes;
    float4* ca[2];               // cluster boxes: (mn.x, mn.y, mn.z, mx.x)
    float2* cb[2];               //                (mx.y, mx.z)
    int32_t* nb;                 // nearest neighbour per cluster (n)
    int32_t* left;               // 2n-1
    int32_t* right;
    int32_t* parent;
    int32_t* first;
    int32_t* last;
    int32_t* split;              // n-1
    int32_t* counters;           // n-1, cleared for the climb
    int32_t* end_at;             // n, set to -1 for the climb
    int32_t* frontier_count;     // cleared for the climb
    int radius, max_radius, adaptive;
};
int launch_ploc(Context* c, const PlocArgs& a, int* rounds_out);
// The whole top-down (depth-first) relabelling: segs[k] = (start, count) of
// the nodes whose children step k places (ids start..start+count-1, or
// list[start..] if list is given), parents' steps first; uses n, left, right,
// parent, first, last (sizes in, last ranks out), split, sorted_tri and tri_out
// of a.  Leading small steps run in one workgroup.
int launch_tree_down_steps(Context* c, const PlocArgs& a, const int32_t* list, const std::vector<int2>& segs);

// hx_atr.hip: ATRBVH treelet restructuring of a Karras tree prepared by the
// climb (ClimbArgs::atr_*), then the depth-first relabelling.  n >= 2.
struct AtrArgs {
    int64_t n;
    int treelet_size;
    int iterations;
    float ct, ci;
    int32_t* left;
    int32_t* right;
    int32_t* parent;
    float4* nmin;
    float4* nmax;
    float* area;         // 2n-1
    float* cost;         // 2n-1
    int32_t* refs;       // n-1
    int32_t* height;     // n-1
    int32_t* list;       // n-1: internal nodes grouped by height
    int32_t* hist;       // bucket counts / cursors (2 * kAtrMaxLevels + 1 each)
    int32_t* cursor;
    int32_t* accepted;   // 1: treelets whose restructuring was kept
};
constexpr int kAtrMaxLevels = 1024;
// relabel: fields of a PLOC argument block used by launch_tree_down.
int launch_atr(Context* c, const AtrArgs& a, const PlocArgs& relabel, int64_t* accepted_out);
// TR (TRBuilder.cpp:140-191): the same level driver with the optimal 7-leaf
// treelet restructuring; a.treelet_size must be 7.
int launch_tr(Context* c, const AtrArgs& a, const PlocArgs& relabel, int64_t* accepted_out);
// Depth-first relabelling of any binary tree whose internal nodes carry leaf
// counts (a.refs) and heights (a.height): nodes are bucketed by height and
// placed top down.  Uses a.n, refs, height, list, hist, cursor.
int launch_relabel_by_height(Context* c, const AtrArgs& a, const PlocArgs& relabel);

// hx_presplit.hip: budgeted presplitting of n triangles into references.
// Phase 1 (priorities, D per mesh, the reference counts) returns the counts
// so that the caller can size the reference arrays (n_refs each) and the two
// task queues (queue_cap = max(1, n_refs - n) entries each, enough since every
// queued task holds at least one unit of the budget); phase 2 writes them.
// Without mesh_of everything is one mesh framed by box.
struct PsArgs {
    int64_t n;
    const float* verts;
    const int32_t* tris;
    const float* box;           // device scene box (mn.xyz, mx.xyz)
    const int32_t* mesh_of;     // batched: mesh of each triangle (else NULL)
    const float* mesh_box;      // batched: 6 floats per mesh
    float* Dm;                  // per-mesh multiplier (M floats, device)
    float* prio;                // n
    unsigned long long* macc;   // 2 per mesh: priority fixed-point sum, budget sum
    int32_t* partial;           // one entry per 4096 triangles
    float4* ref_min;            // n_refs
    float4* ref_max;
    int32_t* ref_tri;
    int4* task[2];              // (triangle, budget, first slot, 0)
    float4* tmn[2];
    float4* tmx[2];
    int32_t* qcount;            // 2
    int64_t queue_cap;
};
// Host float -> int with the device convert's saturation (NaN -> 0).
inline int32_t f2i_host(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 0x7FFFFFFF;
    if (x <= -2147483648.0f) return (int32_t)0x80000000;
    return (int32_t)x;
}
int launch_presplit_count(Context* c, const PsArgs& a, float beta, int M, const int64_t* tri_counts,
                          int64_t* ref_counts, float* D_out);
int launch_presplit_refs(Context* c, const PsArgs& a, int64_t n_refs);
// ids[p] = ref_tri[ids[p]]: reference ids to triangle ids after a build.
int launch_refs_to_tris(Context* c, int32_t* ids, const int32_t* ref_tri, int64_t n);

// hx_insertion.hip: insertion optimisation of a binary tree with boxes (root
// 0, leaves n-1..2n-2), the reference's step loop run from the host.  On
// return the tree is the best one seen, refitted, with refs and heights of
// its internal nodes.  n >= 2.
constexpr int kInsMaxSteps = 1000;
struct InsArgs {
    int64_t n;
    int mod;
    float ct, ci;
    int32_t* left;
    int32_t* right;
    int32_t* parent;
    float4* nmin;
    float4* nmax;
    int32_t* counters;        // n-1 climb counters
    int32_t* out;             // 2n-1 best new sibling per node
    float* red;               // 2n-1 its area reduction
    unsigned long long* locks;  // 2n-1
    int32_t* best_left;       // 2n-1 each: the best tree so far
    int32_t* best_right;
    int32_t* best_parent;
    unsigned long long* acc;  // [0] cost (2^-32 units), [1] moves found, [2] moves applied
    int32_t* refs;            // n-1 leaf counts (filled by the final refit)
    int32_t* height;          // n-1 heights
    int4* rec;                // 2n-1 packed walk records, 4 int4 each
};
int launch_insertion(Context* c, const InsArgs& a, int64_t* steps_out, int64_t* reinserted_out);
int launch_ins_refit(Context* c, const InsArgs& a, bool sizes);

// hx_hlbvh.hip: clusters of equal top-sah_bits Morton prefixes (Karras
// subtrees), binned-SAH upper levels over them, depth-first relabelling.
// Cluster arrays hold up to 2^sah_bits entries.  n >= 2.
constexpr int kHlbvhMaxClusters = 1 << 15;
struct HlbvhArgs {
    int64_t n;
    int bits, sah_bits, nbins;
    const void* keys;            // sorted Morton keys (u64, or u32 when bits <= 32)
    const int32_t* sorted_tri;
    int32_t* tri_out;            // depth-first leaf order
    int32_t* left;
    int32_t* right;
    int32_t* parent;
    int32_t* first;
    int32_t* last;
    int32_t* split;
    float4* nmin;
    float4* nmax;
    int32_t* cl_of;              // n: cluster of each sorted position
    int32_t* tmp;                // n
    int32_t* ca;                 // clusters: first / last position, root node
    int32_t* cb;
    int32_t* croot;
    int32_t* cl_off;             // depth-first offset of each cluster
    int32_t* cl_par;             // upper node above each cluster
    int4* coord;                 // bin coordinates x, y, z and size
    int32_t* members[2];         // cluster lists of the tasks, ping-pong
    int32_t* upper;              // ids of the Karras nodes spanning clusters
    int32_t* task_size;          // by upper index
    int4* tasks[2];              // {upper index, first member, end member, dfs offset}
    int4* decision;              // per task {axis, bin, left count, -}
    int32_t* child_counts;       // per task
};
int launch_hlbvh(Context* c, const HlbvhArgs& a, bool keys64, int* clusters_out);

// hx_batch.hip: batched-build helpers (mesh of each triangle, per-mesh
// boxes, key gathers, mesh-local triangle ids).
int launch_mesh_of(Context* c, const int32_t* off, int M, int64_t n, int32_t* mesh_of);
int launch_mesh_boxes(Context* c, const float* verts, const int32_t* tris, int64_t n, const int32_t* mesh_of, int M,
                      uint32_t* bits, float* boxes);
int launch_gather_mesh_keys(Context* c, const int32_t* vals, const int32_t* off, int M, int64_t n, uint32_t* keys);
int launch_gather_keys(Context* c, const int32_t* vals, const uint64_t* src, int64_t n, uint64_t* dst);
int launch_tri_local(Context* c, const int32_t* sorted_tri, const int32_t* mesh_of, const int32_t* off, int64_t n,
                     int32_t* tri_local);

// hx_sort.hip
// Number of passes a sort of n keys over bits [0, end_bit) takes (digit width
// is the "sort_digit_bits" tunable; a split sort counts its four high passes
// and the local pass, and its result lands in buffer 1 whether or not its
// gated fallback passes run); the result lands in buffer (passes & 1).
int sort_pass_count(const Context* c, int end_bit, int64_t n, bool keys64);
// Caller-provided sort status memory (look-back words, histograms, tile
// counters) instead of the context's grow-only buffer; no split sort.
struct SortWorkspace {
    uint32_t* status;
    int64_t status_words;
};
// Status words a sort of n keys over [0, end_bit) needs with the context's
// digit width and keys per thread (no split sort).
int sort_status_words_max(const Context* c, int64_t n, int end_bit);
// Stable LSD sort of (key, value) over bits [0, end_bit).  keys/vals ping-pong
// between [0] and [1]; returns in *result_index which buffer holds the output.
// vals_identity: the input values are 0..n-1 and are not read.  allow_split:
// long keys may take the split sort (sort_local_run), which pays off when runs
// of equal high bits are short (Morton codes of geometry), not for ray keys
// whose origins coincide (a pinhole camera's).  ws: status memory of the
// caller (nothing is allocated).
int launch_sort_pairs(Context* c, void* keys[2], int32_t* vals[2], int64_t n, int end_bit, bool keys64,
                      bool vals_identity, int* result_index, bool allow_split = true,
                      const SortWorkspace* ws = nullptr);

// hx_wide.hip: Quad / Oct layouts from a built binary tree (internal nodes
// 0..n-2 with left/right/parent and inclusive leaf ranges first/last, leaves
// n-1+p, boxes of all 2n-1 nodes) and their traversal.
struct WideArgs {
    int64_t n;
    int N;                    // children per node: 4 or 8
    const int32_t* left;
    const int32_t* right;
    const int32_t* parent;
    const int32_t* first;
    const int32_t* last;
    const float4* nmin;
    const float4* nmax;
    const int32_t* root_id;   // device; NULL: node 0
    float ct, ci;
    int max_leaf;             // references a collapsed leaf may hold (INT_MAX: adaptive)
    // Workspace: costs and slot splits (N-1 per binary node), arrival counters.
    float* cost;
    uint16_t* cnt;
    int32_t* counters;
    // Wide nodes in breadth-first order (<= 2n entries): binary node, slots,
    // leaf flag, parent entry, rank among interiors / leaves, first child
    // entry and child count.
    int32_t* e_b;
    int8_t* e_budget;
    uint8_t* e_leaf;
    int32_t* e_parent;
    int32_t* e_rank;
    int32_t* e_first;
    int32_t* e_nchild;
    int2* scan;               // per-level counts, 2n + 1
    int2* scan_part;          // their per-block sums, 2n / 2048 + 2
    int4* nodes;              // out: 2N int4 per node
};
int launch_wide_build(Context* c, const WideArgs& a, int64_t* n_interior, int64_t* n_leaves, int64_t* n_entries);
int launch_wide_emit(Context* c, const WideArgs& a, int64_t n_entries, int64_t n_interior);
int wide_cost_partials(int64_t n_nodes);
int launch_wide_cost(Context* c, int N, const int4* nodes, int64_t n_nodes, float scene_area, float ct, float ci,
                     double* partial, double* out);

// hx_raygen.hip: ray generation, occlusion reconstruction, interpolation.
int launch_init_seeds(Context* c, int64_t n, uint32_t frame, uint32_t* seeds);
int launch_primary_rays(Context* c, int sample_index, const float origin[3], const float m[16], int width, int height,
                        float max_dist, const int32_t* index_to_pixel, float* rays);
// ao: AO rays (tris/verts used, light unused), else shadow rays.
int launch_secondary_rays(Context* c, bool ao, int64_t begin, int64_t size, int n_samples, const uint32_t* seeds,
                          float light_radius, const float light[3], float max_dist, const float* in_rays,
                          const int32_t* in_results, const int32_t* tris, const float* verts, float* out_rays);
struct PathRayArgs {
    int64_t n;
    int russian_roulette;
    uint32_t* seeds;
    const int32_t* in_pixel;
    const int32_t* mat_indices;
    const int32_t* tris;
    const float* normals;
    const void* materials;
    const float* in_rays;
    const int32_t* in_results;
    float* out_rays;
    int32_t* out_pixel;
    float* decreases;
    int32_t* n_out;
};
int launch_path_rays(Context* c, const PathRayArgs& p);
int launch_reconstruct_occlusion(Context* c, int64_t begin, int64_t size, int n_samples, bool replace,
                                 const int32_t* results, const int32_t* index_to_pixel, const int32_t* index_to_slot,
                                 const float* in_pixels, float* out_pixels);
int launch_interpolate(Context* c, int64_t n_static, int64_t n, float t, const float* a, const float* b, float* out,
                       bool normalize);

// hx_noalloc.hip: the fixed per-device scratch of hx_trace_sorted_noalloc.
struct TraceScratch {
    void* keys[2];          // u64 ray keys, ping-pong
    int32_t* order[2];      // sorted ray order, ping-pong
    int32_t* control;       // trace counters (2 int32)
    SortWorkspace sort;     // radix-sort status words
};
// Looks up the scratch array's address on `device` (the current device) once;
// hx_context_create calls it.
int noalloc_setup(int device);
// The scratch of `device` carved for n_rays rays; HX_ERR_INVALID_ARG beyond
// its fixed maximum or before noalloc_setup ran on the device.
int noalloc_trace_scratch(int device, int64_t n_rays, TraceScratch* s);

// hx_trace.hip
struct TraceArgs {
    int64_t n_rays;
    int any_hit;
    const float4* rays;
    int4* results;
    const int4* nodes;
    const float4* woop;
    const int32_t* tri_idx;
    int2* stats;            // optional
    const int32_t* order;   // optional: ray slot -> ray index (sorted trace)
    int32_t* control;       // [0] ray counter, [1] overflow counter
    int fetch_threshold;    // dynamic fetch: busy lanes below which a wave refills
    // Batched trace (else NULL): each ray traverses its own mesh's BVH.
    const int32_t* ray_mesh;
    const int4* mesh_tab;   // per mesh {node base, ...}
    const int32_t* mesh_off;
    int n_meshes;
};
int launch_trace(Context* c, const TraceArgs& a);
int launch_trace_wide(Context* c, int N, const TraceArgs& a);
int launch_ray_keys(Context* c, const float4* rays, int64_t n, int method, int bits, float ray_length,
                    uint64_t* keys);

}  // namespace hx
