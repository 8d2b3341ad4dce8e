/*
 * hx_oracle.h -- CPU restatement of hippie's LBVH build and Bin-layout trace.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle and the CPU
 * baseline; it is linked or loaded only by tests/, by __graft_entry__.smoke()
 * and by bench.py's cpu_baseline leg.  The product (hippie_amd/) never calls it.
 *
 * Every function follows the reference file:line named in its comment, op for
 * op, compiled with -ffp-contract=off.  Where the reference is nondeterministic
 * (atomic slot assignment) the oracle fixes one order, stated per function.
 *
 * Parity status per function (how each is pinned) is listed in hx_oracle.cpp's
 * header and in the tests under tests/test_oracle_*.py.
 */
#ifndef HX_ORACLE_H
#define HX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Number of OpenMP threads the oracle uses (0 = library default). */
void hxo_set_threads(int threads);
int hxo_get_threads(void);

/* Scene box = union of triangle boxes, triangles in order.
 * rt/scene/SceneLoader.cpp:143-153. out = {mn.x,mn.y,mn.z,mx.x,mx.y,mx.z}. */
void hxo_normalize(float* verts, int64_t nv, float box6[6]);
void hxo_scene_box(const float* verts, const int32_t* tris, int64_t n, float out[6]);

/* Per-triangle boxes: rt/bvh/LBVHBuilderKernels.cu:65-98 (setupBoxes).
 * box_min/box_max are float4 arrays (w = 0). */
void hxo_setup_boxes(const float* verts, const int32_t* tris, int64_t n,
                     float* box_min, float* box_max);

/* Morton keys from reference boxes: LBVHBuilderKernels.cu:131-160
 * (computeMortonCodes60) with rt/HipUtil.h:35-106; code >> (60 - bits). */
void hxo_morton_codes(const float* box_min, const float* box_max, int64_t n,
                      const float scene_box[6], int bits, uint64_t* keys);

/* Stable LSD radix sort of (key, value) over bits [0, end_bit):
 * semantics of src/radix_sort/RadixSort.cpp:181-238.  In place. */
void hxo_sort_pairs(uint64_t* keys, int32_t* vals, int64_t n, int end_bit);

/* Karras hierarchy: LBVHBuilderKernels.cu:11-63,226-252 and setupLeaves
 * (:162-195).  Arrays of 2n-1 entries; parent[0] = -1. */
void hxo_construct(const uint64_t* sorted_keys, int64_t n,
                   int32_t* left, int32_t* right, int32_t* parent);

/* Leaf boxes from sorted references + bottom-up refit:
 * LBVHBuilderKernels.cu:162-195 and :254-305.  node_min/node_max: float4[2n-1]. */
void hxo_refit(const int32_t* sorted_vals, const float* ref_min, const float* ref_max,
               int64_t n, const int32_t* left, const int32_t* right, const int32_t* parent,
               float* node_min, float* node_max);

/* Binary -> Bin collapse: rt/bvh/BVHCollapser.cpp:521-675 with
 * BVHCollapserKernels.cu:11-526.  adaptive != 0 uses the SAH rule
 * (computeNodeStatesAdaptive), else max_leaf_size (computeNodeStates).
 * nodes: int32[16 * (2n-1)] (Bin layout, HipBVHNode.h:117-216);
 * tri_out: int32[n].  Returns number of nodes; n_interior/n_leaves set. */
int64_t hxo_collapse(int64_t n, const int32_t* sorted_vals,
                     const int32_t* left, const int32_t* right, const int32_t* parent,
                     const float* node_min, const float* node_max,
                     int adaptive, int max_leaf_size, float ct, float ci,
                     int32_t* nodes, int32_t* tri_out,
                     int64_t* n_interior, int64_t* n_leaves);

/* Woop triangles: rt/bvh/HipBVHKernels.cu:12-65 with util/Math.h:318-369.
 * woop: float[12 * n] (3 x float4 per reference slot). */
void hxo_woopify(const float* verts, const int32_t* tris, const int32_t* tri_idx,
                 int64_t n, float* woop);

/* Refit of the final Bin layout in place (HipBVH::refit, HipBVH.cpp:36-93;
 * refitLeaves / refitInteriors, HipBVHKernels.cu:240-326). */
void hxo_refit_final(int32_t* nodes, int64_t n_interior, int64_t n_leaves, const int32_t* tri_idx,
                     const float* verts, const int32_t* tris);

/* SAH cost: HipBVHKernels.cu:184-238 per-node terms in FP32, summed in FP64
 * in node order (the reference sums with a float atomicAdd). */
double hxo_cost(const int32_t* nodes, int64_t n_nodes, float scene_area, float ct, float ci);

/* Trace: rt/tracer/HipTracerBinKernels.cu:19-286 for one lane executing alone
 * (every postponed leaf processed at once).  rays: float4[2R];
 * results: int32[4R] (miss writes only id=-1, t=tmax); stats: int32[2R] or NULL.
 * Closest-hit ties (equal t) go to the smaller original triangle id.
 * Returns the number of rays whose traversal stack overflowed. */
int64_t hxo_trace(const int32_t* nodes, const float* woop, const int32_t* tri_idx,
                  const float* rays, int64_t n_rays, int any_hit,
                  int32_t* results, int32_t* stats);

/* Brute force closest hit over all triangles with the same Woop test and the
 * same tie rule; independent of any BVH.  results: int32[4R]. */
void hxo_trace_brute(const float* verts, const int32_t* tris, int64_t n,
                     const float* rays, int64_t n_rays, int32_t* results);

/* Canonical form of a Bin BVH: DFS from node 0, child 0 first.  Interior:
 * [0, 12 child-box words]; leaf: [1, 6 box words, size, sorted triangle ids].
 * Call with out == NULL to get the length. */
int64_t hxo_bin_canonical(const int32_t* nodes, int64_t n_interior, int64_t n_leaves,
                          const int32_t* tri_idx, uint32_t* out, int64_t capacity);

/* Structural checks in the spirit of HipBVH::validate (HipBVH.cpp:157-307).
 * Returns 0 if valid, else a code naming the first failed check. */
int hxo_validate_bin(const int32_t* nodes, int64_t n_interior, int64_t n_leaves,
                     const int32_t* tri_idx, int64_t n_refs,
                     const float* verts, const int32_t* tris);
/* As hxo_validate_bin for presplit references (n_refs >= n_tris): a leaf box
 * only has to meet its triangles' boxes, and every triangle is referenced at
 * least once. */
int hxo_validate_bin_refs(const int32_t* nodes, int64_t n_interior, int64_t n_leaves, const int32_t* tri_idx,
                          int64_t n_refs, int64_t n_tris, const float* verts, const int32_t* tris);

/* The presplitter's x^y (double-precision series, see hx_oracle.cpp). */
float hxo_ps_pow(float x, float y);
/* One presplitter split (PresplitterKernels.cu:189-419). */
void hxo_ps_split(const float* v, const float* box, int32_t count, const float* sb, float* lbox, float* rbox,
                  int32_t* lcount, int32_t* rcount);

/* Presplitter (Presplitter.cpp:124-326, PresplitterKernels.cu:11-419): returns
 * the number of references; writes them (box corners as float4, triangle id)
 * when ref_tri is given and the count fits cap.  D_out, priorities optional. */
int64_t hxo_presplit(const float* verts, const int32_t* tris, int64_t n, const float* scene_box, float beta,
                     float* ref_min, float* ref_max, int32_t* ref_tri, int64_t cap, float* D_out,
                     float* priorities);

/* Full LBVH build (setup boxes, Morton, sort, construct, refit, collapse,
 * woopify) with per-stage wall-clock seconds in stage_seconds[8].
 * scene_box may be NULL (then computed by hxo_scene_box, timed as stage 0). */
int64_t hxo_build_lbvh(const float* verts, const int32_t* tris, int64_t n_tri,
                       const float* scene_box, int bits, int adaptive, int max_leaf_size,
                       float ct, float ci,
                       int32_t* nodes, int32_t* tri_out, float* woop,
                       int64_t* n_interior, int64_t* n_leaves, double stage_seconds[8], float presplit_beta);

/* PLOC clustering (PLOCBuilder.cpp:21-159, PLOCBuilderKernels.cu:116-389) over
 * the sorted leaves whose boxes are in nmin/nmax[n-1+i]; fills left/right/
 * parent (2n-1) and the internal boxes.  Nodes a round creates are numbered
 * in cluster order from (clusters - 2) down.  adaptive: the radius grows by
 * one before every round up to max_radius.  Returns the number of rounds, or
 * -1 if a round merged nothing (no finite union area). */
int hxo_ploc(int64_t n, int radius, int max_radius, int adaptive, int32_t* left, int32_t* right, int32_t* parent,
             float* nmin, float* nmax);

/* PLOCBuilder::build (PLOCBuilder.cpp:161-208). */
int64_t hxo_build_ploc(const float* verts, const int32_t* tris, int64_t n_tri, const float* scene_box, int bits,
                       int radius, int max_radius, int ploc_adaptive, int adaptive, int max_leaf_size, float ct,
                       float ci, int32_t* nodes, int32_t* tri_out, float* woop, int64_t* n_interior,
                       int64_t* n_leaves, double stage_seconds[8], float presplit_beta);

/* ATRBVH optimisation (ATRBuilder.cpp:23-177, ATRBuilderKernels.cu:11-577) of
 * the binary tree in place: left/right/parent and the boxes of restructured
 * nodes change.  The reference's warp-32 tie rules are emulated.  Returns the
 * number of treelets whose restructuring was kept. */
int64_t hxo_atr(int64_t n, int treelet_size, int iterations, float ct, float ci, int32_t* left, int32_t* right,
                int32_t* parent, float* nmin, float* nmax);

/* ATRBuilder::build (ATRBuilder.cpp:179-236). */
int64_t hxo_build_atr(const float* verts, const int32_t* tris, int64_t n_tri, const float* scene_box, int bits,
                      int treelet_size, int iterations, int adaptive, int max_leaf_size, float ct, float ci,
                      int32_t* nodes, int32_t* tri_out, float* woop, int64_t* n_interior, int64_t* n_leaves,
                      int64_t* accepted, float presplit_beta);

/* TR optimisation (TRBuilder.cpp:140-191, TRBuilderKernels.cu:11-476) of the
 * binary tree in place with treelets of 7 leaves: the optimal restructuring of
 * each treelet by the subset DP, the reference's warp-32 tie rules emulated.
 * Returns the number of treelets whose restructuring was kept. */
int64_t hxo_tr(int64_t n, int iterations, float ct, float ci, int32_t* left, int32_t* right, int32_t* parent,
               float* nmin, float* nmax);

/* TRBuilder::build (TRBuilder.cpp:193-245). */
int64_t hxo_build_tr(const float* verts, const int32_t* tris, int64_t n_tri, const float* scene_box, int bits,
                     int iterations, int adaptive, int max_leaf_size, float ct, float ci, int32_t* nodes,
                     int32_t* tri_out, float* woop, int64_t* n_interior, int64_t* n_leaves, int64_t* accepted);

/* SBVHBuilder (SBVHBuilder.cpp:14-401) + BVHConverter::convertAdaptive: the
 * spatial-split BVH of n_tri triangles as Bin nodes (interiors first, depth
 * first), reference slots (triangle ids, duplicates from spatial splits) and
 * Woop rows.  Returns the node count, -1 when node_cap or ref_cap is too
 * small, -2 for fewer than two triangles. */
int64_t hxo_build_sbvh(const float* verts, const int32_t* tris, int64_t n_tri, float alpha, int max_leaf_size, float ct,
                       float ci, int32_t* nodes_out, int64_t node_cap, int32_t* tri_out, int64_t ref_cap, float* woop,
                       int64_t* n_interior, int64_t* n_leaves, int64_t* n_refs);

/* HLBVH upper levels over a Karras tree (left/right/parent, boxes 2n-1) of the
 * sorted keys, in place; returns the number of clusters (1: unchanged). */
int64_t hxo_hlbvh(int64_t n, int bits, int sah_bits, const uint64_t* sorted_keys, int32_t* left, int32_t* right,
                  int32_t* parent, float* nmin, float* nmax);

/* HLBVHBuilder::build (HLBVHBuilder.cpp:61-437, HLBVHBuilderKernels.cu:35-704):
 * LBVH below clusters sharing the top sah_bits Morton bits, binned SAH above
 * them.  *clusters receives the number of clusters. */
int64_t hxo_build_hlbvh(const float* verts, const int32_t* tris, int64_t n_tri, const float* scene_box, int bits,
                        int sah_bits, int adaptive, int max_leaf_size, float ct, float ci, int32_t* nodes,
                        int32_t* tri_out, float* woop, int64_t* n_interior, int64_t* n_leaves, int64_t* clusters, float presplit_beta);

/* findBestNode (InsertionBuilderKernels.cu:11-131) for every node 1..2n-2 of a
 * binary tree: out[i] the best new sibling (-1 if none), red[i] its area
 * reduction. */
void hxo_ins_find(int64_t n, const int32_t* left, const int32_t* right, const int32_t* parent, const float* nmin,
                  const float* nmax, int32_t* out, float* red);

/* Insertion optimisation (InsertionBuilder.cpp:22-274, InsertionBuilderKernels.cu
 * :11-515) of a binary tree in place; returns the number of reinsertions,
 * *steps_out the number of steps. */
int64_t hxo_insertion(int64_t n, int mod, float ct, float ci, int32_t* left, int32_t* right, int32_t* parent,
                      float* nmin, float* nmax, int64_t* steps_out);

/* InsertionBuilder::build with the LBVH start (InsertionBuilder.cpp:400-440). */
int64_t hxo_build_insertion(const float* verts, const int32_t* tris, int64_t n_tri, const float* scene_box, int bits,
                            int mod, int adaptive, int max_leaf_size, float ct, float ci, int32_t* nodes,
                            int32_t* tri_out, float* woop, int64_t* n_interior, int64_t* n_leaves,
                            int64_t* steps_out, float presplit_beta);

/* Ray Morton keys: rt/ray/RayBufferKernels.cu:556-614 (Aila64) and
 * :492-554 (TwoPoint64) with the ray box fixed to [0,1]^3
 * (rt/ray/RayBuffer.cpp:63-66). */
void hxo_ray_keys_aila64(const float* rays, int64_t n_rays, int bits, uint64_t* keys);
void hxo_ray_keys_twopoint64(const float* rays, int64_t n_rays, int bits,
                             float ray_length, uint64_t* keys);

#ifdef __cplusplus
}
#endif

#endif
