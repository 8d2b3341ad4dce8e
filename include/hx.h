/*
 * hx.h -- C ABI of libhx, the MI355X (gfx950) BVH construction and ray
 * traversal core.  Plain C: no HIP or torch types in any signature; device
 * pointers are passed as plain pointers, sizes as int64_t.
 *
 * Every entry point returns an int status (HX_OK == 0) and never exits the
 * process.  hx_last_error() gives a message for the calling thread.
 *
 * Reference interfaces replaced (paths relative to src/hippie/ of hippie):
 *   BVHBuilder::build / rebuild            rt/bvh/BVHBuilder.h:37-39
 *   LBVHBuilder::build / rebuild           rt/bvh/LBVHBuilder.cpp:183-263
 *   HipBVH (nodes, Woop triangles, triangle indices, getCost, update)
 *                                          rt/bvh/HipBVH.h:19-106
 *   HipTracer::setBVH / trace / traceStats / traceSort
 *                                          rt/tracer/HipTracer.h:16-44
 *   TRACE_FUNC kernel ABI                  rt/tracer/HipTracerKernels.h:29-39
 *   RadixSort::sort                        radix_sort/RadixSort.h:64-87
 *   RayBuffer::mortonSort / computeMortonCodes
 *                                          rt/ray/RayBuffer.cpp:31-131
 *
 * Conventions (differences from the reference are deliberate):
 *   - one context per host thread; calls are asynchronous on the context's
 *     stream unless they return host data (counts, cost, timings);
 *   - geometry, rays and results are caller-owned device memory; BVH storage
 *     (nodes, Woop triangles, triangle indices) is owned by the hx_bvh;
 *   - sizes are 64-bit; the number of references per build is < 2^30.
 *
 * Memory layouts (identical to the reference):
 *   vertices   float[3*V]   (Vec3f)
 *   triangles  int32[3*F]   (Vec3i)
 *   Bin node   64 bytes     int32[16]: words 0-11 child boxes
 *              (c0.lo.x,c0.hi.x,c0.lo.y,c0.hi.y, c1.lo.x,c1.hi.x,c1.lo.y,c1.hi.y,
 *               c0.lo.z,c0.hi.z,c1.lo.z,c1.hi.z), 12/13 child indices (leaf
 *              child = ~leafNodeIndex) or, in a leaf, begin/end, 14 size (leaf:
 *              ~size), 15 parent.  Interior nodes first (root = 0), then leaves.
 *              (rt/bvh/HipBVHNode.h:117-216)
 *   Woop       48 bytes per reference slot: 3 x float4 (HipBVHKernels.cu:59-61)
 *   ray        32 bytes: float4 (origin, tmin), float4 (direction, tmax)
 *   result     16 bytes: int4 (triangle id, t bits, u bits, v bits); a miss
 *              writes only (-1, tmax) to the first 8 bytes (HipTracerKernels.h:109)
 *   stats      8 bytes: int2 (traversed nodes, tested triangles)
 */
#ifndef HX_H
#define HX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define HX_OK 0
#define HX_ERR_INVALID_ARG 1
#define HX_ERR_HIP 2
#define HX_ERR_OUT_OF_MEMORY 3
#define HX_ERR_UNSUPPORTED 4
#define HX_ERR_STACK_OVERFLOW 5

/* ---- builder methods: the Bvh.method strings of Benchmark.cpp:567-589 ---- */
#define HX_METHOD_LBVH 0
#define HX_METHOD_HLBVH 1
#define HX_METHOD_PLOC 2
#define HX_METHOD_ATRBVH 3
#define HX_METHOD_INSERTION 4
#define HX_METHOD_TR 5         /* "tr": TRBuilder, treelets of 7 (TRBuilder.cpp:140-347) */

#define HX_LAYOUT_BIN 2
#define HX_LAYOUT_QUAD 4   /* HipBVHNodeQuad, 128 B (HipBVHNode.h:13-114), collapsed as BVH::collapseWide */
#define HX_LAYOUT_OCT 8    /* HipBVHNodeOct, 256 B */

/* Ray key methods: RayBuffer::MortonCodeMethod (rt/ray/RayBuffer.h). */
#define HX_RAYKEY_TWOPOINT 0
#define HX_RAYKEY_AILA 1

typedef struct hx_context hx_context;
typedef struct hx_bvh hx_bvh;

/* Builder parameters.  Field names follow the Bvh.* env keys of
 * environment/AppEnvironment.cpp:52-75; defaults from hx_build_params_default
 * equal the registered defaults there. */
typedef struct hx_build_params {
    int32_t method;              /* Bvh.method                                */
    int32_t morton_bits;         /* Bvh.{lbvh,ploc,atr,tr,...}MortonCodeBits [6,60] */
    int32_t max_leaf_size;       /* Bvh.maxLeafSize (0,1024]                  */
    int32_t adaptive_leaf_size;  /* Bvh.adaptiveLeafSize                      */
    float ct;                    /* Bvh.ct                                    */
    float ci;                    /* Bvh.ci                                    */
    int32_t presplitting;        /* Bvh.presplitting                          */
    float presplitter_beta;      /* Bvh.presplitterBeta                       */
    int32_t ploc_radius;         /* Bvh.plocRadius                            */
    int32_t ploc_max_radius;     /* Bvh.plocMaxRadius                         */
    int32_t ploc_adaptive;       /* Bvh.plocAdaptive                          */
    int32_t atr_iterations;      /* Bvh.atrIterations [0,25]                  */
    int32_t atr_treelet_size;    /* ATRBuilder::setTreeletSize [4,30]         */
    int32_t tr_iterations;       /* Bvh.trIterations [0,25]                   */
    int32_t hlbvh_sah_bits;      /* Bvh.hlbvhMortonCodeSAHBits                */
    int32_t insertion_mod;       /* Bvh.insertionMod                          */
    int32_t layout;              /* Bvh.layout: HX_LAYOUT_BIN / _QUAD / _OCT  */
    int32_t use_scene_box;       /* 1: scene_box below is the Morton frame    */
    float scene_box[6];          /* mn.xyz, mx.xyz                            */
    int32_t keep_debug;          /* 1: keep keys, binary arrays for hx_bvh_debug */
} hx_build_params;

/* Device pointers and counts of a built BVH (HipBVH accessors). */
typedef struct hx_bvh_view {
    void* nodes;             /* Bin nodes, 64 B each (Quad 128 B, Oct 256 B)  */
    void* woop;              /* 48 B per reference slot                       */
    int32_t* tri_indices;    /* reference slot -> triangle id                 */
    int64_t n_interior;
    int64_t n_leaves;
    int64_t n_references;
    float scene_box[6];
    int32_t layout;          /* HX_LAYOUT_BIN, HX_LAYOUT_QUAD or HX_LAYOUT_OCT */
} hx_bvh_view;

/* Intermediate arrays of the last build (requires keep_debug), for parity
 * checks against the reference's own buffers (LBVHBuilder.h members). */
typedef struct hx_bvh_debug_view {
    void* sorted_keys;       /* uint64 or uint32 per sorted reference          */
    int32_t keys_are_64bit;
    int32_t* sorted_tri;     /* triangle id per sorted position                */
    int32_t* left;           /* children of the n-1 internal nodes: leaves are  */
    int32_t* right;          /* n-1+p; LBVH builds number internal nodes by     */
                             /* split position (the Karras node whose left      */
                             /* child index is s), the other builders' binary   */
                             /* start trees by Karras index                     */
    int32_t* parent;         /* 2n-1 (leaves at n-1..2n-2)                      */
    float* node_min;         /* float4[2n-1]                                   */
    float* node_max;
    int32_t* state;          /* n-1 collapse states of internal nodes (0 leaf) */
} hx_bvh_debug_view;

/* ---- errors --------------------------------------------------------------- */
const char* hx_status_string(int status);
const char* hx_last_error(void);

/* ---- context (HipModule::staticInit, HipModule.cpp:121-149) ------------------ */
int hx_context_create(int device, hx_context** out);
int hx_context_destroy(hx_context* ctx);
/* Use an existing hipStream_t (passed as void*); NULL = the null stream. */
int hx_context_set_stream(hx_context* ctx, void* stream);
int hx_context_synchronize(hx_context* ctx);
/* Per-launch event timing (HipKernel::launchTimed, HipKernel.cpp:296-325, but
 * with no synchronisation between launches). */
int hx_context_set_profiling(hx_context* ctx, int enabled);
int hx_context_profile_count(hx_context* ctx, int32_t* count);
int hx_context_profile_entry(hx_context* ctx, int32_t index, char* name, int32_t name_len,
                             int64_t* launches, double* total_ms);
int hx_context_profile_reset(hx_context* ctx);
/* Launch-shape tunables chosen by the autotuner (no reference counterpart;
 * the reference hard-codes them, HipTracerBinKernels.cu:5-6,10-15):
 *   "trace_lds_stack"      traversal-stack entries held in LDS (0,16,24,32,64)
 *   "trace_fetch"          refill a wave when fewer of its 64 lanes are busy
 *   "trace_blocks_per_cu"  persistent trace workgroups per CU (0 = occupancy)
 *   "trace_postpone"       leaves a lane holds back while its wave traverses (1 to 4; Bin layout,
 *                          closest hit; any-hit traces keep 1)
 *   "trace_woop_ahead"     Woop rows of a leaf's next triangle read while the current one is
 *                          tested (1 or 2; Bin layout)
 *   "sort_digit_bits"      radix digit width of every sort (8 or 10)
 *   "sort_items"           keys per thread in a sort pass (8, 16 or 32)
 *   "sort_lookback"        tiles read per sort look-back round trip (4, 8, 16)
 *   "sort_local_run"       the builds' Morton sorts and hx_sort_pairs_u64 of
 *                          > 40-bit keys with at least 2^18 keys run four
 *                          digit passes over the high bits, then rank each run
 *                          of equal high bits in place when no run is longer
 *                          than this (else the remaining passes run); 0 =
 *                          every pass (ray sorts always).  Same result. */
int hx_context_set_tuning(hx_context* ctx, const char* name, int32_t value);
int hx_context_get_tuning(hx_context* ctx, const char* name, int32_t* value);

/* ---- builders (BVHBuilder.h:37-39) ---------------------------------------- */
int hx_build_params_default(hx_build_params* p);
/* BVHBuilder::build: allocates a new BVH owned by the caller (hx_bvh_destroy). */
int hx_build(hx_context* ctx, const hx_build_params* params, const float* d_vertices, int64_t n_vertices,
             const int32_t* d_triangles, int64_t n_triangles, hx_bvh** out);
/* BVHBuilder::rebuild: rebuild in place from (possibly moved) geometry with
 * the vertex and triangle counts of the first build (else
 * HX_ERR_INVALID_ARG: the BVH's buffers are sized for them). */
int hx_rebuild(hx_context* ctx, hx_bvh* bvh, const float* d_vertices, int64_t n_vertices, const int32_t* d_triangles,
               int64_t n_triangles);
/* HipBVH::update (HipBVH.cpp:472-477; refitLeaves/refitInteriors,
 * HipBVHKernels.cu:240-326): refit the final layout and re-woopify from new
 * vertex positions, topology unchanged; the counts must equal the build's. */
int hx_refit(hx_context* ctx, hx_bvh* bvh, const float* d_vertices, int64_t n_vertices, const int32_t* d_triangles,
             int64_t n_triangles);
int hx_bvh_destroy(hx_bvh* bvh);
/* Synchronises on first call after a build to read the node counts. */
int hx_bvh_get_view(hx_bvh* bvh, hx_bvh_view* out);
int hx_bvh_debug(hx_bvh* bvh, hx_bvh_debug_view* out);
/* Statistics of the last build (no reference counterpart; the reference
 * logs them, PLOCBuilder.cpp:148, ATRBuilder.cpp:166): "ploc_rounds",
 * "atr_accepted" (treelets kept), "method". */
int hx_bvh_info(hx_bvh* bvh, const char* key, int64_t* value);
/* HipBVH::getCost (HipBVH.cpp:346-370): SAH cost, FP64 deterministic sum. */
int hx_bvh_cost(hx_bvh* bvh, double* out);
/* HipBVH::getAvgLeafSize / getLeafSizeHistogram (HipBVH.cpp:380-442): counts
 * of leaves of each size 0..32 (host int32[33]) and the sum of all leaf sizes;
 * divide by the leaf count for the reference's fractions.  Synchronises. */
int hx_bvh_leaf_stats(hx_bvh* bvh, int32_t* histogram33, int64_t* sum_of_leaf_sizes);
/* HipBVH::validate (HipBVH.cpp:157-310) on the device against the geometry of
 * the last build or refit: *failed = 0 if valid, else the OR of 1 child index,
 * 2 parent, 4 size, 8 leaf range, 16 leaf box misses a triangle, 32 triangle
 * index, 64 a node referenced other than once, 128 an unreferenced triangle.
 * Synchronises. */
int hx_bvh_validate(hx_bvh* bvh, int32_t* failed);

/* ---- stages (the reference's individually launched kernels) --------------- */
/* Scene normalisation at load (SceneLoader::loadStaticScene,
 * SceneLoader.cpp:134-141): in place, v = (v - mn) * (1 / max extent) over the
 * box of all n_vertices vertices (AssimpImporter.cpp:99-113); d_box6 (device,
 * may be NULL) receives that box. */
int hx_normalize_vertices(hx_context* ctx, float* d_vertices, int64_t n_vertices, float* d_box6);
/* ---- the steps either side of a trace ----------------------------------------
 * Rays are float4 (origin, tmin) + float4 (direction, tmax); results int4
 * (id, t, u, v); pixels float4.  Every buffer is caller-owned device memory. */
/* initSeeds (RayGenKernels.cu:520-535): seeds[i] = tea<16>(i, frame_index). */
int hx_init_seeds(hx_context* ctx, int64_t n, int32_t frame_index, uint32_t* d_seeds);
/* generatePrimaryRays (RayGenKernels.cu:106-153; RayGen::primary, RayGen.cpp:35-63):
 * pixel p = index_to_pixel[i] (NULL: p = i) of a width x height image, offset
 * by halton2D(sample_index), through screen_to_world (host, row-major 4x4:
 * element (r, c) at 4r + c, the inverse of the camera's projection * view);
 * origin (host, 3 floats), tmin 0, tmax max_dist.  Writes width*height rays. */
int hx_primary_rays(hx_context* ctx, int32_t sample_index, const float* origin3, const float* screen_to_world16,
                    int32_t width, int32_t height, float max_dist, const int32_t* d_index_to_pixel, float* d_rays);
/* generateShadowRays (RayGenKernels.cu:155-229): n_samples rays per input ray
 * batch_begin + i (i < batch_size) towards a light of radius light_radius
 * around light3 (host), Sobol/Hammersley points with a Cranley-Patterson shift
 * hashed from seeds[i]; written at n_samples * i + s.  A missed input gives
 * rays with tmax -1. */
int hx_shadow_rays(hx_context* ctx, int64_t batch_begin, int64_t batch_size, int32_t n_samples,
                   const uint32_t* d_seeds, float light_radius, const float* light3, const float* d_in_rays,
                   const int32_t* d_in_results, float* d_out_rays);
/* generateAORays (RayGenKernels.cu:232-349): n_samples cosine-weighted Halton
 * directions about the hit triangle's normal (flipped to face the ray), rotated
 * by an angle hashed from seeds[i], from 1e-4 before the hit, tmax max_dist
 * (-1 for a missed input). */
int hx_ao_rays(hx_context* ctx, int64_t batch_begin, int64_t batch_size, int32_t n_samples, const uint32_t* d_seeds,
               float max_dist, const float* d_in_rays, const int32_t* d_in_results, const int32_t* d_triangles,
               const float* d_vertices, float* d_out_rays);
/* generatePathRays (RayGenKernels.cu:354-519): one continuation per hit path
 * (glossy or cosine lobe by the material's 32-byte record {int tex_index;
 * float shininess; float diffuse[3]; float specular[3]}, smooth normals,
 * LCG draws from seeds, which advance), written in input order (the reference
 * appends in warp-atomic order); decreases (float3 per pixel) *= the path
 * weight; *d_n_out = number of output rays.  Texture lookups are not part of
 * this library: a material's diffuse colour is used as is. */
int hx_path_rays(hx_context* ctx, int32_t russian_roulette, int64_t n_in, uint32_t* d_seeds,
                 const int32_t* d_in_index_to_pixel, const int32_t* d_mat_indices, const int32_t* d_triangles,
                 const float* d_normals, const void* d_materials, const float* d_in_rays, const int32_t* d_in_results,
                 float* d_out_rays, int32_t* d_out_index_to_pixel, float* d_decreases, int32_t* d_n_out);
/* reconstructAO / reconstructShadow (RendererKernels.cu:302-372): pixel
 * index_to_pixel[batch_begin + i] gets in_pixels' colour times the fraction of
 * its n_samples results (at index_to_slot[n_samples * i + s], NULL: identity)
 * that miss; replace != 0 stores it, else adds it (w included). */
int hx_reconstruct_occlusion(hx_context* ctx, int64_t batch_begin, int64_t batch_size, int32_t n_samples,
                             int32_t replace, const int32_t* d_results, const int32_t* d_index_to_pixel,
                             const int32_t* d_index_to_slot, const float* d_in_pixels, float* d_out_pixels);
/* interpolateVertices / interpolateNormals (InterpolatorKernels.cu:10-46;
 * Interpolator::updateAdaptive, Interpolator.cpp:21-95): out[n_static + i] =
 * frame0[i] (1 - t) + frame1[i] t for i < n_dynamic (float3), normalised when
 * normalize != 0; the refit input of an animated scene. */
int hx_interpolate_vertices(hx_context* ctx, int64_t n_static, int64_t n_dynamic, float t, const float* d_frame0,
                            const float* d_frame1, float* d_out, int32_t normalize);
/* Scene box: exact min/max over the triangles' vertices (SceneLoader.cpp:143-153).
 * d_box6 receives mn.xyz, mx.xyz on the device. */
int hx_scene_box(hx_context* ctx, const float* d_vertices, const int32_t* d_triangles, int64_t n,
                 float* d_box6);
/* computeMortonCodes60 (LBVHBuilderKernels.cu:131-160): code >> (60 - bits),
 * uint64 per triangle in triangle order.  d_box6: scene box on the device. */
int hx_morton_codes(hx_context* ctx, const float* d_vertices, const int32_t* d_triangles, int64_t n,
                    const float* d_box6, int32_t bits, uint64_t* d_keys);
/* RadixSort::sort (RadixSort.cpp:181-238): stable sort of (uint64 key, int32
 * value) over bits [0, end_bit), in place. */
int hx_sort_pairs_u64(hx_context* ctx, uint64_t* d_keys, int32_t* d_values, int64_t n, int32_t end_bit);

/* ---- trace (HipTracer.h:16-44, TRACE_FUNC) ---------------------------------- */
int hx_trace(hx_context* ctx, hx_bvh* bvh, const float* d_rays, int64_t n_rays, int32_t any_hit,
             int32_t* d_results, int32_t* d_stats);
/* HipTracer::traceSort (HipTracer.cpp:114-121): ray keys (method, bits),
 * radix sort, then trace in sorted order writing results at the original
 * index.  d_order (optional, int32[n_rays]) receives the sorted ray order. */
int hx_trace_sorted(hx_context* ctx, hx_bvh* bvh, const float* d_rays, int64_t n_rays, int32_t any_hit,
                    int32_t key_method, int32_t key_bits, float ray_length, int32_t* d_results,
                    int32_t* d_stats, int32_t* d_order);
/* HipTracer::traceSort(rays, sortTime, traceTime) (HipTracer.h:41,
 * HipTracer.cpp:114-121): hx_trace_sorted, then waits for it and returns the
 * seconds spent in the ray keys + sort and in the trace, from events recorded
 * on the context's stream around each half. */
int hx_trace_sorted_timed(hx_context* ctx, hx_bvh* bvh, const float* d_rays, int64_t n_rays, int32_t any_hit,
                          int32_t key_method, int32_t key_bits, float ray_length, int32_t* d_results,
                          int32_t* d_stats, int32_t* d_order, float* sort_seconds, float* trace_seconds);
/* The kernel ABI TRACE_FUNC (HipTracerKernels.h:29-39: numRays, anyHit, rays,
 * results, nodes, tris, stats, indices, triIndices) over caller-held Bin BVH
 * arrays, e.g. a HipBVH's buffers built by other code: d_nodes (n_nodes Bin
 * nodes, root at 0), d_woop (48 B per reference slot), d_tri_indices (slot ->
 * triangle id, n_references entries).  The arrays must form a valid Bin BVH
 * (hx_bvh_validate's checks); they are not checked on the device.  d_indices
 * (optional, int32[n_rays]) is traceSort's ray order: slot i traces ray
 * d_indices[i] and writes its result (and stats) at that index. */
int hx_trace_view(hx_context* ctx, const void* d_nodes, int64_t n_nodes, const void* d_woop,
                  const int32_t* d_tri_indices, int64_t n_references, const float* d_rays, int64_t n_rays,
                  int32_t any_hit, int32_t* d_results, int32_t* d_stats, const int32_t* d_indices);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc): hx_trace_sorted with its ray keys, sorted order,
 * radix-sort status and trace counters in a fixed 32 MiB + 64 KiB array per
 * GPU (held from the first context on that GPU until the process ends), so no
 * call allocates.  At most HX_NOALLOC_MAX_RAYS rays per call
 * (HX_ERR_INVALID_ARG beyond).  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (several streams or host threads) race.  Its
 * stack-overflow count is not reported by hx_trace_overflow_count. */
#define HX_NOALLOC_MAX_RAYS (1 << 20)
int hx_trace_sorted_noalloc(hx_context* ctx, hx_bvh* bvh, const float* d_rays, int64_t n_rays, int32_t any_hit,
                            int32_t key_method, int32_t key_bits, float ray_length, int32_t* d_results,
                            int32_t* d_stats, int32_t* d_order);
/* RayBuffer::computeMortonCodes (RayBuffer.cpp:31-85) for the [0,1]^3 ray box. */
int hx_ray_keys(hx_context* ctx, const float* d_rays, int64_t n_rays, int32_t key_method, int32_t key_bits,
                float ray_length, uint64_t* d_keys);
/* ---- differentiable closest hit (no reference counterpart) ------------------
 * Adjoint of the closest-hit (t, u, v) of each ray with respect to the vertex
 * positions, for the triangle ids in d_results (as written by hx_trace).  The
 * hit is recomputed from the id; upstream gradients dL/dt, dL/du, dL/dv are
 * float[n_rays] each (any may be NULL); the vertex gradients are ADDED into
 * d_dl_dvertices (float[3V]; the caller zeroes it).  Topology stays discrete:
 * the hit triangle is fixed, as is the BVH between refits. */
int hx_backward(hx_context* ctx, const float* d_vertices, const int32_t* d_triangles, const float* d_rays,
                const int32_t* d_results, int64_t n_rays, const float* d_dl_dt, const float* d_dl_du,
                const float* d_dl_dv, float* d_dl_dvertices);
/* Forward companion of hx_backward: (t, u, v) of each hit recomputed from its
 * triangle id by the function hx_backward differentiates (solved in FP64,
 * rounded once); misses get t = tmax and u = v = 0.  Outputs float[n_rays]. */
int hx_recompute_hits(hx_context* ctx, const float* d_vertices, const int32_t* d_triangles, const float* d_rays,
                      const int32_t* d_results, int64_t n_rays, float* d_t, float* d_u, float* d_v);

/* Synchronous copy between any two host/device addresses on the context's
 * stream (Buffer migration, gpu/Buffer.cpp:437-498). */
int hx_memcpy(hx_context* ctx, void* dst, const void* src, int64_t bytes);

/* Number of rays whose traversal stack overflowed in the last trace call. */
int hx_trace_overflow_count(hx_context* ctx, int64_t* out);

/* ---- batched meshes (no reference counterpart) -----------------------------
 * Many independent meshes built, refit, traced and differentiated as one:
 * the callers are workloads with thousands of small scenes (per-object BVHs,
 * differentiable rendering of many meshes), which the reference serves one
 * HipBVH at a time.  Triangles of mesh m are [off[m], off[m+1]) of
 * d_triangles (vertex indices into the shared d_vertices); each mesh gets the
 * BVH a separate build of it would give (params->method LBVH or PLOC: same
 * Bin tree, Woop rows and mesh-local triangle ids).  Every mesh needs >= 2 triangles.  With
 * params->presplitting each mesh is presplit alone (its own box and budget
 * beta * its triangle count), so a mesh has more references than triangles. */
typedef struct hx_batch hx_batch;
typedef struct hx_batch_mesh_view {
    const int32_t* nodes;        /* device: this mesh's Bin nodes, 16 int32 each  */
    const float* woop;           /* device: 12 floats per reference              */
    const int32_t* tri_indices;  /* device: mesh-local triangle ids              */
    int64_t n_interior;
    int64_t n_leaves;
    int64_t n_references;
    int64_t tri_offset;          /* first reference slot of the mesh (off[m]
                                    unless presplit)                             */
} hx_batch_mesh_view;
int hx_build_batch(hx_context* ctx, const hx_build_params* params, const float* d_vertices, int64_t n_vertices,
                   const int32_t* d_triangles, const int64_t* h_mesh_offsets, int32_t n_meshes, hx_batch** out);
/* Rebuild every mesh from new vertices/triangles (same offsets and counts),
 * reusing the batch's memory. */
int hx_batch_rebuild(hx_context* ctx, hx_batch* batch, const float* d_vertices, int64_t n_vertices,
                     const int32_t* d_triangles, int64_t n_triangles);
int hx_batch_refit(hx_context* ctx, hx_batch* batch, const float* d_vertices, int64_t n_vertices,
                   const int32_t* d_triangles, int64_t n_triangles);
int hx_batch_destroy(hx_batch* batch);
/* Synchronises on first call after a build to read the mesh tables. */
int hx_batch_mesh(hx_batch* batch, int32_t mesh, hx_batch_mesh_view* out);
/* Rays as hx_trace; d_ray_mesh[r] names the mesh ray r traverses (a ray with
 * no such mesh misses).  Hit ids are mesh-local. */
int hx_batch_trace(hx_context* ctx, hx_batch* batch, const float* d_rays, const int32_t* d_ray_mesh,
                   int64_t n_rays, int32_t any_hit, int32_t* d_results, int32_t* d_stats);
/* hx_backward / hx_recompute_hits for a batched trace's results. */
int hx_batch_backward(hx_context* ctx, hx_batch* batch, const float* d_vertices, const int32_t* d_triangles,
                      const float* d_rays, const int32_t* d_ray_mesh, const int32_t* d_results, int64_t n_rays,
                      const float* d_dl_dt, const float* d_dl_du, const float* d_dl_dv, float* d_dl_dvertices);
int hx_batch_recompute_hits(hx_context* ctx, hx_batch* batch, const float* d_vertices, const int32_t* d_triangles,
                            const float* d_rays, const int32_t* d_ray_mesh, const int32_t* d_results, int64_t n_rays,
                            float* d_t, float* d_u, float* d_v);

#ifdef __cplusplus
}
#endif

#endif /* HX_H */
