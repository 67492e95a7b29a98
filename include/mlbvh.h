/*
 * mlbvh.h -- C ABI of the MI355X (gfx950) MinimumLBVH library, libmlbvh.so.
 *
 * Plain C: pointers, sizes and status codes only.  Device pointers are HIP
 * device allocations; `stream` is a hipStream_t passed as void* (NULL = the
 * null stream).  Every entry point is asynchronous on `stream` unless noted
 * and returns MLBVH_OK or a negative status; mlbvh_last_error() then holds a
 * message (thread-local).
 *
 * Each entry point names the reference interface of Ushio/MinimumLBVH it
 * replaces (file:line).  The C++ facade in include/minimum_lbvh.h and
 * include/tinyhiponesweep.h keeps the reference's class/method names on top
 * of these functions.
 *
 * Binary layouts (little-endian, identical to the reference's structs):
 *   Triangle      36 B  float v[3][3]                       minimum_lbvh.h:207
 *   InternalNode  64 B  u32 children[2] @0 (bit31 = leaf, 0x7FFFFFFF = invalid),
 *                       AABB aabbs[2] @8 (lower xyz, upper xyz),
 *                       u32 parent @56, u32 context @60     minimum_lbvh.h:254-282
 *   NodeIndex      4 B  u32, bits 0..30 index, bit 31 leaf
 *   Hit           32 B  f32 t @0, f32 u,v @8, f32 ng[3] @16, u32 triangleIndex @28
 *                                                            minimum_lbvh.h:890-896
 *   IndexedMorton  8 B  u32 morton @0, u32 index @4 (a u64 with the key in
 *                       the low word)                        minimum_lbvh.h:502-506
 *   RayGenerator  36 B  f32 origin[3], right[3], up[3]       camera.h:4-8
 *   TriangleAttrib 60 B f32 shadingNormals[3][3], reflectance[3], emissive[3]
 *                                                            main_gpu.cu:10-15
 *   Ray (trace)   32 B  f32 ox,oy,oz,tmax, dx,dy,dz, pad (this library's own
 *                       batch format; tmax seeds Hit.t like main_gpu.cu:167)
 */
#ifndef MLBVH_H
#define MLBVH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__) || defined(__clang__)
#define MLBVH_API __attribute__((visibility("default")))
#else
#define MLBVH_API
#endif

#define MLBVH_ABI_VERSION 1

#define MLBVH_OK 0
#define MLBVH_E_INVALID_ARG (-1) /* bad size / pointer / bit range */
#define MLBVH_E_HIP (-2)         /* a HIP runtime call failed */
#define MLBVH_E_NO_DEVICE (-3)   /* no usable gfx950 device */

/* build options, minimum_lbvh.h:42-46 */
#define MLBVH_BUILD_OPTION_USE_NORMAL 1u
#define MLBVH_BUILD_OPTION_CPU_PARALLEL 2u /* accepted and ignored on the GPU */
/* No reference counterpart: key every triangle by the full 63-bit Morton code
 * of its centroid instead of its top 32 bits (minimum_lbvh.h:550), so that
 * centroids closer than 1/2^10 of the scene box still get distinct keys at
 * scale.  Deltas become 64 - clz64; sorted records keep the 32-bit layout
 * (index high, code >> 31 low), so mlbvh_refit works unchanged.  Not
 * combinable with MLBVH_BUILD_OPTION_USE_NORMAL. */
#define MLBVH_BUILD_OPTION_MORTON64 4u

/* ray queries, minimum_lbvh.h:975-979 */
#define MLBVH_RAY_QUERY_CLOSEST 0
#define MLBVH_RAY_QUERY_ANY 1

typedef struct mlbvh_ctx mlbvh_ctx;

typedef struct mlbvh_raygen {
    float origin[3];
    float right[3];
    float up[3];
} mlbvh_raygen;

/* Library identity. */
MLBVH_API int mlbvh_abi_version(void);
MLBVH_API const char* mlbvh_last_error(void);

/* Per-device context owning the sorter's and builder's scratch (look-back
 * table, digit histograms, records, deltas).  Replaces the resources created
 * by OnesweepSort(oroDevice) (tinyhiponesweep.h:575-579) and
 * BVHGPUBuilder(kernel, includeDir) (minimum_lbvh.h:729-775); kernels are
 * compiled ahead of time for gfx950, so nothing is compiled at run time.
 * Not re-entrant: calls sharing one context must be ordered on one stream. */
MLBVH_API int mlbvh_ctx_create(int device, mlbvh_ctx** out);
MLBVH_API int mlbvh_ctx_destroy(mlbvh_ctx* ctx);

/* Stable LSD radix sort, OnesweepSort::sort(KeyValueSoA32|64, tmp, n,
 * startBit, endBit, stream) (tinyhiponesweep.h:592-599).  Sorts ascending by
 * the key bits [startBit, startBit + 8*ceil((endBit-startBit)/8)) (bits past
 * the key width read as 0); values == NULL means key-only.  The result ends in
 * keys/values; keys_tmp/values_tmp are clobbered.  n may be 0. */
MLBVH_API int mlbvh_sort_u32(mlbvh_ctx* ctx, uint32_t* keys, uint32_t* values, uint32_t* keys_tmp, uint32_t* values_tmp,
                   uint32_t n, int startBit, int endBit, void* stream);
MLBVH_API int mlbvh_sort_u64(mlbvh_ctx* ctx, uint64_t* keys, uint32_t* values, uint64_t* keys_tmp, uint32_t* values_tmp,
                   uint32_t n, int startBit, int endBit, void* stream);

/* LBVH build, BVHGPUBuilder::build(triangles, n, buildOption, sorter, stream)
 * (minimum_lbvh.h:783-871): scene AABB -> 32-bit Morton keys -> stable sort of
 * (morton, index) -> deltas -> agglomerative build.  Writes n-1 InternalNode
 * to d_nodes and the root NodeIndex to d_root.  2 <= n <= 2^31-1.
 * No host synchronisation and no timing printf (the reference syncs at :869). */
MLBVH_API int mlbvh_build(mlbvh_ctx* ctx, const void* d_triangles, int n, uint32_t buildOption, void* d_nodes, void* d_root,
                void* stream);

/* mlbvh_build plus optional copies of the sorted IndexedMorton records
 * (n x u64) and the uint8 deltas (n-1), which mlbvh_refit consumes.
 * No reference counterpart beyond the locals of minimum_lbvh.h:790-796. */
MLBVH_API int mlbvh_build_ex(mlbvh_ctx* ctx, const void* d_triangles, int n, uint32_t buildOption, void* d_nodes, void* d_root,
                   uint64_t* d_sorted_records, uint8_t* d_deltas, void* stream);

/* mlbvh_build_ex without any device allocation.  No reference counterpart:
 * it exists for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Results are bit-identical
 * to mlbvh_build_ex.  Its scratch is one static device array per GPU (about
 * 44 MiB, held from library load to process exit), set up by
 * mlbvh_ctx_create; n is limited to 2^20 triangles (MLBVH_E_INVALID_ARG
 * above).  All calls on one GPU share that array: calls that overlap in time
 * on one GPU race, so order them on one stream. */
MLBVH_API int mlbvh_build_ex_noalloc(mlbvh_ctx* ctx, const void* d_triangles, int n, uint32_t buildOption,
                                     void* d_nodes, void* d_root, uint64_t* d_sorted_records, uint8_t* d_deltas,
                                     void* stream);

/* Refit: recompute every node AABB for moved vertices with the topology
 * fixed by an earlier build's sorted records and deltas (re-runs the
 * agglomerative climb of minimum_lbvh.h:284-361 on unchanged deltas).  No
 * reference counterpart (the reference always rebuilds). */
MLBVH_API int mlbvh_refit(mlbvh_ctx* ctx, const void* d_triangles, int n, const uint64_t* d_sorted_records,
                const uint8_t* d_deltas, void* d_nodes, void* d_root, void* stream);

/* Indexed-mesh build: the triangles are index triples d_indices
 * (nTriangles x 3 u32) into d_vertices (nVertices x 3 fp32 xyz).  No
 * reference counterpart: BVHGPUBuilder::build (minimum_lbvh.h:783-871) takes
 * a Triangle array, so a caller with an indexed mesh first writes the 36-byte
 * de-indexed copy.  Nodes, root and the optional sorted records / deltas are
 * bit-identical to mlbvh_build_ex on that copy; indices >= nVertices are
 * clamped to the last vertex (memory-safe, results unspecified). */
MLBVH_API int mlbvh_build_indexed(mlbvh_ctx* ctx, const float* d_vertices, int nVertices, const uint32_t* d_indices,
                                  int nTriangles, uint32_t buildOption, void* d_nodes, void* d_root,
                                  uint64_t* d_sorted_records, uint8_t* d_deltas, void* stream);

/* mlbvh_refit for an indexed mesh (arguments as mlbvh_build_indexed). */
MLBVH_API int mlbvh_refit_indexed(mlbvh_ctx* ctx, const float* d_vertices, int nVertices, const uint32_t* d_indices,
                                  int nTriangles, const uint64_t* d_sorted_records, const uint8_t* d_deltas,
                                  void* d_nodes, void* d_root, void* stream);

/* Batch trace replacing per-ray intersect_stackfree (minimum_lbvh.h:981-1061).
 * d_rays: nRays x 32 B Ray; d_hits: nRays x 32 B Hit, bit-identical to
 * intersect_stackfree's (ties included).  The walk keeps a 16-entry stack of
 * far children per lane in LDS and re-tests each popped entry's entry
 * distance against the current hit, which visits the leaves in the stack-free
 * walk's order while fetching each internal node once per visit (a deeper
 * stack finishes the ray stack-free).  d_counters (nullable): 2 x u64
 * accumulated [internal-node fetches, triangle tests] of this walk. */
MLBVH_API int mlbvh_trace(mlbvh_ctx* ctx, const void* d_nodes, const void* d_triangles, const void* d_root,
                const float* d_rays, void* d_hits, int nRays, int rayQuery, unsigned long long* d_counters,
                void* stream);

/* Batch trace with intersect_stackfree's own parent-pointer walk (with an
 * exact early exit once no far child is owed): the same arguments and Hit
 * records as mlbvh_trace; d_counters count this walk's fetches. */
MLBVH_API int mlbvh_trace_stackfree(mlbvh_ctx* ctx, const void* d_nodes, const void* d_triangles,
                                    const void* d_root, const float* d_rays, void* d_hits, int nRays, int rayQuery,
                                    unsigned long long* d_counters, void* stream);

/* Batch trace with intersect_escape_link (minimum_lbvh.h:1063-1137): the
 * same arguments and records as mlbvh_trace.  Closest-hit t equals
 * mlbvh_trace's; on equal-t ties the triangle can differ (this walk always
 * takes the left child first).  d_counters[0] counts internal-node visits
 * (each visit tests one child box). */
/* mlbvh_trace over an indexed mesh (vertex / index buffers as
 * mlbvh_build_indexed): Hit records bit-identical to mlbvh_trace on the
 * de-indexed triangles.  No reference counterpart (the reference traverses a
 * Triangle array). */
MLBVH_API int mlbvh_trace_indexed(mlbvh_ctx* ctx, const void* d_nodes, const float* d_vertices, int nVertices,
                                  const uint32_t* d_indices, const void* d_root, const float* d_rays, void* d_hits,
                                  int nRays, int rayQuery, unsigned long long* d_counters, void* stream);

MLBVH_API int mlbvh_trace_escape_link(mlbvh_ctx* ctx, const void* d_nodes, const void* d_triangles,
                                      const void* d_root, const float* d_rays, void* d_hits, int nRays,
                                      int rayQuery, unsigned long long* d_counters, void* stream);

/* Batched build of many meshes in one pipeline.  No reference counterpart:
 * the reference builds one mesh per BVHGPUBuilder::build call
 * (minimum_lbvh.h:783-871).  Meshes are concatenated in d_triangles; mesh m
 * owns triangles [off[m], off[m+1]) of the device array d_triOffsets
 * (nMeshes+1 entries, off[0] = 0, every mesh >= 2 triangles); the caller
 * states totalTriangles (= off[nMeshes]) and an upper bound on the triangles
 * of one mesh.  Mesh m gets exactly the BVH mlbvh_build would give it alone,
 * with mesh-local indices, in the n_m - 1 nodes at d_nodes + (off[m] - m);
 * its root goes to d_roots[m].  Needs ceil(log2(maxTrianglesPerMesh)) + 32 +
 * ceil(log2(nMeshes)) <= 64. */
MLBVH_API int mlbvh_build_batched(mlbvh_ctx* ctx, const void* d_triangles, const uint32_t* d_triOffsets, int nMeshes,
                                  uint32_t totalTriangles, uint32_t maxTrianglesPerMesh, uint32_t buildOption,
                                  void* d_nodes, uint32_t* d_roots, void* stream);

/* Persistent work-queue trace over a batched build: each Ray names its mesh
 * in its pad word (u32); hits hold mesh-local triangle indices, and a ray
 * naming no mesh of the batch gets a miss record (t = its tmax, triangle
 * 0xFFFFFFFF).  Same traversal as mlbvh_trace (minimum_lbvh.h:981-1061).
 * d_counters as in mlbvh_trace.  The work counter lives in ctx: calls on one
 * ctx must not overlap. */
MLBVH_API int mlbvh_trace_batched(mlbvh_ctx* ctx, const void* d_nodes, const void* d_triangles,
                                  const uint32_t* d_triOffsets, const uint32_t* d_roots, int nMeshes,
                                  const float* d_rays, void* d_hits, int nRays, int rayQuery,
                                  unsigned long long* d_counters, void* stream);

/* The plain per-mesh path over the same batch, for comparison: one
 * mlbvh_build and one mlbvh_trace per mesh, launched asynchronously.  Offsets
 * are host arrays here (tri offsets nMeshes+1, ray offsets nMeshes+1); output
 * layout as mlbvh_build_batched / mlbvh_trace_batched. */
MLBVH_API int mlbvh_build_trace_per_mesh(mlbvh_ctx* ctx, const void* d_triangles, const uint32_t* h_triOffsets,
                                         int nMeshes, uint32_t buildOption, void* d_nodes, uint32_t* d_roots,
                                         const float* d_rays, const uint32_t* h_rayOffsets, void* d_hits,
                                         int rayQuery, void* stream);

/* Launch shapes chosen by the autotuner (minimumlbvh_amd/autotune.py):
 * sortConfig selects the radix-sort tile (0: 256x16, 1: 256x32, 2: 512x16,
 * 3: 512x24, 4: 1024x16 threads x keys per thread); the persistent trace runs
 * CUs x traceBlocksPerCU blocks of 256 threads and each wave takes traceChunk
 * rays (a multiple of 64) per work-queue pull.  Results do not depend on
 * these settings. */
MLBVH_API int mlbvh_set_tuning(mlbvh_ctx* ctx, int sortConfig, int traceBlocksPerCU, int traceChunk);
MLBVH_API int mlbvh_get_tuning(mlbvh_ctx* ctx, int* sortConfig, int* traceBlocksPerCU, int* traceChunk);

/* Differentiable primary-hit rendering (no reference counterpart; built on
 * intersectRayTriangle, minimum_lbvh.h:363-413, and the camera of camera.h).
 * gather: Triangle records from an indexed mesh (fp32 xyz vertices, 3 u32
 * indices per triangle). */
MLBVH_API int mlbvh_gather_triangles(mlbvh_ctx* ctx, const float* d_vertices, const uint32_t* d_indices, int nTris,
                                     void* d_triangles, void* stream);
/* render_primary: per pixel (x + y*width) the closest hit of the camera ray
 * (intersect_stackfree order); d_out = (t, max(dot(n, L), 0)) as 2 floats, n
 * the geometric normal normalised and turned toward the viewer; (0, 0) on a
 * miss.  d_hitTri receives the hit triangle (0xFFFFFFFF on a miss). */
MLBVH_API int mlbvh_render_primary(mlbvh_ctx* ctx, float* d_out, uint32_t* d_hitTri, int width, int height,
                                   const mlbvh_raygen* camera, const float lightDir[3], const void* d_root,
                                   const void* d_nodes, const void* d_triangles, void* stream);
/* primary_adjoint: d_gradVertices += sum over hit pixels of
 * d_gradOut[p] . d(t, s)/d(vertices of the hit triangle), with the hit
 * triangle of each pixel held fixed (fp32 atomics; zero the gradient first). */
MLBVH_API int mlbvh_primary_adjoint(mlbvh_ctx* ctx, const float* d_gradOut, const uint32_t* d_hitTri, int width,
                                    int height, const mlbvh_raygen* camera, const float lightDir[3],
                                    const float* d_vertices, const uint32_t* d_indices, float* d_gradVertices,
                                    void* stream);

/* Host-side RayGenerator::lookat (camera.h:10-22); fovy in radians. */
MLBVH_API int mlbvh_raygen_lookat(mlbvh_raygen* out, const float eye[3], const float center[3], const float up[3], float fovy,
                        int width, int height);

/* Render kernels of main_gpu.cu.  `camera` is a host pointer read at launch. */
/* normal (main_gpu.cu:84-109) */
MLBVH_API int mlbvh_render_normal(mlbvh_ctx* ctx, uint32_t* d_pixels, int width, int height, const mlbvh_raygen* camera,
                        const void* d_root, const void* d_nodes, const void* d_triangles, void* stream);
/* ao (main_gpu.cu:111-177), with the sample count, ray length and Sobol seed
 * as parameters (the reference hard-codes 16, 1.0f and 12345).  d_ao
 * (nullable) receives the occlusion fraction per pixel, -1 on a primary miss.
 * d_counters (nullable): 3 x u64 accumulated [rays traced, node fetches,
 * triangle tests]. */
MLBVH_API int mlbvh_render_ao(mlbvh_ctx* ctx, uint32_t* d_pixels, float* d_ao, int width, int height,
                    const mlbvh_raygen* camera, const void* d_root, const void* d_nodes, const void* d_triangles,
                    int useSobol, int nSample, float radius, uint32_t seed, unsigned long long* d_counters,
                    void* stream);
/* pt (main_gpu.cu:179-271): one sample per pixel per call, accumulated in
 * d_accumulators (float4 per pixel, .w = samples so far); maxDepth bounds the
 * bounce loop (the reference uses 1024). */
MLBVH_API int mlbvh_render_pt(mlbvh_ctx* ctx, float* d_accumulators, int width, int height, const mlbvh_raygen* camera,
                    const void* d_root, const void* d_nodes, const void* d_triangles, const void* d_attribs,
                    int useSobol, int maxSPP, int maxDepth, void* stream);
/* pack (main_gpu.cu:273-286) */
MLBVH_API int mlbvh_pack(mlbvh_ctx* ctx, uint32_t* d_pixels, const float* d_accumulators, int n, void* stream);

#ifdef __cplusplus
}
#endif
#endif
