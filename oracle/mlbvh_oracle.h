/*
 * CPU oracle for the MinimumLBVH GPU path -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference's CPU algorithm
 * (minimum_lbvh.h, main_gpu.cu, sobol.h, camera.h of Ushio/MinimumLBVH).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker / the timed CPU baseline.  The product
 * library (minimumlbvh_amd) never calls into it.
 *
 * Every function cites the reference file:line it restates.  Pinning (what
 * checks each function against something other than itself) is listed in the
 * header comment of mlbvh_oracle.c.
 *
 * Arithmetic follows the reference's host semantics exactly (helper_math.h
 * host branch): fminf(a,b) = a<b?a:b, fmaxf(a,b) = a>b?a:b, rsqrtf = 1/sqrtf,
 * float->int truncation with x86 cvttss2si semantics, no FMA contraction
 * (compiled with -ffp-contract=off), IEEE single precision throughout.
 */
#ifndef MLBVH_ORACLE_H
#define MLBVH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float x, y, z; } orc_v3;
typedef struct { orc_v3 vs[3]; } orc_tri;                 /* 36 B, minimum_lbvh.h:207 */
typedef struct { orc_v3 lower, upper; } orc_aabb;         /* 24 B, minimum_lbvh.h:217 */
typedef struct {                                          /* 64 B, minimum_lbvh.h:276 */
    uint32_t children[2];   /* NodeIndex: bits 0..30 index, bit 31 leaf */
    orc_aabb aabbs[2];
    uint32_t parent;
    uint32_t context;
} orc_node;
typedef struct {                                          /* 32 B, minimum_lbvh.h:890 */
    float t;
    uint32_t pad;            /* float2 uv is alignas(8) */
    float u, v;
    float ngx, ngy, ngz;
    uint32_t triangleIndex;
} orc_hit;
typedef struct { orc_v3 origin, right, up; } orc_raygen;  /* camera.h:4-8 */
typedef struct {                                          /* main_gpu.cu:10-15 */
    orc_v3 shadingNormals[3];
    orc_v3 reflectance;
    orc_v3 emissive;
} orc_attrib;

#define ORC_NODE_INVALID 0x7FFFFFFFu
#define ORC_LEAF_BIT 0x80000000u
#define ORC_OPT_USE_NORMAL 1u
#define ORC_OPT_CPU_PARALLEL 2u
#define ORC_OPT_MORTON64 4u /* no reference counterpart: full 63-bit keys */
/* sort the records with the reference CPU builder's std::sort on the key
   alone (minimum_lbvh.h:554; unstable) instead of the stable radix sort:
   the timed CPU baseline in the reference's own structure */
#define ORC_OPT_COMPARISON_SORT 8u
#define ORC_QUERY_CLOSEST 0
#define ORC_QUERY_ANY 1

void orc_sort_records_comparison(uint64_t* records, int n); /* sort_comparison.cpp */
/* the box of n points (AABB::extend), then AABB::encodeMortonCode of each point, minimum_lbvh.h:246-251 */
void orc_points_morton(const float* pts, int n, float* boxOut, uint64_t* codes);

/* bit utilities, minimum_lbvh.h:99-205 */
int orc_clz(uint32_t x);
int orc_clz64(uint64_t x);
int orc_ffs(uint32_t x);
int orc_ffs64(uint64_t x);
int orc_delta32(uint32_t a, uint32_t b);
uint64_t orc_splat3(uint32_t a);
uint32_t orc_compact3(uint64_t m);
uint64_t orc_encode_morton(uint32_t x, uint32_t y, uint32_t z);
uint64_t orc_encode_morton_naive(uint32_t x, uint32_t y, uint32_t z);
void orc_decode_morton(uint64_t m, uint32_t* x, uint32_t* y, uint32_t* z);
void orc_morton_batch(const uint32_t* xyz, uint64_t* out, int n, int naive);
void orc_decode_batch(const uint64_t* codes, uint32_t* xyz, int n);
int orc_f2i_trunc(float f);

/* build pipeline, minimum_lbvh.h:512-592 */
void orc_scene_aabb(const orc_tri* tris, int n, orc_aabb* out);
uint64_t orc_aabb_encode_morton(const orc_aabb* box, orc_v3 p);
void orc_morton_records(const orc_tri* tris, int n, const orc_aabb* scene, uint32_t opt, uint64_t* records);
void orc_sort_records(uint64_t* records, int n);
void orc_deltas(const uint64_t* sorted, int n, uint8_t* deltas);
void orc_build_leaf(uint32_t* root, orc_node* nodes, const orc_tri* tris, int n,
                    const uint8_t* deltas, uint32_t i_leaf, uint32_t triIndex);
/* full build: returns 0 on success, -1 on invalid n.  sorted/deltas may be NULL. */
int orc_build(const orc_tri* tris, int n, uint32_t opt, orc_node* nodes, uint32_t* root,
              uint64_t* sorted_out, uint8_t* deltas_out, int nthreads);
int orc_validate(uint32_t root, const orc_node* nodes, const uint8_t* deltas, int maxDelta);

/* stable LSD radix semantics of OnesweepSort::sort, tinyhiponesweep.h:592-668 */
int orc_radix_sort_u32(uint32_t* keys, uint32_t* vals, uint32_t n, int startBit, int endBit);
int orc_radix_sort_u64(uint64_t* keys, uint32_t* vals, uint32_t n, int startBit, int endBit);

/* traversal, minimum_lbvh.h:82-96, 363-413, 898-1137 */
orc_v3 orc_inv_rd(orc_v3 rd);
void orc_slabs(orc_v3 ro, orc_v3 inv, orc_v3 lo, orc_v3 hi, float tHit, float* rmin, float* rmax);
int orc_intersect_tri(float* t, float* u, float* v, orc_v3* ng, float tmin, float tmax,
                      orc_v3 ro, orc_v3 rd, orc_v3 v0, orc_v3 v1, orc_v3 v2);
void orc_intersect_stackfree(orc_hit* hit, const orc_node* nodes, const orc_tri* tris, uint32_t root,
                             orc_v3 ro, orc_v3 rd, orc_v3 inv, int query, uint64_t* counters);
void orc_intersect_stackfree_ex(orc_hit* hit, const orc_node* nodes, const orc_tri* tris, uint32_t root,
                                orc_v3 ro, orc_v3 rd, orc_v3 inv, int query, uint64_t* counters, int earlyExit);
void orc_intersect_stackfull(orc_hit* hit, const orc_node* nodes, const orc_tri* tris, uint32_t root,
                             orc_v3 ro, orc_v3 rd, orc_v3 inv, uint32_t* stack);
void orc_intersect_escape_link(orc_hit* hit, const orc_node* nodes, const orc_tri* tris, uint32_t root,
                               orc_v3 ro, orc_v3 rd, orc_v3 inv, int query, uint64_t* counters);
/* rays: 8 floats each (ox,oy,oz,tmax,dx,dy,dz,pad); method 0 stackfree, 1 stackfull, 2 escape link,
   3 stackfree with the pending-count early exit, 4 orc_closest_hit_stack (ORC_LANE_STACK entries) */
void orc_trace(const orc_node* nodes, const orc_tri* tris, uint32_t root, const float* rays, int nRays,
               int query, int method, orc_hit* hits, uint64_t* counters, int nthreads);
void orc_brute_force(const orc_tri* tris, int n, const float* rays, int nRays, orc_hit* hits);

/* sampling + camera, sobol.h:18-129, camera.h:10-31, minimum_lbvh.h:904-909 */
uint32_t orc_hash_pcg(uint32_t v);
/* PCG32 of main_gpu.cu:51-82: the first n outputs, and uniformf of them */
void orc_pcg32(uint64_t seed, uint64_t seq, int n, uint32_t* out);
void orc_pcg32_uniformf(uint64_t seed, uint64_t seq, int n, float* out);
/* cosine-weighted hemisphere direction of main_gpu.cu:24-36 (y = up) */
void orc_hemisphere_cos(float xi0, float xi1, float* out);
/* Duff et al. 2017 orthonormal basis of main_gpu.cu:39-45: out = xaxis[3], yaxis[3] for the unit normal z */
void orc_onb(const float* z, float* out);
/* the world direction the AO and PT kernels build for unit normal n from (xi0, xi1): the ONB of n and the
   cosine-hemisphere sample, x*dl.x + y*dl.z + n*dl.y (main_gpu.cu:139-153, :231-235) */
void orc_cosine_dir(const float* n, float xi0, float xi1, float* out);
/* sin / cos / log2 / exp2 / pow in place of libm's (main_gpu.cu:30-31, :174, :283); pow for y > 0 */
void orc_sincos(float x, float* s, float* c);
float orc_log2_pos(float x);
float orc_exp2_fin(float z);
float orc_pow_pos(float x, float y);
void orc_sincos_batch(const float* x, int n, float* s, float* c);
void orc_pow_batch(const float* x, int n, float y, float* out);
void orc_sobol_2d(float* x, float* y, uint32_t index, uint32_t p0, uint32_t p1, uint32_t p2);
void orc_raygen_lookat(orc_raygen* g, orc_v3 eye, orc_v3 center, orc_v3 up, float fovy, int w, int h);
void orc_raygen_shoot(const orc_raygen* g, orc_v3* ro, orc_v3* rd, float u, float v);

/* render kernels, main_gpu.cu:17-286 */
uint32_t orc_pack_rgba(float r, float g, float b, float a);
void orc_render_normal(uint32_t* pixels, int w, int h, const orc_raygen* g, uint32_t root,
                       const orc_node* nodes, const orc_tri* tris, int nthreads);
/* aoOut (nullable): occlusion fraction per pixel, -1 for primary miss.
   rows [row0,row1) only (for bounded CPU baselines); counters: rays traced, node fetches, tri tests */
/* walk: 0 reference stack-free walk, 1 stack-free with early exit, 2 early-exit primary + stack any-hit AO rays,
   3 stack closest-hit primary + stack any-hit AO rays */
#define ORC_AO_STACK 12   /* the AO kernel's per-lane stack */
#define ORC_LANE_STACK 16 /* the trace and render kernels' per-lane stack */
int orc_any_hit_stack(const orc_node* nodes, const orc_tri* tris, uint32_t root, orc_v3 ro, orc_v3 rd, orc_v3 inv,
                      float tmax, uint64_t* cnt, int cap);
void orc_closest_hit_stack(orc_hit* hit, const orc_node* nodes, const orc_tri* tris, uint32_t root, orc_v3 ro,
                           orc_v3 rd, orc_v3 inv, int query, uint64_t* cnt, int cap);
void orc_closest_hit_stack_batch(const orc_node* nodes, const orc_tri* tris, uint32_t root, const float* rays,
                                 int nRays, int query, int cap, orc_hit* hits, uint64_t* counters, int nthreads);
void orc_any_hit_stack_batch(const orc_node* nodes, const orc_tri* tris, uint32_t root, const float* rays, int nRays,
                             int cap, uint8_t* occluded, uint64_t* counters, int nthreads);
void orc_render_ao(uint32_t* pixels, float* aoOut, int w, int h, const orc_raygen* g, uint32_t root,
                   const orc_node* nodes, const orc_tri* tris, int useSobol, int nSample, float radius,
                   uint32_t seed, int row0, int row1, uint64_t* counters, int nthreads, int walk);
void orc_render_pt(float* acc /* float4 per pixel */, int w, int h, const orc_raygen* g, uint32_t root,
                   const orc_node* nodes, const orc_tri* tris, const orc_attrib* attribs, int useSobol,
                   int maxSPP, int maxDepth, int nthreads);
void orc_pack(uint32_t* pixels, const float* acc, int n);

/* Differentiable primary-hit shading, fp64 reference (no reference
   counterpart).  Per pixel p = x + y*w with its hit triangle hitTri[p] held
   fixed (0xFFFFFFFF = miss): out[2p] = t = dot(v0-ro, ng)/dot(ng, rd),
   out[2p+1] = max(sgn dot(ng/|ng|, L), 0) with sgn turning the normal toward
   the viewer; rays from orc_raygen_shoot.  grad += sum_p gout[p] . d out[p] / dV. */
void orc_primary_eval64(const double* verts, const uint32_t* idx, const uint32_t* hitTri, int w, int h,
                        const orc_raygen* g, const double* light, double* out);
void orc_primary_grad64(const double* verts, const uint32_t* idx, const uint32_t* hitTri, int w, int h,
                        const orc_raygen* g, const double* light, const double* gout, double* grad);

double orc_now(void);
int orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
