/*
 * CPU oracle for the MinimumLBVH GPU path -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference CPU path of Ushio/MinimumLBVH.  It is
 * the checker for the HIP library and the timed CPU baseline; the product
 * never links or calls it.  Build: oracle/Makefile (gcc -O2 -fopenmp
 * -ffp-contract=off).
 *
 * How each part is pinned (checked against something other than itself; the
 * tests live in tests/test_oracle_*.py):
 *   clz/ffs/delta ........ closed-form edge cases of unittest.cpp:39-105
 *   splat3/compact3/morton exhaustive 2^21 bit positions + round trip and
 *                          naive==fast (unittest.cpp:6-37); numpy bit loops
 *   scene AABB / keys .... numpy float32 restatement on seeded meshes
 *                          (min/max are exact; the key formula is re-derived
 *                          with numpy's IEEE float32 ops, incl. the x86
 *                          cvttss2si truncation rule)
 *   sort ................. numpy stable argsort (a library routine) on the
 *                          same (key,index) records; radix-sort semantics vs
 *                          numpy on masked digit fields
 *   deltas ............... closed form 32-clz(a^b) via numpy
 *   build topology ....... known-answer toy tree of main.cpp:60-109, the
 *                          validate_lbvh invariant, every-leaf-once, parent /
 *                          child consistency, exact subtree-union AABBs,
 *                          serial == parallel, brute-force Karras split rule
 *   traversal ............ brute force over all triangles (closest t exact),
 *                          stackfree == stackfull == escape-link
 *   sobol / camera ....... golden vectors from the reference's own sobol.h and
 *                          camera.h compiled with g++ (oracle/_ref)
 *   AO / PT shading ...... per-step closed forms (ONB orthonormality, packRGBA
 *                          rounding), the AO/normal images against traced
 *                          hits; PCG32 by the published pcg32-demo outputs
 *                          (seed 42, sequence 54); the cosine hemisphere by
 *                          closed-form points and its moments; the PT loop
 *                          by a closed furnace (every path's radiance is
 *                          1, 1.5, 1.75 at depth 1-3, 1.75 or 2 at depth 4).
 */
#include "mlbvh_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define FLT_MAX_H 3.402823466e+38F /* MINIMUM_LBVH_FLT_MAX, minimum_lbvh.h:37 */
#define MORTON_MAX 0x1FFFFFu       /* MORTON_MAX_VALUE_3D, minimum_lbvh.h:38 */
static const float ORC_PI = 3.14159265358979323846f; /* main_gpu.cu:8 */

/* ---------------------------------------------------------------------- */
/* host float semantics (helper_math.h:134-157, 1079-1102, 1437-1441, 1548) */
/* ---------------------------------------------------------------------- */
static inline float hmin(float a, float b) { return a < b ? a : b; }
static inline float hmax(float a, float b) { return a > b ? a : b; }
static inline orc_v3 V(float x, float y, float z) { orc_v3 r = {x, y, z}; return r; }
static inline orc_v3 vadd(orc_v3 a, orc_v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline orc_v3 vsub(orc_v3 a, orc_v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline orc_v3 vmul(orc_v3 a, orc_v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline orc_v3 vmuls(orc_v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline orc_v3 vdiv(orc_v3 a, orc_v3 b) { return V(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline orc_v3 vdivs(orc_v3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
static inline orc_v3 vneg(orc_v3 a) { return V(-a.x, -a.y, -a.z); }
static inline orc_v3 vmin(orc_v3 a, orc_v3 b) { return V(hmin(a.x, b.x), hmin(a.y, b.y), hmin(a.z, b.z)); }
static inline orc_v3 vmax(orc_v3 a, orc_v3 b) { return V(hmax(a.x, b.x), hmax(a.y, b.y), hmax(a.z, b.z)); }
static inline float vdot(orc_v3 a, orc_v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline orc_v3 vcross(orc_v3 a, orc_v3 b)
{
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline orc_v3 vnormalize(orc_v3 v)
{
    float invLen = 1.0f / sqrtf(vdot(v, v)); /* host rsqrtf, helper_math.h:154-157 */
    return vmuls(v, invLen);
}
/* lerp(a,b,t) = a + t*(b-a), helper_math.h:1266-1269 */
static inline orc_v3 vlerp(orc_v3 a, orc_v3 b, float t) { return vadd(a, vmuls(vsub(b, a), t)); }

/* make_int3(float3) truncation as compiled for x86-64 (cvttss2si):
   out-of-range and NaN give INT_MIN.  helper_math.h:252-255 */
int orc_f2i_trunc(float f)
{
    if (f >= -2147483648.0f && f < 2147483648.0f) return (int)f;
    return (int)0x80000000u;
}
static inline int iclamp(int f, int a, int b)
{
    int m = f < b ? f : b; /* min(f, b) */
    return a > m ? a : m;  /* max(a, min(f,b)), helper_math.h:1326-1329 */
}

double orc_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static int pick_threads(int nthreads)
{
#ifdef _OPENMP
    return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

/* ---------------------------------------------------------------------- */
/* bit utilities, minimum_lbvh.h:99-205                                    */
/* ---------------------------------------------------------------------- */
int orc_clz(uint32_t x) { return x ? __builtin_clz(x) : 32; }        /* :99-106 (__lzcnt) */
int orc_clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }    /* :107-114 */
int orc_ffs(uint32_t x) { return x ? __builtin_ctz(x) + 1 : 0; }     /* :117-127 */
int orc_ffs64(uint64_t x) { return x ? __builtin_ctzll(x) + 1 : 0; } /* :128-138 */
int orc_delta32(uint32_t a, uint32_t b) { return 32 - orc_clz(a ^ b); } /* :140-143 */

uint64_t orc_encode_morton_naive(uint32_t x, uint32_t y, uint32_t z) /* :149-160 */
{
    uint64_t code = 0;
    for (uint64_t i = 0; i < 64 / 3; ++i) {
        code |= ((uint64_t)(x & (1u << i)) << (2 * i + 0)) | ((uint64_t)(y & (1u << i)) << (2 * i + 1)) |
                ((uint64_t)(z & (1u << i)) << (2 * i + 2));
    }
    return code;
}

/* Spread the low 21 bits of a to every third bit (bit k -> bit 3k) by the
   usual shift-and-mask ladder.  minimum_lbvh.h:173-182 */
uint64_t orc_splat3(uint32_t a)
{
    uint64_t x = a & 0x1fffffu;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

/* Inverse of splat3: gather every third bit.  minimum_lbvh.h:162-171 */
uint32_t orc_compact3(uint64_t m)
{
    uint64_t x = m & 0x1249249249249249ull;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
    x = (x ^ (x >> 8)) & 0x1f0000ff0000ffull;
    x = (x ^ (x >> 16)) & 0x1f00000000ffffull;
    x = (x ^ (x >> 32)) & 0x1fffffull;
    return (uint32_t)x;
}

uint64_t orc_encode_morton(uint32_t x, uint32_t y, uint32_t z) /* :190-195 */
{
    return orc_splat3(x) | orc_splat3(y) << 1 | orc_splat3(z) << 2;
}

void orc_decode_morton(uint64_t m, uint32_t* x, uint32_t* y, uint32_t* z) /* :184-189 */
{
    *x = orc_compact3(m);
    *y = orc_compact3(m >> 1);
    *z = orc_compact3(m >> 2);
}

void orc_morton_batch(const uint32_t* xyz, uint64_t* out, int n, int naive)
{
    for (int i = 0; i < n; i++) {
        uint32_t x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        out[i] = naive ? orc_encode_morton_naive(x, y, z) : orc_encode_morton(x, y, z);
    }
}

void orc_decode_batch(const uint64_t* codes, uint32_t* xyz, int n)
{
    for (int i = 0; i < n; i++) orc_decode_morton(codes[i], &xyz[3 * i], &xyz[3 * i + 1], &xyz[3 * i + 2]);
}

/* ---------------------------------------------------------------------- */
/* AABB + Morton keys, minimum_lbvh.h:212-252, 522-553                     */
/* ---------------------------------------------------------------------- */
static inline float f_bits_u(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static inline uint32_t bits_u(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }
static inline void aabb_empty(orc_aabb* b)
{
    b->lower = V(FLT_MAX_H, FLT_MAX_H, FLT_MAX_H);
    b->upper = V(-FLT_MAX_H, -FLT_MAX_H, -FLT_MAX_H);
}
static inline void aabb_extend_p(orc_aabb* b, orc_v3 p)
{
    b->lower = vmin(b->lower, p);
    b->upper = vmax(b->upper, p);
}
/* The climb's merge of a child box into the other (AABB::extend(AABB),
   minimum_lbvh.h:205-209 with the host fminf/fmaxf), except on ties: the
   host rule keeps the second operand, so -0 against +0 would give a sign
   that depends on which child merges (order-dependent in the reference's GPU
   and parallel CPU builds).  Ties keep -0 for a lower bound and +0 for an
   upper one, which makes the merge symmetric; every other input gives the
   host rule's bits. */
static inline float umin(float a, float b)
{
    if (a < b) return a;
    if (b < a) return b;
    return f_bits_u(bits_u(a) | bits_u(b));
}
static inline float umax(float a, float b)
{
    if (a > b) return a;
    if (b > a) return b;
    return f_bits_u(bits_u(a) & bits_u(b));
}
static inline void aabb_extend_b(orc_aabb* b, const orc_aabb* o)
{
    b->lower = V(umin(b->lower.x, o->lower.x), umin(b->lower.y, o->lower.y), umin(b->lower.z, o->lower.z));
    b->upper = V(umax(b->upper.x, o->upper.x), umax(b->upper.y, o->upper.y), umax(b->upper.z, o->upper.z));
}

void orc_scene_aabb(const orc_tri* tris, int n, orc_aabb* out) /* :522-532 */
{
    orc_aabb s;
    aabb_empty(&s);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < 3; ++j) aabb_extend_p(&s, tris[i].vs[j]);
    *out = s;
}

/* int3 coord = (p - lower) / (upper - lower) * 2^21, truncated, clamped to
   [0, 2^21-1], then interleaved.  minimum_lbvh.h:246-251 */
uint64_t orc_aabb_encode_morton(const orc_aabb* box, orc_v3 p)
{
    orc_v3 f = vmuls(vdiv(vsub(p, box->lower), vsub(box->upper, box->lower)), (float)(MORTON_MAX + 1));
    int cx = iclamp(orc_f2i_trunc(f.x), 0, (int)MORTON_MAX);
    int cy = iclamp(orc_f2i_trunc(f.y), 0, (int)MORTON_MAX);
    int cz = iclamp(orc_f2i_trunc(f.z), 0, (int)MORTON_MAX);
    return orc_encode_morton((uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
}

void orc_points_morton(const float* pts, int n, float* boxOut, uint64_t* codes) /* scene box of the points, then
                                                                                  AABB::encodeMortonCode of each */
{
    orc_aabb box;
    aabb_empty(&box);
    for (int i = 0; i < n; i++) aabb_extend_p(&box, V(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
    for (int i = 0; i < n; i++) codes[i] = orc_aabb_encode_morton(&box, V(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
    boxOut[0] = box.lower.x, boxOut[1] = box.lower.y, boxOut[2] = box.lower.z;
    boxOut[3] = box.upper.x, boxOut[4] = box.upper.y, boxOut[5] = box.upper.z;
}

/* records: low 32 bits = morton key, high 32 bits = triangle index, the
   little-endian view of IndexedMorton{morton, index}.  minimum_lbvh.h:534-553 */
void orc_morton_records(const orc_tri* tris, int n, const orc_aabb* scene, uint32_t opt, uint64_t* records)
{
    for (int i = 0; i < n; i++) {
        orc_tri tri = tris[i];
        orc_v3 center = vdivs(vadd(vadd(tri.vs[0], tri.vs[1]), tri.vs[2]), 3.0f);
        uint32_t morton;
        if (opt & ORC_OPT_USE_NORMAL) {
            /* normalOf, minimum_lbvh.h:212-215 */
            orc_v3 ng = vnormalize(vcross(vsub(tri.vs[1], tri.vs[0]), vsub(tri.vs[2], tri.vs[0])));
            orc_aabb unit = {{-1, -1, -1}, {+1, +1, +1}};
            uint32_t nMorton = (uint32_t)(orc_aabb_encode_morton(&unit, ng) >> 31);
            uint32_t pMorton = (uint32_t)(orc_aabb_encode_morton(scene, center) >> 31);
            morton = (pMorton & 0xFFFFFFC0u) | (nMorton >> 26);
        } else {
            morton = (uint32_t)(orc_aabb_encode_morton(scene, center) >> 31);
        }
        records[i] = (uint64_t)morton | ((uint64_t)(uint32_t)i << 32);
    }
}

/* Stable order by (morton, original index).  The reference CPU uses an
   unstable std::sort (minimum_lbvh.h:554); the reference GPU path sorts the
   same records with a stable LSD radix sort over bits [0,32)
   (minimum_lbvh.h:833), which is what is restated here: a 4-pass LSD. */
void orc_sort_records(uint64_t* records, int n)
{
    orc_radix_sort_u64(records, NULL, (uint32_t)n, 0, 32);
}

void orc_deltas(const uint64_t* sorted, int n, uint8_t* deltas) /* :556-561 */
{
    for (int i = 0; i + 1 < n; i++)
        deltas[i] = (uint8_t)orc_delta32((uint32_t)sorted[i], (uint32_t)sorted[i + 1]);
}

/* ---------------------------------------------------------------------- */
/* Agglomerative bottom-up climb for one leaf, minimum_lbvh.h:284-361.     */
/* ---------------------------------------------------------------------- */
void orc_build_leaf(uint32_t* root, orc_node* nodes, const orc_tri* tris, int n, const uint8_t* deltas,
                    uint32_t i_leaf, uint32_t triIndex)
{
    int nInternals = n - 1;
    int nDeltas = nInternals;
    uint32_t leaf_lower = i_leaf;
    uint32_t leaf_upper = i_leaf;
    uint32_t node = triIndex | ORC_LEAF_BIT;

    orc_aabb aabb;
    aabb_empty(&aabb);
    if (tris)
        for (int k = 0; k < 3; k++) aabb_extend_p(&aabb, tris[triIndex].vs[k]);

    int isRoot = 1;
    while (leaf_upper - leaf_lower < (uint32_t)nInternals) {
        uint32_t deltaL = 0 < leaf_lower ? deltas[leaf_lower - 1] : 0xFFFFFFFFu;
        uint32_t deltaR = leaf_upper < (uint32_t)nDeltas ? deltas[leaf_upper] : 0xFFFFFFFFu;
        int goLeft = deltaL < deltaR ? 1 : 0;
        uint32_t parent = goLeft ? (leaf_lower - 1) : leaf_upper;

        nodes[parent].children[goLeft] = node;
        nodes[parent].aabbs[goLeft] = aabb;
        if (!(node & ORC_LEAF_BIT)) nodes[node].parent = parent; /* NodeIndex(parent, false) */

        uint32_t index = goLeft ? leaf_upper : leaf_lower;
        index = __atomic_exchange_n(&nodes[parent].context, index, __ATOMIC_SEQ_CST);
        if (index == 0xFFFFFFFFu) {
            isRoot = 0;
            break;
        }
        leaf_lower = leaf_lower < index ? leaf_lower : index; /* ss_min */
        leaf_upper = leaf_upper < index ? index : leaf_upper; /* ss_max */
        nodes[parent].context = leaf_upper < (uint32_t)nInternals ? leaf_upper : 0xFFFFFFFFu;
        node = parent;
        orc_aabb other = nodes[parent].aabbs[goLeft ^ 1];
        aabb_extend_b(&aabb, &other);
    }
    if (isRoot) {
        *root = node;
        nodes[node & ~ORC_LEAF_BIT].parent = ORC_NODE_INVALID;
    }
}

/* The 63-bit key variant (build option ORC_OPT_MORTON64, no reference
   counterpart).  The reference keeps the top 32 of the 63 Morton bits
   (minimum_lbvh.h:550), so at scale many centroids share a key.  Here the
   full code of aabb.encodeMortonCode(center) is the key: pairs (code, index)
   sorted stably by code over bits [0, 63) (8 digit passes), deltas
   64 - clz64(a ^ b) (0 for equal codes, at most 63).  The exported records
   keep the layout of the 32-bit variant: high word the triangle index, low
   word code >> 31. */
static void orc_morton64_sorted(const orc_tri* tris, int n, const orc_aabb* scene, uint64_t* rec, uint8_t* deltas)
{
    uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n);
    uint32_t* idx = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    for (int i = 0; i < n; i++) {
        orc_tri tri = tris[i];
        orc_v3 center = vdivs(vadd(vadd(tri.vs[0], tri.vs[1]), tri.vs[2]), 3.0f);
        keys[i] = orc_aabb_encode_morton(scene, center);
        idx[i] = (uint32_t)i;
    }
    orc_radix_sort_u64(keys, idx, (uint32_t)n, 0, 63);
    for (int i = 0; i + 1 < n; i++) {
        uint64_t x = keys[i] ^ keys[i + 1];
        deltas[i] = (uint8_t)(x ? 64 - __builtin_clzll(x) : 0);
    }
    for (int i = 0; i < n; i++) rec[i] = (uint64_t)(uint32_t)(keys[i] >> 31) | ((uint64_t)idx[i] << 32);
    free(keys);
    free(idx);
}

/* BVHCPUBuilder::build, minimum_lbvh.h:512-592.  n < 2 is rejected (the
   reference writes out of bounds for n == 1 at :359), and so is
   ORC_OPT_MORTON64 together with ORC_OPT_USE_NORMAL. */
int orc_build(const orc_tri* tris, int n, uint32_t opt, orc_node* nodes, uint32_t* root, uint64_t* sorted_out,
              uint8_t* deltas_out, int nthreads)
{
    if (n < 2) return -1;
    if ((opt & ORC_OPT_MORTON64) && (opt & ORC_OPT_USE_NORMAL)) return -1;
    for (int i = 0; i < n - 1; i++) {
        memset(&nodes[i], 0, sizeof(orc_node));
        nodes[i].context = 0xFFFFFFFFu;
    }
    orc_aabb scene;
    orc_scene_aabb(tris, n, &scene);
    uint64_t* rec = sorted_out ? sorted_out : (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n);
    uint8_t* deltas = deltas_out ? deltas_out : (uint8_t*)malloc((size_t)n);
    if (opt & ORC_OPT_MORTON64) {
        orc_morton64_sorted(tris, n, &scene, rec, deltas);
    } else {
        orc_morton_records(tris, n, &scene, opt, rec);
        if (opt & ORC_OPT_COMPARISON_SORT)
            orc_sort_records_comparison(rec, n);
        else
            orc_sort_records(rec, n);
        orc_deltas(rec, n, deltas);
    }
    if (opt & ORC_OPT_CPU_PARALLEL) {
        int nt = pick_threads(nthreads);
#pragma omp parallel for schedule(static) num_threads(nt)
        for (int i = 0; i < n; i++) orc_build_leaf(root, nodes, tris, n, deltas, (uint32_t)i, (uint32_t)(rec[i] >> 32));
    } else {
        for (int i = 0; i < n; i++) orc_build_leaf(root, nodes, tris, n, deltas, (uint32_t)i, (uint32_t)(rec[i] >> 32));
    }
    if (!sorted_out) free(rec);
    if (!deltas_out) free(deltas);
    return 0;
}

/* validate_lbvh, minimum_lbvh.h:415-426, with an explicit stack (a chain of
   equal keys makes the tree n-1 deep).  Returns 1 when valid. */
int orc_validate(uint32_t root, const orc_node* nodes, const uint8_t* deltas, int maxDelta)
{
    if (root & ORC_LEAF_BIT) return 1;
    size_t cap = 1024, sp = 0;
    uint32_t* st = (uint32_t*)malloc(cap * 2 * sizeof(uint32_t));
    st[0] = root;
    st[1] = (uint32_t)maxDelta;
    sp = 1;
    int ok = 1;
    while (sp) {
        sp--;
        uint32_t node = st[2 * sp];
        int md = (int)st[2 * sp + 1];
        if (node & ORC_LEAF_BIT) continue;
        int d = deltas[node];
        if (!(d <= md)) {
            ok = 0;
            break;
        }
        if (sp + 2 > cap) {
            cap *= 2;
            st = (uint32_t*)realloc(st, cap * 2 * sizeof(uint32_t));
        }
        for (int c = 0; c < 2; c++) {
            st[2 * sp] = nodes[node].children[c];
            st[2 * sp + 1] = (uint32_t)d;
            sp++;
        }
    }
    free(st);
    return ok;
}

/* ---------------------------------------------------------------------- */
/* Stable LSD radix sort with the digit schedule of OnesweepSort::sort:    */
/* ceil((endBit-startBit)/8) passes of 8-bit digits starting at startBit   */
/* (tinyhiponesweep.h:606, :227 bitLocation = startBits + 8*iteration).    */
/* Digits at or beyond the key width read as 0.  Returns the pass count,   */
/* or -1 on a bad bit range.                                               */
/* ---------------------------------------------------------------------- */
#define RADIX_SORT_BODY(KT, NBITS)                                                                     \
    if (endBit <= startBit || startBit < 0) return -1;                                                 \
    int passes = (endBit - startBit + 7) / 8;                                                          \
    if (n == 0) return passes;                                                                         \
    KT* k2 = (KT*)malloc(sizeof(KT) * (size_t)n);                                                      \
    uint32_t* v2 = vals ? (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n) : NULL;                      \
    KT* ks = keys;                                                                                     \
    uint32_t* vs = vals;                                                                               \
    KT* kd = k2;                                                                                       \
    uint32_t* vd = v2;                                                                                 \
    for (int p = 0; p < passes; p++) {                                                                 \
        int shift = startBit + 8 * p;                                                                  \
        size_t cnt[257];                                                                               \
        memset(cnt, 0, sizeof(cnt));                                                                   \
        for (uint32_t i = 0; i < n; i++) {                                                             \
            uint32_t d = shift < NBITS ? (uint32_t)((ks[i] >> shift) & 0xFF) : 0;                      \
            cnt[d + 1]++;                                                                              \
        }                                                                                              \
        for (int d = 0; d < 256; d++) cnt[d + 1] += cnt[d];                                            \
        for (uint32_t i = 0; i < n; i++) {                                                             \
            uint32_t d = shift < NBITS ? (uint32_t)((ks[i] >> shift) & 0xFF) : 0;                      \
            size_t o = cnt[d]++;                                                                       \
            kd[o] = ks[i];                                                                             \
            if (vs) vd[o] = vs[i];                                                                     \
        }                                                                                              \
        KT* tk = ks; ks = kd; kd = tk;                                                                 \
        uint32_t* tv = vs; vs = vd; vd = tv;                                                           \
    }                                                                                                  \
    if (ks != keys) {                                                                                  \
        memcpy(keys, ks, sizeof(KT) * (size_t)n);                                                      \
        if (vals) memcpy(vals, vs, sizeof(uint32_t) * (size_t)n);                                      \
    }                                                                                                  \
    free(k2);                                                                                          \
    free(v2);                                                                                          \
    return passes;

int orc_radix_sort_u32(uint32_t* keys, uint32_t* vals, uint32_t n, int startBit, int endBit)
{
    RADIX_SORT_BODY(uint32_t, 32)
}
int orc_radix_sort_u64(uint64_t* keys, uint32_t* vals, uint32_t n, int startBit, int endBit)
{
    RADIX_SORT_BODY(uint64_t, 64)
}

/* ---------------------------------------------------------------------- */
/* Ray/box, ray/triangle, traversals.  minimum_lbvh.h:82-96, 363-413,      */
/* 890-1137.                                                               */
/* ---------------------------------------------------------------------- */
orc_v3 orc_inv_rd(orc_v3 rd) /* :898-901, clamp = max(a, min(f, b)) */
{
    orc_v3 r = V(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    r.x = hmax(-FLT_MAX_H, hmin(r.x, FLT_MAX_H));
    r.y = hmax(-FLT_MAX_H, hmin(r.y, FLT_MAX_H));
    r.z = hmax(-FLT_MAX_H, hmin(r.z, FLT_MAX_H));
    return r;
}

void orc_slabs(orc_v3 ro, orc_v3 inv, orc_v3 lower, orc_v3 upper, float knownHitT, float* rmin, float* rmax)
{
    orc_v3 t0 = vmul(vsub(lower, ro), inv);
    orc_v3 t1 = vmul(vsub(upper, ro), inv);
    orc_v3 tmin = vmin(t0, t1);
    orc_v3 tmax = vmax(t0, t1);
    float region_min = hmax(hmax(tmin.x, tmin.y), tmin.z);            /* compMax */
    float region_max = hmin(hmin(tmax.x, tmax.y), tmax.z) * 1.00000024f; /* compMin * robust factor */
    region_min = hmax(region_min, 0.0f);
    region_max = hmin(region_max, knownHitT);
    *rmin = region_min;
    *rmax = region_max;
}

int orc_intersect_tri(float* tOut, float* uOut, float* vOut, orc_v3* ngOut, float t_min, float t_max, orc_v3 ro,
                      orc_v3 rd, orc_v3 v0, orc_v3 v1, orc_v3 v2)
{
    orc_v3 e0 = vsub(v1, v0);
    orc_v3 e1 = vsub(v2, v1);
    orc_v3 ng = vcross(e0, e1);
    float t = vdot(vsub(v0, ro), ng) / vdot(ng, rd);
    if (t_min <= t && t <= t_max) {
        orc_v3 e2 = vsub(v0, v2);
        orc_v3 ro2 = vmuls(ro, 2.0f);
        float u_vol = vdot(vcross(e2, vsub(vadd(v0, v2), ro2)), rd);
        float v_vol = vdot(vcross(e0, vsub(vadd(v1, v0), ro2)), rd);
        float w_vol = vdot(vcross(e1, vsub(vadd(v2, v1), ro2)), rd);
        if (hmin(hmin(u_vol, v_vol), w_vol) < 0.0f && 0.0f < hmax(hmax(u_vol, v_vol), w_vol)) return 0;
        float vol = u_vol + v_vol + w_vol;
        *tOut = t;
        *uOut = u_vol / vol;
        *vOut = v_vol / vol;
        *ngOut = ng;
        return 1;
    }
    return 0;
}

static inline void hit_reset(orc_hit* h, float t)
{
    memset(h, 0, sizeof(*h));
    h->t = t;
    h->triangleIndex = 0xFFFFFFFFu;
}

/* Parent-pointer stack-free traversal, minimum_lbvh.h:981-1061.
   counters (nullable): [0] += internal-node fetches, [1] += triangle tests.
   earlyExit: also track the far children still owed on the current path and
   stop once the walk would climb with none owed (the remaining steps would
   only climb to the root); the leaves tested and the hit are unchanged. */
void orc_intersect_stackfree_ex(orc_hit* hit, const orc_node* nodes, const orc_tri* tris, uint32_t root, orc_v3 ro,
                                orc_v3 rd, orc_v3 inv, int query, uint64_t* counters, int earlyExit)
{
    uint32_t curr = root;
    uint32_t prev = ORC_NODE_INVALID;
    uint64_t nFetch = 0, nTest = 0;
    uint32_t pend = 0;
    while (curr != ORC_NODE_INVALID) {
        if (curr & ORC_LEAF_BIT) {
            float t, u, v;
            orc_v3 ng;
            const orc_tri* tri = &tris[curr & ~ORC_LEAF_BIT];
            nTest++;
            if (orc_intersect_tri(&t, &u, &v, &ng, 0.0f, hit->t, ro, rd, tri->vs[0], tri->vs[1], tri->vs[2])) {
                hit->t = t;
                hit->u = u;
                hit->v = v;
                hit->triangleIndex = curr & ~ORC_LEAF_BIT;
                hit->ngx = ng.x;
                hit->ngy = ng.y;
                hit->ngz = ng.z;
                if (query == ORC_QUERY_ANY) break;
            }
            uint32_t tmp = curr;
            curr = prev;
            prev = tmp;
            continue;
        }
        const orc_node* nd = &nodes[curr];
        nFetch++;
        float lmin, lmax, rmin, rmax;
        orc_slabs(ro, inv, nd->aabbs[0].lower, nd->aabbs[0].upper, hit->t, &lmin, &lmax);
        orc_slabs(ro, inv, nd->aabbs[1].lower, nd->aabbs[1].upper, hit->t, &rmin, &rmax);
        int hitL = lmin <= lmax;
        int hitR = rmin <= rmax;
        uint32_t parent_node = nd->parent;
        uint32_t near_node = nd->children[0];
        uint32_t far_node = nd->children[1];
        int nHits = 0;
        if (hitL && hitR) {
            if (rmin < lmin) {
                uint32_t t = near_node;
                near_node = far_node;
                far_node = t;
            }
            nHits = 2;
        } else if (hitL || hitR) {
            nHits = 1;
            near_node = hitR ? far_node : near_node;
        }
        uint32_t next;
        if (prev == parent_node) {
            next = 0 < nHits ? near_node : parent_node;
            if (nHits == 2) pend++;
        } else if (prev == near_node) {
            next = nHits == 2 ? far_node : parent_node;
            if (nHits == 2) pend--;
        } else {
            next = parent_node;
        }
        if (earlyExit && next == parent_node && pend == 0) break;
        prev = curr;
        curr = next;
    }
    if (counters) {
        counters[0] += nFetch;
        counters[1] += nTest;
    }
}

void orc_intersect_stackfree(orc_hit* hit, const orc_node* nodes, const orc_tri* tris, uint32_t root, orc_v3 ro,
                             orc_v3 rd, orc_v3 inv, int query, uint64_t* counters)
{
    orc_intersect_stackfree_ex(hit, nodes, tris, root, ro, rd, inv, query, counters, 0);
}

/* Explicit-stack traversal, minimum_lbvh.h:912-973 (closest hit). */
void orc_intersect_stackfull(orc_hit* hit, const orc_node* nodes, const orc_tri* tris, uint32_t root, orc_v3 ro,
                             orc_v3 rd, orc_v3 inv, uint32_t* stack)
{
    int sp = 0;
    uint32_t node = root;
    while (node != ORC_NODE_INVALID) {
        if (node & ORC_LEAF_BIT) {
            float t, u, v;
            orc_v3 ng;
            const orc_tri* tri = &tris[node & ~ORC_LEAF_BIT];
            if (orc_intersect_tri(&t, &u, &v, &ng, 0.0f, hit->t, ro, rd, tri->vs[0], tri->vs[1], tri->vs[2])) {
                hit->t = t;
                hit->u = u;
                hit->v = v;
                hit->ngx = ng.x;
                hit->ngy = ng.y;
                hit->ngz = ng.z;
                hit->triangleIndex = node & ~ORC_LEAF_BIT;
            }
            node = sp == 0 ? ORC_NODE_INVALID : stack[--sp];
            continue;
        }
        const orc_node* nd = &nodes[node];
        float lmin, lmax, rmin, rmax;
        orc_slabs(ro, inv, nd->aabbs[0].lower, nd->aabbs[0].upper, hit->t, &lmin, &lmax);
        orc_slabs(ro, inv, nd->aabbs[1].lower, nd->aabbs[1].upper, hit->t, &rmin, &rmax);
        int hitL = lmin <= lmax;
        int hitR = rmin <= rmax;
        if (hitL && hitR) {
            uint32_t nearN = nd->children[0], farN = nd->children[1];
            if (rmin < lmin) {
                uint32_t t = nearN;
                nearN = farN;
                farN = t;
            }
            node = nearN;
            stack[sp++] = farN;
        } else if (hitL || hitR) {
            node = nd->children[hitL ? 0 : 1];
        } else {
            node = sp == 0 ? ORC_NODE_INVALID : stack[--sp];
        }
    }
}

/* Escape-link traversal using `context` as the miss link,
   minimum_lbvh.h:1063-1137.  counters (may be NULL): [0] internal-node
   visits, [1] triangle tests. */
void orc_intersect_escape_link(orc_hit* hit, const orc_node* nodes, const orc_tri* tris, uint32_t root, orc_v3 ro,
                               orc_v3 rd, orc_v3 inv, int query, uint64_t* counters)
{
    uint32_t curr = root;
    uint32_t prev = ORC_NODE_INVALID;
    while (curr != ORC_NODE_INVALID) {
        if (curr & ORC_LEAF_BIT) {
            float t, u, v;
            orc_v3 ng;
            const orc_tri* tri = &tris[curr & ~ORC_LEAF_BIT];
            if (counters) counters[1]++;
            if (orc_intersect_tri(&t, &u, &v, &ng, 0.0f, hit->t, ro, rd, tri->vs[0], tri->vs[1], tri->vs[2])) {
                hit->t = t;
                hit->u = u;
                hit->v = v;
                hit->triangleIndex = curr & ~ORC_LEAF_BIT;
                hit->ngx = ng.x;
                hit->ngy = ng.y;
                hit->ngz = ng.z;
                if (query == ORC_QUERY_ANY) break;
            }
            if (nodes[prev & ~ORC_LEAF_BIT].children[0] == curr) {
                uint32_t tmp = curr;
                curr = prev;
                prev = tmp;
                continue;
            }
            uint32_t idx = nodes[prev & ~ORC_LEAF_BIT].context;
            uint32_t next = idx != 0xFFFFFFFFu ? idx : ORC_NODE_INVALID;
            prev = curr;
            curr = next;
            continue;
        }
        const orc_node* nd = &nodes[curr];
        if (counters) counters[0]++;
        uint32_t parent_node = nd->parent;
        orc_aabb box;
        uint32_t hitChild, miss;
        if (prev == parent_node) {
            box = nd->aabbs[0];
            hitChild = nd->children[0];
            miss = curr;
        } else {
            box = nd->aabbs[1];
            hitChild = nd->children[1];
            uint32_t idx = nd->context;
            miss = idx != 0xFFFFFFFFu ? idx : ORC_NODE_INVALID;
        }
        float rmin, rmax;
        orc_slabs(ro, inv, box.lower, box.upper, hit->t, &rmin, &rmax);
        uint32_t next = rmin <= rmax ? hitChild : miss;
        prev = curr;
        curr = next;
    }
}

void orc_trace(const orc_node* nodes, const orc_tri* tris, uint32_t root, const float* rays, int nRays, int query,
               int method, orc_hit* hits, uint64_t* counters, int nthreads)
{
    int nt = pick_threads(nthreads);
    uint64_t c0 = 0, c1 = 0;
#pragma omp parallel num_threads(nt) reduction(+ : c0, c1)
    {
        uint32_t* stack = NULL;
        if (method == 1) stack = (uint32_t*)malloc(sizeof(uint32_t) * 4096u * 64u);
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < nRays; i++) {
            const float* r = rays + 8 * (size_t)i;
            orc_v3 ro = V(r[0], r[1], r[2]);
            orc_v3 rd = V(r[4], r[5], r[6]);
            orc_hit h;
            hit_reset(&h, r[3]);
            uint64_t cnt[2] = {0, 0};
            if (method == 0 || method == 3)
                orc_intersect_stackfree_ex(&h, nodes, tris, root, ro, rd, orc_inv_rd(rd), query, cnt, method == 3);
            else if (method == 1)
                orc_intersect_stackfull(&h, nodes, tris, root, ro, rd, orc_inv_rd(rd), stack);
            else if (method == 4)
                orc_closest_hit_stack(&h, nodes, tris, root, ro, rd, orc_inv_rd(rd), query, cnt, ORC_LANE_STACK);
            else
                orc_intersect_escape_link(&h, nodes, tris, root, ro, rd, orc_inv_rd(rd), query, cnt);
            hits[i] = h;
            c0 += cnt[0];
            c1 += cnt[1];
        }
        free(stack);
    }
    if (counters) {
        counters[0] += c0;
        counters[1] += c1;
    }
}

/* Closest hit by testing every triangle in index order (the last of equal-t
   hits in index order wins). */
void orc_brute_force(const orc_tri* tris, int n, const float* rays, int nRays, orc_hit* hits)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < nRays; i++) {
        const float* r = rays + 8 * (size_t)i;
        orc_v3 ro = V(r[0], r[1], r[2]);
        orc_v3 rd = V(r[4], r[5], r[6]);
        orc_hit h;
        hit_reset(&h, r[3]);
        for (int k = 0; k < n; k++) {
            float t, u, v;
            orc_v3 ng;
            if (orc_intersect_tri(&t, &u, &v, &ng, 0.0f, h.t, ro, rd, tris[k].vs[0], tris[k].vs[1], tris[k].vs[2])) {
                h.t = t;
                h.u = u;
                h.v = v;
                h.ngx = ng.x;
                h.ngy = ng.y;
                h.ngz = ng.z;
                h.triangleIndex = (uint32_t)k;
            }
        }
        hits[i] = h;
    }
}

/* ---------------------------------------------------------------------- */
/* Sampling + camera.                                                      */
/* ---------------------------------------------------------------------- */
uint32_t orc_hash_pcg(uint32_t v) /* minimum_lbvh.h:904-909, sobol.h:76-81 */
{
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28) + 4)) ^ state) * 277803737u;
    return (word >> 22) ^ word;
}

static inline uint32_t rev32(uint32_t v) /* sobol.h:18-30 */
{
    v = ((v >> 1) & 0x55555555u) | ((v & 0x55555555u) << 1);
    v = ((v >> 2) & 0x33333333u) | ((v & 0x33333333u) << 2);
    v = ((v >> 4) & 0x0F0F0F0Fu) | ((v & 0x0F0F0F0Fu) << 4);
    v = ((v >> 8) & 0x00FF00FFu) | ((v & 0x00FF00FFu) << 8);
    return (v >> 16) | (v << 16);
}
static inline uint32_t lk_perm(uint32_t x, uint32_t seed) /* Laine-Karras, sobol.h:31-39 */
{
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}
static inline uint32_t nested_scramble(uint32_t x, uint32_t seed) /* sobol.h:40-46 */
{
    return rev32(lk_perm(rev32(x), seed));
}
static inline uint32_t pascal(uint32_t v) /* Pascal matrix, sobol.h:51-59 */
{
    v ^= v << 16;
    v ^= (v & 0x00FF00FFu) << 8;
    v ^= (v & 0x0F0F0F0Fu) << 4;
    v ^= (v & 0x33333333u) << 2;
    v ^= (v & 0x55555555u) << 1;
    return v;
}
static inline float u2f(uint32_t u) { return (float)(u >> 8) / (float)((0xFFFFFFFFu >> 8) + 1); } /* :62-72 */
static inline void pcg3d(uint32_t* x, uint32_t* y, uint32_t* z) /* sobol.h:83-107 */
{
    uint32_t vx = *x * 1664525u + 1013904223u;
    uint32_t vy = *y * 1664525u + 1013904223u;
    uint32_t vz = *z * 1664525u + 1013904223u;
    vx += vy * vz;
    vy += vz * vx;
    vz += vx * vy;
    vx ^= vx >> 16u;
    vy ^= vy >> 16u;
    vz ^= vz >> 16u;
    vx += vy * vz;
    vy += vz * vx;
    vz += vx * vy;
    *x = vx;
    *y = vy;
    *z = vz;
}

void orc_sobol_2d(float* x, float* y, uint32_t index, uint32_t p0, uint32_t p1, uint32_t p2) /* sobol.h:119-129 */
{
    pcg3d(&p0, &p1, &p2);
    uint32_t v = lk_perm(rev32(index), p0);
    *x = u2f(nested_scramble(v, p1));
    *y = u2f(nested_scramble(pascal(v), p2));
}

void orc_raygen_lookat(orc_raygen* g, orc_v3 eye, orc_v3 center, orc_v3 up, float fovy, int w, int h) /* camera.h:10-22 */
{
    orc_v3 f = vnormalize(vsub(center, eye));
    orc_v3 s = vnormalize(vcross(f, up));
    orc_v3 u = vcross(s, f);
    float tanThetaY = tanf(fovy * 0.5f);
    float tanThetaX = tanThetaY / (float)h * (float)w;
    g->origin = eye;
    g->right = vmuls(s, tanThetaX);
    g->up = vmuls(u, tanThetaY);
}

void orc_raygen_shoot(const orc_raygen* g, orc_v3* ro, orc_v3* rd, float u, float v) /* camera.h:24-31 */
{
    orc_v3 from = g->origin;
    orc_v3 forward = vnormalize(vcross(g->up, g->right));
    orc_v3 to = vadd(vadd(vadd(g->origin, forward), vlerp(vneg(g->right), g->right, u)), vlerp(g->up, vneg(g->up), v));
    *ro = from;
    *rd = vsub(to, from);
}

/* ---------------------------------------------------------------------- */
/* Render kernels restated for the host, main_gpu.cu:17-286.  The bounds   */
/* test uses || (the reference's && at :116/:184 admits out-of-range       */
/* pixels).                                                                */
/* ---------------------------------------------------------------------- */
uint32_t orc_pack_rgba(float r, float g, float b, float a) /* main_gpu.cu:17-22 */
{
    (void)a;
    int x = iclamp(orc_f2i_trunc(r * 255.0f + 0.5f), 0, 255);
    int y = iclamp(orc_f2i_trunc(g * 255.0f + 0.5f), 0, 255);
    int z = iclamp(orc_f2i_trunc(b * 255.0f + 0.5f), 0, 255);
    return ((uint32_t)z << 16) | ((uint32_t)y << 8) | (uint32_t)x;
}

/* ---------------------------------------------------------------------- */
/* sin / cos / pow in place of libm's (main_gpu.cu:30-31, :174, :283).      */
/* Same operations in the same order as the kernels' functions, so host and */
/* GPU agree bit for bit.  Published methods:                               */
/*  - sin/cos: Cody-Waite reduction (Cody & Waite 1980) by pi/2 split into  */
/*    P1 + P2 + P3 (P1, P2: 13 significant bits, so k*P1, k*P2 are exact    */
/*    for |k| < 2^11), then the Taylor polynomials of sin to r^9 and cos to */
/*    r^10 on |r| <= pi/4, Horner form, quadrant swap/sign.                 */
/*  - pow(x, y), x >= 0, y > 0: 2^(y log2 x); log2 x = e + ln(m)/ln 2 with  */
/*    m in [sqrt(1/2), sqrt 2] and ln m = 2 atanh((m-1)/(m+1)) by its odd   */
/*    series to s^9; y*e split exactly (y cut to 12 significant bits);      */
/*    2^f, |f| <= 1/2, by the Taylor series to f^7 with coefficients        */
/*    (ln 2)^n / n!; scaled by 2^k in two exact steps.                      */
/* Pinned by tests/test_oracle_math.py: the constants against their exact   */
/* values, and the functions against float64 sin / cos / pow.               */
/* ---------------------------------------------------------------------- */
static inline float f_bits(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static inline uint32_t bits_f(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }

void orc_sincos(float x, float* s, float* c)
{
    const float P1 = 0x1.921p+0f, P2 = 0x1.f6ap-13f, P3 = 0x1.110b46p-26f;
    const float kf = rintf(x * 0x1.45f306p-1f);
    const int k = (int)kf;
    const float r = ((x - kf * P1) - kf * P2) - kf * P3;
    const float r2 = r * r;
    const float sp = r + (r * r2) * (-0x1.555556p-3f + r2 * (0x1.111112p-7f + r2 * (-0x1.a01a02p-13f +
                                                                                       r2 * 0x1.71de3ap-19f)));
    const float cp = 1.0f + r2 * (-0x1p-1f + r2 * (0x1.555556p-5f + r2 * (-0x1.6c16c2p-10f +
                                                                           r2 * (0x1.a01a02p-16f +
                                                                                 r2 * -0x1.27e4fcp-22f))));
    switch (k & 3) {
        case 0: *s = sp; *c = cp; break;
        case 1: *s = cp; *c = -sp; break;
        case 2: *s = -sp; *c = -cp; break;
        default: *s = -cp; *c = sp; break;
    }
}

static float log2_parts(float x, int* eOut) /* x > 0, finite: log2 x = e + L, |L| <= 1/2 */
{
    uint32_t b = bits_f(x);
    int e = 0;
    if (b < 0x00800000u) {
        x = x * 0x1p24f;
        b = bits_f(x);
        e = -24;
    }
    e += (int)(b >> 23) - 127;
    float m = f_bits((b & 0x007FFFFFu) | 0x3F800000u);
    if (m > 0x1.6a09e6p+0f) {
        m = m * 0.5f;
        e += 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    const float lnm = s * (2.0f + s2 * (0x1.555556p-1f + s2 * (0x1.99999ap-2f + s2 * (0x1.24924ap-2f +
                                                                                       s2 * 0x1.c71c72p-3f))));
    *eOut = e;
    return lnm * 0x1.715476p+0f;
}

static float exp2_parts(float kf, float f) /* 2^(k + f), k integer-valued, |f| <= 1/2 */
{
    if (kf > 128.0f) return f_bits(0x7F800000u);
    if (kf < -151.0f) return 0.0f;
    const float p = 1.0f + f * (0x1.62e430p-1f + f * (0x1.ebfbe0p-3f + f * (0x1.c6b08ep-5f + f * (0x1.3b2ab6p-7f +
                    f * (0x1.5d87fep-10f + f * (0x1.430912p-13f + f * 0x1.ffcbfcp-17f))))));
    const int k = (int)kf;
    const int k1 = k >> 1, k2 = k - k1;
    return (p * f_bits((uint32_t)(k1 + 127) << 23)) * f_bits((uint32_t)(k2 + 127) << 23);
}

float orc_log2_pos(float x) /* x > 0, finite */
{
    int e;
    const float L = log2_parts(x, &e);
    return (float)e + L;
}

float orc_exp2_fin(float z) /* finite z */
{
    const float kf = rintf(z);
    return exp2_parts(kf, z - kf);
}

float orc_pow_pos(float x, float y) /* y > 0; x <= 0 gives 0 (libm: NaN for x < 0, which packs to the same 0) */
{
    if (x != x) return x;
    if (!(x > 0.0f)) return 0.0f;
    if (x == f_bits(0x7F800000u)) return x;
    int e;
    const float L = log2_parts(x, &e);
    const float yh = f_bits(bits_f(y) & 0xFFFFF000u); /* 12 significant bits: yh * e is exact */
    const float yl = y - yh;
    const float fe = (float)e;
    const float zh = yh * fe;
    float kf = rintf(zh);
    float f = (zh - kf) + (yl * fe + y * L);
    const float k2 = rintf(f);
    f = f - k2;
    kf = kf + k2;
    return exp2_parts(kf, f);
}

void orc_sincos_batch(const float* x, int n, float* s, float* c)
{
    for (int i = 0; i < n; i++) orc_sincos(x[i], &s[i], &c[i]);
}

void orc_pow_batch(const float* x, int n, float y, float* out)
{
    for (int i = 0; i < n; i++) out[i] = orc_pow_pos(x[i], y);
}

static orc_v3 hemisphere_cos(float xi0, float xi1) /* main_gpu.cu:24-36 */
{
    float phi = xi0 * 2.0f * ORC_PI;
    float r = sqrtf(xi1);
    float sp, cp;
    orc_sincos(phi, &sp, &cp);
    float x = cp * r;
    float z = sp * r;
    float y = sqrtf(fmaxf(1.0f - r * r, 0.0f));
    return V(x, y, z);
}

static void onb(orc_v3 z, orc_v3* xaxis, orc_v3* yaxis) /* Duff et al. 2017, main_gpu.cu:39-45 */
{
    const float sign = copysignf(1.0f, z.z);
    const float a = -1.0f / (sign + z.z);
    const float b = z.x * z.y * a;
    *xaxis = V(1.0f + sign * z.x * z.x * a, sign * b, -sign * z.x);
    *yaxis = V(b, sign + z.y * z.y * a, -z.y);
}

/* dir = xaxis * dl.x + zaxis * dl.z + n * dl.y, main_gpu.cu:153 and :235 */
static inline orc_v3 sample_dir(orc_v3 n, orc_v3 xaxis, orc_v3 zaxis, float r0, float r1)
{
    orc_v3 dl = hemisphere_cos(r0, r1);
    return vadd(vadd(vmuls(xaxis, dl.x), vmuls(zaxis, dl.z)), vmuls(n, dl.y));
}

void orc_onb(const float* z, float* out)
{
    orc_v3 x, y;
    onb(V(z[0], z[1], z[2]), &x, &y);
    out[0] = x.x, out[1] = x.y, out[2] = x.z;
    out[3] = y.x, out[4] = y.y, out[5] = y.z;
}

void orc_cosine_dir(const float* n, float xi0, float xi1, float* out)
{
    orc_v3 nn = V(n[0], n[1], n[2]), x, y;
    onb(nn, &x, &y);
    orc_v3 d = sample_dir(nn, x, y, xi0, xi1);
    out[0] = d.x, out[1] = d.y, out[2] = d.z;
}

typedef struct { uint64_t state, inc; } orc_pcg; /* main_gpu.cu:51-82 */
static uint32_t pcg_next(orc_pcg* p)
{
    uint64_t old = p->state;
    p->state = old * 6364136223846793005ULL + p->inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
}
static void pcg_init(orc_pcg* p, uint64_t seed, uint64_t seq)
{
    p->state = 0u;
    p->inc = (seq << 1u) | 1u;
    pcg_next(p);
    p->state += seed;
    pcg_next(p);
}
static float pcg_uniformf(orc_pcg* p)
{
    uint32_t bits = (pcg_next(p) >> 9) | 0x3f800000u;
    float f;
    memcpy(&f, &bits, 4);
    return f - 1.0f;
}

void orc_pcg32(uint64_t seed, uint64_t seq, int n, uint32_t* out)
{
    orc_pcg p;
    pcg_init(&p, seed, seq);
    for (int i = 0; i < n; i++) out[i] = pcg_next(&p);
}

void orc_pcg32_uniformf(uint64_t seed, uint64_t seq, int n, float* out)
{
    orc_pcg p;
    pcg_init(&p, seed, seq);
    for (int i = 0; i < n; i++) out[i] = pcg_uniformf(&p);
}

void orc_hemisphere_cos(float xi0, float xi1, float* out)
{
    orc_v3 v = hemisphere_cos(xi0, xi1);
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
}

void orc_render_normal(uint32_t* pixels, int w, int h, const orc_raygen* g, uint32_t root, const orc_node* nodes,
                       const orc_tri* tris, int nthreads) /* main_gpu.cu:84-109 */
{
    int nt = pick_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int yi = 0; yi < h; yi++) {
        for (int xi = 0; xi < w; xi++) {
            orc_v3 ro, rd;
            orc_raygen_shoot(g, &ro, &rd, (float)xi / (float)w, (float)yi / (float)h);
            orc_hit hit;
            hit_reset(&hit, FLT_MAX_H);
            orc_intersect_stackfree(&hit, nodes, tris, root, ro, rd, orc_inv_rd(rd), ORC_QUERY_CLOSEST, NULL);
            uint32_t px;
            if (hit.t != FLT_MAX_H) {
                orc_v3 n = vnormalize(V(hit.ngx, hit.ngy, hit.ngz));
                orc_v3 c = vmuls(vadd(n, V(1.0f, 1.0f, 1.0f)), 0.5f);
                px = orc_pack_rgba(c.x, c.y, c.z, 1.0f);
            } else {
                px = orc_pack_rgba(0, 0, 0, 1);
            }
            pixels[xi + yi * w] = px;
        }
    }
}

/* Any-hit walk with an explicit stack of at most `cap` entries: at a node whose
   two children are both entered the near child (smaller entry distance, the
   left one on ties) is taken and the far child pushed; a leaf or a node with
   no child entered pops.  A push past `cap` restarts the ray with the
   stack-free walk with early exit.  No reference counterpart: the reference
   traces its AO rays with intersect_stackfree (minimum_lbvh.h:981-1061); the
   answer (occluded or not) does not depend on the visit order, and the tests
   check it against that walk.  cnt: node fetches, triangle tests. */
int orc_any_hit_stack(const orc_node* nodes, const orc_tri* tris, uint32_t root, orc_v3 ro, orc_v3 rd, orc_v3 inv,
                      float tmax, uint64_t* cnt, int cap)
{
    uint32_t stack[64];
    int sp = 0;
    uint32_t node = root;
    if (cap > 64) cap = 64;
    for (;;) {
        if (node & ORC_LEAF_BIT) {
            float t, u, v;
            orc_v3 ng;
            const orc_tri* tri = &tris[node & ~ORC_LEAF_BIT];
            cnt[1]++;
            if (orc_intersect_tri(&t, &u, &v, &ng, 0.0f, tmax, ro, rd, tri->vs[0], tri->vs[1], tri->vs[2])) return 1;
            if (sp == 0) return 0;
            node = stack[--sp];
            continue;
        }
        const orc_node* nd = &nodes[node];
        cnt[0]++;
        float lmin, lmax, rmin, rmax;
        orc_slabs(ro, inv, nd->aabbs[0].lower, nd->aabbs[0].upper, tmax, &lmin, &lmax);
        orc_slabs(ro, inv, nd->aabbs[1].lower, nd->aabbs[1].upper, tmax, &rmin, &rmax);
        int hitL = lmin <= lmax;
        int hitR = rmin <= rmax;
        if (hitL && hitR) {
            uint32_t nearN = nd->children[0], farN = nd->children[1];
            if (rmin < lmin) {
                nearN = nd->children[1];
                farN = nd->children[0];
            }
            if (sp == cap) {
                orc_hit h;
                hit_reset(&h, tmax);
                orc_intersect_stackfree_ex(&h, nodes, tris, root, ro, rd, inv, ORC_QUERY_ANY, cnt, 1);
                return h.t != tmax;
            }
            stack[sp++] = farN;
            node = nearN;
        } else if (hitL || hitR) {
            node = nd->children[hitL ? 0 : 1];
        } else {
            if (sp == 0) return 0;
            node = stack[--sp];
        }
    }
}

/* Closest (or any) hit with an explicit stack of `cap` (<= 64) entries, each
   a far child and its entry distance.  It tests the leaves of
   intersect_stackfree (minimum_lbvh.h:981-1061) in the same order and so ends
   on the same Hit, ties included:
   - descending, both walks order the children by entry distance (rmin),
     which does not depend on hit.t;
   - back at a node from its near child, the stack-free walk re-runs both
     slab tests with the current hit.t and takes the far child iff it is still
     entered.  Its box's own exit bound held when it was pushed, so that is
     rmin_far <= hit.t, the test a popped entry takes here.  (A near child
     culled on the way back implies a culled far child: rmin_far >= rmin_near.)
   A push past cap restarts the walk stack-free from the root with the hit
   found so far.  That walk also ends on the same Hit: every triangle at the
   final t is tested in both walks, in the same depth-first order, and the
   last of them wins.  query ANY stops at the first hit.
   cnt: [0] internal-node fetches, [1] triangle tests. */
void orc_closest_hit_stack(orc_hit* hit, const orc_node* nodes, const orc_tri* tris, uint32_t root, orc_v3 ro,
                           orc_v3 rd, orc_v3 inv, int query, uint64_t* cnt, int cap)
{
    uint32_t stack[64];
    float entry[64];
    int sp = 0;
    uint32_t node = root;
    if (cap > 64) cap = 64;
    for (;;) {
        int pop = 0;
        if (node & ORC_LEAF_BIT) {
            float t, u, v;
            orc_v3 ng;
            const orc_tri* tri = &tris[node & ~ORC_LEAF_BIT];
            cnt[1]++;
            if (orc_intersect_tri(&t, &u, &v, &ng, 0.0f, hit->t, ro, rd, tri->vs[0], tri->vs[1], tri->vs[2])) {
                hit->t = t;
                hit->u = u;
                hit->v = v;
                hit->triangleIndex = node & ~ORC_LEAF_BIT;
                hit->ngx = ng.x;
                hit->ngy = ng.y;
                hit->ngz = ng.z;
                if (query == ORC_QUERY_ANY) return;
            }
            pop = 1;
        } else {
            const orc_node* nd = &nodes[node];
            cnt[0]++;
            float lmin, lmax, rmin, rmax;
            orc_slabs(ro, inv, nd->aabbs[0].lower, nd->aabbs[0].upper, hit->t, &lmin, &lmax);
            orc_slabs(ro, inv, nd->aabbs[1].lower, nd->aabbs[1].upper, hit->t, &rmin, &rmax);
            int hitL = lmin <= lmax;
            int hitR = rmin <= rmax;
            if (hitL && hitR) {
                uint32_t nearN = nd->children[0], farN = nd->children[1];
                float farMin = rmin;
                if (rmin < lmin) {
                    nearN = nd->children[1];
                    farN = nd->children[0];
                    farMin = lmin;
                }
                if (sp == cap) {
                    orc_intersect_stackfree_ex(hit, nodes, tris, root, ro, rd, inv, query, cnt, 1);
                    return;
                }
                stack[sp] = farN;
                entry[sp] = farMin;
                sp++;
                node = nearN;
            } else if (hitL || hitR) {
                node = nd->children[hitL ? 0 : 1];
            } else {
                pop = 1;
            }
        }
        if (pop) {
            for (;;) {
                if (sp == 0) return;
                sp--;
                if (entry[sp] <= hit->t) break;
            }
            node = stack[sp];
        }
    }
}

/* orc_closest_hit_stack over a batch of rays (8 floats each, tmax in slot 3). */
void orc_closest_hit_stack_batch(const orc_node* nodes, const orc_tri* tris, uint32_t root, const float* rays,
                                 int nRays, int query, int cap, orc_hit* hits, uint64_t* counters, int nthreads)
{
    int nt = pick_threads(nthreads);
    uint64_t c0 = 0, c1 = 0;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt) reduction(+ : c0, c1)
    for (int i = 0; i < nRays; i++) {
        const float* r = rays + 8 * (size_t)i;
        orc_v3 ro = V(r[0], r[1], r[2]), rd = V(r[4], r[5], r[6]);
        uint64_t cnt[2] = {0, 0};
        orc_hit h;
        hit_reset(&h, r[3]);
        orc_closest_hit_stack(&h, nodes, tris, root, ro, rd, orc_inv_rd(rd), query, cnt, cap);
        hits[i] = h;
        c0 += cnt[0];
        c1 += cnt[1];
    }
    if (counters) {
        counters[0] += c0;
        counters[1] += c1;
    }
}

/* orc_any_hit_stack over a batch of rays (8 floats each, tmax in slot 3). */
void orc_any_hit_stack_batch(const orc_node* nodes, const orc_tri* tris, uint32_t root, const float* rays, int nRays,
                             int cap, uint8_t* occluded, uint64_t* counters, int nthreads)
{
    int nt = pick_threads(nthreads);
    uint64_t c0 = 0, c1 = 0;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt) reduction(+ : c0, c1)
    for (int i = 0; i < nRays; i++) {
        const float* r = rays + 8 * (size_t)i;
        orc_v3 ro = V(r[0], r[1], r[2]), rd = V(r[4], r[5], r[6]);
        uint64_t cnt[2] = {0, 0};
        occluded[i] = (uint8_t)orc_any_hit_stack(nodes, tris, root, ro, rd, orc_inv_rd(rd), r[3], cnt, cap);
        c0 += cnt[0];
        c1 += cnt[1];
    }
    if (counters) {
        counters[0] += c0;
        counters[1] += c1;
    }
}

/* walk: 0 = the reference stack-free walk for every ray; 1 = stack-free with
   the early exit; 2 = stack-free with early exit for the primary ray and
   orc_any_hit_stack (stack of ORC_AO_STACK entries) for the AO rays; 3 = as
   2 with orc_closest_hit_stack for the primary ray. */
void orc_render_ao(uint32_t* pixels, float* aoOut, int w, int h, const orc_raygen* g, uint32_t root,
                   const orc_node* nodes, const orc_tri* tris, int useSobol, int nSample, float radius, uint32_t seed,
                   int row0, int row1, uint64_t* counters, int nthreads, int walk) /* main_gpu.cu:111-177 */
{
    const int earlyExit = walk != 0;
    int nt = pick_threads(nthreads);
    uint64_t cRays = 0, cFetch = 0, cTest = 0;
    if (row0 < 0) row0 = 0;
    if (row1 > h) row1 = h;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt) reduction(+ : cRays, cFetch, cTest)
    for (int yi = row0; yi < row1; yi++) {
        for (int xi = 0; xi < w; xi++) {
            int pixel = xi + yi * w;
            uint64_t cnt[2] = {0, 0};
            orc_v3 ro, rd;
            orc_raygen_shoot(g, &ro, &rd, (float)xi / (float)w, (float)yi / (float)h);
            orc_hit hit;
            hit_reset(&hit, FLT_MAX_H);
            if (walk == 3)
                orc_closest_hit_stack(&hit, nodes, tris, root, ro, rd, orc_inv_rd(rd), ORC_QUERY_CLOSEST, cnt,
                                      ORC_AO_STACK);
            else
                orc_intersect_stackfree_ex(&hit, nodes, tris, root, ro, rd, orc_inv_rd(rd), ORC_QUERY_CLOSEST, cnt,
                                           earlyExit);
            cRays++;
            if (hit.t == FLT_MAX_H) {
                pixels[pixel] = orc_pack_rgba(0, 0, 0, 1);
                if (aoOut) aoOut[pixel] = -1.0f;
                cFetch += cnt[0];
                cTest += cnt[1];
                continue;
            }
            orc_v3 n = vnormalize(V(hit.ngx, hit.ngy, hit.ngz));
            if (0.0f < vdot(n, rd)) n = vneg(n);
            orc_v3 xaxis, zaxis;
            onb(n, &xaxis, &zaxis);
            orc_v3 ao_origin = vadd(vadd(ro, vmuls(rd, hit.t)), vmuls(n, 0.001f));
            orc_pcg rng;
            pcg_init(&rng, orc_hash_pcg((uint32_t)pixel), 0);
            float ao = 0.0f;
            for (int i = 0; i < nSample; i++) {
                float r0, r1;
                if (useSobol) {
                    orc_sobol_2d(&r0, &r1, (uint32_t)i, (uint32_t)xi, (uint32_t)yi, seed);
                } else {
                    r0 = pcg_uniformf(&rng);
                    r1 = pcg_uniformf(&rng);
                }
                orc_v3 dir = sample_dir(n, xaxis, zaxis, r0, r1);
                int occluded;
                if (walk >= 2) {
                    occluded = orc_any_hit_stack(nodes, tris, root, ao_origin, dir, orc_inv_rd(dir), radius, cnt,
                                                 ORC_AO_STACK);
                } else {
                    orc_hit h2;
                    hit_reset(&h2, radius);
                    orc_intersect_stackfree_ex(&h2, nodes, tris, root, ao_origin, dir, orc_inv_rd(dir), ORC_QUERY_ANY,
                                               cnt, earlyExit);
                    occluded = h2.t != radius;
                }
                cRays++;
                if (occluded) ao += 1.0f / (float)nSample;
            }
            cFetch += cnt[0];
            cTest += cnt[1];
            float brightness = orc_pow_pos(1.0f - ao, 1.0f / 2.2f);
            pixels[pixel] = orc_pack_rgba(brightness, brightness, brightness, 1.0f);
            if (aoOut) aoOut[pixel] = ao;
        }
    }
    if (counters) {
        counters[0] += cRays;
        counters[1] += cFetch;
        counters[2] += cTest;
    }
}

void orc_render_pt(float* acc4, int w, int h, const orc_raygen* g, uint32_t root, const orc_node* nodes,
                   const orc_tri* tris, const orc_attrib* attribs, int useSobol, int maxSPP, int maxDepth,
                   int nthreads) /* main_gpu.cu:179-271 */
{
    int nt = pick_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int yi = 0; yi < h; yi++) {
        for (int xi = 0; xi < w; xi++) {
            int pixel = xi + yi * w;
            float* acc = acc4 + 4 * (size_t)pixel;
            int iteration = (int)acc[3];
            orc_pcg rng;
            pcg_init(&rng, orc_hash_pcg(orc_hash_pcg((uint32_t)pixel) + (uint32_t)iteration), 0);
            uint32_t dimLevel = 0;
#define NEXT_RANDOM(R0, R1)                                                                                  \
    do {                                                                                                     \
        if (useSobol) {                                                                                      \
            orc_sobol_2d(&(R0), &(R1), (uint32_t)iteration, (uint32_t)xi, (uint32_t)yi, dimLevel++);         \
        } else {                                                                                             \
            R0 = pcg_uniformf(&rng);                                                                         \
            R1 = pcg_uniformf(&rng);                                                                         \
        }                                                                                                    \
    } while (0)
            float r0, r1;
            NEXT_RANDOM(r0, r1);
            orc_v3 ro, rd;
            orc_raygen_shoot(g, &ro, &rd, ((float)xi + r0) / (float)w, ((float)yi + r1) / (float)h);
            orc_v3 color = V(0, 0, 0);
            orc_v3 tp = V(1.0f, 1.0f, 1.0f);
            orc_hit hit;
            hit_reset(&hit, FLT_MAX_H);
            orc_intersect_stackfree(&hit, nodes, tris, root, ro, rd, orc_inv_rd(rd), ORC_QUERY_CLOSEST, NULL);
            if (hit.t != FLT_MAX_H) {
                for (int depth = 0; depth < maxDepth; depth++) {
                    NEXT_RANDOM(r0, r1);
                    orc_v3 hp = vadd(ro, vmuls(rd, hit.t));
                    orc_v3 hn = vnormalize(V(hit.ngx, hit.ngy, hit.ngz));
                    if (0.0f < vdot(hn, rd)) hn = vneg(hn);
                    const orc_attrib* at = &attribs[hit.triangleIndex];
                    color = vadd(color, vmul(tp, at->emissive));
                    orc_v3 xaxis, zaxis;
                    onb(hn, &xaxis, &zaxis);
                    orc_v3 nro = vadd(hp, vmuls(hn, 0.0001f));
                    orc_v3 nrd = sample_dir(hn, xaxis, zaxis, r0, r1);
                    tp = vmul(tp, at->reflectance);
                    float maxComp = hmax(hmax(tp.x, tp.y), tp.z);
                    float p = hmin(maxComp / 0.25f, 1.0f);
                    float q0, q1;
                    NEXT_RANDOM(q0, q1);
                    (void)q1;
                    if (p < q0) break;
                    tp = vdivs(tp, p);
                    hit_reset(&hit, FLT_MAX_H);
                    orc_intersect_stackfree(&hit, nodes, tris, root, nro, nrd, orc_inv_rd(nrd), ORC_QUERY_CLOSEST,
                                            NULL);
                    if (hit.t == FLT_MAX_H) break;
                    ro = nro;
                    rd = nrd;
                }
            }
#undef NEXT_RANDOM
            if (iteration < maxSPP) {
                acc[0] += color.x;
                acc[1] += color.y;
                acc[2] += color.z;
                acc[3] += 1.0f;
            }
        }
    }
}

void orc_pack(uint32_t* pixels, const float* acc, int n) /* main_gpu.cu:273-286 */
{
    for (int i = 0; i < n; i++) {
        const float* a = acc + 4 * (size_t)i;
        pixels[i] = orc_pack_rgba(orc_pow_pos(a[0] / a[3], 1.0f / 2.2f), orc_pow_pos(a[1] / a[3], 1.0f / 2.2f),
                                  orc_pow_pos(a[2] / a[3], 1.0f / 2.2f), 1.0f);
    }
}

/* ---------------------------------------------------------------------- */
/* Differentiable primary shading in fp64 (no reference counterpart).      */
/* ---------------------------------------------------------------------- */
typedef struct { double x, y, z; } d3;
static inline d3 D(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static inline d3 dsub(d3 a, d3 b) { return D(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline d3 dadd(d3 a, d3 b) { return D(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline d3 dmul(d3 a, double s) { return D(a.x * s, a.y * s, a.z * s); }
static inline double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline d3 dcross(d3 a, d3 b) { return D(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

static void pixel_ray64(const orc_raygen* g, int xi, int yi, int w, int h, d3* ro, d3* rd)
{
    orc_v3 o, d;
    orc_raygen_shoot(g, &o, &d, (float)xi / (float)w, (float)yi / (float)h);
    *ro = D(o.x, o.y, o.z);
    *rd = D(d.x, d.y, d.z);
}

static void tri64(const double* v, const uint32_t* idx, uint32_t tri, d3* v0, d3* v1, d3* v2, uint32_t j[3])
{
    for (int k = 0; k < 3; k++) j[k] = idx[3 * tri + k];
    *v0 = D(v[3 * j[0]], v[3 * j[0] + 1], v[3 * j[0] + 2]);
    *v1 = D(v[3 * j[1]], v[3 * j[1] + 1], v[3 * j[1] + 2]);
    *v2 = D(v[3 * j[2]], v[3 * j[2] + 1], v[3 * j[2] + 2]);
}

void orc_primary_eval64(const double* verts, const uint32_t* idx, const uint32_t* hitTri, int w, int h,
                        const orc_raygen* g, const double* light, double* out)
{
    d3 L = D(light[0], light[1], light[2]);
#pragma omp parallel for schedule(static)
    for (int p = 0; p < w * h; p++) {
        uint32_t tri = hitTri[p];
        if (tri == 0xFFFFFFFFu) {
            out[2 * p] = 0.0;
            out[2 * p + 1] = 0.0;
            continue;
        }
        d3 ro, rd, v0, v1, v2;
        uint32_t j[3];
        pixel_ray64(g, p % w, p / w, w, h, &ro, &rd);
        tri64(verts, idx, tri, &v0, &v1, &v2, j);
        d3 ng = dcross(dsub(v1, v0), dsub(v2, v1));
        double t = ddot(dsub(v0, ro), ng) / ddot(ng, rd);
        double len = sqrt(ddot(ng, ng));
        d3 nh = dmul(ng, 1.0 / len);
        double sgn = ddot(nh, rd) > 0.0 ? -1.0 : 1.0;
        double s = sgn * ddot(nh, L);
        out[2 * p] = t;
        out[2 * p + 1] = s > 0.0 ? s : 0.0;
    }
}

void orc_primary_grad64(const double* verts, const uint32_t* idx, const uint32_t* hitTri, int w, int h,
                        const orc_raygen* g, const double* light, const double* gout, double* grad)
{
    d3 L = D(light[0], light[1], light[2]);
    for (int p = 0; p < w * h; p++) {
        uint32_t tri = hitTri[p];
        if (tri == 0xFFFFFFFFu) continue;
        double gt = gout[2 * p], gs = gout[2 * p + 1];
        d3 ro, rd, v0, v1, v2;
        uint32_t j[3];
        pixel_ray64(g, p % w, p / w, w, h, &ro, &rd);
        tri64(verts, idx, tri, &v0, &v1, &v2, j);
        d3 e0 = dsub(v1, v0), e1 = dsub(v2, v1);
        d3 ng = dcross(e0, e1);
        double den = ddot(ng, rd);
        double t = ddot(dsub(v0, ro), ng) / den;
        /* t: d t = dv0.ng/den + w.dng, w = (v0 - ro - t rd)/den, dng = de0 x e1 + e0 x de1 */
        d3 wv = dmul(dsub(dsub(v0, ro), dmul(rd, t)), 1.0 / den);
        d3 e1w = dcross(e1, wv), we0 = dcross(wv, e0);
        d3 g0 = dmul(dsub(dmul(ng, 1.0 / den), e1w), gt);
        d3 g1 = dmul(dsub(e1w, we0), gt);
        d3 g2 = dmul(we0, gt);
        double len = sqrt(ddot(ng, ng));
        d3 nh = dmul(ng, 1.0 / len);
        double sgn = ddot(nh, rd) > 0.0 ? -1.0 : 1.0;
        if (sgn * ddot(nh, L) > 0.0) {
            /* s = sgn nh.L: ds = q.dng, q = sgn (L - (nh.L) nh)/|ng| */
            d3 q = dmul(dsub(L, dmul(nh, ddot(nh, L))), sgn / len);
            d3 e1q = dcross(e1, q), qe0 = dcross(q, e0);
            g0 = dsub(g0, dmul(e1q, gs));
            g1 = dadd(g1, dmul(dsub(e1q, qe0), gs));
            g2 = dadd(g2, dmul(qe0, gs));
        }
        d3 gg[3] = {g0, g1, g2};
        for (int k = 0; k < 3; k++) {
            grad[3 * j[k]] += gg[k].x;
            grad[3 * j[k] + 1] += gg[k].y;
            grad[3 * j[k] + 2] += gg[k].z;
        }
    }
}
