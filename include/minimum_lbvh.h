// minimum_lbvh.h -- C++ facade of the MI355X (gfx950) MinimumLBVH library.
//
// Drop-in for the reference header minimum_lbvh.h of Ushio/MinimumLBVH on the
// GPU path: same namespace, types (identical layouts), host+device inline
// functions and the BVHGPUBuilder class, compiled offline with hipcc instead
// of through Orochi/orortc.  BVHGPUBuilder runs on libmlbvh.so (include/mlbvh.h).
// BVHCPUBuilder builds on the host with the same build_lbvh (leaf-parallel
// under OpenMP when compiled with -fopenmp); the reference's Embree builder is
// not provided.  The functions restate the reference API functions (cited
// file:line) with the reference's signatures.
//
// Arithmetic: every inline function evaluates in the reference's operation
// order; compile with -ffp-contract=off to keep results bit-identical to the
// reference's host arithmetic (HIP's fp32 divide/sqrt are correctly rounded by
// default).  Point-into-box extension uses the reference's host min/max
// (a<b?a:b); box-into-box merges keep -0 below +0 (see umin); slab tests use
// the hardware min/max (equal for the NaN-free values of finite rays).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "mlbvh.h"
#include "tinyhiponesweep.h"

#define MINIMUM_LBVH_FLT_MAX 3.402823466e+38F
#define MORTON_MAX_VALUE_3D 0x1FFFFF
#define MINIMUM_LBVH_DEVICE __host__ __device__
#define MINIMUM_LBVH_ASSERT(ExpectTrue) \
    if ((ExpectTrue) == 0) {            \
        abort();                        \
    }

namespace minimum_lbvh {

// minimum_lbvh.h:42-46
enum BUILD_OPTION { BUILD_OPTION_USE_NORMAL = 1, BUILD_OPTION_CPU_PARALLEL = 2 };

// --- small vector helpers (helper_math.h semantics, host min/max) ----------
MINIMUM_LBVH_DEVICE inline float3 v3(float s) { return make_float3(s, s, s); }
MINIMUM_LBVH_DEVICE inline float3 add(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
MINIMUM_LBVH_DEVICE inline float3 sub(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
MINIMUM_LBVH_DEVICE inline float3 mul(float3 a, float3 b) { return make_float3(a.x * b.x, a.y * b.y, a.z * b.z); }
MINIMUM_LBVH_DEVICE inline float3 mul(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
MINIMUM_LBVH_DEVICE inline float3 div(float3 a, float3 b) { return make_float3(a.x / b.x, a.y / b.y, a.z / b.z); }
MINIMUM_LBVH_DEVICE inline float3 div(float3 a, float s) { return make_float3(a.x / s, a.y / s, a.z / s); }
MINIMUM_LBVH_DEVICE inline float3 neg(float3 a) { return make_float3(-a.x, -a.y, -a.z); }
MINIMUM_LBVH_DEVICE inline float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MINIMUM_LBVH_DEVICE inline float3 cross(float3 a, float3 b)
{
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
MINIMUM_LBVH_DEVICE inline float3 normalize(float3 v) { return mul(v, 1.0f / sqrtf(dot(v, v))); }
MINIMUM_LBVH_DEVICE inline float hmin(float a, float b) { return a < b ? a : b; }
MINIMUM_LBVH_DEVICE inline float hmax(float a, float b) { return a > b ? a : b; }
MINIMUM_LBVH_DEVICE inline float3 hmin(float3 a, float3 b)
{
    return make_float3(hmin(a.x, b.x), hmin(a.y, b.y), hmin(a.z, b.z));
}
MINIMUM_LBVH_DEVICE inline float3 hmax(float3 a, float3 b)
{
    return make_float3(hmax(a.x, b.x), hmax(a.y, b.y), hmax(a.z, b.z));
}
// Merge of two boxes: the host rule above keeps the second operand on ties,
// so -0 against +0 would store a sign that depends on which child merges
// first.  Ties keep -0 for a lower bound and +0 for an upper one (the
// libmlbvh kernels and its checker do the same), so a node's box has the same
// bits whatever the order; any other input gives the host rule's bits.
MINIMUM_LBVH_DEVICE inline float umin(float a, float b)
{
    uint32_t ua, ub;
    memcpy(&ua, &a, 4);
    memcpy(&ub, &b, 4);
    uint32_t u = ua | ub;
    float t;
    memcpy(&t, &u, 4);
    const float m = a < b ? a : b;
    return a == b ? t : m;
}
MINIMUM_LBVH_DEVICE inline float umax(float a, float b)
{
    uint32_t ua, ub;
    memcpy(&ua, &a, 4);
    memcpy(&ub, &b, 4);
    uint32_t u = ua & ub;
    float t;
    memcpy(&t, &u, 4);
    const float m = a > b ? a : b;
    return a == b ? t : m;
}
// float -> int truncation with the reference host's (x86 cvttss2si) rule
MINIMUM_LBVH_DEVICE inline int f2i(float f)
{
    return (f >= -2147483648.0f && f < 2147483648.0f) ? (int)f : (int)0x80000000u;
}
MINIMUM_LBVH_DEVICE inline int iclamp(int f, int a, int b)
{
    int m = f < b ? f : b;
    return a > m ? a : m;
}

MINIMUM_LBVH_DEVICE inline uint64_t div_round_up64(uint64_t val, uint64_t divisor) noexcept
{
    return (val + divisor - 1) / divisor;
}
MINIMUM_LBVH_DEVICE inline uint64_t next_multiple64(uint64_t val, uint64_t divisor) noexcept
{
    return div_round_up64(val, divisor) * divisor;
}
template <class T>
MINIMUM_LBVH_DEVICE void swap(T* a, T* b)
{
    T tmp = *a;
    *a = *b;
    *b = tmp;
}
template <class T>
MINIMUM_LBVH_DEVICE inline T ss_max(T x, T y)
{
    return (x < y) ? y : x;
}
template <class T>
MINIMUM_LBVH_DEVICE inline T ss_min(T x, T y)
{
    return (y < x) ? y : x;
}
MINIMUM_LBVH_DEVICE inline float compMin(float3 v) { return fminf(fminf(v.x, v.y), v.z); }
MINIMUM_LBVH_DEVICE inline float compMax(float3 v) { return fmaxf(fmaxf(v.x, v.y), v.z); }

// minimum_lbvh.h:82-96
MINIMUM_LBVH_DEVICE inline float2 slabs(float3 ro, float3 one_over_rd, float3 lower, float3 upper, float knownHitT)
{
    float3 t0 = mul(sub(lower, ro), one_over_rd);
    float3 t1 = mul(sub(upper, ro), one_over_rd);
    float3 tmin = make_float3(fminf(t0.x, t1.x), fminf(t0.y, t1.y), fminf(t0.z, t1.z));
    float3 tmax = make_float3(fmaxf(t0.x, t1.x), fmaxf(t0.y, t1.y), fmaxf(t0.z, t1.z));
    float region_min = compMax(tmin);
    float region_max = compMin(tmax) * 1.00000024f;  // robust traversal factor
    region_min = fmaxf(region_min, 0.0f);
    region_max = fminf(region_max, knownHitT);
    return make_float2(region_min, region_max);
}

// minimum_lbvh.h:99-147 (clz(0) = 32, ffs(0) = 0 on host and device)
MINIMUM_LBVH_DEVICE inline int clz(uint32_t x) { return x ? __builtin_clz(x) : 32; }
MINIMUM_LBVH_DEVICE inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }
MINIMUM_LBVH_DEVICE inline int ffs(uint32_t x) { return x ? __builtin_ctz(x) + 1 : 0; }
MINIMUM_LBVH_DEVICE inline int ffs64(uint64_t x) { return x ? __builtin_ctzll(x) + 1 : 0; }
MINIMUM_LBVH_DEVICE inline int delta(uint32_t a, uint32_t b) { return 32 - clz(a ^ b); }
MINIMUM_LBVH_DEVICE inline int delta(uint64_t a, uint64_t b) { return 64 - clz64(a ^ b); }

// minimum_lbvh.h:149-205
MINIMUM_LBVH_DEVICE inline uint64_t encodeMortonCode_Naive(uint32_t x, uint32_t y, uint32_t z)
{
    uint64_t code = 0;
    for (uint64_t i = 0; i < 64 / 3; ++i)
        code |= ((uint64_t)(x & (1u << i)) << (2 * i + 0)) | ((uint64_t)(y & (1u << i)) << (2 * i + 1)) |
                ((uint64_t)(z & (1u << i)) << (2 * i + 2));
    return code;
}
MINIMUM_LBVH_DEVICE inline uint32_t compact3(uint64_t m)
{
    uint64_t x = m & 0x1249249249249249ull;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
    x = (x ^ (x >> 8)) & 0x1f0000ff0000ffull;
    x = (x ^ (x >> 16)) & 0x1f00000000ffffull;
    x = (x ^ (x >> 32)) & 0x1fffffull;
    return (uint32_t)x;
}
MINIMUM_LBVH_DEVICE inline uint64_t splat3(uint32_t a)
{
    uint64_t x = a & 0x1fffffu;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}
MINIMUM_LBVH_DEVICE inline void decodeMortonCode(uint64_t morton, uint32_t* x, uint32_t* y, uint32_t* z)
{
    *x = compact3(morton);
    *y = compact3(morton >> 1);
    *z = compact3(morton >> 2);
}
MINIMUM_LBVH_DEVICE inline uint64_t encodeMortonCode(uint32_t x, uint32_t y, uint32_t z)
{
    return splat3(x) | splat3(y) << 1 | splat3(z) << 2;
}

// minimum_lbvh.h:207-282 (layouts are the ABI)
struct Triangle {
    float3 vs[3];
};
// minimum_lbvh.h:212-215
MINIMUM_LBVH_DEVICE inline float3 normalOf(const Triangle& tri)
{
    return normalize(cross(sub(tri.vs[1], tri.vs[0]), sub(tri.vs[2], tri.vs[0])));
}

struct AABB {
    float3 lower;
    float3 upper;
    MINIMUM_LBVH_DEVICE static AABB empty()
    {
        AABB aabb;
        aabb.setEmpty();
        return aabb;
    }
    MINIMUM_LBVH_DEVICE void setEmpty()
    {
        lower = v3(+MINIMUM_LBVH_FLT_MAX);
        upper = v3(-MINIMUM_LBVH_FLT_MAX);
    }
    MINIMUM_LBVH_DEVICE void extend(const float3& p)
    {
        lower = hmin(lower, p);
        upper = hmax(upper, p);
    }
    MINIMUM_LBVH_DEVICE void extend(const AABB& b)
    {
        lower = make_float3(umin(lower.x, b.lower.x), umin(lower.y, b.lower.y), umin(lower.z, b.lower.z));
        upper = make_float3(umax(upper.x, b.upper.x), umax(upper.y, b.upper.y), umax(upper.z, b.upper.z));
    }
    MINIMUM_LBVH_DEVICE uint64_t encodeMortonCode(float3 p) const
    {
        float3 f = mul(div(sub(p, lower), sub(upper, lower)), (float)(MORTON_MAX_VALUE_3D + 1));
        uint32_t x = (uint32_t)iclamp(f2i(f.x), 0, MORTON_MAX_VALUE_3D);
        uint32_t y = (uint32_t)iclamp(f2i(f.y), 0, MORTON_MAX_VALUE_3D);
        uint32_t z = (uint32_t)iclamp(f2i(f.z), 0, MORTON_MAX_VALUE_3D);
        return minimum_lbvh::encodeMortonCode(x, y, z);
    }
};

struct NodeIndex {
    MINIMUM_LBVH_DEVICE NodeIndex() : m_index(0), m_isLeaf(0) {}
    MINIMUM_LBVH_DEVICE NodeIndex(uint32_t index, bool isLeaf) : m_index(index), m_isLeaf(isLeaf) {}
    uint32_t m_index : 31;
    uint32_t m_isLeaf : 1;
    MINIMUM_LBVH_DEVICE static NodeIndex invalid() { return NodeIndex(0x7FFFFFFF, false); }
};
MINIMUM_LBVH_DEVICE inline bool operator==(NodeIndex a, NodeIndex b)
{
    return a.m_index == b.m_index && a.m_isLeaf == b.m_isLeaf;
}
MINIMUM_LBVH_DEVICE inline bool operator!=(NodeIndex a, NodeIndex b) { return !(a == b); }

struct InternalNode {
    NodeIndex children[2];
    AABB aabbs[2];
    NodeIndex parent;
    uint32_t context;  // upper bound of the range after building, or 0xFFFFFFFF
};

struct IndexedMorton {
    uint32_t morton;
    uint32_t index;
};

struct Hit {
    float t = MINIMUM_LBVH_FLT_MAX;
    float2 uv = {};
    float3 ng = {};
    uint32_t triangleIndex = 0xFFFFFFFF;
};

static_assert(sizeof(Triangle) == 36 && sizeof(AABB) == 24 && sizeof(NodeIndex) == 4, "ABI");
static_assert(sizeof(InternalNode) == 64 && sizeof(IndexedMorton) == 8 && sizeof(Hit) == 32, "ABI");

// Agglomerative climb for one leaf, minimum_lbvh.h:284-361.  Device callers
// must pass nodes whose `context` is 0xFFFFFFFF; the exchange is an
// agent-scope acq_rel swap so the hand-off is visible across XCDs.
MINIMUM_LBVH_DEVICE inline void build_lbvh(NodeIndex* rootNode, InternalNode* internals, const Triangle* triangles,
                                           int nTriangles, const uint8_t* deltas, uint32_t i_leaf,
                                           uint32_t triangleIndex)
{
    int nInternals = nTriangles - 1;
    int nDeltas = nInternals;
    uint32_t leaf_lower = i_leaf;
    uint32_t leaf_upper = i_leaf;
    NodeIndex node(triangleIndex, true);
    AABB aabb;
    aabb.setEmpty();
    if (triangles)
        for (int k = 0; k < 3; k++) aabb.extend(triangles[triangleIndex].vs[k]);
    bool isRoot = true;
    while (leaf_upper - leaf_lower < (uint32_t)nInternals) {
        uint32_t deltaL = 0 < leaf_lower ? deltas[leaf_lower - 1] : 0xFFFFFFFF;
        uint32_t deltaR = leaf_upper < (uint32_t)nDeltas ? deltas[leaf_upper] : 0xFFFFFFFF;
        int goLeft = deltaL < deltaR ? 1 : 0;
        int parent = goLeft ? (leaf_lower - 1) : leaf_upper;
        internals[parent].children[goLeft] = node;
        internals[parent].aabbs[goLeft] = aabb;
        if (!node.m_isLeaf) internals[node.m_index].parent = NodeIndex(parent, false);
        uint32_t index = goLeft ? leaf_upper : leaf_lower;
#if defined(__HIP_DEVICE_COMPILE__)
        index = __hip_atomic_exchange(&internals[parent].context, index, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
#else
        index = __atomic_exchange_n(&internals[parent].context, index, __ATOMIC_SEQ_CST);
#endif
        if (index == 0xFFFFFFFF) {
            isRoot = false;
            break;
        }
        leaf_lower = ss_min(leaf_lower, index);
        leaf_upper = ss_max(leaf_upper, index);
        internals[parent].context = leaf_upper < (uint32_t)nInternals ? leaf_upper : 0xFFFFFFFF;
        node = NodeIndex(parent, false);
        AABB otherAABB = internals[parent].aabbs[goLeft ^ 0x1];
        aabb.extend(otherAABB);
    }
    if (isRoot) {
        *rootNode = node;
        internals[node.m_index].parent = NodeIndex::invalid();
    }
}

// minimum_lbvh.h:363-413
MINIMUM_LBVH_DEVICE inline bool intersectRayTriangle(float* tOut, float* uOut, float* vOut, float3* ngOut,
                                                     float t_min, float t_max, float3 ro, float3 rd, float3 v0,
                                                     float3 v1, float3 v2)
{
    float3 e0 = sub(v1, v0);
    float3 e1 = sub(v2, v1);
    float3 ng = cross(e0, e1);
    float t = dot(sub(v0, ro), ng) / dot(ng, rd);
    if (t_min <= t && t <= t_max) {
        float3 e2 = sub(v0, v2);
        float3 ro2 = mul(ro, 2.0f);
        float u_vol = dot(cross(e2, sub(add(v0, v2), ro2)), rd);
        float v_vol = dot(cross(e0, sub(add(v1, v0), ro2)), rd);
        float w_vol = dot(cross(e1, sub(add(v2, v1), ro2)), rd);
        if (fminf(fminf(u_vol, v_vol), w_vol) < 0.0f && 0.0f < fmaxf(fmaxf(u_vol, v_vol), w_vol)) return false;
        float vol = u_vol + v_vol + w_vol;
        *tOut = t;
        *uOut = u_vol / vol;
        *vOut = v_vol / vol;
        *ngOut = ng;
        return true;
    }
    return false;
}

// minimum_lbvh.h:415-426 (recursive; host use)
MINIMUM_LBVH_DEVICE inline void validate_lbvh(NodeIndex node, const InternalNode* internals, const uint8_t* deltas,
                                              int maxDelta)
{
    if (node.m_isLeaf) return;
    auto d = deltas[node.m_index];
#if !defined(__HIP_DEVICE_COMPILE__)
    MINIMUM_LBVH_ASSERT(d <= maxDelta);
#endif
    validate_lbvh(internals[node.m_index].children[0], internals, deltas, d);
    validate_lbvh(internals[node.m_index].children[1], internals, deltas, d);
}

// minimum_lbvh.h:898-909
MINIMUM_LBVH_DEVICE inline float3 invRd(float3 rd)
{
    float3 r = make_float3(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    return make_float3(fmaxf(-MINIMUM_LBVH_FLT_MAX, fminf(r.x, MINIMUM_LBVH_FLT_MAX)),
                       fmaxf(-MINIMUM_LBVH_FLT_MAX, fminf(r.y, MINIMUM_LBVH_FLT_MAX)),
                       fmaxf(-MINIMUM_LBVH_FLT_MAX, fminf(r.z, MINIMUM_LBVH_FLT_MAX)));
}
MINIMUM_LBVH_DEVICE inline uint32_t hashPCG(uint32_t v)
{
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28) + 4)) ^ state) * 277803737u;
    return (word >> 22) ^ word;
}

MINIMUM_LBVH_DEVICE inline void hit_record(Hit* hit, float t, float u, float v, float3 ng, uint32_t tri)
{
    hit->t = t;
    hit->uv = make_float2(u, v);
    hit->ng = ng;
    hit->triangleIndex = tri;
}

// minimum_lbvh.h:912-973
MINIMUM_LBVH_DEVICE inline void intersect_stackfull(Hit* hit, const InternalNode* nodes, const Triangle* triangles,
                                                    NodeIndex rootNode, float3 ro, float3 rd, float3 one_over_rd,
                                                    NodeIndex* stack)
{
    int sp = 0;
    NodeIndex node = rootNode;
    while (node != NodeIndex::invalid()) {
        if (node.m_isLeaf) {
            float t, u, v;
            float3 ng;
            const Triangle& tri = triangles[node.m_index];
            if (intersectRayTriangle(&t, &u, &v, &ng, 0.0f, hit->t, ro, rd, tri.vs[0], tri.vs[1], tri.vs[2]))
                hit_record(hit, t, u, v, ng, node.m_index);
            node = sp == 0 ? NodeIndex::invalid() : stack[--sp];
            continue;
        }
        const AABB& L = nodes[node.m_index].aabbs[0];
        const AABB& R = nodes[node.m_index].aabbs[1];
        float2 rangeL = slabs(ro, one_over_rd, L.lower, L.upper, hit->t);
        float2 rangeR = slabs(ro, one_over_rd, R.lower, R.upper, hit->t);
        bool hitL = rangeL.x <= rangeL.y;
        bool hitR = rangeR.x <= rangeR.y;
        if (hitL && hitR) {
            NodeIndex near_node = nodes[node.m_index].children[0];
            NodeIndex far_node = nodes[node.m_index].children[1];
            if (rangeR.x < rangeL.x) swap(&near_node, &far_node);
            node = near_node;
            stack[sp++] = far_node;
        } else if (hitL || hitR) {
            node = nodes[node.m_index].children[hitL ? 0 : 1];
        } else {
            node = sp == 0 ? NodeIndex::invalid() : stack[--sp];
        }
    }
}

enum RAY_QUERY { RAY_QUERY_CLOSEST, RAY_QUERY_ANY };

// minimum_lbvh.h:981-1061
MINIMUM_LBVH_DEVICE inline void intersect_stackfree(Hit* hit, const InternalNode* nodes, const Triangle* triangles,
                                                    NodeIndex node, float3 ro, float3 rd, float3 one_over_rd,
                                                    RAY_QUERY rayQuery = RAY_QUERY_CLOSEST)
{
    NodeIndex curr_node = node;
    NodeIndex prev_node = NodeIndex::invalid();
    while (curr_node != NodeIndex::invalid()) {
        if (curr_node.m_isLeaf) {
            float t, u, v;
            float3 ng;
            const Triangle& tri = triangles[curr_node.m_index];
            if (intersectRayTriangle(&t, &u, &v, &ng, 0.0f, hit->t, ro, rd, tri.vs[0], tri.vs[1], tri.vs[2])) {
                hit_record(hit, t, u, v, ng, curr_node.m_index);
                if (rayQuery == RAY_QUERY_ANY) break;
            }
            swap(&curr_node, &prev_node);
            continue;
        }
        AABB L = nodes[curr_node.m_index].aabbs[0];
        AABB R = nodes[curr_node.m_index].aabbs[1];
        float2 rangeL = slabs(ro, one_over_rd, L.lower, L.upper, hit->t);
        float2 rangeR = slabs(ro, one_over_rd, R.lower, R.upper, hit->t);
        bool hitL = rangeL.x <= rangeL.y;
        bool hitR = rangeR.x <= rangeR.y;
        NodeIndex parent_node = nodes[curr_node.m_index].parent;
        NodeIndex near_node = nodes[curr_node.m_index].children[0];
        NodeIndex far_node = nodes[curr_node.m_index].children[1];
        int nHits = 0;
        if (hitL && hitR) {
            if (rangeR.x < rangeL.x) swap(&near_node, &far_node);
            nHits = 2;
        } else if (hitL || hitR) {
            nHits = 1;
            near_node = hitR ? far_node : near_node;
        }
        NodeIndex next_node;
        if (prev_node == parent_node) next_node = 0 < nHits ? near_node : parent_node;
        else if (prev_node == near_node) next_node = nHits == 2 ? far_node : parent_node;
        else next_node = parent_node;
        prev_node = curr_node;
        curr_node = next_node;
    }
}

// minimum_lbvh.h:1063-1137
MINIMUM_LBVH_DEVICE inline void intersect_escape_link(Hit* hit, const InternalNode* nodes, const Triangle* triangles,
                                                      NodeIndex node, float3 ro, float3 rd, float3 one_over_rd,
                                                      RAY_QUERY rayQuery = RAY_QUERY_CLOSEST)
{
    NodeIndex curr_node = node;
    NodeIndex prev_node = NodeIndex::invalid();
    while (curr_node != NodeIndex::invalid()) {
        if (curr_node.m_isLeaf) {
            float t, u, v;
            float3 ng;
            const Triangle& tri = triangles[curr_node.m_index];
            if (intersectRayTriangle(&t, &u, &v, &ng, 0.0f, hit->t, ro, rd, tri.vs[0], tri.vs[1], tri.vs[2])) {
                hit_record(hit, t, u, v, ng, curr_node.m_index);
                if (rayQuery == RAY_QUERY_ANY) break;
            }
            if (nodes[prev_node.m_index].children[0] == curr_node) {
                swap(&curr_node, &prev_node);
                continue;
            }
            uint32_t link = nodes[prev_node.m_index].context;
            NodeIndex next_node = link != 0xFFFFFFFF ? NodeIndex(link, false) : NodeIndex::invalid();
            prev_node = curr_node;
            curr_node = next_node;
            continue;
        }
        NodeIndex parent_node = nodes[curr_node.m_index].parent;
        AABB aabb;
        NodeIndex hitChild, miss;
        if (prev_node == parent_node) {
            aabb = nodes[curr_node.m_index].aabbs[0];
            hitChild = nodes[curr_node.m_index].children[0];
            miss = curr_node;
        } else {
            aabb = nodes[curr_node.m_index].aabbs[1];
            hitChild = nodes[curr_node.m_index].children[1];
            uint32_t link = nodes[curr_node.m_index].context;
            miss = link != 0xFFFFFFFF ? NodeIndex(link, false) : NodeIndex::invalid();
        }
        float2 range = slabs(ro, one_over_rd, aabb.lower, aabb.upper, hit->t);
        NodeIndex next_node = range.x <= range.y ? hitChild : miss;
        prev_node = curr_node;
        curr_node = next_node;
    }
}

// BVHCPUBuilder, minimum_lbvh.h:509-669 (without the Embree builder): the
// host build over build_lbvh.  Records are ordered by (morton, index), the
// stable order the GPU builder produces; the reference's std::sort on the
// key alone leaves equal keys in an unspecified order.  With
// BUILD_OPTION_CPU_PARALLEL the leaves climb in parallel when the caller is
// compiled with OpenMP, else one after another (the same tree either way).
// n < 2 aborts (the reference writes out of bounds for n == 1).
class BVHCPUBuilder {
public:
    void build(const Triangle* triangles, int nTriangles, uint32_t buildOption)
    {
        MINIMUM_LBVH_ASSERT(2 <= nTriangles);
        m_internals.assign((size_t)nTriangles - 1, InternalNode());
        for (InternalNode& node : m_internals) node.context = 0xFFFFFFFF;
        m_deltas.assign((size_t)nTriangles - 1, 0);

        AABB sceneAABB;
        sceneAABB.setEmpty();
        for (int i = 0; i < nTriangles; i++)
            for (int j = 0; j < 3; ++j) sceneAABB.extend(triangles[i].vs[j]);

        std::vector<IndexedMorton> mortons((size_t)nTriangles);
        for (int i = 0; i < nTriangles; i++) {
            const Triangle& tri = triangles[i];
            float3 center = div(add(add(tri.vs[0], tri.vs[1]), tri.vs[2]), 3.0f);
            uint32_t pMorton = (uint32_t)(sceneAABB.encodeMortonCode(center) >> 31);
            if (buildOption & BUILD_OPTION_USE_NORMAL) {
                AABB unit;
                unit.lower = v3(-1.0f);
                unit.upper = v3(1.0f);
                uint32_t nMorton = (uint32_t)(unit.encodeMortonCode(normalOf(tri)) >> 31);
                mortons[i].morton = (pMorton & 0xFFFFFFC0) | (nMorton >> 26);
            } else {
                mortons[i].morton = pMorton;
            }
            mortons[i].index = (uint32_t)i;
        }
        std::sort(mortons.begin(), mortons.end(), [](IndexedMorton a, IndexedMorton b) {
            return a.morton != b.morton ? a.morton < b.morton : a.index < b.index;
        });
        for (size_t i = 0; i < m_deltas.size(); i++) m_deltas[i] = (uint8_t)delta(mortons[i].morton, mortons[i + 1].morton);

        const int n = nTriangles;
        if (buildOption & BUILD_OPTION_CPU_PARALLEL) {
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; i++)
                build_lbvh(&m_rootNode, m_internals.data(), triangles, n, m_deltas.data(), (uint32_t)i,
                           mortons[i].index);
        } else {
            for (int i = 0; i < n; i++)
                build_lbvh(&m_rootNode, m_internals.data(), triangles, n, m_deltas.data(), (uint32_t)i,
                           mortons[i].index);
        }
    }
    bool empty() const { return m_internals.empty(); }
    void validate() const { validate_lbvh(m_rootNode, m_internals.data(), m_deltas.data(), 255); }

    NodeIndex m_rootNode;
    std::vector<InternalNode> m_internals;
    std::vector<uint8_t> m_deltas;
};

// DeviceStopwatch, minimum_lbvh.h:693-724 (hipEvent pair).
class DeviceStopwatch {
public:
    explicit DeviceStopwatch(hipStream_t stream) : m_stream(stream)
    {
        (void)hipEventCreateWithFlags(&m_start, hipEventDefault);
        (void)hipEventCreateWithFlags(&m_stop, hipEventDefault);
    }
    ~DeviceStopwatch()
    {
        (void)hipEventDestroy(m_start);
        (void)hipEventDestroy(m_stop);
    }
    DeviceStopwatch(const DeviceStopwatch&) = delete;
    void operator=(const DeviceStopwatch&) = delete;
    void start() { (void)hipEventRecord(m_start, m_stream); }
    void stop() { (void)hipEventRecord(m_stop, m_stream); }
    float getElapsedMs() const
    {
        (void)hipEventSynchronize(m_stop);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, m_start, m_stop);
        return ms;
    }

private:
    hipStream_t m_stream;
    hipEvent_t m_start;
    hipEvent_t m_stop;
};

// BVHGPUBuilder, minimum_lbvh.h:726-885, on libmlbvh.  The constructor keeps
// the reference signature; its arguments (kernel source path and include
// directory for run-time compilation) are unused because the kernels are
// built ahead of time.  build() is asynchronous; set m_printTiming to get the
// reference's synchronising "%f ms" printout.
class BVHGPUBuilder {
public:
    BVHGPUBuilder(const char* kernel = nullptr, const char* includeDir = nullptr)
    {
        (void)kernel;
        (void)includeDir;
        check(hipMalloc((void**)&m_rootNode, sizeof(NodeIndex)), "hipMalloc(root)");
    }
    ~BVHGPUBuilder()
    {
        if (m_internals) (void)hipFree(m_internals);
        if (m_rootNode) (void)hipFree(m_rootNode);
    }
    BVHGPUBuilder(const BVHGPUBuilder&) = delete;
    void operator=(const BVHGPUBuilder&) = delete;

    void build(const Triangle* triangles, int nTriangles, uint32_t buildOption, tinyhiponesweep::OnesweepSort& sorter,
               hipStream_t stream)
    {
        DeviceStopwatch sw(stream);
        if (m_printTiming) sw.start();
        m_nTriangles = nTriangles;
        if (m_capacity < nTriangles - 1) {
            if (m_internals) check(hipFree(m_internals), "hipFree(internals)");
            check(hipMalloc((void**)&m_internals, sizeof(InternalNode) * (size_t)(nTriangles - 1)),
                  "hipMalloc(internals)");
            m_capacity = nTriangles - 1;
        }
        int rc = mlbvh_build(sorter.context(), triangles, nTriangles, buildOption, m_internals, m_rootNode, stream);
        if (rc != MLBVH_OK) {
            fprintf(stderr, "mlbvh_build: %s\n", mlbvh_last_error());
            abort();
        }
        if (m_printTiming) {
            sw.stop();
            printf("%f ms\n", sw.getElapsedMs());
        }
    }
    bool empty() const { return m_nTriangles == 0; }

    InternalNode* m_internals = nullptr;
    NodeIndex* m_rootNode = nullptr;
    int m_nTriangles = 0;
    bool m_printTiming = false;

private:
    static void check(hipError_t e, const char* what)
    {
        if (e != hipSuccess) {
            fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
            abort();
        }
    }
    int m_capacity = 0;
};

}  // namespace minimum_lbvh
