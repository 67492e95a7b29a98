// Device-side types and IEEE-exact vector math shared by the gfx950 kernels.
//
// Layouts are the ABI of the reference's minimum_lbvh.h (Triangle 36 B,
// AABB 24 B, NodeIndex 4 B, InternalNode 64 B, Hit 32 B; minimum_lbvh.h:207-282,
// 890-896).  Arithmetic is written operation by operation in the order the
// reference's helper_math.h evaluates it, and every file is compiled with
// -ffp-contract=off and HIP's default correctly-rounded fp32 divide/sqrt, so
// results are bit-identical to the reference's host arithmetic for finite
// inputs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mlbvh.h"

namespace mlbvh {

constexpr float kFltMax = 3.402823466e+38F;  // MINIMUM_LBVH_FLT_MAX
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kInvalidNode = 0x7FFFFFFFu;  // NodeIndex::invalid()
constexpr uint32_t kMortonMax = 0x1FFFFFu;      // MORTON_MAX_VALUE_3D

struct f3 {
    float x, y, z;
};
struct Tri {
    f3 v[3];
};
struct Box {
    f3 lo, hi;
};
struct Node {
    uint32_t child[2];
    Box box[2];
    uint32_t parent;
    uint32_t context;
};
struct HitRec {
    float t;
    uint32_t pad;
    float u, v;
    float ngx, ngy, ngz;
    uint32_t tri;
};
static_assert(sizeof(Tri) == 36, "Triangle is 36 bytes");
static_assert(sizeof(Box) == 24, "AABB is 24 bytes");
static_assert(sizeof(Node) == 64, "InternalNode is 64 bytes");
static_assert(sizeof(HitRec) == 32, "Hit is 32 bytes");

__host__ __device__ inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ inline f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline f3 operator*(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ inline f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ inline f3 operator/(f3 a, f3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ inline f3 operator/(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ inline f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline f3 cross(f3 a, f3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// normalize with 1/sqrtf (the reference's host rsqrtf), not the approximate rsq.
__device__ inline f3 normalize(f3 v)
{
    float inv = 1.0f / sqrtf(dot(v, v));
    return v * inv;
}
// Host helper_math semantics of fminf/fmaxf (a<b?a:b, a>b?a:b): used where the
// chosen operand is stored (AABB unions), so even the sign of zero matches.
__device__ inline float hmin(float a, float b) { return a < b ? a : b; }
__device__ inline float hmax(float a, float b) { return a > b ? a : b; }
__device__ inline f3 hmin3(f3 a, f3 b) { return mk(hmin(a.x, b.x), hmin(a.y, b.y), hmin(a.z, b.z)); }
__device__ inline f3 hmax3(f3 a, f3 b) { return mk(hmax(a.x, b.x), hmax(a.y, b.y), hmax(a.z, b.z)); }
// The climb's merge of two child boxes.  The host rule keeps the second
// operand on ties, so with -0 on one side and +0 on the other the stored sign
// would depend on which child's thread merges first (the reference's GPU and
// parallel CPU builds are order-dependent there).  Ties keep -0 for a lower
// bound and +0 for an upper one (bitwise OR / AND of equal values): every
// other input gives the host rule's bits, and the merge is symmetric.
// (Written as two selects: the nested form compiled to branches.)
__device__ inline float umin(float a, float b)
{
    const float m = a < b ? a : b;
    const float tie = __uint_as_float(__float_as_uint(a) | __float_as_uint(b));
    return a == b ? tie : m;
}
__device__ inline float umax(float a, float b)
{
    const float m = a > b ? a : b;
    const float tie = __uint_as_float(__float_as_uint(a) & __float_as_uint(b));
    return a == b ? tie : m;
}
__device__ inline f3 umin3(f3 a, f3 b) { return mk(umin(a.x, b.x), umin(a.y, b.y), umin(a.z, b.z)); }
__device__ inline f3 umax3(f3 a, f3 b) { return mk(umax(a.x, b.x), umax(a.y, b.y), umax(a.z, b.z)); }

// float -> int truncation with x86-64 cvttss2si semantics (out of range and
// NaN give INT_MIN), the behaviour of the reference's host make_int3(float3).
__device__ inline int f2i_trunc(float f)
{
    return (f >= -2147483648.0f && f < 2147483648.0f) ? (int)f : (int)0x80000000u;
}
__device__ inline int iclamp(int f, int a, int b)
{
    int m = f < b ? f : b;
    return a > m ? a : m;
}

// Spread the low 21 bits to every third bit (bit k -> 3k).
__device__ inline uint64_t splat3(uint32_t a)
{
    uint64_t x = a & 0x1fffffu;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

// 63-bit Morton code of p inside box (x -> bit 0, y -> bit 1, z -> bit 2).
__device__ inline uint64_t morton_in_box(f3 lo, f3 hi, f3 p)
{
    f3 f = ((p - lo) / (hi - lo)) * (float)(kMortonMax + 1);
    uint32_t cx = (uint32_t)iclamp(f2i_trunc(f.x), 0, (int)kMortonMax);
    uint32_t cy = (uint32_t)iclamp(f2i_trunc(f.y), 0, (int)kMortonMax);
    uint32_t cz = (uint32_t)iclamp(f2i_trunc(f.z), 0, (int)kMortonMax);
    return splat3(cx) | splat3(cy) << 1 | splat3(cz) << 2;
}

// Block barrier for LDS data only: unlike __syncthreads it does not wait for
// the wave's outstanding global loads and stores.
__device__ inline void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Order-preserving float <-> uint map for atomic min/max of floats.
__device__ inline uint32_t f2ord(float f)
{
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ inline float ord2f(uint32_t k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// Pinhole camera of the reference's camera.h (origin, right, up; rd is not
// normalised).
__device__ inline void shoot(const mlbvh_raygen& g, f3* ro, f3* rd, float u, float v)
{
    f3 origin = mk(g.origin[0], g.origin[1], g.origin[2]);
    f3 right = mk(g.right[0], g.right[1], g.right[2]);
    f3 up = mk(g.up[0], g.up[1], g.up[2]);
    f3 forward = normalize(cross(up, right));
    // lerp(a, b, t) = a + t * (b - a)
    f3 nr = -right;
    f3 lr = nr + (right - nr) * u;
    f3 nu = -up;
    f3 lu = up + (nu - up) * v;
    f3 to = ((origin + forward) + lr) + lu;
    *ro = origin;
    *rd = to - origin;
}

// ---------------------------------------------------------------------------
// Agglomerative climb (build_lbvh, minimum_lbvh.h:284-361) through device
// memory, on mesh-local indices: the deltas, nodes and arrival words of one
// n-triangle mesh.
//
// Cross-XCD hand-off without L2 flushes: the only data one thread reads from
// another is the sibling subtree's AABB.  It is written with device-scope
// (sc1) stores, which complete at the device coherence point; the wait on the
// vector memory counter makes them complete before the device-scope exchange
// on the arrival word; the second arriver reads the box back with
// device-scope loads issued after the exchange returned.  All other node
// fields are read only after the kernel ends.  Compiler signal fences keep
// the stores, the wait, the exchange and the loads in this order.  This
// relies on gfx950's sc1 semantics rather than on an acquire/release pair:
// an agent-scope acq_rel exchange is correct by the memory model, but emits
// an L2 write-back and invalidate per hand-off and made the 5.24M-triangle
// build 1.03 ms instead of 0.565 ms.
// ---------------------------------------------------------------------------
__device__ inline void store_box_coherent(Box* dst, const Box& b)
{
    uint64_t* d = reinterpret_cast<uint64_t*>(dst);  // 8-byte aligned (offset 8 or 32 in the node)
    __hip_atomic_store(d + 0, (uint64_t)__float_as_uint(b.lo.x) | ((uint64_t)__float_as_uint(b.lo.y) << 32),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(d + 1, (uint64_t)__float_as_uint(b.lo.z) | ((uint64_t)__float_as_uint(b.hi.x) << 32),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(d + 2, (uint64_t)__float_as_uint(b.hi.y) | ((uint64_t)__float_as_uint(b.hi.z) << 32),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline Box load_box_coherent(Box* src)
{
    uint64_t* s = reinterpret_cast<uint64_t*>(src);
    uint64_t a = __hip_atomic_load(s + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint64_t b = __hip_atomic_load(s + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint64_t c = __hip_atomic_load(s + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    Box r;
    r.lo = mk(__uint_as_float((uint32_t)a), __uint_as_float((uint32_t)(a >> 32)), __uint_as_float((uint32_t)b));
    r.hi = mk(__uint_as_float((uint32_t)(b >> 32)), __uint_as_float((uint32_t)c), __uint_as_float((uint32_t)(c >> 32)));
    return r;
}

// Wait until this lane's earlier vector memory operations (the sc1 box
// stores) have completed: s_waitcnt vmcnt(0), gfx9 encoding.
__device__ inline void wait_vector_memory()
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// An indexed mesh as a triangle source: u32 index triples into fp32 xyz
// vertices.  `tris[i]` yields the same Triangle a de-indexed soup holds at i,
// so every kernel templated on its triangle source (Tri* or IndexedTris)
// computes bit-identical results from either form.  An index past the last
// vertex is clamped to it, so a malformed index array cannot read outside the
// vertex buffer.
struct IndexedTris {
    const float* __restrict__ v;
    const uint32_t* __restrict__ f;
    uint32_t vMax;  // nVertices - 1
    __device__ inline Tri operator[](uint32_t i) const
    {
        const uint32_t a = f[3 * (size_t)i], b = f[3 * (size_t)i + 1], c = f[3 * (size_t)i + 2];
        const uint32_t j[3] = {a < vMax ? a : vMax, b < vMax ? b : vMax, c < vMax ? c : vMax};
        Tri r;
#pragma unroll
        for (int k = 0; k < 3; k++) r.v[k] = mk(v[3 * (size_t)j[k]], v[3 * (size_t)j[k] + 1], v[3 * (size_t)j[k] + 2]);
        return r;
    }
};

__device__ inline Box tri_box(const Tri& tri)
{
    Box box;
    box.lo = mk(kFltMax, kFltMax, kFltMax);
    box.hi = mk(-kFltMax, -kFltMax, -kFltMax);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        box.lo = hmin3(box.lo, tri.v[k]);
        box.hi = hmax3(box.hi, tri.v[k]);
    }
    return box;
}

// Continue the agglomerative climb from `node`, which covers the sorted
// leaves [lo, hi] and is bounded by `box`, through the device-scope arrival
// words: at each parent the first of the two children's threads stops and
// the second merges and goes on.  Returns true for the thread that
// completed the root (it writes *root).
__device__ inline bool climb_from(uint32_t n, uint32_t lo, uint32_t hi, uint32_t node, Box box,
                                  const uint8_t* __restrict__ deltas, Node* nodes, uint32_t* flags, uint32_t* root)
{
    const uint32_t nInternals = n - 1;
    while (hi - lo < nInternals) {
        uint32_t dL = 0 < lo ? (uint32_t)deltas[lo - 1] : 0xFFFFFFFFu;
        uint32_t dR = hi < nInternals ? (uint32_t)deltas[hi] : 0xFFFFFFFFu;
        int goLeft = dL < dR ? 1 : 0;
        uint32_t parent = goLeft ? lo - 1 : hi;
        Node* p = nodes + parent;
        store_box_coherent(&p->box[goLeft], box);
        p->child[goLeft] = node;
        if (!(node & kLeafBit)) nodes[node].parent = parent;
        uint32_t idx = goLeft ? hi : lo;
        wait_vector_memory();
        idx = __hip_atomic_exchange(&flags[parent], idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);  // the box loads below stay after the exchange
        if (idx == 0xFFFFFFFFu) return false;
        lo = lo < idx ? lo : idx;
        hi = hi < idx ? idx : hi;
        p->context = hi < nInternals ? hi : 0xFFFFFFFFu;
        node = parent;
        Box other = load_box_coherent(&p->box[goLeft ^ 1]);
        box.lo = umin3(box.lo, other.lo);
        box.hi = umax3(box.hi, other.hi);
    }
    *root = node;
    nodes[node & ~kLeafBit].parent = kInvalidNode;
    return true;
}

// Error plumbing shared by the entry points.
void set_error(const char* fmt, ...);
int hip_status(hipError_t e, const char* what);

}  // namespace mlbvh
