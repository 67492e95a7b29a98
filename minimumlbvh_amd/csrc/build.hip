// LBVH build for gfx950: scene AABB -> Morton records (+ fused digit
// histograms) -> Onesweep sort of (morton, index) -> deltas -> agglomerative
// bottom-up climb.  Reference: BVHGPUBuilder::build (minimum_lbvh.h:783-871),
// kernels of minimum_lbvh.cu:38-120 and the climb build_lbvh
// (minimum_lbvh.h:284-361).
//
// Differences in structure from the reference (results are identical):
//  * the Morton kernel also builds the four 8-bit digit histograms the sort
//    needs and initialises the climb's arrival words, so the sort starts with
//    its first reorder pass and no memset of the node array is needed;
//  * the climb runs in two phases: nodes whose leaves all lie in one block of
//    1024 sorted leaves are merged in LDS, the rest (about 1% of the nodes of
//    a surface mesh) through device-scope arrival words, one u32 per internal
//    node; `context` is written once, after the merge, with the reference's
//    final value;
//  * one thread per leaf for all n leaves (the reference launches
//    ceil((n-1)/256) blocks and drops the last leaf when (n-1) % 256 == 0).
#include "mlbvh_ctx.h"

namespace mlbvh {

constexpr int kBuildBlock = 256;
// The scene box is reduced into kBoxCopies copies of its 6 order-mapped
// words (block b uses copy b % kBoxCopies), so that the blocks' atomics
// spread over many L2 lines; the Morton kernels fold the copies.
#ifndef MLBVH_BOX_COPIES
#define MLBVH_BOX_COPIES 32
#endif
constexpr int kBoxCopies = MLBVH_BOX_COPIES;
constexpr int kBoxWords = 6 * kBoxCopies;
static_assert(kBoxWords <= 256, "k_build_init clears the box copies with one block");

// The scene box from the kBoxCopies partial boxes (all threads of the block
// call it; the result is the same in every thread).
__device__ inline void fold_scene_box(const uint32_t* sceneBox, f3* lo, f3* hi)
{
    __shared__ uint32_t s_box[6];
    if (threadIdx.x < 6) {
        uint32_t r = sceneBox[threadIdx.x];
        for (int c = 1; c < kBoxCopies; c++) {
            uint32_t v = sceneBox[6 * c + threadIdx.x];
            r = threadIdx.x < 3 ? (v < r ? v : r) : (v > r ? v : r);
        }
        s_box[threadIdx.x] = r;
    }
    __syncthreads();
    *lo = mk(ord2f(s_box[0]), ord2f(s_box[1]), ord2f(s_box[2]));
    *hi = mk(ord2f(s_box[3]), ord2f(s_box[4]), ord2f(s_box[5]));
}

// Scene-box slots plus the sort state of the 4-pass record sort.
__global__ void k_build_init(uint32_t* sceneBox, uint32_t* ctrl, uint32_t* hist, uint64_t* status,
                             size_t statusWords)
{
    int t = threadIdx.x;
    if (t < kBoxWords) sceneBox[t] = t % 6 < 3 ? 0xFFFFFFFFu : 0u;  // atomicMin / atomicMax slots (order-mapped)
    sort_state_init(ctrl, 4, hist, 4, status, statusWords);
}

// Scene bounds over all vertices: per-thread running min/max, wave and block
// reduction, then 6 order-mapped integer atomics per block.  TS, here and in
// the kernels below, is the triangle source: `const Tri*` (a soup) or
// IndexedTris.
template <class TS>
__global__ void __launch_bounds__(kBuildBlock) k_scene_aabb(TS tris, int n, uint32_t* sceneBox)
{
    f3 lo = mk(kFltMax, kFltMax, kFltMax), hi = mk(-kFltMax, -kFltMax, -kFltMax);
    for (int i = blockIdx.x * kBuildBlock + threadIdx.x; i < n; i += gridDim.x * kBuildBlock) {
        Tri tri = tris[i];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            lo = mk(fminf(lo.x, tri.v[k].x), fminf(lo.y, tri.v[k].y), fminf(lo.z, tri.v[k].z));
            hi = mk(fmaxf(hi.x, tri.v[k].x), fmaxf(hi.y, tri.v[k].y), fmaxf(hi.z, tri.v[k].z));
        }
    }
    float v[6] = {lo.x, lo.y, lo.z, hi.x, hi.y, hi.z};
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int c = 0; c < 3; c++) v[c] = fminf(v[c], __shfl_xor(v[c], o, 64));
#pragma unroll
        for (int c = 3; c < 6; c++) v[c] = fmaxf(v[c], __shfl_xor(v[c], o, 64));
    }
    __shared__ float s_v[kBuildBlock / 64][6];
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int c = 0; c < 6; c++) s_v[wave][c] = v[c];
    __syncthreads();
    if (threadIdx.x < 6) {
        int c = threadIdx.x;
        float r = s_v[0][c];
        for (int w = 1; w < kBuildBlock / 64; w++) r = c < 3 ? fminf(r, s_v[w][c]) : fmaxf(r, s_v[w][c]);
        uint32_t* box = sceneBox + 6 * (blockIdx.x % kBoxCopies);
        if (c < 3) atomicMin(&box[c], f2ord(r));
        else atomicMax(&box[c], f2ord(r));
    }
}

// Morton record per triangle (minimum_lbvh.cu:68-92 / minimum_lbvh.h:534-553),
// the 4 digit histograms of the 32-bit key, and the climb's arrival words.
// A few blocks per CU, so that the histograms are merged into global memory
// by few blocks (global atomics on 1024 shared counters are the cost
// otherwise); each block takes whole sort tiles (tileKeys triangles,
// tile-strided) and writes every tile's counts of the first digit,
// tileCounts[digit * nTiles + tile], so the record sort's first pass can skip
// its look-back.
constexpr int kMortonBlock = 1024;  // one block per sort tile still fills the CUs
#ifndef MLBVH_MORTON_UNROLL
#define MLBVH_MORTON_UNROLL 2
#endif
// triangles loaded at once per thread: 2 took 57 us at 5.24M triangles, 1
// 59 us, 4 63 us (76 VGPRs, one block per CU)
constexpr int kMortonUnroll = MLBVH_MORTON_UNROLL;
template <class TS>
__global__ void __launch_bounds__(kMortonBlock)
    k_morton(TS tris, int n, const uint32_t* __restrict__ sceneBox, uint32_t opt,
             uint64_t* __restrict__ rec, uint32_t* __restrict__ partials, uint32_t* __restrict__ flags,
             uint32_t* __restrict__ ctrl, uint32_t tileKeys, uint32_t nTiles, uint32_t* __restrict__ tileCounts)
{
    __shared__ uint32_t s_h[4][256];
    __shared__ uint32_t s_t[256];  // this tile's counts of the first digit
    const int t = threadIdx.x;
    for (int i = t; i < 4 * 256; i += kMortonBlock) (&s_h[0][0])[i] = 0;
    if (t < 256) s_t[t] = 0;
    f3 lo, hi;
    fold_scene_box(sceneBox, &lo, &hi);  // its barrier also orders the clears above
    auto record = [&](const Tri& tri, uint32_t i) {
        f3 center = ((tri.v[0] + tri.v[1]) + tri.v[2]) / 3.0f;
        uint32_t morton;
        if (opt & MLBVH_BUILD_OPTION_USE_NORMAL) {
            f3 ng = normalize(cross(tri.v[1] - tri.v[0], tri.v[2] - tri.v[0]));
            uint32_t nMorton = (uint32_t)(morton_in_box(mk(-1, -1, -1), mk(1, 1, 1), ng) >> 31);
            uint32_t pMorton = (uint32_t)(morton_in_box(lo, hi, center) >> 31);
            morton = (pMorton & 0xFFFFFFC0u) | (nMorton >> 26);
        } else {
            morton = (uint32_t)(morton_in_box(lo, hi, center) >> 31);
        }
        rec[i] = (uint64_t)morton | ((uint64_t)i << 32);
        if (i < (uint32_t)n - 1) flags[i] = 0xFFFFFFFFu;
        if (i == 0) ctrl[kCtrlFrontier] = 0;
        atomicAdd(&s_t[morton & 0xFFu], 1u);
#pragma unroll
        for (int p = 1; p < 4; p++) atomicAdd(&s_h[p][(morton >> (8 * p)) & 0xFFu], 1u);
    };
    for (uint32_t tile = blockIdx.x; tile < nTiles; tile += gridDim.x) {
        const uint32_t t0 = tile * tileKeys;
        const uint32_t t1 = (uint32_t)n - t0 < tileKeys ? (uint32_t)n : t0 + tileKeys;
        uint32_t i = t0 + (uint32_t)t;
        for (; i + (kMortonUnroll - 1) * kMortonBlock < t1; i += kMortonUnroll * kMortonBlock) {
            Tri tr[kMortonUnroll];
#pragma unroll
            for (int u = 0; u < kMortonUnroll; u++) tr[u] = tris[i + u * kMortonBlock];
#pragma unroll
            for (int u = 0; u < kMortonUnroll; u++) record(tr[u], i + u * kMortonBlock);
        }
        for (; i < t1; i += kMortonBlock) record(tris[i], i);
        __syncthreads();
        if (t < 256) {
            const uint32_t c = s_t[t];
            tileCounts[(size_t)t * nTiles + tile] = c;
            s_h[0][t] += c;
            s_t[t] = 0;
        }
        __syncthreads();
    }
    if (t < 256)
#pragma unroll
        for (int p = 0; p < 4; p++) partials[((size_t)blockIdx.x * 4 + p) * 256 + t] = s_h[p][t];
}

// The 63-bit key variant (MLBVH_BUILD_OPTION_MORTON64, no reference
// counterpart): the full Morton code of the centroid as a u64 key beside the
// triangle index, for a key-value sort over bits [0, 63).  Also opens the
// climb's arrival words, as k_morton does.
template <class TS>
__global__ void __launch_bounds__(kBuildBlock)
    k_morton64(TS tris, int n, const uint32_t* __restrict__ sceneBox, uint64_t* __restrict__ keys,
               uint32_t* __restrict__ idx, uint32_t* __restrict__ flags, uint32_t* __restrict__ ctrl)
{
    f3 lo, hi;
    fold_scene_box(sceneBox, &lo, &hi);
    for (int i = blockIdx.x * kBuildBlock + threadIdx.x; i < n; i += gridDim.x * kBuildBlock) {
        Tri tri = tris[i];
        f3 center = ((tri.v[0] + tri.v[1]) + tri.v[2]) / 3.0f;
        keys[i] = morton_in_box(lo, hi, center);
        idx[i] = (uint32_t)i;
        if (i < n - 1) flags[i] = 0xFFFFFFFFu;
        if (i == 0) ctrl[kCtrlFrontier] = 0;
    }
}

// Deltas 64 - clz64(a ^ b) of the sorted 63-bit keys (0 for equal keys), and
// the records the climb and mlbvh_refit read: index high, code >> 31 low.
__global__ void __launch_bounds__(kBuildBlock)
    k_deltas64(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ idx, int n,
               uint8_t* __restrict__ deltas, uint64_t* __restrict__ rec)
{
    int i = blockIdx.x * kBuildBlock + threadIdx.x;
    if (i < n) {
        const uint64_t k = keys[i];
        rec[i] = (uint64_t)(uint32_t)(k >> 31) | ((uint64_t)idx[i] << 32);
        if (i < n - 1) {
            const uint64_t x = k ^ keys[i + 1];
            deltas[i] = (uint8_t)(x ? 64 - __clzll(x) : 0);
        }
    }
}

// Agglomerative climb (build_lbvh, minimum_lbvh.h:284-361) in two phases.
//
// Internal node p (the split between sorted leaves p and p+1) covers the
// leaves [Lp, Rp] with Lp = 1 + (last q < p with d[q] >= d[p], else -1) and
// Rp = (first q > p with d[q] > d[p], else n-1): the climb merges across the
// smaller boundary delta first and goes right on ties.  So p lies entirely
// inside a block's leaves [b0, b1] iff
//   (b0 == 0 or max d[b0-1 .. p-1] >= d[p]) and (b1 == n-1 or max d[p+1 .. b1] > d[p]),
// which both of p's children evaluate alike from a prefix and a suffix
// maximum of the block's deltas.
//
// Phase 1 (k_climb_local): a block of kLocalLeaves threads climbs its leaves
// through such block-local nodes, handing the sibling bounds over in LDS.  A
// thread whose next parent is not block-local keeps (node, lo, hi) for the
// frontier list and stops.  The block-local nodes are assembled in LDS
// (children, boxes, parent, context) and written at the end as whole 64-byte
// nodes with coalesced 16-byte stores; a block-local node whose parent is not
// block-local gets its parent word from phase 2.
//
// Phase 2 (k_climb_upper): the frontier entries (about 1% of the nodes for
// surface meshes) climb on through the device-scope arrival words
// (climb_from).  The result is the same tree as one thread per leaf climbing
// through device memory all the way.
#ifndef MLBVH_CLIMB_LEAVES
#define MLBVH_CLIMB_LEAVES 1024
#endif
constexpr int kLocalLeaves = MLBVH_CLIMB_LEAVES;

// kFuseDeltas: the deltas come from the block's own sorted records
// (delta(a, b) = 32 - clz(a ^ b) of neighbouring keys, minimum_lbvh.cu:94-102)
// and the block writes d[b0 .. b1] to `deltasOut` for the upper phase and the
// caller; otherwise they are read from `deltas` (63-bit keys, refit).
template <class TS, bool kFuseDeltas>
__global__ void __launch_bounds__(kLocalLeaves)
    k_climb_local(TS tris, uint32_t n, const uint64_t* __restrict__ rec, const uint8_t* __restrict__ deltas,
                  uint8_t* __restrict__ deltasOut, Node* __restrict__ nodes, uint32_t* __restrict__ frontier,
                  uint32_t* __restrict__ ctrl, uint32_t* __restrict__ root)
{
    constexpr int kWaves = kLocalLeaves / 64;
    // Local delta j is d[b0 - 1 + j], j in [0, m] (m = leaves in the block).
    __shared__ uint8_t s_d[kLocalLeaves + 1];
    __shared__ uint8_t s_pm[kLocalLeaves];      // s_pm[j] = max s_d[0 .. j]
    __shared__ uint8_t s_sm[kLocalLeaves + 1];  // s_sm[j] = max s_d[j .. kLocalLeaves]
    __shared__ uint8_t s_loc[kLocalLeaves];     // internal node b0 + j lies inside the block (the predicate above)
    __shared__ uint32_t s_wmax[2][kWaves];
    __shared__ uint32_t s_flag[kLocalLeaves];   // arrival word of local node b0 + slot
    __shared__ Box s_box[kLocalLeaves][2];      // the node's child boxes
    __shared__ uint32_t s_child[kLocalLeaves][2];
    __shared__ uint32_t s_parent[kLocalLeaves];
    __shared__ uint32_t s_context[kLocalLeaves];
    __shared__ uint32_t s_wcount[kWaves];
    __shared__ uint32_t s_base;

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t b0 = blockIdx.x * kLocalLeaves;
    const uint32_t m = n - b0 < (uint32_t)kLocalLeaves ? n - b0 : (uint32_t)kLocalLeaves;
    const uint32_t b1 = b0 + m - 1;
    const uint32_t nInternals = n - 1;

    const uint32_t pos = b0 + (uint32_t)t;
    const uint64_t myRec = (uint32_t)t < m ? rec[pos] : 0ull;
    // deltas that exist and lie in [b0-1, b1]; others read as 0 (they never
    // decide a comparison: the b0 == 0 / b1 == n-1 cases are tested apart)
    uint32_t e0, e1;  // local deltas j = t (prefix element) and j = t + 1 (suffix element)
    if (kFuseDeltas) {
        __shared__ uint32_t s_k[kLocalLeaves + 2];  // s_k[j] = key of sorted position b0 - 1 + j
        if ((uint32_t)t < m) s_k[t + 1] = (uint32_t)myRec;
        if (t == 0 && b0 > 0) s_k[0] = (uint32_t)rec[b0 - 1];
        if ((uint32_t)t == m - 1 && b0 + m < n) s_k[m + 1] = (uint32_t)rec[b0 + m];
        __syncthreads();
        auto fused_delta = [&](uint32_t j) -> uint32_t {
            const uint32_t g = b0 + j;  // global delta index + 1
            return (j <= m && g >= 1 && g - 1 < nInternals) ? (uint32_t)(32 - __clz(s_k[j] ^ s_k[j + 1])) : 0u;
        };
        e0 = fused_delta((uint32_t)t);
        e1 = fused_delta((uint32_t)t + 1);
        if ((uint32_t)t < m && pos < nInternals) deltasOut[pos] = (uint8_t)e1;  // d[b0 + t]
    } else {
        auto local_delta = [&](uint32_t j) -> uint32_t {
            uint32_t g = b0 + j;  // global delta index + 1
            return (j <= m && g >= 1 && g - 1 < nInternals) ? (uint32_t)deltas[g - 1] : 0u;
        };
        e0 = local_delta((uint32_t)t);
        e1 = local_delta((uint32_t)t + 1);
    }
    s_d[t] = (uint8_t)e0;
    if (t == kLocalLeaves - 1) s_d[kLocalLeaves] = (uint8_t)e1;
    s_flag[t] = 0xFFFFFFFFu;
    s_parent[t] = kInvalidNode;
    uint32_t pm = e0, sm = e1;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t up = __shfl_up(pm, o, 64);
        uint32_t dn = __shfl_down(sm, o, 64);
        if (lane >= o) pm = pm > up ? pm : up;
        if (lane + o < 64) sm = sm > dn ? sm : dn;
    }
    if (lane == 63) s_wmax[0][wave] = pm;
    if (lane == 0) s_wmax[1][wave] = sm;
    __syncthreads();
#pragma unroll
    for (int w = 0; w < kWaves; w++) {  // all reads at once, no per-wave trip count
        const uint32_t a = s_wmax[0][w], b = s_wmax[1][w];
        if (w < wave) pm = pm > a ? pm : a;
        if (w > wave) sm = sm > b ? sm : b;
    }
    s_pm[t] = (uint8_t)pm;
    s_sm[t + 1] = (uint8_t)sm;
    __syncthreads();
    // the block-local test of each internal node of the block, once, so that a
    // climb step reads it together with its two deltas
    if ((uint32_t)t + 1 < m) {
        const uint32_t dp = s_d[t + 1];
        s_loc[t] = (uint8_t)((b0 == 0 || (uint32_t)s_pm[t] >= dp) && (b1 == nInternals || (uint32_t)s_sm[t + 2] > dp));
    }
    __syncthreads();

    bool rootHere = false, emit = false;
    uint32_t node = 0, lo = pos, hi = pos;
    if ((uint32_t)t < m) {
        const uint32_t triIndex = (uint32_t)(myRec >> 32);
        Box box = tri_box(tris[triIndex]);
        node = triIndex | kLeafBit;
        while (true) {
            if (hi - lo >= nInternals) {
                rootHere = true;
                break;
            }
            // four independent LDS reads (indices stay inside the block: the
            // range [lo, hi] only grows through block-local parents)
            const uint32_t vL = s_d[lo - b0], vR = s_d[hi - b0 + 1];
            const uint32_t lL = s_loc[lo > b0 ? lo - 1 - b0 : 0], lR = s_loc[hi - b0];
            const uint32_t dL = 0 < lo ? vL : 0xFFFFFFFFu;
            const uint32_t dR = hi < nInternals ? vR : 0xFFFFFFFFu;
            const bool locL = lo > b0 && lL != 0;  // parent lo - 1
            const bool locR = hi < b1 && lR != 0;  // parent hi
            int goLeft = dL < dR ? 1 : 0;
            uint32_t parent = goLeft ? lo - 1 : hi;
            const bool local = goLeft ? locL : locR;
            if (!local) {
                emit = true;
                break;
            }
            const uint32_t slot = parent - b0;
            if (!(node & kLeafBit)) s_parent[node - b0] = parent;
            s_child[slot][goLeft] = node;
            s_box[slot][goLeft] = box;
            uint32_t idx = __hip_atomic_exchange(&s_flag[slot], goLeft ? hi : lo, __ATOMIC_ACQ_REL,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP);
            if (idx == 0xFFFFFFFFu) break;
            lo = lo < idx ? lo : idx;
            hi = hi < idx ? idx : hi;
            s_context[slot] = hi < nInternals ? hi : 0xFFFFFFFFu;
            node = parent;
            Box other = s_box[slot][goLeft ^ 1];
            box.lo = umin3(box.lo, other.lo);
            box.hi = umax3(box.hi, other.hi);
        }
    }
    if (rootHere) *root = node;

    // frontier entries: a block-wide compaction, one global atomic per block
    const uint64_t emitMask = __ballot(emit);
    if (lane == 0) s_wcount[wave] = (uint32_t)__popcll(emitMask);
    __syncthreads();
    if (t == 0) {
        uint32_t total = 0;
        for (int w = 0; w < kWaves; w++) {
            uint32_t c = s_wcount[w];
            s_wcount[w] = total;
            total += c;
        }
        s_base = total ? atomicAdd(&ctrl[kCtrlFrontier], total) : 0u;
    }
    __syncthreads();
    if (emit) {
        const uint32_t k = s_base + s_wcount[wave] + (uint32_t)__popcll(emitMask & ((1ull << lane) - 1ull));
        frontier[3 * (size_t)k] = node;
        frontier[3 * (size_t)k + 1] = lo;
        frontier[3 * (size_t)k + 2] = hi;
    }

    // whole block-local nodes (both arrivals seen: the arrival word holds the
    // second one's bound), four 16-byte stores each
    const uint32_t nSlots = m - 1;  // internal nodes b0 .. b1-1
    for (uint32_t q = (uint32_t)t; q < 4 * nSlots; q += kLocalLeaves) {
        const uint32_t slot = q >> 2, part = q & 3;
        if (s_flag[slot] == 0xFFFFFFFFu) continue;
        const Box& l = s_box[slot][0];
        const Box& r = s_box[slot][1];
        uint4 v;
        if (part == 0) {
            v = make_uint4(s_child[slot][0], s_child[slot][1], __float_as_uint(l.lo.x), __float_as_uint(l.lo.y));
        } else if (part == 1) {
            v = make_uint4(__float_as_uint(l.lo.z), __float_as_uint(l.hi.x), __float_as_uint(l.hi.y),
                           __float_as_uint(l.hi.z));
        } else if (part == 2) {
            v = make_uint4(__float_as_uint(r.lo.x), __float_as_uint(r.lo.y), __float_as_uint(r.lo.z),
                           __float_as_uint(r.hi.x));
        } else {
            v = make_uint4(__float_as_uint(r.hi.y), __float_as_uint(r.hi.z), s_parent[slot], s_context[slot]);
        }
        reinterpret_cast<uint4*>(nodes + b0 + slot)[part] = v;
    }
}

template <class TS>
__global__ void __launch_bounds__(kBuildBlock)
    k_climb_upper(TS tris, uint32_t n, const uint8_t* __restrict__ deltas, Node* nodes,
                  uint32_t* flags, const uint32_t* __restrict__ frontier, const uint32_t* __restrict__ ctrl,
                  uint32_t* root)
{
    const uint32_t count = ctrl[kCtrlFrontier];
    for (uint32_t i = blockIdx.x * kBuildBlock + threadIdx.x; i < count; i += gridDim.x * kBuildBlock) {
        const uint32_t node = frontier[3 * (size_t)i], lo = frontier[3 * (size_t)i + 1],
                       hi = frontier[3 * (size_t)i + 2];
        Box box;
        if (node & kLeafBit) {
            box = tri_box(tris[node & ~kLeafBit]);
        } else {
            const Node& nd = nodes[node];
            box.lo = umin3(nd.box[0].lo, nd.box[1].lo);
            box.hi = umax3(nd.box[0].hi, nd.box[1].hi);
        }
        climb_from(n, lo, hi, node, box, deltas, nodes, flags, root);
    }
}

// Zero the frontier count and open every internal node's arrival word (the
// build does both in k_morton).
__global__ void __launch_bounds__(kBuildBlock) k_climb_reset(uint32_t* flags, int nInternals, uint32_t* ctrl)
{
    int i = blockIdx.x * kBuildBlock + threadIdx.x;
    if (i < nInternals) flags[i] = 0xFFFFFFFFu;
    if (i == 0) ctrl[kCtrlFrontier] = 0;
}

// fuseDeltas: compute the deltas from `rec` in the local phase and write
// them to `deltas` (the 32-bit build); else read them from `deltas`.
template <class TS>
static int launch_climb(mlbvh_ctx* ctx, TS tris, int n, const uint64_t* rec, uint8_t* deltas, bool fuseDeltas,
                        Node* nodes, uint32_t* root, hipStream_t stream)
{
    int rc = ensure(ctx, &ctx->frontier, &ctx->frontierCap, (size_t)n * 3, "hipMalloc(climb frontier)");
    if (rc) return rc;
    unsigned blocks = (unsigned)(((size_t)n + kLocalLeaves - 1) / kLocalLeaves);
    if (fuseDeltas)
        k_climb_local<TS, true><<<blocks, kLocalLeaves, 0, stream>>>(tris, (uint32_t)n, rec, nullptr, deltas, nodes,
                                                                     ctx->frontier, ctx->ctrl, root);
    else
        k_climb_local<TS, false><<<blocks, kLocalLeaves, 0, stream>>>(tris, (uint32_t)n, rec, deltas, nullptr, nodes,
                                                                      ctx->frontier, ctx->ctrl, root);
    // one lane per expected entry (about n/64 for surface meshes), at most 8
    // blocks per CU; lanes stride over a larger list
    size_t want = ((size_t)n / 64 + kBuildBlock - 1) / kBuildBlock + 1;
    unsigned cap = (unsigned)ctx->cuCount * 8;
    unsigned upper = want < cap ? (unsigned)want : cap;
    k_climb_upper<<<upper, kBuildBlock, 0, stream>>>(tris, (uint32_t)n, deltas, nodes, ctx->flags, ctx->frontier,
                                                     ctx->ctrl, root);
    return MLBVH_OK;
}

static inline unsigned grid1(size_t n) { return (unsigned)((n + kBuildBlock - 1) / kBuildBlock); }

template <class TS>
static int build_impl(mlbvh_ctx* ctx, TS tris, int n, uint32_t opt, void* d_nodes, void* d_root, uint64_t* d_sorted,
                      uint8_t* d_deltas, hipStream_t stream)
{
    if (!ctx || !d_nodes || !d_root) {
        set_error("build: null argument");
        return MLBVH_E_INVALID_ARG;
    }
    if (n < 2) {
        set_error("build: n = %d; at least 2 triangles are required", n);
        return MLBVH_E_INVALID_ARG;
    }
    const bool wide = (opt & MLBVH_BUILD_OPTION_MORTON64) != 0;
    if (wide && (opt & MLBVH_BUILD_OPTION_USE_NORMAL)) {
        set_error("build: MORTON64 and USE_NORMAL cannot be combined");
        return MLBVH_E_INVALID_ARG;
    }
    int rc = set_device(ctx);
    if (rc) return rc;
    size_t nn = (size_t)n;
    if ((rc = ensure(ctx, &ctx->recA, &ctx->recCap, nn, "hipMalloc(records)"))) return rc;
    if ((rc = ensure(ctx, &ctx->recB, &ctx->recBCap, nn, "hipMalloc(records tmp)"))) return rc;
    if ((rc = ensure(ctx, &ctx->deltas, &ctx->deltasCap, nn, "hipMalloc(deltas)"))) return rc;
    if ((rc = ensure(ctx, &ctx->flags, &ctx->flagsCap, nn, "hipMalloc(flags)"))) return rc;

    Node* nodes = (Node*)d_nodes;
    uint32_t* root = (uint32_t*)d_root;

    if ((rc = ensure_status(ctx, sort_tiles(ctx, (uint32_t)n)))) return rc;
    k_build_init<<<1, 256, 0, stream>>>(ctx->sceneBox, ctx->ctrl, ctx->hist, ctx->status, ctx->statusWords);
    unsigned g = grid1(nn);
    unsigned gAabb = g < (unsigned)ctx->cuCount * 8 ? g : (unsigned)ctx->cuCount * 8;
    k_scene_aabb<<<gAabb, kBuildBlock, 0, stream>>>(tris, n, ctx->sceneBox);
    unsigned gMorton = g < (unsigned)ctx->cuCount * 4 ? g : (unsigned)ctx->cuCount * 4;
    if (wide) {
        if ((rc = ensure(ctx, &ctx->idxA, &ctx->idxCap, nn, "hipMalloc(indices)"))) return rc;
        if ((rc = ensure(ctx, &ctx->idxB, &ctx->idxBCap, nn, "hipMalloc(indices tmp)"))) return rc;
        k_morton64<<<gMorton, kBuildBlock, 0, stream>>>(tris, n, ctx->sceneBox, ctx->recB, ctx->idxA, ctx->flags,
                                                        ctx->ctrl);
        // keys in recB (tmp: recA), records written to recA afterwards
        rc = sort_pairs_u64(ctx, ctx->recB, ctx->idxA, ctx->recA, ctx->idxB, (uint32_t)n, 0, 63, stream);
        if (rc) return rc;
        k_deltas64<<<g, kBuildBlock, 0, stream>>>(ctx->recB, ctx->idxA, n, ctx->deltas, ctx->recA);
    } else {
        const uint32_t tiles = sort_tiles(ctx, (uint32_t)n);
        const unsigned gm = gMorton < tiles ? gMorton : tiles;
        // partials: gm blocks x 4 histograms, then the per-tile counts
        if ((rc = ensure_partials(ctx, (int)(gm * 4 + tiles), 1))) return rc;
        uint32_t* tileCounts = ctx->partials + (size_t)gm * 4 * 256;
        k_morton<<<gm, kMortonBlock, 0, stream>>>(tris, n, ctx->sceneBox, opt, ctx->recA, ctx->partials, ctx->flags,
                                                 ctx->ctrl, (uint32_t)sort_tile_keys(ctx), tiles, tileCounts);
        launch_hist_reduce(ctx, (int)gm, 4, stream);
        rc = sort_records_u64(ctx, ctx->recA, ctx->recB, (uint32_t)n, 0, 32, true, stream, tileCounts);
        if (rc) return rc;
    }
    // 32-bit keys: the local climb computes the deltas from the records
    if ((rc = launch_climb(ctx, tris, n, ctx->recA, ctx->deltas, !wide, nodes, root, stream))) return rc;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_status(e, "build kernel launch");
    if (d_sorted) {
        e = hipMemcpyAsync(d_sorted, ctx->recA, nn * sizeof(uint64_t), hipMemcpyDeviceToDevice, stream);
        if (e != hipSuccess) return hip_status(e, "hipMemcpyAsync(sorted)");
    }
    if (d_deltas) {
        e = hipMemcpyAsync(d_deltas, ctx->deltas, nn - 1, hipMemcpyDeviceToDevice, stream);
        if (e != hipSuccess) return hip_status(e, "hipMemcpyAsync(deltas)");
    }
    return MLBVH_OK;
}

int build_one(mlbvh_ctx* ctx, const void* d_tris, int n, uint32_t opt, void* d_nodes, void* d_root,
              uint64_t* d_sorted, uint8_t* d_deltas, hipStream_t stream)
{
    if (!d_tris) {
        set_error("build: null argument");
        return MLBVH_E_INVALID_ARG;
    }
    return build_impl(ctx, (const Tri*)d_tris, n, opt, d_nodes, d_root, d_sorted, d_deltas, stream);
}

// Indexed-mesh source from the C-ABI arguments, or an error.
static int indexed_source(const float* d_vertices, int nVertices, const uint32_t* d_indices, const char* who,
                          IndexedTris* out)
{
    if (!d_vertices || !d_indices || nVertices < 1) {
        set_error("%s: null vertex/index buffer or nVertices = %d", who, nVertices);
        return MLBVH_E_INVALID_ARG;
    }
    out->v = d_vertices;
    out->f = d_indices;
    out->vMax = (uint32_t)nVertices - 1u;
    return MLBVH_OK;
}

template <class TS>
static int refit_impl(mlbvh_ctx* ctx, TS tris, int n, const uint64_t* d_sorted_records, const uint8_t* d_deltas,
                      void* d_nodes, void* d_root, hipStream_t s)
{
    if (!ctx || !d_sorted_records || !d_deltas || !d_nodes || !d_root) {
        set_error("refit: null argument");
        return MLBVH_E_INVALID_ARG;
    }
    if (n < 2) {
        set_error("refit: n = %d; at least 2 triangles are required", n);
        return MLBVH_E_INVALID_ARG;
    }
    int rc = set_device(ctx);
    if (rc) return rc;
    if ((rc = ensure(ctx, &ctx->flags, &ctx->flagsCap, (size_t)n, "hipMalloc(flags)"))) return rc;
    unsigned g = grid1((size_t)n);
    k_climb_reset<<<g, kBuildBlock, 0, s>>>(ctx->flags, n - 1, ctx->ctrl);
    if ((rc = launch_climb(ctx, tris, n, d_sorted_records, const_cast<uint8_t*>(d_deltas), false, (Node*)d_nodes,
                           (uint32_t*)d_root, s)))
        return rc;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_status(e, "refit kernel launch");
    return MLBVH_OK;
}

}  // namespace mlbvh

using namespace mlbvh;

extern "C" int mlbvh_build(mlbvh_ctx* ctx, const void* d_triangles, int n, uint32_t buildOption, void* d_nodes,
                           void* d_root, void* stream)
{
    return build_one(ctx, d_triangles, n, buildOption, d_nodes, d_root, nullptr, nullptr, (hipStream_t)stream);
}

extern "C" int mlbvh_build_ex(mlbvh_ctx* ctx, const void* d_triangles, int n, uint32_t buildOption, void* d_nodes,
                              void* d_root, uint64_t* d_sorted_records, uint8_t* d_deltas, void* stream)
{
    return build_one(ctx, d_triangles, n, buildOption, d_nodes, d_root, d_sorted_records, d_deltas,
                      (hipStream_t)stream);
}

extern "C" int mlbvh_refit(mlbvh_ctx* ctx, const void* d_triangles, int n, const uint64_t* d_sorted_records,
                           const uint8_t* d_deltas, void* d_nodes, void* d_root, void* stream)
{
    if (!d_triangles) {
        set_error("refit: null argument");
        return MLBVH_E_INVALID_ARG;
    }
    return refit_impl(ctx, (const Tri*)d_triangles, n, d_sorted_records, d_deltas, d_nodes, d_root,
                      (hipStream_t)stream);
}

extern "C" int mlbvh_build_indexed(mlbvh_ctx* ctx, const float* d_vertices, int nVertices, const uint32_t* d_indices,
                                   int nTriangles, uint32_t buildOption, void* d_nodes, void* d_root,
                                   uint64_t* d_sorted_records, uint8_t* d_deltas, void* stream)
{
    IndexedTris src;
    int rc = indexed_source(d_vertices, nVertices, d_indices, "build_indexed", &src);
    if (rc) return rc;
    return build_impl(ctx, src, nTriangles, buildOption, d_nodes, d_root, d_sorted_records, d_deltas,
                      (hipStream_t)stream);
}

extern "C" int mlbvh_refit_indexed(mlbvh_ctx* ctx, const float* d_vertices, int nVertices, const uint32_t* d_indices,
                                   int nTriangles, const uint64_t* d_sorted_records, const uint8_t* d_deltas,
                                   void* d_nodes, void* d_root, void* stream)
{
    IndexedTris src;
    int rc = indexed_source(d_vertices, nVertices, d_indices, "refit_indexed", &src);
    if (rc) return rc;
    return refit_impl(ctx, src, nTriangles, d_sorted_records, d_deltas, d_nodes, d_root, (hipStream_t)stream);
}
