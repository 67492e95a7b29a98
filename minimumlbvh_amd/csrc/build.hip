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
// Phase 1 (k_climb_local): a block of kLocalLeaves threads builds its
// block-local nodes without climbing, each node on its own.  For internal
// node p, Lp and Rp come from two nearest-greater searches (binary lifting
// over a sparse table of the block's deltas in LDS).  Its parent is Lp - 1
// if d[Lp - 1] < d[Rp], else Rp (the climb's direction rule), and every
// node and leaf writes its index into its parent's child slot.  Its two
// child boxes are the unions of the leaf boxes over [Lp, p] and [p + 1, Rp],
// read from a segment tree of the block's leaf boxes in LDS.  The box merge
// (umin / umax: min and max under a total order that puts -0 below +0) is
// associative and commutative, so these unions equal the climb's pairwise
// merges bit for bit.  The tree holds each coordinate as its order-
// preserving integer key (f2ord), on which umin / umax are integer min / max
// for every non-NaN value: a box merge is six integer instructions.  A leaf or block-local node whose parent is not
// block-local goes to the frontier list with its range: the block's own
// stretch of the list, entries b0 .. b0 + m - 1 (a block has at most one
// entry per leaf), closed by an end mark when shorter.  Each block-local
// node is written whole (64 bytes) by its thread; one whose parent is not
// block-local gets its parent word again from phase 2.  The block's barriers
// order LDS only, so its global stores are never waited for.
//
// Phase 2 (k_climb_upper): the frontier entries (about 1% of the nodes for
// surface meshes) climb on through the device-scope arrival words
// (climb_from).  The result is the same tree as one thread per leaf climbing
// through device memory all the way.
#ifndef MLBVH_CLIMB_LEAVES
#define MLBVH_CLIMB_LEAVES 1024
#endif
constexpr int kLocalLeaves = MLBVH_CLIMB_LEAVES;
constexpr uint32_t kDeltaInf = 255;  // the missing d[-1] and d[n-1]: above every delta
constexpr int log2_exact(int v) { return v <= 1 ? 0 : 1 + log2_exact(v / 2); }
constexpr int kLiftLevels = log2_exact(kLocalLeaves) + 1;  // windows of 1 .. kLocalLeaves deltas
constexpr uint32_t kFrontierEnd = 0xFFFFFFFFu;  // no node or leaf index (n < 2^31)

// A box as order-preserving keys of its six coordinates.
struct KeyBox {
    uint32_t lo[3], hi[3];
};
__device__ inline KeyBox key_box(const Box& b)
{
    return {{f2ord(b.lo.x), f2ord(b.lo.y), f2ord(b.lo.z)}, {f2ord(b.hi.x), f2ord(b.hi.y), f2ord(b.hi.z)}};
}
__device__ inline void key_merge(KeyBox& a, const KeyBox& b)
{
#pragma unroll
    for (int k = 0; k < 3; k++) {
        a.lo[k] = a.lo[k] < b.lo[k] ? a.lo[k] : b.lo[k];
        a.hi[k] = a.hi[k] > b.hi[k] ? a.hi[k] : b.hi[k];
    }
}

// kFuseDeltas: the deltas come from the block's own sorted records
// (delta(a, b) = 32 - clz(a ^ b) of neighbouring keys, minimum_lbvh.cu:94-102)
// and the block writes d[b0 .. b1] to `deltasOut` for the upper phase and the
// caller; otherwise they are read from `deltas` (63-bit keys, refit).
template <class TS, bool kFuseDeltas>
__global__ void __launch_bounds__(kLocalLeaves)
    k_climb_local(TS tris, uint32_t n, const uint64_t* __restrict__ rec, const uint8_t* __restrict__ deltas,
                  uint8_t* __restrict__ deltasOut, Node* __restrict__ nodes, uint32_t* __restrict__ frontier,
                  uint32_t* __restrict__ root)
{
    static_assert(kLocalLeaves == 1 << (kLiftLevels - 1) && kLocalLeaves >= 256 && kLocalLeaves <= 1024,
                  "a power-of-two block of 256 .. 1024 leaves");
    constexpr int kWaves = kLocalLeaves / 64;
    constexpr int kRow = kLocalLeaves + 16;
    // Local delta j is d[b0 - 1 + j], j in [0, m]; s_st[k][j] = max of the
    // local deltas j .. j + 2^k - 1.
    __shared__ uint8_t s_st[kLiftLevels][kRow];
    // Segment tree of the leaf boxes: leaf j at kLocalLeaves + j, node i the
    // union of nodes 2i and 2i + 1.
    __shared__ KeyBox s_tree[2 * kLocalLeaves];
    __shared__ uint32_t s_child[kLocalLeaves][2];
    __shared__ uint8_t s_loc[kLocalLeaves];  // internal node b0 + j lies inside the block
    __shared__ uint32_t s_wcount[kWaves];

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t b0 = blockIdx.x * kLocalLeaves;
    const uint32_t m = n - b0 < (uint32_t)kLocalLeaves ? n - b0 : (uint32_t)kLocalLeaves;
    const uint32_t nInternals = n - 1;
    const uint32_t pos = b0 + (uint32_t)t;
    const bool isLeaf = (uint32_t)t < m;
    const bool isNode = (uint32_t)t + 1 < m;  // internal node pos: the split between leaves pos and pos + 1

    const uint64_t myRec = isLeaf ? rec[pos] : 0ull;
    const uint32_t triIndex = (uint32_t)(myRec >> 32);
    Box leafBox;
    if (isLeaf) {
        leafBox = tri_box(tris[triIndex]);
    } else {
        leafBox.lo = mk(kFltMax, kFltMax, kFltMax);
        leafBox.hi = mk(-kFltMax, -kFltMax, -kFltMax);
    }
    s_tree[kLocalLeaves + t] = key_box(leafBox);

    if (kFuseDeltas) {
        __shared__ uint32_t s_k[kLocalLeaves + 2];  // s_k[j] = key of sorted position b0 - 1 + j
        if (isLeaf) s_k[t + 1] = (uint32_t)myRec;
        if (t == 0 && b0 > 0) s_k[0] = (uint32_t)rec[b0 - 1];
        if ((uint32_t)t == m - 1 && b0 + m < n) s_k[m + 1] = (uint32_t)rec[b0 + m];
        lds_barrier();
        auto delta_at = [&](uint32_t j) -> uint32_t {
            const uint32_t g = b0 + j;  // global delta index + 1
            return (g == 0 || g - 1 >= nInternals) ? kDeltaInf : (uint32_t)(32 - __clz(s_k[j] ^ s_k[j + 1]));
        };
        s_st[0][t] = (uint8_t)delta_at((uint32_t)t);
        if (t == kLocalLeaves - 1) s_st[0][kLocalLeaves] = (uint8_t)delta_at(kLocalLeaves);
    } else {
        auto delta_at = [&](uint32_t j) -> uint32_t {
            const uint32_t g = b0 + j;
            return (g == 0 || g - 1 >= nInternals) ? kDeltaInf : (uint32_t)deltas[g - 1];
        };
        s_st[0][t] = (uint8_t)delta_at((uint32_t)t);
        if (t == kLocalLeaves - 1) s_st[0][kLocalLeaves] = (uint8_t)delta_at(kLocalLeaves);
    }
    lds_barrier();
    // Two levels per step: level k from k - 1, level k + 1 from the four
    // windows of level k - 1 (the box merge is exact, so the order does not
    // matter).
#pragma unroll
    for (int k = 1; k < kLiftLevels; k += 2) {
        const int h = 1 << (k - 1);
        if (t + 2 * h - 1 <= kLocalLeaves) {
            const uint32_t a = s_st[k - 1][t], b = s_st[k - 1][t + h];
            s_st[k][t] = (uint8_t)(a > b ? a : b);
        }
        if (k + 1 < kLiftLevels && t + 4 * h - 1 <= kLocalLeaves) {
            uint32_t a = s_st[k - 1][t], b = s_st[k - 1][t + h], c = s_st[k - 1][t + 2 * h], d = s_st[k - 1][t + 3 * h];
            a = a > b ? a : b;
            c = c > d ? c : d;
            s_st[k + 1][t] = (uint8_t)(a > c ? a : c);
        }
        const int first = kLocalLeaves >> k;  // level k of the tree: nodes first .. 2 first - 1
        if (t < first) {
            KeyBox u = s_tree[2 * (first + t)];
            key_merge(u, s_tree[2 * (first + t) + 1]);
            s_tree[first + t] = u;
        } else if (t >= kLocalLeaves / 2 && t - kLocalLeaves / 2 < first / 2) {  // level k + 1 from level k - 1
            const int i = first / 2 + (t - kLocalLeaves / 2);
            KeyBox u = s_tree[4 * i], v = s_tree[4 * i + 2];
            key_merge(u, s_tree[4 * i + 1]);
            key_merge(v, s_tree[4 * i + 3]);
            key_merge(u, v);
            s_tree[i] = u;
        }
        lds_barrier();
    }

    // Node pos has local delta jp = t + 1.  jL: the last j < jp with
    // d_j >= d_jp, so Lp = b0 + jL; jR: the first j > jp with d_j > d_jp, so
    // Rp = b0 - 1 + jR.  The node is block-local iff both lie in [0, m].
    int jL = -1, jR = (int)m + 1;
    bool local = false;
    if (isNode) {
        const uint32_t x = s_st[0][t + 1];
        jL = t;
        jR = t + 2;
#pragma unroll
        for (int k = kLiftLevels - 1; k >= 0; k--) {
            const int w = 1 << k;
            if (jL - w + 1 >= 0 && s_st[k][jL - w + 1] < x) jL -= w;
            if (jR + w - 1 <= (int)m && s_st[k][jR] <= x) jR += w;
        }
        local = jL >= 0 && jR <= (int)m;
        s_loc[t] = (uint8_t)local;
    }
    lds_barrier();

    // Parents, child slots and frontier entries.
    bool emitLeaf = false, emitNode = false, isRoot = false;
    uint32_t parent = kInvalidNode;
    const uint32_t Lp = b0 + (uint32_t)jL, Rp = b0 - 1 + (uint32_t)jR;
    if (isLeaf) {
        const uint32_t dL = s_st[0][t], dR = s_st[0][t + 1];
        const int goLeft = dL < dR ? 1 : 0;  // parent pos - 1 (slot t - 1) or pos (slot t)
        const bool pLocal = goLeft ? (t > 0 && s_loc[t - 1] != 0) : (isNode && s_loc[t] != 0);
        if (pLocal)
            s_child[goLeft ? t - 1 : t][goLeft] = triIndex | kLeafBit;
        else
            emitLeaf = true;
    }
    if (local) {
        isRoot = Lp == 0 && Rp == nInternals;
        if (isRoot) {
            *root = pos;
        } else {
            const uint32_t dL = s_st[0][jL], dR = s_st[0][jR];
            const int goLeft = dL < dR ? 1 : 0;
            parent = goLeft ? Lp - 1 : Rp;
            const uint32_t slot = parent - b0;  // wraps to a huge value for the parent b0 - 1
            const bool pLocal = slot < m - 1u && s_loc[slot] != 0;
            if (pLocal)
                s_child[slot][goLeft] = pos;
            else
                emitNode = true;
        }
    }

    // frontier entries: compacted into the block's stretch of the list
    const uint64_t leafMask = __ballot(emitLeaf), nodeMask = __ballot(emitNode);
    if (lane == 0) s_wcount[wave] = (uint32_t)(__popcll(leafMask) + __popcll(nodeMask));
    lds_barrier();  // also publishes the child slots
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) {
        const uint32_t c = s_wcount[w];
        before += w < wave ? c : 0u;
        total += c;
    }
    if (emitLeaf || emitNode) {
        const uint64_t below = (1ull << lane) - 1ull;
        uint32_t k = b0 + before + (uint32_t)(__popcll(leafMask & below) + __popcll(nodeMask & below));
        if (emitLeaf) {
            frontier[3 * (size_t)k] = triIndex | kLeafBit;
            frontier[3 * (size_t)k + 1] = pos;
            frontier[3 * (size_t)k + 2] = pos;
            k++;
        }
        if (emitNode) {
            frontier[3 * (size_t)k] = pos;
            frontier[3 * (size_t)k + 1] = Lp;
            frontier[3 * (size_t)k + 2] = Rp;
        }
    }
    if (t == 0 && total < m) frontier[3 * (size_t)(b0 + total)] = kFrontierEnd;

    // The block-local node, whole: child boxes over [Lp, pos] and [pos + 1, Rp].
    if (local) {
        auto range_box = [&](uint32_t a, uint32_t b) -> Box {  // local leaves a .. b
            KeyBox acc = {{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, {0u, 0u, 0u}};
            uint32_t l = a + kLocalLeaves, r = b + kLocalLeaves + 1;
            while (l < r) {
                if (l & 1) key_merge(acc, s_tree[l++]);
                if (r & 1) key_merge(acc, s_tree[--r]);
                l >>= 1;
                r >>= 1;
            }
            Box out;
            out.lo = mk(ord2f(acc.lo[0]), ord2f(acc.lo[1]), ord2f(acc.lo[2]));
            out.hi = mk(ord2f(acc.hi[0]), ord2f(acc.hi[1]), ord2f(acc.hi[2]));
            return out;
        };
        const Box l = range_box(Lp - b0, (uint32_t)t), r = range_box((uint32_t)t + 1, Rp - b0);
        const uint32_t context = Rp < nInternals ? Rp : 0xFFFFFFFFu;
        uint4* dst = reinterpret_cast<uint4*>(nodes + pos);
        dst[0] = make_uint4(s_child[t][0], s_child[t][1], __float_as_uint(l.lo.x), __float_as_uint(l.lo.y));
        dst[1] = make_uint4(__float_as_uint(l.lo.z), __float_as_uint(l.hi.x), __float_as_uint(l.hi.y),
                            __float_as_uint(l.hi.z));
        dst[2] = make_uint4(__float_as_uint(r.lo.x), __float_as_uint(r.lo.y), __float_as_uint(r.lo.z),
                            __float_as_uint(r.hi.x));
        dst[3] = make_uint4(__float_as_uint(r.hi.y), __float_as_uint(r.hi.z), parent, context);
    }
    // d[b0 + t], stored last: nothing in the block waits for a global store
    if (kFuseDeltas && isLeaf && pos < nInternals) deltasOut[pos] = s_st[0][t + 1];
}

// One wave per block of phase 1, over that block's stretch of the frontier
// list up to its end mark.
template <class TS>
__global__ void __launch_bounds__(64)
    k_climb_upper(TS tris, uint32_t n, const uint8_t* __restrict__ deltas, Node* nodes, uint32_t* flags,
                  const uint32_t* __restrict__ frontier, uint32_t* root)
{
    const uint32_t b0 = blockIdx.x * kLocalLeaves;
    const uint32_t m = n - b0 < (uint32_t)kLocalLeaves ? n - b0 : (uint32_t)kLocalLeaves;
    for (uint32_t base = 0; base < m; base += 64) {
        const uint32_t k = base + threadIdx.x;
        const size_t e = 3 * (size_t)(b0 + k);
        const uint32_t node = k < m ? frontier[e] : kFrontierEnd;
        const uint64_t endMask = __ballot(node == kFrontierEnd);
        const uint32_t valid = endMask ? (uint32_t)__ffsll((unsigned long long)endMask) - 1u : 64u;
        if (threadIdx.x < valid) {
            const uint32_t lo = frontier[e + 1], hi = frontier[e + 2];
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
        if (endMask) break;
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
                                                                     ctx->frontier, root);
    else
        k_climb_local<TS, false><<<blocks, kLocalLeaves, 0, stream>>>(tris, (uint32_t)n, rec, deltas, nullptr, nodes,
                                                                      ctx->frontier, root);
    k_climb_upper<<<blocks, 64, 0, stream>>>(tris, (uint32_t)n, deltas, nodes, ctx->flags, ctx->frontier, root);
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
