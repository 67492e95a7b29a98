// LBVH build for gfx950: scene AABB -> Morton records (+ fused digit
// histograms) -> Onesweep sort of (morton, index) -> deltas -> agglomerative
// bottom-up climb.  Reference: BVHGPUBuilder::build (minimum_lbvh.h:783-871),
// kernels of minimum_lbvh.cu:38-120 and the climb build_lbvh
// (minimum_lbvh.h:284-361).
//
// Differences in structure from the reference (results are identical):
//  * the Morton kernel also builds the four 8-bit digit histograms the sort
//    needs, so the sort starts with its first reorder pass; no memset of the
//    node array is needed (every node is written whole once);
//  * no climb inside blocks: nodes whose leaves lie in one block of 256
//    sorted leaves are built each on its own from the deltas (ranges by
//    nearest-greater searches, boxes by range unions in LDS), then the same
//    over blocks of 1024 of the resulting subtrees; the rest (about 0.05% of
//    the nodes of a surface mesh) climbs inside one workgroup with its
//    hand-offs in LDS (k_top_tree), or, when there are more of them than it
//    holds, through device-scope arrival words; `context` is written once,
//    with the reference's final value;
//  * one thread per leaf for all n leaves (the reference launches
//    ceil((n-1)/256) blocks and drops the last leaf when (n-1) % 256 == 0).
#include <algorithm>

#include "mlbvh_ctx.h"

namespace mlbvh {

constexpr int kBuildBlock = 256;
// The scene box is reduced into kBoxCopies copies of its 6 coordinates
// (block b uses copy b % kBoxCopies), so that the blocks' atomics spread over
// many L2 lines; the Morton kernels fold the copies.  Each slot is a 64-bit
// word, tag << 32 | key, merged with atomic max: key = f2ord(hi) for the
// upper bounds and ~f2ord(lo) for the lower ones, tag = the build's tag
// (kCtrlBuildTag + 1).  A newer tag wins over whatever an earlier build left,
// so the slots are never cleared (zeroed slots carry tag 0, which no build
// uses; level 1 advances the tag for the next build, and zeroes the slots
// when it would wrap).
#ifndef MLBVH_BOX_COPIES
#define MLBVH_BOX_COPIES 32
#endif
constexpr int kBoxCopies = MLBVH_BOX_COPIES;
constexpr int kBoxWords = 6 * kBoxCopies;
static_assert(kBoxWords <= kSceneBoxWords, "the box copies fit the context's scene-box slots");

// The scene box from the kBoxCopies partial boxes (all threads of the block
// call it; the result is the same in every thread).
__device__ inline void fold_scene_box(const uint64_t* sceneBox, f3* lo, f3* hi)
{
    __shared__ uint32_t s_box[6];
    if (threadIdx.x < 6) {
        uint64_t r = sceneBox[threadIdx.x];
        for (int c = 1; c < kBoxCopies; c++) {
            const uint64_t v = sceneBox[6 * c + threadIdx.x];
            r = v > r ? v : r;
        }
        s_box[threadIdx.x] = threadIdx.x < 3 ? ~(uint32_t)r : (uint32_t)r;
    }
    __syncthreads();
    *lo = mk(ord2f(s_box[0]), ord2f(s_box[1]), ord2f(s_box[2]));
    *hi = mk(ord2f(s_box[3]), ord2f(s_box[4]), ord2f(s_box[5]));
}

// Scene bounds over all vertices: per-thread running min/max, wave and block
// reduction, then 6 atomics per block on the tagged slots above.  Block 0
// also sets up the state of the 4-pass record sort.  TS, here and in the
// kernels below, is the triangle source: `const Tri*` (a soup) or
// IndexedTris.
template <class TS>
__global__ void __launch_bounds__(kBuildBlock)
    k_scene_aabb(TS tris, int n, unsigned long long* sceneBox, uint32_t* ctrl, uint32_t* hist, uint64_t* status,
                 size_t statusWords)
{
    if (blockIdx.x == 0) sort_state_init(ctrl, 4, hist, 4, status, statusWords);
    const unsigned long long tag = (unsigned long long)(ctrl[kCtrlBuildTag] + 1u) << 32;
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
        unsigned long long* box = sceneBox + 6 * (blockIdx.x % kBoxCopies);
        atomicMax(&box[c], tag | (c < 3 ? ~f2ord(r) : f2ord(r)));
    }
}

// Morton record per triangle (minimum_lbvh.cu:68-92 / minimum_lbvh.h:534-553),
// and the 4 digit histograms of the 32-bit key.
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
    k_morton(TS tris, int n, const uint64_t* __restrict__ sceneBox, uint32_t opt,
             uint64_t* __restrict__ rec, uint32_t* __restrict__ partials, uint32_t tileKeys, uint32_t nTiles,
             uint32_t* __restrict__ tileCounts)
{
    __shared__ uint32_t s_h[4][256];
    __shared__ uint32_t s_t[256];  // this tile's counts of the first digit
    const int t = threadIdx.x;
    for (int i = t; i < 4 * 256; i += kMortonBlock) (&s_h[0][0])[i] = 0;
    if (t < 256) s_t[t] = 0;
    f3 lo, hi;
    fold_scene_box(sceneBox, &lo, &hi);  // its barrier also orders the clears above
    // One LDS atomic per wave when the active lanes share the digit (the high
    // digits of spatially ordered triangles mostly do), else one per lane.
    auto hist_add = [&](uint32_t* h, uint32_t d) {
        const uint64_t active = __ballot(1);
        const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
        const uint64_t same = __ballot(d == d0);
        if (same == active) {
            if ((uint32_t)__lane_id() == (uint32_t)__ffsll((unsigned long long)active) - 1u)
                atomicAdd(&h[d0], (uint32_t)__popcll(active));
        } else {
            atomicAdd(&h[d], 1u);
        }
    };
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
        hist_add(s_t, morton & 0xFFu);
#pragma unroll
        for (int p = 1; p < 4; p++) hist_add(s_h[p], (morton >> (8 * p)) & 0xFFu);
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
// triangle index, for a key-value sort over bits [0, 63).
template <class TS>
__global__ void __launch_bounds__(kBuildBlock)
    k_morton64(TS tris, int n, const uint64_t* __restrict__ sceneBox, uint64_t* __restrict__ keys,
               uint32_t* __restrict__ idx)
{
    f3 lo, hi;
    fold_scene_box(sceneBox, &lo, &hi);
    for (int i = blockIdx.x * kBuildBlock + threadIdx.x; i < n; i += gridDim.x * kBuildBlock) {
        Tri tri = tris[i];
        f3 center = ((tri.v[0] + tri.v[1]) + tri.v[2]) / 3.0f;
        keys[i] = morton_in_box(lo, hi, center);
        idx[i] = (uint32_t)i;
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

// Agglomerative climb (build_lbvh, minimum_lbvh.h:284-361), built level by
// level without climbing where the nodes fit a block.
//
// Internal node p (the split between sorted leaves p and p+1) covers the
// leaves [Lp, Rp] with Lp = 1 + (last q < p with d[q] >= d[p], else -1) and
// Rp = (first q > p with d[q] > d[p], else n-1): the climb merges across the
// smaller boundary delta first and goes right on ties.  Its parent is Lp - 1
// if d[Lp - 1] < d[Rp], else Rp (a missing d[-1] or d[n-1] counts as above
// every delta).
//
// k_local_nodes builds, for a block of consecutive elements of a sequence,
// every node whose range lies inside the block, each node on its own:
//   - the block's deltas go into a sparse table of range maxima in LDS, and
//     the elements' boxes into a segment tree in LDS;
//   - node p finds the two ends of its range by binary lifting over the
//     sparse table; it is block-local iff both lie in the block;
//   - every element and block-local node writes its index into its parent's
//     child slot (when that parent is block-local), else it becomes a
//     frontier entry (node, first leaf, last leaf);
//   - the child boxes are unions over the element ranges [Lp, p] and
//     [p + 1, Rp] from the segment tree.  The box merge (umin / umax: min
//     and max under a total order that puts -0 below +0) is associative and
//     commutative, so these unions equal the climb's pairwise merges bit for
//     bit.  The tree holds each coordinate as its order-preserving integer
//     key (f2ord), on which umin / umax are integer min / max for every
//     non-NaN value;
//   - each block-local node is written whole (64 bytes) by its thread, and
//     the block's frontier entries go to its own stretch of the frontier
//     list (entries b0 .. b0 + m - 1: at most one per element) in element
//     order, closed by an end mark when shorter, with their count.
// Level 1 runs it over the sorted leaves in blocks of kLeaves1.  Its frontier
// entries, in leaf order over all blocks, are the maximal level-1 subtrees:
// a sequence of "super-leaves" that tiles the leaves.  Between super-leaves
// j and j + 1 lies internal node hi_j with delta d[hi_j], and the nodes above
// level 1 are exactly the tree these splits form under the same rule, so
// level 2 runs the same kernel over the super-leaves in blocks of kLeaves2.
// Level 1's frontier stretches are laid out as the super-leaf list by
// k_frontier_compact.  What crosses a level-2 block (a few thousand nodes for
// surface meshes) is built by k_top_tree from level 2's frontier records, by
// the same rule applied to the climb of one thread per entry through LDS; if
// there are more than kTopMax entries, the same launch climbs them through the
// device-scope arrival words (climb_from) instead.  The result is the tree of
// one thread per leaf climbing through device memory all the way.  A node
// whose parent is built by a later level gets its parent word from its own
// level (the rule above gives it) and again, identically, from the climb.
// The blocks' barriers order LDS only, so their global stores are never
// waited for.
#ifndef MLBVH_LEVEL1_LEAVES
#define MLBVH_LEVEL1_LEAVES 256
#endif
constexpr int kLeaves1 = MLBVH_LEVEL1_LEAVES;
constexpr int kLeaves2 = 1024;
constexpr uint32_t kDeltaInf = 255;  // the missing d[-1] and d[n-1]: above every delta
constexpr int log2_exact(int v) { return v <= 1 ? 0 : 1 + log2_exact(v / 2); }
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

// Sequences k_local_nodes runs over.
enum LocalMode {
    kLeavesFusedDeltas = 0,  // sorted leaves; deltas from the records' keys (32 - clz(a ^ b),
                             // minimum_lbvh.cu:94-102), written to deltasOut
    kLeavesGivenDeltas = 1,  // sorted leaves; deltas read from `deltas` (63-bit keys, refit)
    kSuperLeaves = 2,        // the super-leaves (node, lo, hi) of `elems`, ctrl[kCtrlSuperCount] of them
};

// Single-word decoupled look-back of the build's ordered workgroups (the
// level-1 frontier compaction, the level-2 blocks).  Word of workgroup j:
// tag << 32 | prefix flag << 31 | value, value = its own count (flag clear)
// or the count of all workgroups up to and including it (flag set).  The
// tag changes with every build, so the words are never cleared (a zeroed
// word carries tag 0, which no build uses).
constexpr uint64_t kLbPrefix = 1ull << 31;
__device__ inline uint64_t lb_word(uint32_t tag, bool prefix, uint32_t value)
{
    return ((uint64_t)tag << 32) | (prefix ? kLbPrefix : 0ull) | (uint64_t)(value & 0x7FFFFFFFu);
}

// Sum of the counts of workgroups 0 .. idx - 1, by the 64 lanes of one wave
// (the loop is the same in every lane): lane l reads workgroup base - l;
// the lanes up to the nearest published prefix (lane 0 is the nearest) must
// all be published, then their values are summed.  Workgroups before the
// first count as a published prefix of 0.  The caller orders its workgroups
// by ticket, so every workgroup waited on is running and publishes its count
// before it waits on anything.
__device__ inline uint32_t lb_exclusive(const uint64_t* status, uint32_t idx, uint32_t tag, int lane)
{
    uint32_t excl = 0;
    int64_t base = (int64_t)idx - 1;
    while (base >= 0) {
        const int64_t j = base - lane;
        const uint64_t w =
            j >= 0 ? __hip_atomic_load(status + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : lb_word(tag, true, 0u);
        const bool pub = (uint32_t)(w >> 32) == tag;
        const uint64_t pubMask = __ballot(pub);
        const uint64_t preMask = __ballot(pub && (w & kLbPrefix) != 0);
        const int first = preMask ? __ffsll((unsigned long long)preMask) - 1 : 64;
        const uint64_t need = first >= 63 ? ~0ull : ((2ull << first) - 1ull);
        if ((pubMask & need) == need) {
            uint32_t v = lane <= first ? (uint32_t)(w & 0x7FFFFFFFu) : 0u;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            excl += v;
            base = first < 64 ? -1 : base - 64;
        } else {
            __builtin_amdgcn_s_sleep(1);
        }
    }
    return excl;
}

// lb_exclusive by all kWaves waves of a block at once (kWaves * 64
// workgroups per round); every thread calls it, and it returns the same value
// in every thread.  s_pub / s_pre / s_sum: kWaves words of LDS each.
template <int kWaves>
__device__ inline uint32_t lb_exclusive_block(const uint64_t* status, uint32_t idx, uint32_t tag, uint64_t* s_pub,
                                              uint64_t* s_pre, uint32_t* s_sum)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t excl = 0;
    int64_t base = (int64_t)idx - 1;
    while (base >= 0) {  // the same in every thread
        const int64_t j = base - (int64_t)threadIdx.x;
        const uint64_t w =
            j >= 0 ? __hip_atomic_load(status + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : lb_word(tag, true, 0u);
        const bool pub = (uint32_t)(w >> 32) == tag;
        const uint64_t pubMask = __ballot(pub), preMask = __ballot(pub && (w & kLbPrefix) != 0);
        if (lane == 0) {
            s_pub[wave] = pubMask;
            s_pre[wave] = preMask;
        }
        lds_barrier();
        // the nearest prefix over the block's window, thread t reading base - t
        int first = kWaves * 64;
        bool ready = true;
#pragma unroll
        for (int k = 0; k < kWaves; k++) {
            const uint64_t pre = s_pre[k], pubk = s_pub[k];
            if (first == kWaves * 64) {
                const int f = pre ? __ffsll((unsigned long long)pre) - 1 : 64;
                const uint64_t need = f >= 63 ? ~0ull : ((2ull << f) - 1ull);
                if ((pubk & need) != need) ready = false;
                if (pre) first = 64 * k + f;
            }
        }
        uint32_t v = ready && (int)threadIdx.x <= first ? (uint32_t)(w & 0x7FFFFFFFu) : 0u;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) s_sum[wave] = v;
        lds_barrier();
        if (ready) {
            uint32_t sum = 0;
#pragma unroll
            for (int k = 0; k < kWaves; k++) sum += s_sum[k];
            excl += sum;
            base = first < kWaves * 64 ? -1 : base - kWaves * 64;
        } else {
            __builtin_amdgcn_s_sleep(1);
        }
        lds_barrier();  // s_pub / s_pre / s_sum are rewritten by the next round
    }
    return excl;
}

template <class TS, int kLeaves, int kMode>
__global__ void __launch_bounds__(kLeaves) __attribute__((amdgpu_waves_per_eu(kLeaves <= 256 ? 8 : 1)))
    k_local_nodes(TS tris, uint32_t n, const uint64_t* __restrict__ rec, const uint8_t* __restrict__ deltas,
                  uint8_t* __restrict__ deltasOut, const uint32_t* __restrict__ elems, uint32_t* __restrict__ ctrl,
                  Node* __restrict__ nodes, uint32_t* __restrict__ frontier, uint32_t* __restrict__ blockCounts,
                  uint32_t* __restrict__ flags, uint32_t* __restrict__ root, uint64_t* __restrict__ status2,
                  uint32_t* __restrict__ topList, uint64_t* __restrict__ sceneBox)
{
    constexpr int kLift = log2_exact(kLeaves) + 1;  // windows of 1 .. kLeaves deltas
    static_assert(kLeaves == 1 << (kLift - 1) && kLeaves >= 256 && kLeaves <= 1024,
                  "a power-of-two block of 256 .. 1024 elements");
    constexpr int kWaves = kLeaves / 64;
    constexpr int kRow = kLeaves + 16;
    // Local delta j belongs to the split between elements b0 - 1 + j and
    // b0 + j, j in [0, m]; s_st[k][j] = max of the local deltas j .. j + 2^k - 1.
    __shared__ uint8_t s_st[kLift][kRow];
    // Segment tree of the element boxes: element j at kLeaves + j, node i the
    // union of nodes 2i and 2i + 1.
    __shared__ KeyBox s_tree[2 * kLeaves];
    __shared__ uint32_t s_child[kLeaves][2];
    __shared__ uint8_t s_loc[kLeaves];   // split b0 + j lies inside the block
    __shared__ uint8_t s_mark[kLeaves];  // a frontier entry starts at element b0 + j
    // kLeavesFusedDeltas: s_aux[j] = key of sorted position b0 - 1 + j;
    // kSuperLeaves: s_aux[j] = first leaf of element b0 + j, s_aux[m] = last leaf of the block + 1
    __shared__ uint32_t s_aux[kLeaves + 2];
    __shared__ uint32_t s_wcount[kWaves];
    __shared__ uint32_t s_rank[kMode == kSuperLeaves ? kLeaves : 1];  // level 2: rank of the entry at element j
    __shared__ uint32_t s_blk;  // level 2: this block (by ticket)
    __shared__ uint64_t s_lbPub[kWaves], s_lbPre[kWaves];  // level 2: the look-back's window, per wave
    __shared__ uint32_t s_lbSum[kWaves];
    // The frontier entry starting at element j: its node and last leaf (its
    // first leaf is element j's), in the sparse table's LDS once that is dead.
    uint2* s_ent = reinterpret_cast<uint2*>(&s_st[0][0]);
    static_assert(sizeof(uint2) * kLeaves <= sizeof(s_st), "the entries fit the sparse table's LDS");

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    constexpr bool kLevel1 = kMode != kSuperLeaves;
    if (kLevel1 && blockIdx.x == 0) {
        // This call's later kernels start from ticket 0 with a fresh look-back
        // tag (their words carry tag + 1 != 0), and the next build's scene box
        // slots get a fresh tag (on wrap the slots restart from zero).  No
        // kernel before this one in the call reads these words, and none after
        // it reads them before it ends.
        const uint32_t boxTag = ctrl[kCtrlBuildTag] + 1u;
        if (t == 0) {
            ctrl[kCtrlCompactTicket] = 0u;
            ctrl[kCtrlL2Ticket] = 0u;
            const uint32_t tag = ctrl[kCtrlCompactTag] + 1u;
            ctrl[kCtrlCompactTag] = tag == 0xFFFFFFFFu ? 0u : tag;
        }
        if (boxTag == 0xFFFFFFFFu)
            for (int k = t; k < kSceneBoxWords; k += kLeaves) sceneBox[k] = 0ull;
        lds_barrier();  // every thread has read the box tag
        if (t == 0) ctrl[kCtrlBuildTag] = boxTag == 0xFFFFFFFFu ? 0u : boxTag;
    }
    const uint32_t count = kLevel1 ? n : ctrl[kCtrlSuperCount];
    const uint32_t nInternals = n - 1;
    // Level 1 has one block per kLeaves leaves; level 2 takes its blocks by
    // ticket (their number is known on the device only, and a block looks
    // back only at blocks that are running).
    uint32_t blk = blockIdx.x, tag = 0;
    if (!kLevel1) {
        if (t == 0) s_blk = atomicAdd(&ctrl[kCtrlL2Ticket], 1u);
        tag = ctrl[kCtrlCompactTag] + 1u;
        lds_barrier();
        blk = s_blk;
    }
    const uint32_t nBlocks = (uint32_t)(((size_t)count + kLeaves - 1) / kLeaves);
    while (blk < nBlocks) {
        const uint32_t b0 = blk * kLeaves;
        const uint32_t m = count - b0 < (uint32_t)kLeaves ? count - b0 : (uint32_t)kLeaves;
        const uint32_t s = b0 + (uint32_t)t;
        const bool isElem = (uint32_t)t < m;
        const bool isSplit = (uint32_t)t + 1 < m;  // split s: between elements s and s + 1

        // Element t: node, first and last leaf, box.
        uint32_t elemNode = 0, elemLo = s, elemHi = s;
        KeyBox kb = {{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, {0u, 0u, 0u}};
        uint64_t myRec = 0;
        if (kMode != kSuperLeaves) {
            myRec = isElem ? rec[s] : 0ull;
            elemNode = (uint32_t)(myRec >> 32) | kLeafBit;
            if (isElem) kb = key_box(tri_box(tris[(uint32_t)(myRec >> 32)]));
        } else if (isElem) {
            elemNode = elems[3 * (size_t)s];
            elemLo = elems[3 * (size_t)s + 1];
            elemHi = elems[3 * (size_t)s + 2];
            if (elemNode & kLeafBit) {
                kb = key_box(tri_box(tris[elemNode & ~kLeafBit]));
            } else {
                const Node& nd = nodes[elemNode];
                kb = key_box(nd.box[0]);
                key_merge(kb, key_box(nd.box[1]));
            }
        }
        s_tree[kLeaves + t] = kb;
        s_mark[t] = 0;

        // Local deltas.  Split b0 - 1 + t exists for t < m unless b0 + t == 0
        // (for t == m it is the sequence's last split's successor, or lies past
        // the block); its node is leaf position b0 - 1 + t (level 1) or the
        // last leaf of element b0 - 1 + t = elemLo - 1 (level 2).
        const bool jValid = isElem && s > 0;
        uint32_t dj = kDeltaInf, dLast = kDeltaInf;  // local deltas t and (thread kLeaves - 1) kLeaves
        const bool lastValid = t == kLeaves - 1 && m == (uint32_t)kLeaves && (size_t)b0 + kLeaves < count;
        if (kMode == kLeavesFusedDeltas) {
            if (isElem) s_aux[t + 1] = (uint32_t)myRec;
            if (t == 0 && b0 > 0) s_aux[0] = (uint32_t)rec[b0 - 1];
            if ((uint32_t)t == m - 1 && b0 + m < count) s_aux[m + 1] = (uint32_t)rec[b0 + m];
            lds_barrier();
            if (jValid) dj = 32 - __clz(s_aux[t] ^ s_aux[t + 1]);
            if (lastValid) dLast = 32 - __clz(s_aux[kLeaves] ^ s_aux[kLeaves + 1]);
        } else if (kMode == kLeavesGivenDeltas) {
            if (jValid) dj = deltas[s - 1];
            if (lastValid) dLast = deltas[s];
        } else {
            if (isElem) s_aux[t] = elemLo;
            if ((uint32_t)t == m - 1) s_aux[m] = elemHi + 1;
            if (jValid) dj = deltas[elemLo - 1];
            if (lastValid) dLast = deltas[elemHi];
        }
        s_st[0][t] = (uint8_t)dj;
        if (t == kLeaves - 1) s_st[0][kLeaves] = (uint8_t)dLast;
        lds_barrier();
        // Two levels per step: level k from k - 1, level k + 1 from the four
        // windows of level k - 1 (the merge is exact, so grouping does not matter).
#pragma unroll
        for (int k = 1; k < kLift; k += 2) {
            const int h = 1 << (k - 1);
            if (t + 2 * h - 1 <= kLeaves) {
                const uint32_t a = s_st[k - 1][t], b = s_st[k - 1][t + h];
                s_st[k][t] = (uint8_t)(a > b ? a : b);
            }
            if (k + 1 < kLift && t + 4 * h - 1 <= kLeaves) {
                uint32_t a = s_st[k - 1][t], b = s_st[k - 1][t + h], c = s_st[k - 1][t + 2 * h],
                         d = s_st[k - 1][t + 3 * h];
                a = a > b ? a : b;
                c = c > d ? c : d;
                s_st[k + 1][t] = (uint8_t)(a > c ? a : c);
            }
            const int first = kLeaves >> k;  // level k of the tree: nodes first .. 2 first - 1
            if (t < first) {
                KeyBox u = s_tree[2 * (first + t)];
                key_merge(u, s_tree[2 * (first + t) + 1]);
                s_tree[first + t] = u;
            } else if (t >= kLeaves / 2 && t - kLeaves / 2 < first / 2) {  // level k + 1 from level k - 1
                const int i = first / 2 + (t - kLeaves / 2);
                KeyBox u = s_tree[4 * i], v = s_tree[4 * i + 2];
                key_merge(u, s_tree[4 * i + 1]);
                key_merge(v, s_tree[4 * i + 3]);
                key_merge(u, v);
                s_tree[i] = u;
            }
            lds_barrier();
        }

        // Split s has local delta jp = t + 1.  jL: the last j < jp with
        // d_j >= d_jp; jR: the first j > jp with d_j > d_jp.  Its range is
        // elements b0 + jL .. b0 - 1 + jR; it is block-local iff both lie in [0, m].
        int jL = -1, jR = (int)m + 1;
        bool local = false;
        if (isSplit) {
            const uint32_t x = s_st[0][t + 1];
            jL = t;
            jR = t + 2;
#pragma unroll
            for (int k = kLift - 1; k >= 0; k--) {
                const int w = 1 << k;
                if (jL - w + 1 >= 0 && s_st[k][jL - w + 1] < x) jL -= w;
                if (jR + w - 1 <= (int)m && s_st[k][jR] <= x) jR += w;
            }
            local = jL >= 0 && jR <= (int)m;
            s_loc[t] = (uint8_t)local;
        }
        // Level 2: every split it does not build (one crossing its block, or
        // the one after its last element) is built by the upper climb, which
        // needs its arrival word open; no other node uses one.
        if (kMode == kSuperLeaves && ((isSplit && !local) || ((uint32_t)t == m - 1 && (size_t)b0 + m < count)))
            flags[elemHi] = 0xFFFFFFFFu;
        lds_barrier();

        // Parents, child slots and frontier entries.
        bool emitElem = false, emitNode = false, isRoot = false;
        uint32_t parent = kInvalidNode, gNode = 0, gLo = 0, gHi = 0, dAfter = kDeltaInf;
        const uint32_t dNext = s_st[0][t + 1];  // d[b0 + t] (level 1), kept for deltasOut
        if (isElem) {
            const uint32_t dL = s_st[0][t], dR = dNext;
            const int goLeft = dL < dR ? 1 : 0;  // parent: split s - 1 (slot t - 1) or split s (slot t)
            const bool pLocal = goLeft ? (t > 0 && s_loc[t - 1] != 0) : (isSplit && s_loc[t] != 0);
            if (pLocal)
                s_child[goLeft ? t - 1 : t][goLeft] = elemNode;
            else
                emitElem = true;
        }
        if (local) {
            // the split's node: leaf position s (level 1) or the last leaf of element s (level 2)
            gNode = kMode == kSuperLeaves ? elemHi : s;
            gLo = kMode == kSuperLeaves ? s_aux[jL] : b0 + (uint32_t)jL;
            gHi = kMode == kSuperLeaves ? s_aux[jR] - 1u : b0 - 1u + (uint32_t)jR;
            isRoot = gLo == 0 && gHi == nInternals;
            if (isRoot) {
                *root = gNode;
            } else {
                dAfter = s_st[0][jR];
                const int goLeft = s_st[0][jL] < dAfter ? 1 : 0;
                parent = goLeft ? gLo - 1 : gHi;  // the last leaf left of the range, or the range's last leaf
                const uint32_t slot = goLeft ? (uint32_t)jL - 1u : (uint32_t)jR - 1u;  // wraps for jL == 0
                const bool pLocal = slot < m - 1u && s_loc[slot] != 0;
                if (pLocal)
                    s_child[slot][goLeft] = gNode;
                else
                    emitNode = true;
            }
        }
        if (emitElem) s_mark[t] = 1;
        if (emitNode) s_mark[jL] = 1;
        lds_barrier();  // publishes the child slots and the marks; the sparse table is dead from here

        // Frontier entries in element order: the entry starting at element j
        // is stored at j, then ranked by the marks before it.
        if (emitElem) s_ent[t] = make_uint2(elemNode, elemHi);
        if (emitNode) s_ent[jL] = make_uint2(gNode, gHi);
        const bool marked = isElem && s_mark[t] != 0;
        const uint64_t markMask = __ballot(marked);
        if (lane == 0) s_wcount[wave] = (uint32_t)__popcll(markMask);
        lds_barrier();
        uint32_t before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < kWaves; w++) {
            const uint32_t c = s_wcount[w];
            before += w < wave ? c : 0u;
            total += c;
        }
        const uint32_t rank = before + (uint32_t)__popcll(markMask & ((1ull << lane) - 1ull));
        if (!kLevel1) {  // the count goes out now, for the successors' look-back
            if (t == 0)
                __hip_atomic_store(status2 + blk, lb_word(tag, blk == 0, total), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            if (marked) s_rank[t] = rank;
        }
        if (marked) {
            const uint32_t k = b0 + rank;
            const uint2 e = s_ent[t];
            frontier[3 * (size_t)k] = e.x;
            frontier[3 * (size_t)k + 1] = elemLo;  // the entry's first leaf is element t's
            frontier[3 * (size_t)k + 2] = e.y;
        }
        if (t == 0) {
            if (total < m) frontier[3 * (size_t)(b0 + total)] = kFrontierEnd;
            blockCounts[blk] = total;
        }

        // The block-local node, whole: child boxes over elements [jL, t] and [t + 1, jR - 1].
        Box entBox;  // level 2: the subtree box of this thread's node entry
        if (local) {
            auto range_box = [&](uint32_t a, uint32_t b) -> Box {  // local elements a .. b
                KeyBox acc = {{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, {0u, 0u, 0u}};
                uint32_t l = a + kLeaves, r = b + kLeaves + 1;
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
            const Box l = range_box((uint32_t)jL, (uint32_t)t), r = range_box((uint32_t)t + 1, (uint32_t)jR - 1u);
            entBox.lo = umin3(l.lo, r.lo);
            entBox.hi = umax3(l.hi, r.hi);
            const uint32_t context = gHi < nInternals ? gHi : 0xFFFFFFFFu;
            uint4* dst = reinterpret_cast<uint4*>(nodes + gNode);
            dst[0] = make_uint4(s_child[t][0], s_child[t][1], __float_as_uint(l.lo.x), __float_as_uint(l.lo.y));
            dst[1] = make_uint4(__float_as_uint(l.lo.z), __float_as_uint(l.hi.x), __float_as_uint(l.hi.y),
                                __float_as_uint(l.hi.z));
            dst[2] = make_uint4(__float_as_uint(r.lo.x), __float_as_uint(r.lo.y), __float_as_uint(r.lo.z),
                                __float_as_uint(r.hi.x));
            dst[3] = make_uint4(__float_as_uint(r.hi.y), __float_as_uint(r.hi.z), parent, context);
        }
        // d[b0 + t], stored last: nothing in the block waits for a global store
        if (kMode == kLeavesFusedDeltas && isElem && s < nInternals) deltasOut[s] = (uint8_t)dNext;
        if (kLevel1) break;  // one block per workgroup: no loop code
        // Level 2: this block's entries go to the top list, after those of
        // the blocks before it (found by the look-back); the last block
        // writes the list's length, and whether it overflows the top tree.
        const uint32_t excl = lb_exclusive_block<kWaves>(status2, blk, tag, s_lbPub, s_lbPre, s_lbSum);
        if (t == 0) {
            __hip_atomic_store(status2 + blk, lb_word(tag, true, excl + total), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            if (blk == nBlocks - 1) {
                ctrl[kCtrlTopCount] = excl + total;
                ctrl[kCtrlTopFallback] = excl + total > kTopMax ? 1u : 0u;
            }
        }
        auto put_record = [&](uint32_t k, uint32_t node, uint32_t hi, uint32_t d, const Box& b) {
            if (k >= kTopMax) return;
            uint4* dst = reinterpret_cast<uint4*>(topList + (size_t)k * kTopWords);
            dst[0] = make_uint4(node, hi, d, 0u);
            dst[1] = make_uint4(__float_as_uint(b.lo.x), __float_as_uint(b.lo.y), __float_as_uint(b.lo.z),
                                __float_as_uint(b.hi.x));
            dst[2] = make_uint4(__float_as_uint(b.hi.y), __float_as_uint(b.hi.z), 0u, 0u);
        };
        if (emitElem) {
            Box eb;
            eb.lo = mk(ord2f(kb.lo[0]), ord2f(kb.lo[1]), ord2f(kb.lo[2]));
            eb.hi = mk(ord2f(kb.hi[0]), ord2f(kb.hi[1]), ord2f(kb.hi[2]));
            put_record(excl + rank, elemNode, elemHi, dNext, eb);
        }
        if (emitNode) put_record(excl + s_rank[jL], gNode, gHi, dAfter, entBox);
        lds_barrier();  // the next block of the loop reuses the LDS
        if (t == 0) s_blk = atomicAdd(&ctrl[kCtrlL2Ticket], 1u);
        lds_barrier();
        blk = s_blk;
    }
}

// The level-1 frontier stretches (entries of block b at b * leavesPerBlock,
// counts[b] of them, each stretch in leaf order) laid out in leaf order as
// the super-leaf list, and its length in ctrl[kCtrlSuperCount].  A
// workgroup takes kCompactBlocks level-1 blocks (by ticket, so it waits only
// on workgroups that are running): it scans their counts, publishes its total
// at once, finds its offset by summing its predecessors' published totals
// (or a published prefix) 64 at a time, and copies its entries with all its
// threads, a few per thread (lb_exclusive).
constexpr int kCompactBlocks = 256;    // level-1 blocks per workgroup
constexpr int kCompactThreads = 1024;  // about 9 entries per level-1 block: about 2 per thread
__global__ void __launch_bounds__(kCompactThreads)
    k_frontier_compact(const uint32_t* __restrict__ counts, uint32_t nBlocks, const uint32_t* __restrict__ frontier,
                       uint32_t* __restrict__ superList, uint32_t leavesPerBlock, uint64_t* __restrict__ status,
                       uint32_t* __restrict__ ctrl)
{
    constexpr int kScanWaves = kCompactBlocks / 64;
    __shared__ uint32_t s_off[kCompactBlocks];
    __shared__ uint32_t s_w[kScanWaves];
    __shared__ uint32_t s_tile, s_excl;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t == 0) s_tile = atomicAdd(&ctrl[kCtrlCompactTicket], 1u);
    const uint32_t tag = ctrl[kCtrlCompactTag] + 1u;
    __syncthreads();
    const uint32_t tile = s_tile;
    if (t < kCompactBlocks) {  // scan of the tile's counts, waves 0 .. kScanWaves - 1
        const uint32_t b = tile * kCompactBlocks + (uint32_t)t;
        const uint32_t c = b < nBlocks ? counts[b] : 0u;
        uint32_t x = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) s_w[wave] = x;
        s_off[t] = x - c;  // within the wave, for now
    }
    __syncthreads();
    uint32_t total = 0;
#pragma unroll
    for (int w = 0; w < kScanWaves; w++) total += s_w[w];
    if (t < kCompactBlocks) {
        uint32_t before = 0;
#pragma unroll
        for (int w = 0; w < kScanWaves; w++) before += w < wave ? s_w[w] : 0u;
        s_off[t] += before;
    }
    if (t == 0)
        __hip_atomic_store(status + tile, lb_word(tag, tile == 0, total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (wave == 0) {
        const uint32_t excl = lb_exclusive(status, tile, tag, lane);
        if (lane == 0) {
            s_excl = excl;
            __hip_atomic_store(status + tile, lb_word(tag, true, excl + total), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            if ((size_t)(tile + 1) * kCompactBlocks >= nBlocks) ctrl[kCtrlSuperCount] = excl + total;
        }
    }
    __syncthreads();
    // Entry e of this tile: the last block k with offset <= e holds it.  Each
    // round of loads is issued for all of the thread's entries before use.
    const uint32_t excl = s_excl;
    constexpr int kUnroll = 4;
    for (uint32_t e0 = 0; e0 < total; e0 += kUnroll * kCompactThreads) {
        size_t src[kUnroll];
        uint32_t v[kUnroll][3];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const uint32_t e = e0 + (uint32_t)(u * kCompactThreads + t);
            src[u] = 0;
            if (e < total) {
                uint32_t lo = 0, hi = kCompactBlocks - 1;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) / 2;
                    if (s_off[mid] <= e) lo = mid;
                    else hi = mid - 1;
                }
                src[u] = 3 * ((size_t)(tile * kCompactBlocks + lo) * leavesPerBlock + (e - s_off[lo]));
            }
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const uint32_t e = e0 + (uint32_t)(u * kCompactThreads + t);
            if (e < total) {
                v[u][0] = frontier[src[u]];
                v[u][1] = frontier[src[u] + 1];
                v[u][2] = frontier[src[u] + 2];
            }
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const uint32_t e = e0 + (uint32_t)(u * kCompactThreads + t);
            if (e < total) {
                const size_t d = 3 * ((size_t)excl + e);
                superList[d] = v[u][0];
                superList[d + 1] = v[u][1];
                superList[d + 2] = v[u][2];
            }
        }
    }
}

// The device-scope climb of level 2's frontier entries, for top lists longer
// than kTopMax: wave w of nWaves takes level-2 blocks w, w + nWaves, ..., and
// each of its lanes climbs one entry of the block's stretch (up to its end
// mark) through the device-scope arrival words (climb_from).
template <class TS>
__device__ inline void climb_upper_wave(uint32_t w, uint32_t nWaves, TS tris, uint32_t n,
                                        const uint8_t* __restrict__ deltas, Node* nodes, uint32_t* flags,
                                        const uint32_t* __restrict__ frontier, uint32_t count, uint32_t* root)
{
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t blk = w; (size_t)blk * kLeaves2 < count; blk += nWaves) {
        const uint32_t b0 = blk * kLeaves2;
        const uint32_t m = count - b0 < (uint32_t)kLeaves2 ? count - b0 : (uint32_t)kLeaves2;
        for (uint32_t base = 0; base < m; base += 64) {
            const uint32_t k = base + lane;
            const size_t e = 3 * (size_t)(b0 + k);
            const uint32_t node = k < m ? frontier[e] : kFrontierEnd;
            const uint64_t endMask = __ballot(node == kFrontierEnd);
            const uint32_t valid = endMask ? (uint32_t)__ffsll((unsigned long long)endMask) - 1u : 64u;
            if (lane < valid) {
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
}

// The nodes above level 2, in one workgroup.  Level 2's frontier entries, in
// leaf order in the top list (node, last leaf, delta of the split after it,
// subtree box), are the elements of a last sequence: entry j covers the
// leaves [lo_j, hi_j], and the split between entries j and j + 1 is internal
// node hi_j with delta d[hi_j].  The climb runs over entry indices with its
// hand-offs in LDS.  A node arriving at split p leaves its subtree box and
// node index in slot p if it is p's left child (its range ends at entry p) or
// in slot p + 1 if it is the right child (its range starts there), then adds
// one arrival and its far end to p's word in a single atomic: the first
// arriver stops, the second finds its sibling's far end in the value the
// atomic returns and its box and index in the sibling's slot, and goes on
// with p.  Slot q holds entry q to start with and is written later by at most
// one ancestor of entry q, which climbs only after entry q's parent is
// complete, so no deposit is overwritten before it is read, and no arriver
// waits.  A completed node is written whole (64 bytes) with the next step,
// which finds its parent and its context; an entry only gets its parent.
// With more than kTopMax entries level 2 has set ctrl[kCtrlTopFallback]
// instead, and the launch's other blocks climb them through device memory
// (climb_upper_wave); they return at once otherwise.  (One launch for both: a
// launch costs about 4.5 us even when it returns at once.)
//
// The loop below is bound by instruction issue: every wave with an active
// lane runs the whole step, so kTopClimbers (fewer than the block's threads,
// which all load the list) trades a few more steps per lane for fewer waves
// per SIMD.
constexpr int kTopThreads = 1024;
#ifndef MLBVH_TOP_CLIMBERS
#define MLBVH_TOP_CLIMBERS 512
#endif
constexpr int kTopClimbers = MLBVH_TOP_CLIMBERS;
static_assert(kTopClimbers % 64 == 0 && kTopClimbers <= kTopThreads, "whole waves of the block climb");
constexpr uint32_t kTopNoSplit = 0x100u;  // a missing split's delta: above every delta
constexpr int kTopArrive = 41;            // split words: last leaf | delta << 32 | (arrivals + 2 far ends) << 41
template <class TS>
__global__ void __launch_bounds__(kTopThreads)
    k_top_tree(uint32_t n, Node* __restrict__ nodes, const uint32_t* __restrict__ topList, uint32_t* ctrl,
               uint32_t* root, TS tris, const uint8_t* __restrict__ deltas, uint32_t* flags,
               const uint32_t* __restrict__ frontier)
{
    if (blockIdx.x > 0) {  // the device-scope climb, only for a long top list (flag set by level 2)
        const uint32_t count = ctrl[kCtrlSuperCount];
        if (count == 0 || ctrl[kCtrlTopFallback] == 0) return;
        const uint32_t wavesPer = kTopThreads / 64;
        climb_upper_wave(((blockIdx.x - 1) * wavesPer) + (threadIdx.x >> 6), (gridDim.x - 1) * wavesPer, tris, n,
                         deltas, nodes, flags, frontier, count, root);
        return;
    }
    __shared__ Box s_box[kTopMax];       // deposit slots: subtree box
    __shared__ uint32_t s_nd[kTopMax];   // deposit slots: node index
    // split j: the last leaf of entry j (its node index) | the delta of split j << 32 (kTopNoSplit
    // for the last entry, which has no split after it) | its arrival count and far ends << kTopArrive
    __shared__ uint64_t s_pk[kTopMax];

    const int t = threadIdx.x;
    // no super-leaves: level 1 built the root, level 2 ran no block and wrote no count
    const bool none = ctrl[kCtrlSuperCount] == 0;
    const uint32_t F = none ? 0u : ctrl[kCtrlTopCount];
    if (t == 0 && none) ctrl[kCtrlTopFallback] = 0u;
    if (F == 0 || F > kTopMax) return;  // the root was built below, or the device-scope climb runs
    for (uint32_t i = (uint32_t)t; i < F; i += kTopThreads) {
        const uint4* src = reinterpret_cast<const uint4*>(topList + (size_t)i * kTopWords);
        const uint4 r0 = src[0], r1 = src[1], r2 = src[2];
        const uint32_t d = r0.z == kDeltaInf ? kTopNoSplit : r0.z;
        s_pk[i] = (uint64_t)r0.y | ((uint64_t)d << 32);
        s_nd[i] = r0.x;
        Box bx;
        bx.lo = mk(__uint_as_float(r1.x), __uint_as_float(r1.y), __uint_as_float(r1.z));
        bx.hi = mk(__uint_as_float(r1.w), __uint_as_float(r2.x), __uint_as_float(r2.y));
        s_box[i] = bx;
    }
    lds_barrier();
    if (t >= kTopClimbers) return;

    // Entries i0 .. i0 + per - 1 of this lane, per = ceil(F / kTopClimbers):
    // consecutive entries are often siblings, so a lane's climbs tend to meet.
    // The loop's condition is a ballot, so the wave runs it together; no step
    // waits on another lane.  R and L are the words of the splits at the
    // range's two ends (b, and a - 1), read ahead.
    constexpr uint64_t kNoSplitWord = (uint64_t)kTopNoSplit << 32;
    auto split_word = [&](uint32_t j) {
        return __hip_atomic_load(&s_pk[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    auto delta_of = [](uint64_t w) { return (uint32_t)(w >> 32) & 0x1FFu; };
    const uint32_t nInternals = n - 1;
    const uint32_t per = (F + kTopClimbers - 1) / kTopClimbers;
    const uint32_t i0 = (uint32_t)t * per;
    uint32_t q = 0;           // the entry being climbed: i0 + q
    bool active = i0 < F;     // an entry of this lane is climbing
    bool pending = false;     // `node` was just completed: it is written with the next step
    uint32_t a = i0, b = i0, node = 0, c0 = 0, c1 = 0;
    Box box = {}, box0 = {}, box1 = {};  // the subtree box; a completed node's child boxes
    uint64_t R = 0, L = kNoSplitWord;
    if (active) {
        node = s_nd[i0];
        box = s_box[i0];
        R = split_word(i0);
        if (i0 > 0) L = split_word(i0 - 1);
    }
    while (__ballot(active) != 0) {
        if (!active) continue;
        const uint32_t myHi = (uint32_t)R;  // the last leaf of [a, b]
        const bool isRoot = a == 0 && b == F - 1;
        const uint32_t goLeft = delta_of(L) < delta_of(R) ? 1u : 0u;  // parent: split a - 1 or b
        const uint32_t pNode = isRoot ? kInvalidNode : goLeft ? (uint32_t)L : myHi;
        if (pending) {
            const uint32_t context = myHi < nInternals ? myHi : 0xFFFFFFFFu;
            uint4* dst = reinterpret_cast<uint4*>(nodes + node);
            dst[0] = make_uint4(c0, c1, __float_as_uint(box0.lo.x), __float_as_uint(box0.lo.y));
            dst[1] = make_uint4(__float_as_uint(box0.lo.z), __float_as_uint(box0.hi.x), __float_as_uint(box0.hi.y),
                                __float_as_uint(box0.hi.z));
            dst[2] = make_uint4(__float_as_uint(box1.lo.x), __float_as_uint(box1.lo.y), __float_as_uint(box1.lo.z),
                                __float_as_uint(box1.hi.x));
            dst[3] = make_uint4(__float_as_uint(box1.hi.y), __float_as_uint(box1.hi.z), pNode, context);
        } else if (!(node & kLeafBit)) {
            nodes[node].parent = pNode;
        }
        pending = false;
        bool done = true;
        if (isRoot) {
            *root = node;
        } else {
            const uint32_t p = goLeft ? a - 1 : b;
            if (a != b) {  // an entry is in its slot already
                const uint32_t slot = goLeft ? a : b;
                s_box[slot] = box;
                s_nd[slot] = node;
            }
            const uint32_t far = goLeft ? b : a;
            const uint64_t prior = __hip_atomic_fetch_add(&s_pk[p], (uint64_t)(1u | (far << 1)) << kTopArrive,
                                                          __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            const uint32_t arr = (uint32_t)(prior >> kTopArrive);
            if (arr & 1u) {  // second: the sibling's far end, box and index are in place
                const uint32_t other = arr >> 1;
                const uint32_t sib = goLeft ? a - 1 : b + 1;
                const Box ob = s_box[sib];
                const uint32_t on = s_nd[sib];
                c0 = goLeft ? on : node;
                c1 = goLeft ? node : on;
                box0 = goLeft ? ob : box;
                box1 = goLeft ? box : ob;
                a = a < other ? a : other;
                b = b > other ? b : other;
                R = split_word(b);
                L = a > 0 ? split_word(a - 1) : kNoSplitWord;
                box.lo = umin3(box.lo, ob.lo);
                box.hi = umax3(box.hi, ob.hi);
                node = pNode;
                pending = true;
                done = false;
            }
        }
        if (done) {  // the next entry of this lane, if any
            q++;
            const uint32_t i = i0 + q;
            active = q < per && i < F;
            if (active) {
                a = b = i;
                node = s_nd[i];
                box = s_box[i];
                R = split_word(i);
                L = split_word(i - 1);  // i > i0 >= 0
            }
        }
    }
}


// fuseDeltas: compute the deltas from `rec` in level 1 and write them to
// `deltas` (the 32-bit build); else read them from `deltas`.
template <class TS>
static int launch_climb(mlbvh_ctx* ctx, TS tris, int n, const uint64_t* rec, uint8_t* deltas, bool fuseDeltas,
                        Node* nodes, uint32_t* root, hipStream_t stream)
{
    const unsigned blocks1 = (unsigned)(((size_t)n + kLeaves1 - 1) / kLeaves1);
    int rc = ensure(ctx, &ctx->frontier, &ctx->frontierCap, (size_t)n * 3, "hipMalloc(climb frontier)");
    if (!rc) rc = ensure(ctx, &ctx->superList, &ctx->superCap, (size_t)n * 3, "hipMalloc(super-leaf list)");
    if (!rc) rc = ensure(ctx, &ctx->blockCounts, &ctx->blockCountsCap, 2 * (size_t)blocks1, "hipMalloc(block counts)");
    const unsigned compactTiles = (blocks1 + kCompactBlocks - 1) / kCompactBlocks;
    const size_t blocks2Max = ((size_t)n + kLeaves2 - 1) / kLeaves2;
    if (!rc && (ctx->l1StatusCap < compactTiles + blocks2Max || !ctx->l1Status)) {
        // a new table starts zeroed: tag 0 is never current
        rc = ensure(ctx, &ctx->l1Status, &ctx->l1StatusCap, compactTiles + blocks2Max, "hipMalloc(build look-back)");
        if (!rc) {
            hipError_t e = hipMemset(ctx->l1Status, 0, ctx->l1StatusCap * sizeof(uint64_t));
            if (e != hipSuccess) rc = hip_status(e, "hipMemset(build look-back)");
        }
    }
    if (!rc) rc = ensure(ctx, &ctx->topList, &ctx->topListCap, (size_t)kTopMax * kTopWords, "hipMalloc(top list)");
    if (rc) return rc;
    // Level 1's block 0 resets the tickets and advances the look-back tag of
    // the kernels after it, so each call starts clean.
    uint32_t* counts1 = ctx->blockCounts;            // level-1 frontier counts
    uint32_t* counts2 = ctx->blockCounts + blocks1;  // level-2 frontier counts
    if (fuseDeltas)
        k_local_nodes<TS, kLeaves1, kLeavesFusedDeltas><<<blocks1, kLeaves1, 0, stream>>>(
            tris, (uint32_t)n, rec, nullptr, deltas, nullptr, ctx->ctrl, nodes, ctx->frontier, counts1, nullptr,
            root, nullptr, nullptr, ctx->sceneBox);
    else
        k_local_nodes<TS, kLeaves1, kLeavesGivenDeltas><<<blocks1, kLeaves1, 0, stream>>>(
            tris, (uint32_t)n, rec, deltas, nullptr, nullptr, ctx->ctrl, nodes, ctx->frontier, counts1, nullptr,
            root, nullptr, nullptr, ctx->sceneBox);
    k_frontier_compact<<<compactTiles, kCompactThreads, 0, stream>>>(counts1, blocks1, ctx->frontier,
                                                                    ctx->superList, kLeaves1, ctx->l1Status, ctx->ctrl);
    // level 2 and the upper climb loop over however many level-2 blocks the
    // super-leaves fill (there are at most n super-leaves)
    const unsigned grid2 = (unsigned)std::min<size_t>(blocks2Max, (size_t)ctx->cuCount * 2);
    k_local_nodes<TS, kLeaves2, kSuperLeaves><<<grid2, kLeaves2, 0, stream>>>(
        tris, (uint32_t)n, nullptr, deltas, nullptr, ctx->superList, ctx->ctrl, nodes, ctx->frontier, counts2,
        ctx->flags, root, ctx->l1Status + compactTiles, ctx->topList, nullptr);
    // block 0: the top tree; the others: the device-scope climb when the top list overflowed
    const unsigned climbBlocks = (unsigned)std::min<size_t>((blocks2Max + 15) / 16, (size_t)ctx->cuCount);
    k_top_tree<TS><<<1 + climbBlocks, kTopThreads, 0, stream>>>((uint32_t)n, nodes, ctx->topList, ctx->ctrl, root,
                                                                 tris, deltas, ctx->flags, ctx->frontier);
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
    unsigned g = grid1(nn);
    unsigned gAabb = g < (unsigned)ctx->cuCount * 8 ? g : (unsigned)ctx->cuCount * 8;
    k_scene_aabb<<<gAabb, kBuildBlock, 0, stream>>>(tris, n, (unsigned long long*)ctx->sceneBox, ctx->ctrl, ctx->hist,
                                                    ctx->status, ctx->statusWords);
    unsigned gMorton = g < (unsigned)ctx->cuCount * 4 ? g : (unsigned)ctx->cuCount * 4;
    if (wide) {
        if ((rc = ensure(ctx, &ctx->idxA, &ctx->idxCap, nn, "hipMalloc(indices)"))) return rc;
        if ((rc = ensure(ctx, &ctx->idxB, &ctx->idxBCap, nn, "hipMalloc(indices tmp)"))) return rc;
        k_morton64<<<gMorton, kBuildBlock, 0, stream>>>(tris, n, ctx->sceneBox, ctx->recB, ctx->idxA);
        // keys in recB (tmp: recA), records written to recA afterwards.  Large
        // sorts take the route of four top-digit passes and the segment
        // finish: on a surface mesh the 63-bit codes' runs of equal 24-bit
        // prefix are often longer than the finish handles (then the three-pass
        // route re-ran every pass), their 32-bit prefixes far less often.
        rc = sort_pairs_u64(ctx, ctx->recB, ctx->idxA, ctx->recA, ctx->idxB, (uint32_t)n, 0, 63, stream, 4, true);
        if (rc) return rc;
        k_deltas64<<<g, kBuildBlock, 0, stream>>>(ctx->recB, ctx->idxA, n, ctx->deltas, ctx->recA);
    } else {
        const uint32_t tiles = sort_tiles(ctx, (uint32_t)n);
        const unsigned gm = gMorton < tiles ? gMorton : tiles;
        // partials: gm blocks x 4 histograms, then the per-tile counts
        if ((rc = ensure_partials(ctx, (int)(gm * 4 + tiles), 1))) return rc;
        uint32_t* tileCounts = ctx->partials + (size_t)gm * 4 * 256;
        k_morton<<<gm, kMortonBlock, 0, stream>>>(tris, n, ctx->sceneBox, opt, ctx->recA, ctx->partials,
                                                 (uint32_t)sort_tile_keys(ctx), tiles, tileCounts);
        rc = sort_records_u64(ctx, ctx->recA, ctx->recB, (uint32_t)n, 0, 32, true, stream, tileCounts, (int)gm);
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
