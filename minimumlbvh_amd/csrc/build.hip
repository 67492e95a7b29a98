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
//    the nodes of a surface mesh) climbs through device-scope arrival words,
//    one u32 per internal node; `context` is written once, with the
//    reference's final value;
//  * one thread per leaf for all n leaves (the reference launches
//    ceil((n-1)/256) blocks and drops the last leaf when (n-1) % 256 == 0).
#include <algorithm>

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
    k_morton(TS tris, int n, const uint32_t* __restrict__ sceneBox, uint32_t opt,
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
    k_morton64(TS tris, int n, const uint32_t* __restrict__ sceneBox, uint64_t* __restrict__ keys,
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
// What crosses a level-2 block (a few thousand nodes for surface meshes)
// climbs on through the device-scope arrival words (k_climb_upper,
// climb_from).  The result is the tree of one thread per leaf climbing
// through device memory all the way.  A node whose parent is built by a
// later level gets its parent word from its own level (the rule above gives
// it) and again, identically, from climb_from.  The blocks' barriers order
// LDS only, so their global stores are never waited for.
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

template <class TS, int kLeaves, int kMode>
__global__ void __launch_bounds__(kLeaves) __attribute__((amdgpu_waves_per_eu(kLeaves <= 256 ? 8 : 1)))
    k_local_nodes(TS tris, uint32_t n, const uint64_t* __restrict__ rec, const uint8_t* __restrict__ deltas,
                  uint8_t* __restrict__ deltasOut, const uint32_t* __restrict__ elems,
                  const uint32_t* __restrict__ ctrl, Node* __restrict__ nodes, uint32_t* __restrict__ frontier,
                  uint32_t* __restrict__ blockCounts, uint32_t* __restrict__ flags, uint32_t* __restrict__ root)
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
    // The frontier entry starting at element j: its node and last leaf (its
    // first leaf is element j's), in the sparse table's LDS once that is dead.
    uint2* s_ent = reinterpret_cast<uint2*>(&s_st[0][0]);
    static_assert(sizeof(uint2) * kLeaves <= sizeof(s_st), "the entries fit the sparse table's LDS");

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t count = kMode == kSuperLeaves ? ctrl[kCtrlSuperCount] : n;
    const uint32_t nInternals = n - 1;
    // Level 1 has one block per kLeaves leaves; level 2 loops over its blocks
    // (their number is known on the device only).
    for (uint32_t blk = blockIdx.x; (size_t)blk * kLeaves < count; blk += gridDim.x) {
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
        uint32_t parent = kInvalidNode, gNode = 0, gLo = 0, gHi = 0;
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
                const int goLeft = s_st[0][jL] < s_st[0][jR] ? 1 : 0;
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
        if (marked) {
            const uint32_t k = b0 + before + (uint32_t)__popcll(markMask & ((1ull << lane) - 1ull));
            const uint2 e = s_ent[t];
            frontier[3 * (size_t)k] = e.x;
            frontier[3 * (size_t)k + 1] = elemLo;  // the entry's first leaf is element t's
            frontier[3 * (size_t)k + 2] = e.y;
        }
        if (t == 0) {
            if (total < m) frontier[3 * (size_t)(b0 + total)] = kFrontierEnd;
            if (blockCounts) blockCounts[blk] = total;
        }

        // The block-local node, whole: child boxes over elements [jL, t] and [t + 1, jR - 1].
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
        if (kMode != kSuperLeaves) break;  // one block per workgroup: no loop code
        lds_barrier();                     // the next block of the loop reuses the LDS
    }
}

// Exclusive scan of the level-1 blocks' frontier counts into offsets, and
// their total (the number of super-leaves) into ctrl[kCtrlSuperCount].  One
// block, in tiles of kScanTile counts: coalesced loads into LDS, each thread
// scans a contiguous run of kScanTile / 1024 there, coalesced stores.
constexpr uint32_t kScanTile = 16384;
__global__ void __launch_bounds__(1024)
    k_frontier_scan(const uint32_t* __restrict__ counts, uint32_t nBlocks, uint32_t* __restrict__ offsets,
                    uint32_t* __restrict__ ctrl)
{
    constexpr int kWaves = 16;
    constexpr uint32_t kPer = kScanTile / 1024;
    __shared__ uint32_t s_c[kScanTile];
    __shared__ uint32_t s_w[kWaves];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nBlocks; base += kScanTile) {
        const uint32_t m = nBlocks - base < kScanTile ? nBlocks - base : kScanTile;
#pragma unroll
        for (uint32_t q = 0; q < kPer; q++) {
            const uint32_t i = q * 1024 + (uint32_t)t;
            s_c[i] = i < m ? counts[base + i] : 0u;
        }
        __syncthreads();
        uint32_t v[kPer], sum = 0;
#pragma unroll
        for (uint32_t q = 0; q < kPer; q++) {
            v[q] = s_c[(uint32_t)t * kPer + q];
            sum += v[q];
        }
        uint32_t x = sum;  // inclusive scan over the wave
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) s_w[wave] = x;
        __syncthreads();
        uint32_t run = carry + x - sum, total = 0;
#pragma unroll
        for (int w = 0; w < kWaves; w++) {
            const uint32_t c = s_w[w];
            run += w < wave ? c : 0u;
            total += c;
        }
#pragma unroll
        for (uint32_t q = 0; q < kPer; q++) {
            s_c[(uint32_t)t * kPer + q] = run;
            run += v[q];
        }
        __syncthreads();
#pragma unroll
        for (uint32_t q = 0; q < kPer; q++) {
            const uint32_t i = q * 1024 + (uint32_t)t;
            if (i < m) offsets[base + i] = s_c[i];
        }
        carry += total;
        __syncthreads();
    }
    if (t == 0) ctrl[kCtrlSuperCount] = carry;
}

// One wave per level-1 block: its frontier stretch, already in leaf order, to
// its place in the super-leaf list.
__global__ void __launch_bounds__(64)
    k_frontier_compact(const uint32_t* __restrict__ frontier, const uint32_t* __restrict__ counts,
                       const uint32_t* __restrict__ offsets, uint32_t* __restrict__ superList, uint32_t leavesPerBlock)
{
    const uint32_t b = blockIdx.x;
    const size_t words = 3 * (size_t)counts[b];
    const size_t src = 3 * (size_t)b * leavesPerBlock, dst = 3 * (size_t)offsets[b];
    for (size_t k = threadIdx.x; k < words; k += 64) superList[dst + k] = frontier[src + k];
}

// One wave per level-2 block (looping over them), over that block's stretch
// of the frontier list up to its end mark: each entry climbs through the
// device-scope arrival words.
template <class TS>
__global__ void __launch_bounds__(64)
    k_climb_upper(TS tris, uint32_t n, const uint8_t* __restrict__ deltas, Node* nodes, uint32_t* flags,
                  const uint32_t* __restrict__ frontier, const uint32_t* __restrict__ ctrl, uint32_t* root)
{
    const uint32_t count = ctrl[kCtrlSuperCount];
    for (uint32_t blk = blockIdx.x; (size_t)blk * kLeaves2 < count; blk += gridDim.x) {
        const uint32_t b0 = blk * kLeaves2;
        const uint32_t m = count - b0 < (uint32_t)kLeaves2 ? count - b0 : (uint32_t)kLeaves2;
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
    if (rc) return rc;
    uint32_t* counts = ctx->blockCounts;
    uint32_t* offsets = ctx->blockCounts + blocks1;
    if (fuseDeltas)
        k_local_nodes<TS, kLeaves1, kLeavesFusedDeltas><<<blocks1, kLeaves1, 0, stream>>>(
            tris, (uint32_t)n, rec, nullptr, deltas, nullptr, nullptr, nodes, ctx->frontier, counts, nullptr, root);
    else
        k_local_nodes<TS, kLeaves1, kLeavesGivenDeltas><<<blocks1, kLeaves1, 0, stream>>>(
            tris, (uint32_t)n, rec, deltas, nullptr, nullptr, nullptr, nodes, ctx->frontier, counts, nullptr, root);
    k_frontier_scan<<<1, 1024, 0, stream>>>(counts, blocks1, offsets, ctx->ctrl);
    k_frontier_compact<<<blocks1, 64, 0, stream>>>(ctx->frontier, counts, offsets, ctx->superList, kLeaves1);
    // level 2 and the upper climb loop over however many level-2 blocks the
    // super-leaves fill (there are at most n super-leaves)
    const size_t blocks2Max = ((size_t)n + kLeaves2 - 1) / kLeaves2;
    const unsigned grid2 = (unsigned)std::min<size_t>(blocks2Max, (size_t)ctx->cuCount * 2);
    k_local_nodes<TS, kLeaves2, kSuperLeaves><<<grid2, kLeaves2, 0, stream>>>(
        tris, (uint32_t)n, nullptr, deltas, nullptr, ctx->superList, ctx->ctrl, nodes, ctx->frontier, nullptr,
        ctx->flags, root);
    const unsigned gridUp = (unsigned)std::min<size_t>(blocks2Max, (size_t)ctx->cuCount * 8);
    k_climb_upper<<<gridUp, 64, 0, stream>>>(tris, (uint32_t)n, deltas, nodes, ctx->flags, ctx->frontier, ctx->ctrl,
                                             root);
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
        k_morton64<<<gMorton, kBuildBlock, 0, stream>>>(tris, n, ctx->sceneBox, ctx->recB, ctx->idxA);
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
        k_morton<<<gm, kMortonBlock, 0, stream>>>(tris, n, ctx->sceneBox, opt, ctx->recA, ctx->partials,
                                                 (uint32_t)sort_tile_keys(ctx), tiles, tileCounts);
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
