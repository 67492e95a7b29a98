// Batched LBVH build and trace over many meshes in one pipeline (no reference
// counterpart; the reference builds one mesh per BVHGPUBuilder::build call,
// minimum_lbvh.h:783-871).
//
// Meshes are concatenated: mesh m owns triangles [off[m], off[m+1]).  Each
// mesh gets exactly the BVH a separate mlbvh_build would give it -- its own
// scene AABB and Morton keys, ties by its own triangle index, mesh-local node
// and triangle indices -- stored at nodes + (off[m] - m) with its root in
// roots[m].  One sort orders all records by the segmented key
//     key = meshId << (32 + L) | morton << L | localIndex      (L = index bits)
// over bits [L, L + 32 + meshBits), so every mesh's records land, sorted, in
// its own segment [off[m], off[m+1]).  The climb then runs in mesh-local
// coordinates.  Tracing uses a persistent kernel (k_trace_ww in trace.hip)
// whose waves pull chunks of rays from an atomic work counter and hand the
// next ray of the chunk to each lane whose walk has ended; each ray names its
// mesh in the Ray's pad word.
#include "mlbvh_ctx.h"

namespace mlbvh {

constexpr int kBatchBlock = 256;

// tileMesh[t] = the mesh of triangle t * kBatchBlock (the first of tile t):
// thread m fills the tiles whose first triangle lies in [off[m], off[m+1]).
__global__ void k_batch_tile_mesh(const uint32_t* __restrict__ off, int nMeshes, uint32_t total,
                                  uint32_t* __restrict__ tileMesh)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nMeshes) return;
    const uint32_t nTiles = (uint32_t)(((uint64_t)total + kBatchBlock - 1) / kBatchBlock);
    uint32_t t1 = (uint32_t)(((uint64_t)off[m + 1] + kBatchBlock - 1) / kBatchBlock);
    t1 = t1 < nTiles ? t1 : nTiles;  // offsets past `total` write nothing outside the table
    for (uint32_t t = (uint32_t)(((uint64_t)off[m] + kBatchBlock - 1) / kBatchBlock); t < t1; t++)
        tileMesh[t] = (uint32_t)m;
}

// The mesh of triangle i from the mesh m0 of an earlier triangle (the mesh
// of its tile's first triangle: the same mesh or one of the next few).
// Clamped to the meshes that exist, so offsets that break the documented
// layout give wrong boxes but no read past `off`.
__device__ inline uint32_t mesh_from(const uint32_t* __restrict__ off, int nMeshes, uint32_t m0, uint32_t i)
{
    const uint32_t last = (uint32_t)nMeshes - 1u;
    m0 = m0 < last ? m0 : last;
    while (m0 < last && off[m0 + 1] <= i) m0++;
    return m0;
}

__global__ void k_batch_init(uint32_t* meshBox, int nMeshes, uint32_t* ctrl, uint32_t* hist, int nPass,
                             uint64_t* status, size_t statusWords)
{
    for (int i = threadIdx.x; i < nMeshes * 6; i += blockDim.x) meshBox[i] = (i % 6) < 3 ? 0xFFFFFFFFu : 0u;
    if (threadIdx.x == 0) ctrl[kCtrlFrontier] = 0;
    sort_state_init(ctrl, nPass, hist, nPass, status, statusWords);
}

// Per-mesh scene AABB.  A block takes kAabbTiles consecutive tiles of 256
// triangles; a block inside one mesh (the usual case) keeps a running box per
// thread, reduces it in LDS and issues 6 atomics.  A block that straddles
// mesh boundaries reduces per wave when the wave's 64 triangles share a mesh,
// else issues per-lane atomics.
constexpr uint32_t kAabbTiles = 8;
__global__ void __launch_bounds__(kBatchBlock)
    k_batch_aabb(const Tri* __restrict__ tris, const uint32_t* __restrict__ off, int nMeshes,
                 const uint32_t* __restrict__ tileMesh, uint32_t total, uint32_t* meshBox)
{
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t tile0 = blockIdx.x * kAabbTiles;
    const uint32_t first = tile0 * kBatchBlock;  // < total
    const uint32_t end = (uint32_t)min((uint64_t)first + kAabbTiles * kBatchBlock, (uint64_t)total);
    const uint32_t mFirst = mesh_from(off, nMeshes, tileMesh[tile0], first);
    const bool blockUniform = off[mFirst + 1] >= end;
    auto reduce_wave = [&](float* v) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
            for (int c = 0; c < 3; c++) v[c] = fminf(v[c], __shfl_xor(v[c], o, 64));
#pragma unroll
            for (int c = 3; c < 6; c++) v[c] = fmaxf(v[c], __shfl_xor(v[c], o, 64));
        }
    };
    auto add_tri = [&](float* v, const Tri& tri) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            v[0] = fminf(v[0], tri.v[k].x);
            v[1] = fminf(v[1], tri.v[k].y);
            v[2] = fminf(v[2], tri.v[k].z);
            v[3] = fmaxf(v[3], tri.v[k].x);
            v[4] = fmaxf(v[4], tri.v[k].y);
            v[5] = fmaxf(v[5], tri.v[k].z);
        }
    };
    if (blockUniform) {
        float v[6] = {kFltMax, kFltMax, kFltMax, -kFltMax, -kFltMax, -kFltMax};
#pragma unroll
        for (uint32_t k = 0; k < kAabbTiles; k++) {
            const uint32_t i = first + k * kBatchBlock + (uint32_t)t;
            if (i < end) add_tri(v, tris[i]);
        }
        reduce_wave(v);
        __shared__ float s_v[kBatchBlock / 64][6];
        if (lane == 0)
            for (int c = 0; c < 6; c++) s_v[wave][c] = v[c];
        __syncthreads();
        if (t < 6) {
            const int c = t;
            float r = s_v[0][c];
            for (int w = 1; w < kBatchBlock / 64; w++) r = c < 3 ? fminf(r, s_v[w][c]) : fmaxf(r, s_v[w][c]);
            if (c < 3) atomicMin(&meshBox[mFirst * 6 + c], f2ord(r));
            else atomicMax(&meshBox[mFirst * 6 + c], f2ord(r));
        }
        return;
    }
    for (uint32_t k = 0; k < kAabbTiles; k++) {
        const uint32_t w0 = first + k * kBatchBlock + (uint32_t)(t - lane);  // the wave's first triangle
        if (w0 >= end) break;                                                // wave-uniform
        const uint32_t i = w0 + (uint32_t)lane;
        const bool valid = i < end;
        float v[6] = {kFltMax, kFltMax, kFltMax, -kFltMax, -kFltMax, -kFltMax};
        uint32_t m = 0xFFFFFFFFu;
        if (valid) {
            add_tri(v, tris[i]);
            m = mesh_from(off, nMeshes, tileMesh[i / kBatchBlock], i);
        }
        const uint32_t m0 = __shfl(m, 0, 64);
        if (__ballot(valid && m != m0) == 0) {
            reduce_wave(v);
            if (lane < 6) {
                const float r = lane == 0 ? v[0] : lane == 1 ? v[1] : lane == 2 ? v[2] : lane == 3 ? v[3]
                                : lane == 4 ? v[4] : v[5];
                if (lane < 3) atomicMin(&meshBox[m0 * 6 + lane], f2ord(r));
                else atomicMax(&meshBox[m0 * 6 + lane], f2ord(r));
            }
        } else if (valid) {
            for (int c = 0; c < 3; c++) atomicMin(&meshBox[m * 6 + c], f2ord(v[c]));
            for (int c = 3; c < 6; c++) atomicMax(&meshBox[m * 6 + c], f2ord(v[c]));
        }
    }
}

// Segmented key per triangle, its digit histograms, and the arrival words.
__global__ void __launch_bounds__(kBatchBlock)
    k_batch_keys(const Tri* __restrict__ tris, const uint32_t* __restrict__ off, int nMeshes,
                 const uint32_t* __restrict__ tileMesh, uint32_t total,
                 const uint32_t* __restrict__ meshBox, uint32_t opt, int localBits, int nPass,
                 uint64_t* __restrict__ keys, uint32_t* __restrict__ partials, uint32_t* __restrict__ flags)
{
    __shared__ uint32_t s_h[8][256];
    for (int p = 0; p < nPass; p++) s_h[p][threadIdx.x] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    for (uint32_t base = blockIdx.x * kBatchBlock; base < total; base += gridDim.x * kBatchBlock) {
        const uint32_t waveFirst = base + (threadIdx.x - lane);  // same in the whole wave
        if (waveFirst >= total) continue;
        const uint32_t i = waveFirst + lane;
        const bool valid = i < total;
        const uint64_t validMask = __ballot(valid);
        if (!valid) continue;
        const uint32_t m = mesh_from(off, nMeshes, tileMesh[i / kBatchBlock], i);
        const uint32_t* b = meshBox + 6 * m;
        f3 lo = mk(ord2f(b[0]), ord2f(b[1]), ord2f(b[2]));
        f3 hi = mk(ord2f(b[3]), ord2f(b[4]), ord2f(b[5]));
        Tri tri = tris[i];
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
        uint32_t local = i - off[m];
        uint64_t key = ((uint64_t)m << (32 + localBits)) | ((uint64_t)morton << localBits) | (uint64_t)local;
        keys[i] = key;
        if (i + 1 < off[m + 1]) flags[i - m] = 0xFFFFFFFFu;
        for (int p = 0; p < nPass; p++) {
            int sh = localBits + 8 * p;
            uint32_t d = sh < 64 ? (uint32_t)(key >> sh) & 0xFFu : 0u;
            // digits of the mesh id and the high Morton bits are mostly the
            // same across a wave: one LDS atomic for the wave then
            const uint32_t d0 = __shfl(d, 0, 64);
            if (__ballot(d == d0) == validMask) {
                if (lane == 0) atomicAdd(&s_h[p][d], (uint32_t)__popcll(validMask));
            } else {
                atomicAdd(&s_h[p][d], 1u);
            }
        }
    }
    __syncthreads();
    for (int p = 0; p < nPass; p++)
        partials[((size_t)blockIdx.x * nPass + p) * 256 + threadIdx.x] = s_h[p][threadIdx.x];
}

// Deltas inside each mesh (the last position of a mesh has no right neighbour
// and is never read).
__global__ void __launch_bounds__(kBatchBlock)
    k_batch_deltas(const uint64_t* __restrict__ keys, uint32_t total, int localBits, uint8_t* __restrict__ deltas)
{
    uint32_t i = blockIdx.x * kBatchBlock + threadIdx.x;
    if (i + 1 < total) {
        uint64_t a = keys[i], b = keys[i + 1];
        uint32_t ma = (uint32_t)(a >> localBits), mb = (uint32_t)(b >> localBits);
        deltas[i] = (uint8_t)(32 - __clz(ma ^ mb));
    }
}

// Mesh-local agglomerative climb (build_lbvh, minimum_lbvh.h:284-361) in two
// phases: a climb with the sibling hand-off in LDS for the nodes inside a
// block, then the device-scope climb of the single-mesh build (climb_from)
// for the rest; the block-local predicate is the one derived in build.hip.
// (The single-mesh build builds its in-block nodes each on its own instead,
// k_local_nodes.)  A block takes
// kBatchLeaves consecutive sorted positions, which may belong to several
// meshes.  In the block's delta array a split between two meshes (or past
// either end of the batch) holds 255, above every Morton delta (<= 32), so no
// node that a block finishes reaches across a mesh.  Node, leaf and range
// values are mesh-local; mesh m's split q is global split off[m] + q and its
// node is stored at nodesAll + (off[m] - m) + q.
constexpr int kBatchLeaves = 1024;
constexpr uint32_t kMeshEdge = 255u;

__global__ void __launch_bounds__(kBatchLeaves)
    k_batch_climb_local(const Tri* __restrict__ tris, const uint32_t* __restrict__ off, uint32_t total,
                        int localBits, const uint64_t* __restrict__ keys, const uint8_t* __restrict__ deltasAll,
                        Node* __restrict__ nodesAll, uint32_t* __restrict__ frontier, uint32_t* __restrict__ ctrl,
                        uint32_t* __restrict__ roots)
{
    constexpr int kL = kBatchLeaves, kWaves = kL / 64;
    // s_d[j] = delta of global split b0 - 1 + j (j in [0, m]); kMeshEdge across meshes or ends
    __shared__ uint8_t s_d[kL + 1];
    __shared__ uint8_t s_pm[kL];
    __shared__ uint8_t s_sm[kL + 1];
    __shared__ uint32_t s_wmax[2][kWaves];
    __shared__ uint32_t s_mesh[kL];  // mesh of position b0 + t
    __shared__ uint32_t s_edge[2];   // mesh of positions b0 - 1 and b1 + 1 (~0: none)
    __shared__ uint32_t s_flag[kL];
    __shared__ Box s_box[kL][2];
    __shared__ uint32_t s_child[kL][2];
    __shared__ uint32_t s_parent[kL];
    __shared__ uint32_t s_context[kL];
    __shared__ uint32_t s_wcount[kWaves];
    __shared__ uint32_t s_base;

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t b0 = blockIdx.x * kL;
    const uint32_t m = total - b0 < (uint32_t)kL ? total - b0 : (uint32_t)kL;
    const uint32_t b1 = b0 + m - 1;
    const int meshShift = 32 + localBits;

    uint64_t key = 0;
    uint32_t mesh = 0;
    if ((uint32_t)t < m) {
        key = keys[b0 + t];
        mesh = (uint32_t)(key >> meshShift);
        s_mesh[t] = mesh;
    }
    if (t == 0) s_edge[0] = b0 > 0 ? (uint32_t)(keys[b0 - 1] >> meshShift) : 0xFFFFFFFFu;
    if (t == 1) s_edge[1] = b1 + 1 < total ? (uint32_t)(keys[b1 + 1] >> meshShift) : 0xFFFFFFFFu;
    s_flag[t] = 0xFFFFFFFFu;
    s_parent[t] = kInvalidNode;
    __syncthreads();
    auto delta_at = [&](uint32_t j) -> uint32_t {
        if (j > m) return 0u;  // past a partial block: never decides anything
        const uint32_t meshA = j == 0 ? s_edge[0] : s_mesh[j - 1];
        const uint32_t meshB = j == m ? s_edge[1] : s_mesh[j];
        if (meshA == 0xFFFFFFFFu || meshB == 0xFFFFFFFFu || meshA != meshB) return kMeshEdge;
        return (uint32_t)deltasAll[b0 - 1 + j];
    };
    uint32_t e0 = delta_at((uint32_t)t), e1 = delta_at((uint32_t)t + 1);
    s_d[t] = (uint8_t)e0;
    if (t == kL - 1) s_d[kL] = (uint8_t)e1;
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
    for (int w = 0; w < wave; w++) pm = pm > s_wmax[0][w] ? pm : s_wmax[0][w];
    for (int w = wave + 1; w < kWaves; w++) sm = sm > s_wmax[1][w] ? sm : s_wmax[1][w];
    s_pm[t] = (uint8_t)pm;
    s_sm[t + 1] = (uint8_t)sm;
    __syncthreads();

    bool rootHere = false, emit = false;
    uint32_t node = 0, lo = 0, hi = 0;
    if ((uint32_t)t < m) {
        const uint32_t base = off[mesh];
        const uint32_t nInternals = off[mesh + 1] - base - 1;
        const uint32_t triIndex = (uint32_t)key & ((1u << localBits) - 1u);
        Box box = tri_box(tris[base + triIndex]);
        lo = hi = b0 + (uint32_t)t - base;
        node = triIndex | kLeafBit;
        const uint32_t gOff = base - b0 + 1;  // s_d index of mesh-local split q is q + gOff
        while (true) {
            if (hi - lo >= nInternals) {
                rootHere = true;
                break;
            }
            uint32_t dL = 0 < lo ? (uint32_t)s_d[lo - 1 + gOff] : 0xFFFFFFFFu;
            uint32_t dR = hi < nInternals ? (uint32_t)s_d[hi + gOff] : 0xFFFFFFFFu;
            int goLeft = dL < dR ? 1 : 0;
            uint32_t parent = goLeft ? lo - 1 : hi;
            bool local = goLeft ? base + lo > b0 : base + hi < b1;
            if (local) {
                uint32_t jp = parent + gOff;
                uint32_t dp = s_d[jp];
                local = (uint32_t)s_pm[jp - 1] >= dp && (uint32_t)s_sm[jp + 1] > dp;
            }
            if (!local) {
                emit = true;
                break;
            }
            const uint32_t slot = parent + gOff - 1;
            if (!(node & kLeafBit)) s_parent[node + gOff - 1] = parent;
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
    if (rootHere) roots[mesh] = node;

    const uint64_t emitMask = __ballot(emit);
    if (lane == 0) s_wcount[wave] = (uint32_t)__popcll(emitMask);
    __syncthreads();
    if (t == 0) {
        uint32_t sum = 0;
        for (int w = 0; w < kWaves; w++) {
            uint32_t c = s_wcount[w];
            s_wcount[w] = sum;
            sum += c;
        }
        s_base = sum ? atomicAdd(&ctrl[kCtrlFrontier], sum) : 0u;
    }
    __syncthreads();
    if (emit) {
        const uint32_t k = s_base + s_wcount[wave] + (uint32_t)__popcll(emitMask & ((1ull << lane) - 1ull));
        uint4* f = reinterpret_cast<uint4*>(frontier) + k;
        *f = make_uint4(node, lo, hi, mesh);
    }

    // whole block-local nodes: slot s is global split b0 + s of mesh s_mesh[s]
    const uint32_t nSlots = m - 1;
    for (uint32_t q = (uint32_t)t; q < 4 * nSlots; q += kL) {
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
        reinterpret_cast<uint4*>(nodesAll + (b0 + slot - s_mesh[slot]))[part] = v;
    }
}

// Phase 2: the frontier entries (node, lo, hi, mesh) climb on through the
// device-scope arrival words of their mesh.
__global__ void __launch_bounds__(kBatchBlock)
    k_batch_climb_upper(const Tri* __restrict__ tris, const uint32_t* __restrict__ off,
                        const uint8_t* __restrict__ deltasAll, Node* nodesAll, uint32_t* flagsAll,
                        const uint32_t* __restrict__ frontier, const uint32_t* __restrict__ ctrl, uint32_t* roots)
{
    const uint32_t count = ctrl[kCtrlFrontier];
    for (uint32_t i = blockIdx.x * kBatchBlock + threadIdx.x; i < count; i += gridDim.x * kBatchBlock) {
        const uint4 e = reinterpret_cast<const uint4*>(frontier)[i];
        const uint32_t node = e.x, lo = e.y, hi = e.z, mesh = e.w;
        const uint32_t base = off[mesh];
        Node* nodes = nodesAll + (base - mesh);
        Box box;
        if (node & kLeafBit) {
            box = tri_box(tris[base + (node & ~kLeafBit)]);
        } else {
            const Node& nd = nodes[node];
            box.lo = umin3(nd.box[0].lo, nd.box[1].lo);
            box.hi = umax3(nd.box[0].hi, nd.box[1].hi);
        }
        climb_from(off[mesh + 1] - base, lo, hi, node, box, deltasAll + base, nodes, flagsAll + (base - mesh),
                   roots + mesh);
    }
}

static int bits_for(uint32_t v)  // smallest b with (1 << b) >= v
{
    int b = 0;
    while (b < 32 && ((uint64_t)1 << b) < (uint64_t)v) b++;
    return b;
}

}  // namespace mlbvh

using namespace mlbvh;

// Declared in include/mlbvh.h.
extern "C" int mlbvh_build_batched(mlbvh_ctx* ctx, const void* d_triangles, const uint32_t* d_triOffsets,
                                   int nMeshes, uint32_t totalTriangles, uint32_t maxTrianglesPerMesh,
                                   uint32_t buildOption, void* d_nodes, uint32_t* d_roots, void* stream)
{
    if (!ctx || !d_triangles || !d_triOffsets || !d_nodes || !d_roots || nMeshes < 1) {
        set_error("build_batched: null argument or no meshes");
        return MLBVH_E_INVALID_ARG;
    }
    if (maxTrianglesPerMesh < 2 || totalTriangles < 2 * (uint32_t)nMeshes || totalTriangles > 0x7FFFFFFFu) {
        set_error("build_batched: every mesh needs >= 2 triangles (total %u, meshes %d, max %u)", totalTriangles,
                  nMeshes, maxTrianglesPerMesh);
        return MLBVH_E_INVALID_ARG;
    }
    const int localBits = bits_for(maxTrianglesPerMesh);
    const int meshBits = bits_for((uint32_t)nMeshes) > 0 ? bits_for((uint32_t)nMeshes) : 1;
    if (localBits + 32 + meshBits > 64) {
        set_error("build_batched: %d meshes of up to %u triangles need %d key bits (> 64)", nMeshes,
                  maxTrianglesPerMesh, localBits + 32 + meshBits);
        return MLBVH_E_INVALID_ARG;
    }
    const int keyBits = 32 + meshBits;
    const int nPass = (keyBits + 7) / 8;
    int rc = set_device(ctx);
    if (rc) return rc;
    size_t total = totalTriangles;
    if ((rc = ensure(ctx, &ctx->recA, &ctx->recCap, total, "hipMalloc(records)"))) return rc;
    if ((rc = ensure(ctx, &ctx->recB, &ctx->recBCap, total, "hipMalloc(records tmp)"))) return rc;
    if ((rc = ensure(ctx, &ctx->deltas, &ctx->deltasCap, total, "hipMalloc(deltas)"))) return rc;
    if ((rc = ensure(ctx, &ctx->flags, &ctx->flagsCap, total, "hipMalloc(flags)"))) return rc;
    if ((rc = ensure(ctx, &ctx->meshBox, &ctx->meshBoxCap, (size_t)nMeshes * 6, "hipMalloc(meshBox)"))) return rc;
    if ((rc = ensure(ctx, &ctx->frontier, &ctx->frontierCap, total * 4, "hipMalloc(climb frontier)"))) return rc;
    if ((rc = ensure_status(ctx, sort_tiles(ctx, totalTriangles)))) return rc;
    unsigned g = (unsigned)((total + kBatchBlock - 1) / kBatchBlock);
// grid-stride key kernel: 8 blocks per CU built the 4096-mesh batch in
// 12.55 ms against 12.87 ms with 4 (16: 12.54)
#ifndef MLBVH_KEYS_BLOCKS_PER_CU
#define MLBVH_KEYS_BLOCKS_PER_CU 8
#endif
    const unsigned keysCap = (unsigned)ctx->cuCount * MLBVH_KEYS_BLOCKS_PER_CU;
    unsigned gKeys = g < keysCap ? g : keysCap;
    if ((rc = ensure_partials(ctx, (int)gKeys, nPass))) return rc;

    hipStream_t s = (hipStream_t)stream;
    const Tri* tris = (const Tri*)d_triangles;
    k_batch_init<<<1, 256, 0, s>>>(ctx->meshBox, nMeshes, ctx->ctrl, ctx->hist, nPass, ctx->status,
                                   ctx->statusWords);
    // the tile -> mesh table lives in the sort's second record buffer until the sort
    uint32_t* tileMesh = (uint32_t*)ctx->recB;
    k_batch_tile_mesh<<<(nMeshes + kBatchBlock - 1) / kBatchBlock, kBatchBlock, 0, s>>>(d_triOffsets, nMeshes,
                                                                                       totalTriangles, tileMesh);
    const unsigned nTiles = (unsigned)((total + kBatchBlock - 1) / kBatchBlock);
    k_batch_aabb<<<(nTiles + kAabbTiles - 1) / kAabbTiles, kBatchBlock, 0, s>>>(tris, d_triOffsets, nMeshes, tileMesh,
                                                                              totalTriangles, ctx->meshBox);
    k_batch_keys<<<gKeys, kBatchBlock, 0, s>>>(tris, d_triOffsets, nMeshes, tileMesh, totalTriangles, ctx->meshBox,
                                               buildOption, localBits, nPass, ctx->recA, ctx->partials, ctx->flags);
    launch_hist_reduce(ctx, (int)gKeys, nPass, s);
    rc = sort_records_u64(ctx, ctx->recA, ctx->recB, totalTriangles, localBits, localBits + keyBits, true, s);
    if (rc) return rc;
    k_batch_deltas<<<g, kBatchBlock, 0, s>>>(ctx->recA, totalTriangles, localBits, ctx->deltas);
    unsigned gLocal = (unsigned)((total + kBatchLeaves - 1) / kBatchLeaves);
    k_batch_climb_local<<<gLocal, kBatchLeaves, 0, s>>>(tris, d_triOffsets, totalTriangles, localBits, ctx->recA,
                                                        ctx->deltas, (Node*)d_nodes, ctx->frontier, ctx->ctrl,
                                                        d_roots);
    // one lane per expected entry (a few percent of the positions), at most 8
    // blocks per CU; lanes stride over a larger list
    size_t want = (total / 32 + kBatchBlock - 1) / kBatchBlock + 1;
    unsigned cap = (unsigned)ctx->cuCount * 8;
    k_batch_climb_upper<<<want < cap ? (unsigned)want : cap, kBatchBlock, 0, s>>>(
        tris, d_triOffsets, ctx->deltas, (Node*)d_nodes, ctx->flags, ctx->frontier, ctx->ctrl, d_roots);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? MLBVH_OK : hip_status(e, "build_batched launch");
}

// The plain path over the same batch: one mlbvh_build + one mlbvh_trace per
// mesh, all launched asynchronously on `stream` (what a reference user does
// with one BVHGPUBuilder::build per mesh).  Output layout is that of
// mlbvh_build_batched / mlbvh_trace_batched.
extern "C" int mlbvh_build_trace_per_mesh(mlbvh_ctx* ctx, const void* d_triangles, const uint32_t* h_triOffsets,
                                          int nMeshes, uint32_t buildOption, void* d_nodes, uint32_t* d_roots,
                                          const float* d_rays, const uint32_t* h_rayOffsets, void* d_hits,
                                          int rayQuery, void* stream)
{
    if (!ctx || !d_triangles || !h_triOffsets || !h_rayOffsets || !d_nodes || !d_roots || nMeshes < 1) {
        set_error("build_trace_per_mesh: null argument");
        return MLBVH_E_INVALID_ARG;
    }
    const Tri* tris = (const Tri*)d_triangles;
    Node* nodes = (Node*)d_nodes;
    for (int m = 0; m < nMeshes; m++) {
        uint32_t b = h_triOffsets[m];
        int n = (int)(h_triOffsets[m + 1] - b);
        int rc = build_one(ctx, tris + b, n, buildOption, nodes + (b - (uint32_t)m), d_roots + m, nullptr, nullptr,
                           (hipStream_t)stream);
        if (rc) return rc;
        uint32_t r0 = h_rayOffsets[m];
        int nr = (int)(h_rayOffsets[m + 1] - r0);
        if (nr > 0) {
            rc = mlbvh_trace(ctx, nodes + (b - (uint32_t)m), tris + b, d_roots + m, d_rays + 8 * (size_t)r0,
                             (char*)d_hits + 32 * (size_t)r0, nr, rayQuery, nullptr, stream);
            if (rc) return rc;
        }
    }
    return MLBVH_OK;
}
