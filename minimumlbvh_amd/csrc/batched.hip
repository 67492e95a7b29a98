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
// coordinates.  Tracing uses a persistent kernel whose waves pull 64-ray
// chunks from an atomic work counter; each ray names its mesh in the Ray's
// pad word.
#include "mlbvh_ctx.h"

namespace mlbvh {

constexpr int kBatchBlock = 256;

__device__ inline uint32_t mesh_of(const uint32_t* __restrict__ off, int nMeshes, uint32_t i)
{
    // last m with off[m] <= i
    int lo = 0, hi = nMeshes - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return (uint32_t)lo;
}

__global__ void k_batch_init(uint32_t* meshBox, int nMeshes, uint32_t* ctrl, uint32_t* hist, int nPass,
                             uint64_t* status, size_t statusWords)
{
    for (int i = threadIdx.x; i < nMeshes * 6; i += blockDim.x) meshBox[i] = (i % 6) < 3 ? 0xFFFFFFFFu : 0u;
    sort_state_init(ctrl, nPass, hist, nPass, status, statusWords);
}

// Per-mesh scene AABB: blocks inside one mesh reduce in LDS and issue 6
// atomics; blocks that straddle mesh boundaries fall back to per-wave (or
// per-lane) atomics.
__global__ void __launch_bounds__(kBatchBlock)
    k_batch_aabb(const Tri* __restrict__ tris, const uint32_t* __restrict__ off, int nMeshes, uint32_t total,
                 uint32_t* meshBox)
{
    uint32_t i = blockIdx.x * kBatchBlock + threadIdx.x;
    uint32_t first = blockIdx.x * kBatchBlock;
    uint32_t last = min(first + kBatchBlock, total) - 1;
    const uint32_t mFirst = mesh_of(off, nMeshes, first);
    const bool blockUniform = off[mFirst + 1] > last;
    float v[6] = {kFltMax, kFltMax, kFltMax, -kFltMax, -kFltMax, -kFltMax};
    uint32_t m = mFirst;
    if (i < total) {
        Tri t = tris[i];
        for (int k = 0; k < 3; k++) {
            v[0] = fminf(v[0], t.v[k].x);
            v[1] = fminf(v[1], t.v[k].y);
            v[2] = fminf(v[2], t.v[k].z);
            v[3] = fmaxf(v[3], t.v[k].x);
            v[4] = fmaxf(v[4], t.v[k].y);
            v[5] = fmaxf(v[5], t.v[k].z);
        }
        if (!blockUniform) m = mesh_of(off, nMeshes, i);
    }
    if (blockUniform) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
            for (int c = 0; c < 3; c++) v[c] = fminf(v[c], __shfl_xor(v[c], o, 64));
#pragma unroll
            for (int c = 3; c < 6; c++) v[c] = fmaxf(v[c], __shfl_xor(v[c], o, 64));
        }
        __shared__ float s_v[kBatchBlock / 64][6];
        int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        if (lane == 0)
            for (int c = 0; c < 6; c++) s_v[wave][c] = v[c];
        __syncthreads();
        if (threadIdx.x < 6) {
            int c = threadIdx.x;
            float r = s_v[0][c];
            for (int w = 1; w < kBatchBlock / 64; w++) r = c < 3 ? fminf(r, s_v[w][c]) : fmaxf(r, s_v[w][c]);
            if (c < 3) atomicMin(&meshBox[mFirst * 6 + c], f2ord(r));
            else atomicMax(&meshBox[mFirst * 6 + c], f2ord(r));
        }
    } else if (i < total) {
        for (int c = 0; c < 3; c++) atomicMin(&meshBox[m * 6 + c], f2ord(v[c]));
        for (int c = 3; c < 6; c++) atomicMax(&meshBox[m * 6 + c], f2ord(v[c]));
    }
}

// Segmented key per triangle, its digit histograms, and the arrival words.
__global__ void __launch_bounds__(kBatchBlock)
    k_batch_keys(const Tri* __restrict__ tris, const uint32_t* __restrict__ off, int nMeshes, uint32_t total,
                 const uint32_t* __restrict__ meshBox, uint32_t opt, int localBits, int nPass,
                 uint64_t* __restrict__ keys, uint32_t* __restrict__ partials, uint32_t* __restrict__ flags)
{
    __shared__ uint32_t s_h[8][256];
    for (int p = 0; p < nPass; p++) s_h[p][threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t i = blockIdx.x * kBatchBlock + threadIdx.x; i < total; i += gridDim.x * kBatchBlock) {
        uint32_t m = mesh_of(off, nMeshes, i);
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
            atomicAdd(&s_h[p][d], 1u);
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

// Mesh-local agglomerative climb (build_lbvh, minimum_lbvh.h:284-361) for the
// sorted position i of every mesh.
__global__ void __launch_bounds__(kBatchBlock)
    k_batch_climb(const Tri* __restrict__ tris, const uint32_t* __restrict__ off, uint32_t total, int localBits,
                  const uint64_t* __restrict__ keys, const uint8_t* __restrict__ deltasAll, Node* nodesAll,
                  uint32_t* flagsAll, uint32_t* roots)
{
    uint32_t i = blockIdx.x * kBatchBlock + threadIdx.x;
    if (i >= total) return;
    const uint64_t key = keys[i];
    const uint32_t m = (uint32_t)(key >> (32 + localBits));
    const uint32_t triIndex = (uint32_t)key & ((1u << localBits) - 1u);
    const uint32_t base = off[m];
    climb_leaf(tris + base, off[m + 1] - base, i - base, triIndex, deltasAll + base, nodesAll + (base - m),
               flagsAll + (base - m), roots + m);
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
    if ((rc = ensure(&ctx->recA, &ctx->recCap, total, "hipMalloc(records)"))) return rc;
    if ((rc = ensure(&ctx->recB, &ctx->recBCap, total, "hipMalloc(records tmp)"))) return rc;
    if ((rc = ensure(&ctx->deltas, &ctx->deltasCap, total, "hipMalloc(deltas)"))) return rc;
    if ((rc = ensure(&ctx->flags, &ctx->flagsCap, total, "hipMalloc(flags)"))) return rc;
    if ((rc = ensure(&ctx->meshBox, &ctx->meshBoxCap, (size_t)nMeshes * 6, "hipMalloc(meshBox)"))) return rc;
    if ((rc = ensure_status(ctx, sort_tiles(ctx, totalTriangles)))) return rc;
    unsigned g = (unsigned)((total + kBatchBlock - 1) / kBatchBlock);
    unsigned gKeys = g < (unsigned)ctx->cuCount * 4 ? g : (unsigned)ctx->cuCount * 4;
    if ((rc = ensure_partials(ctx, (int)gKeys, nPass))) return rc;

    hipStream_t s = (hipStream_t)stream;
    const Tri* tris = (const Tri*)d_triangles;
    k_batch_init<<<1, 256, 0, s>>>(ctx->meshBox, nMeshes, ctx->ctrl, ctx->hist, nPass, ctx->status,
                                   ctx->statusWords);
    k_batch_aabb<<<g, kBatchBlock, 0, s>>>(tris, d_triOffsets, nMeshes, totalTriangles, ctx->meshBox);
    k_batch_keys<<<gKeys, kBatchBlock, 0, s>>>(tris, d_triOffsets, nMeshes, totalTriangles, ctx->meshBox,
                                               buildOption, localBits, nPass, ctx->recA, ctx->partials, ctx->flags);
    launch_hist_reduce(ctx, (int)gKeys, nPass, s);
    rc = sort_records_u64(ctx, ctx->recA, ctx->recB, totalTriangles, localBits, localBits + keyBits, true, s);
    if (rc) return rc;
    k_batch_deltas<<<g, kBatchBlock, 0, s>>>(ctx->recA, totalTriangles, localBits, ctx->deltas);
    k_batch_climb<<<g, kBatchBlock, 0, s>>>(tris, d_triOffsets, totalTriangles, localBits, ctx->recA, ctx->deltas,
                                            (Node*)d_nodes, ctx->flags, d_roots);
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
