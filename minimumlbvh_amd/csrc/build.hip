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
//  * the climb exchanges a separate u32 arrival word per internal node and
//    writes `context` once, after the merge, with the reference's final value;
//  * one thread per leaf for all n leaves (the reference launches
//    ceil((n-1)/256) blocks and drops the last leaf when (n-1) % 256 == 0).
#include "mlbvh_ctx.h"

namespace mlbvh {

constexpr int kBuildBlock = 256;

// Scene-box slots plus the sort state of the 4-pass record sort.
__global__ void k_build_init(uint32_t* sceneBox, uint32_t* ctrl, uint32_t* hist, uint64_t* status,
                             size_t statusWords)
{
    int t = threadIdx.x;
    if (t < 3) sceneBox[t] = 0xFFFFFFFFu;  // atomicMin slots (order-mapped)
    else if (t < 6) sceneBox[t] = 0u;      // atomicMax slots
    sort_state_init(ctrl, 4, hist, 4, status, statusWords);
}

// Scene bounds over all vertices: per-thread running min/max, wave and block
// reduction, then 6 order-mapped integer atomics per block.
__global__ void __launch_bounds__(kBuildBlock) k_scene_aabb(const Tri* __restrict__ tris, int n, uint32_t* sceneBox)
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
        if (c < 3) atomicMin(&sceneBox[c], f2ord(r));
        else atomicMax(&sceneBox[c], f2ord(r));
    }
}

// Morton record per triangle (minimum_lbvh.cu:68-92 / minimum_lbvh.h:534-553),
// the 4 digit histograms of the 32-bit key, and the climb's arrival words.
// Grid-stride over a few blocks per CU so that the histograms are merged into
// global memory by few blocks (global atomics on 1024 shared counters are the
// cost otherwise).
__global__ void __launch_bounds__(kBuildBlock)
    k_morton(const Tri* __restrict__ tris, int n, const uint32_t* __restrict__ sceneBox, uint32_t opt,
             uint64_t* __restrict__ rec, uint32_t* __restrict__ partials, uint32_t* __restrict__ flags)
{
    __shared__ uint32_t s_h[4][256];
    for (int p = 0; p < 4; p++) s_h[p][threadIdx.x] = 0;
    __syncthreads();
    const f3 lo = mk(ord2f(sceneBox[0]), ord2f(sceneBox[1]), ord2f(sceneBox[2]));
    const f3 hi = mk(ord2f(sceneBox[3]), ord2f(sceneBox[4]), ord2f(sceneBox[5]));
    for (int i = blockIdx.x * kBuildBlock + threadIdx.x; i < n; i += gridDim.x * kBuildBlock) {
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
        rec[i] = (uint64_t)morton | ((uint64_t)(uint32_t)i << 32);
        if (i < n - 1) flags[i] = 0xFFFFFFFFu;
#pragma unroll
        for (int p = 0; p < 4; p++) atomicAdd(&s_h[p][(morton >> (8 * p)) & 0xFFu], 1u);
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < 4; p++) partials[((size_t)blockIdx.x * 4 + p) * 256 + threadIdx.x] = s_h[p][threadIdx.x];
}

// delta(a, b) = 32 - clz(a ^ b) of neighbouring sorted keys
// (minimum_lbvh.cu:94-102).
__global__ void __launch_bounds__(kBuildBlock)
    k_deltas(const uint64_t* __restrict__ rec, int n, uint8_t* __restrict__ deltas)
{
    int i = blockIdx.x * kBuildBlock + threadIdx.x;
    if (i < n - 1) {
        uint32_t a = (uint32_t)rec[i], b = (uint32_t)rec[i + 1];
        deltas[i] = (uint8_t)(32 - __clz(a ^ b));
    }
}

__global__ void __launch_bounds__(kBuildBlock) k_fill_u32(uint32_t* p, int n, uint32_t v)
{
    int i = blockIdx.x * kBuildBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

// Agglomerative climb, one thread per sorted leaf (build_lbvh,
// minimum_lbvh.h:284-361): see climb_leaf in mlbvh_device.h.
__global__ void __launch_bounds__(kBuildBlock)
    k_climb(const Tri* __restrict__ tris, int n, const uint64_t* __restrict__ rec, const uint8_t* __restrict__ deltas,
            Node* nodes, uint32_t* flags, uint32_t* root)
{
    int i = blockIdx.x * kBuildBlock + threadIdx.x;
    if (i >= n) return;
    climb_leaf(tris, (uint32_t)n, (uint32_t)i, (uint32_t)(rec[i] >> 32), deltas, nodes, flags, root);
}

static inline unsigned grid1(size_t n) { return (unsigned)((n + kBuildBlock - 1) / kBuildBlock); }

int build_one(mlbvh_ctx* ctx, const void* d_tris, int n, uint32_t opt, void* d_nodes, void* d_root,
                      uint64_t* d_sorted, uint8_t* d_deltas, hipStream_t stream)
{
    if (!ctx || !d_tris || !d_nodes || !d_root) {
        set_error("build: null argument");
        return MLBVH_E_INVALID_ARG;
    }
    if (n < 2) {
        set_error("build: n = %d; at least 2 triangles are required", n);
        return MLBVH_E_INVALID_ARG;
    }
    int rc = set_device(ctx);
    if (rc) return rc;
    size_t nn = (size_t)n;
    if ((rc = ensure(&ctx->recA, &ctx->recCap, nn, "hipMalloc(records)"))) return rc;
    if ((rc = ensure(&ctx->recB, &ctx->recBCap, nn, "hipMalloc(records tmp)"))) return rc;
    if ((rc = ensure(&ctx->deltas, &ctx->deltasCap, nn, "hipMalloc(deltas)"))) return rc;
    if ((rc = ensure(&ctx->flags, &ctx->flagsCap, nn, "hipMalloc(flags)"))) return rc;

    const Tri* tris = (const Tri*)d_tris;
    Node* nodes = (Node*)d_nodes;
    uint32_t* root = (uint32_t*)d_root;

    if ((rc = ensure_status(ctx, sort_tiles(ctx, (uint32_t)n)))) return rc;
    k_build_init<<<1, 256, 0, stream>>>(ctx->sceneBox, ctx->ctrl, ctx->hist, ctx->status, ctx->statusWords);
    unsigned g = grid1(nn);
    unsigned gAabb = g < (unsigned)ctx->cuCount * 8 ? g : (unsigned)ctx->cuCount * 8;
    k_scene_aabb<<<gAabb, kBuildBlock, 0, stream>>>(tris, n, ctx->sceneBox);
    unsigned gMorton = g < (unsigned)ctx->cuCount * 4 ? g : (unsigned)ctx->cuCount * 4;
    if ((rc = ensure_partials(ctx, (int)gMorton, 4))) return rc;
    k_morton<<<gMorton, kBuildBlock, 0, stream>>>(tris, n, ctx->sceneBox, opt, ctx->recA, ctx->partials, ctx->flags);
    launch_hist_reduce(ctx, (int)gMorton, 4, stream);
    rc = sort_records_u64(ctx, ctx->recA, ctx->recB, (uint32_t)n, 0, 32, true, stream);
    if (rc) return rc;
    k_deltas<<<g, kBuildBlock, 0, stream>>>(ctx->recA, n, ctx->deltas);
    k_climb<<<g, kBuildBlock, 0, stream>>>(tris, n, ctx->recA, ctx->deltas, nodes, ctx->flags, root);
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
    if (!ctx || !d_triangles || !d_sorted_records || !d_deltas || !d_nodes || !d_root) {
        set_error("refit: null argument");
        return MLBVH_E_INVALID_ARG;
    }
    if (n < 2) {
        set_error("refit: n = %d; at least 2 triangles are required", n);
        return MLBVH_E_INVALID_ARG;
    }
    int rc = set_device(ctx);
    if (rc) return rc;
    if ((rc = ensure(&ctx->flags, &ctx->flagsCap, (size_t)n, "hipMalloc(flags)"))) return rc;
    hipStream_t s = (hipStream_t)stream;
    unsigned g = grid1((size_t)n);
    k_fill_u32<<<g, kBuildBlock, 0, s>>>(ctx->flags, n - 1, 0xFFFFFFFFu);
    k_climb<<<g, kBuildBlock, 0, s>>>((const Tri*)d_triangles, n, d_sorted_records, d_deltas, (Node*)d_nodes,
                                      ctx->flags, (uint32_t*)d_root);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_status(e, "refit kernel launch");
    return MLBVH_OK;
}
