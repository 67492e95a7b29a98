// Differentiable primary-hit rendering (no reference counterpart; built on
// the reference's intersectRayTriangle, minimum_lbvh.h:363-413, and its
// camera, camera.h:24-31).
//
// Forward (k_render_primary): for each pixel the closest hit of the camera
// ray; out[p] = (t, s) with s = max(dot(n, L), 0), n the geometric normal
// normalised and flipped toward the viewer; miss = (0, 0).  The hit triangle
// id is kept for the backward pass.
//
// Adjoint (k_primary_adjoint): with the hit triangle fixed, t and s are
// smooth functions of its three vertices:
//   t = dot(v0 - ro, ng) / dot(ng, rd),  ng = (v1 - v0) x (v2 - v1)
//   w = (v0 - ro - t rd) / dot(ng, rd)
//   dt/dv0 = ng/den - e1 x w,  dt/dv1 = e1 x w - w x e0,  dt/dv2 = w x e0
//   q = sgn (L - (nh.L) nh) / |ng|  (nh = ng/|ng|, sgn flips toward the viewer;
//                                    q = 0 where s = 0)
//   ds/dv0 = -(e1 x q), ds/dv1 = e1 x q - q x e0, ds/dv2 = q x e0
// and dL/dv accumulates g_t dt/dv + g_s ds/dv into the per-vertex gradient
// with fp32 atomics.  A block takes a 16x16 pixel tile, whose hits fall on a
// compact patch of triangles sharing vertices: the block first sums its
// contributions per vertex in an LDS hash table (LDS fp32 atomics), then adds
// each distinct vertex's sum to the gradient with 3 global fp32 atomics.  The
// global atomics are device-scope read-modify-writes performed beyond the
// XCD's L2; merging in LDS issues about a sixth of the 9 per hit pixel.
#include "mlbvh_ctx.h"
#include "traverse.h"

namespace mlbvh {

constexpr int kDiffBlock = 256;

__global__ void __launch_bounds__(kDiffBlock)
    k_gather(const float* __restrict__ v, const uint32_t* __restrict__ idx, int nTris, Tri* __restrict__ tris)
{
    int f = blockIdx.x * kDiffBlock + threadIdx.x;
    if (f >= nTris) return;
    Tri t;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        uint32_t j = idx[3 * f + k];
        t.v[k] = mk(v[3 * j], v[3 * j + 1], v[3 * j + 2]);
    }
    tris[f] = t;
}

__global__ void __launch_bounds__(kDiffBlock)
    k_render_primary(float2* __restrict__ out, uint32_t* __restrict__ hitTri, int w, int h, mlbvh_raygen cam,
                     f3 light, const uint32_t* __restrict__ rootPtr, const Node* __restrict__ nodes,
                     const Tri* __restrict__ tris)
{
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int xi = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    int yi = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
    if (xi >= w || yi >= h) return;
    int p = xi + yi * w;
    f3 ro, rd;
    shoot(cam, &ro, &rd, (float)xi / (float)w, (float)yi / (float)h);
    Hit hit;
    hit.t = kFltMax;
    hit.u = hit.v = 0.0f;
    hit.ng = mk(0, 0, 0);
    hit.tri = 0xFFFFFFFFu;
    uint32_t nFetch = 0, nTest = 0;
    __shared__ uint32_t s_stack[kLaneStack][kDiffBlock];
    __shared__ float s_entry[kLaneStack][kDiffBlock];
    traverse_stack<false, false, kDiffBlock>(nodes, tris, *rootPtr, ro, rd, inv_rd(rd), hit, nFetch, nTest,
                                             &s_stack[0][threadIdx.x], &s_entry[0][threadIdx.x]);
    const float t = hit.t;
    hitTri[p] = hit.tri;
    if (hit.tri == 0xFFFFFFFFu) {
        out[p] = make_float2(0.0f, 0.0f);
        return;
    }
    f3 n = normalize(hit.ng);
    if (0.0f < dot(n, rd)) n = -n;
    float s = dot(n, light);
    out[p] = make_float2(t, s > 0.0f ? s : 0.0f);
}

constexpr int kAdjSlots = 1024;  // a power of two (the hash keeps 10 bits)
constexpr uint32_t kAdjEmpty = 0xFFFFFFFFu;
static_assert(kAdjSlots > 3 * kDiffBlock, "more slots than insertions, so every probe finds a free slot");

__global__ void __launch_bounds__(kDiffBlock)
    k_primary_adjoint(const float2* __restrict__ gout, const uint32_t* __restrict__ hitTri, int w, int h,
                      mlbvh_raygen cam, f3 light, const float* __restrict__ vtx, const uint32_t* __restrict__ idx,
                      float* __restrict__ grad)
{
    __shared__ uint32_t s_key[kAdjSlots];
    __shared__ float s_val[kAdjSlots][3];
    for (int i = threadIdx.x; i < kAdjSlots; i += kDiffBlock) {
        s_key[i] = kAdjEmpty;
        s_val[i][0] = s_val[i][1] = s_val[i][2] = 0.0f;
    }
    __syncthreads();

    const int xi = blockIdx.x * 16 + (threadIdx.x & 15), yi = blockIdx.y * 16 + (threadIdx.x >> 4);
    uint32_t tri = 0xFFFFFFFFu;
    float2 g = make_float2(0.0f, 0.0f);
    if (xi < w && yi < h) {
        const int p = xi + yi * w;
        tri = hitTri[p];
        if (tri != 0xFFFFFFFFu) g = gout[p];
    }
    if (tri != 0xFFFFFFFFu && (g.x != 0.0f || g.y != 0.0f)) {
        f3 ro, rd;
        shoot(cam, &ro, &rd, (float)xi / (float)w, (float)yi / (float)h);
        uint32_t j[3];
        f3 v[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            j[k] = idx[3 * tri + k];
            v[k] = mk(vtx[3 * j[k]], vtx[3 * j[k] + 1], vtx[3 * j[k] + 2]);
        }
        f3 e0 = v[1] - v[0], e1 = v[2] - v[1];
        f3 ng = cross(e0, e1);
        float den = dot(ng, rd);
        float t = dot(v[0] - ro, ng) / den;
        f3 wv = (v[0] - ro - rd * t) * (1.0f / den);
        f3 e1w = cross(e1, wv), we0 = cross(wv, e0);
        f3 dt0 = ng * (1.0f / den) - e1w, dt1 = e1w - we0, dt2 = we0;
        f3 d0 = dt0 * g.x, d1 = dt1 * g.x, d2 = dt2 * g.x;
        float len = sqrtf(dot(ng, ng));
        f3 nh = ng * (1.0f / len);
        float sgn = dot(nh, rd) > 0.0f ? -1.0f : 1.0f;
        float sh = sgn * dot(nh, light);
        if (sh > 0.0f && g.y != 0.0f) {
            f3 q = (light - nh * dot(nh, light)) * (sgn / len);
            f3 e1q = cross(e1, q), qe0 = cross(q, e0);
            d0 = d0 - e1q * g.y;
            d1 = d1 + (e1q - qe0) * g.y;
            d2 = d2 + qe0 * g.y;
        }
        const f3 d[3] = {d0, d1, d2};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            // open addressing: claim the vertex's slot (or find it), then add
            uint32_t slot = (j[k] * 2654435761u) >> 22;  // 10-bit multiplicative hash
            while (true) {
                const uint32_t old = atomicCAS(&s_key[slot], kAdjEmpty, j[k]);
                if (old == kAdjEmpty || old == j[k]) break;
                slot = (slot + 1) & (kAdjSlots - 1);
            }
            atomicAdd(&s_val[slot][0], d[k].x);
            atomicAdd(&s_val[slot][1], d[k].y);
            atomicAdd(&s_val[slot][2], d[k].z);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kAdjSlots; i += kDiffBlock) {
        const uint32_t vj = s_key[i];
        if (vj == kAdjEmpty) continue;
        atomicAdd(&grad[3 * (size_t)vj + 0], s_val[i][0]);
        atomicAdd(&grad[3 * (size_t)vj + 1], s_val[i][1]);
        atomicAdd(&grad[3 * (size_t)vj + 2], s_val[i][2]);
    }
}

}  // namespace mlbvh

using namespace mlbvh;

extern "C" int mlbvh_gather_triangles(mlbvh_ctx* ctx, const float* d_vertices, const uint32_t* d_indices, int nTris,
                                      void* d_triangles, void* stream)
{
    if (!ctx || nTris < 0 || (nTris > 0 && (!d_vertices || !d_indices || !d_triangles))) {
        set_error("gather_triangles: invalid arguments");
        return MLBVH_E_INVALID_ARG;
    }
    if (nTris == 0) return MLBVH_OK;
    int rc = set_device(ctx);
    if (rc) return rc;
    k_gather<<<(unsigned)((nTris + kDiffBlock - 1) / kDiffBlock), kDiffBlock, 0, (hipStream_t)stream>>>(
        d_vertices, d_indices, nTris, (Tri*)d_triangles);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? MLBVH_OK : hip_status(e, "gather launch");
}

extern "C" int mlbvh_render_primary(mlbvh_ctx* ctx, float* d_out, uint32_t* d_hitTri, int width, int height,
                                    const mlbvh_raygen* camera, const float lightDir[3], const void* d_root,
                                    const void* d_nodes, const void* d_triangles, void* stream)
{
    if (!ctx || !d_out || !d_hitTri || !camera || !lightDir || !d_root || !d_nodes || !d_triangles || width <= 0 ||
        height <= 0) {
        set_error("render_primary: invalid arguments");
        return MLBVH_E_INVALID_ARG;
    }
    int rc = set_device(ctx);
    if (rc) return rc;
    dim3 g((unsigned)((width + 15) / 16), (unsigned)((height + 15) / 16), 1);
    k_render_primary<<<g, kDiffBlock, 0, (hipStream_t)stream>>>(
        (float2*)d_out, d_hitTri, width, height, *camera, mk(lightDir[0], lightDir[1], lightDir[2]),
        (const uint32_t*)d_root, (const Node*)d_nodes, (const Tri*)d_triangles);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? MLBVH_OK : hip_status(e, "render_primary launch");
}

extern "C" int mlbvh_primary_adjoint(mlbvh_ctx* ctx, const float* d_gradOut, const uint32_t* d_hitTri, int width,
                                     int height, const mlbvh_raygen* camera, const float lightDir[3],
                                     const float* d_vertices, const uint32_t* d_indices, float* d_gradVertices,
                                     void* stream)
{
    if (!ctx || !d_gradOut || !d_hitTri || !camera || !lightDir || !d_vertices || !d_indices || !d_gradVertices ||
        width <= 0 || height <= 0) {
        set_error("primary_adjoint: invalid arguments");
        return MLBVH_E_INVALID_ARG;
    }
    int rc = set_device(ctx);
    if (rc) return rc;
    dim3 g((unsigned)((width + 15) / 16), (unsigned)((height + 15) / 16), 1);
    k_primary_adjoint<<<g, kDiffBlock, 0, (hipStream_t)stream>>>(
        (const float2*)d_gradOut, d_hitTri, width, height, *camera, mk(lightDir[0], lightDir[1], lightDir[2]),
        d_vertices, d_indices, d_gradVertices);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? MLBVH_OK : hip_status(e, "primary_adjoint launch");
}
