// A reference-style application written against the C++ facade only
// (include/minimum_lbvh.h, tinyhiponesweep.h, camera.h): the flow of
// main_gpu.cpp (OnesweepSort + BVHGPUBuilder + a user kernel that calls
// intersect_stackfree), the sort self-test that main_gpu.cpp:104-145 keeps
// commented out, the reference's own build kernel shape (one thread per leaf
// calling build_lbvh, minimum_lbvh.cu:104-120) on the library's sorted
// records and deltas, and a user kernel that runs intersect_stackfree
// (closest and any hit), intersect_stackfull and intersect_escape_link on a
// ray file.  tests/test_gpu_facade.py runs it and checks its outputs against
// the CPU oracle.
//
// usage: facade_demo <tris.bin> <out_prefix> <w> <h> ex ey ez cx cy cz ux uy uz fovy_rad <rays.bin>
// rays.bin: 8 floats per ray (ox, oy, oz, tmax, dx, dy, dz, pad).
#include <stdio.h>
#include <stdlib.h>

#include <random>
#include <vector>

#include "camera.h"
#include "minimum_lbvh.h"
#include "tinyhiponesweep.h"

using namespace minimum_lbvh;

#define CHECK(x)                                                              \
    do {                                                                      \
        hipError_t e_ = (x);                                                  \
        if (e_ != hipSuccess) {                                               \
            fprintf(stderr, "%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
            return 3;                                                         \
        }                                                                     \
    } while (0)

__device__ uint32_t packRGBA(float r, float g, float b)
{
    int x = iclamp(f2i(r * 255.0f + 0.5f), 0, 255);
    int y = iclamp(f2i(g * 255.0f + 0.5f), 0, 255);
    int z = iclamp(f2i(b * 255.0f + 0.5f), 0, 255);
    return (z << 16) | (y << 8) | x;
}

// main_gpu.cu:84-109 written against the facade
__global__ void normal(uint32_t* pixels, int2 imageSize, RayGenerator rayGenerator, const NodeIndex* rootNode,
                       const InternalNode* internals, const Triangle* triangles)
{
    int xi = threadIdx.x + blockDim.x * blockIdx.x;
    int yi = threadIdx.y + blockDim.y * blockIdx.y;
    if (xi < imageSize.x && yi < imageSize.y) {
        int pixel = xi + yi * imageSize.x;
        float3 ro, rd;
        rayGenerator.shoot(&ro, &rd, (float)xi / imageSize.x, (float)yi / imageSize.y);
        Hit hit;
        intersect_stackfree(&hit, internals, triangles, *rootNode, ro, rd, invRd(rd));
        if (hit.t != MINIMUM_LBVH_FLT_MAX) {
            float3 n = normalize(hit.ng);
            float3 c = mul(add(n, v3(1.0f)), 0.5f);
            pixels[pixel] = packRGBA(c.x, c.y, c.z);
        } else {
            pixels[pixel] = packRGBA(0, 0, 0);
        }
    }
}

// The reference's build kernel (minimum_lbvh.cu:104-120): one thread per
// sorted leaf; internals[].context must be 0xFFFFFFFF beforehand.
__global__ void build(NodeIndex* rootNode, InternalNode* internals, const Triangle* triangles, int nTriangles,
                      const uint8_t* deltas, const IndexedMorton* indexedMortons)
{
    int i_leaf = blockIdx.x * blockDim.x + threadIdx.x;
    if (nTriangles <= i_leaf) return;
    build_lbvh(rootNode, internals, triangles, nTriangles, deltas, i_leaf, indexedMortons[i_leaf].index);
}

__global__ void init_internals(InternalNode* internals, int nInternals)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nInternals) {
        InternalNode node = {};
        node.context = 0xFFFFFFFF;
        internals[i] = node;
    }
}

constexpr int kStackDepth = 64;  // the host checks the tree depth before launching

// Per ray: intersect_stackfree closest and any hit, intersect_stackfull,
// intersect_escape_link, each from Hit.t = tmax.
__global__ void trace_all(const InternalNode* internals, const Triangle* triangles, const NodeIndex* rootNode,
                          const float* rays, int nRays, Hit* out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (nRays <= i) return;
    const float* r = rays + 8 * (size_t)i;
    float3 ro = make_float3(r[0], r[1], r[2]);
    float3 rd = make_float3(r[4], r[5], r[6]);
    float3 inv = invRd(rd);
    Hit h[4];
    for (int k = 0; k < 4; k++) h[k].t = r[3];
    NodeIndex stack[kStackDepth];
    intersect_stackfree(&h[0], internals, triangles, *rootNode, ro, rd, inv, RAY_QUERY_CLOSEST);
    intersect_stackfree(&h[1], internals, triangles, *rootNode, ro, rd, inv, RAY_QUERY_ANY);
    intersect_stackfull(&h[2], internals, triangles, *rootNode, ro, rd, inv, stack);
    intersect_escape_link(&h[3], internals, triangles, *rootNode, ro, rd, inv, RAY_QUERY_CLOSEST);
    for (int k = 0; k < 4; k++) out[4 * (size_t)i + k] = h[k];
}

// Depth of the tree in internal nodes along the deepest path (host).
static int tree_depth(const std::vector<InternalNode>& nodes, NodeIndex root)
{
    if (root.m_isLeaf) return 0;
    std::vector<std::pair<uint32_t, int>> st{{(uint32_t)root.m_index, 1}};
    int depth = 0;
    while (!st.empty()) {
        auto [i, d] = st.back();
        st.pop_back();
        depth = d > depth ? d : depth;
        for (int c = 0; c < 2; c++)
            if (!nodes[i].children[c].m_isLeaf) st.push_back({(uint32_t)nodes[i].children[c].m_index, d + 1});
    }
    return depth;
}

static bool write_file(const char* path, const void* p, size_t bytes)
{
    FILE* f = fopen(path, "wb");
    if (!f) return false;
    bool ok = fwrite(p, 1, bytes, f) == bytes;
    fclose(f);
    return ok;
}

int main(int argc, char** argv)
{
    if (argc != 16) {
        fprintf(stderr, "usage: %s tris.bin out_prefix w h eye3 center3 up3 fovy rays.bin\n", argv[0]);
        return 2;
    }
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    std::vector<Triangle> tris;
    Triangle t;
    while (fread(&t, sizeof(Triangle), 1, f) == 1) tris.push_back(t);
    fclose(f);
    int n = (int)tris.size();
    int w = atoi(argv[3]), h = atoi(argv[4]);
    float a[10];
    for (int i = 0; i < 10; i++) a[i] = (float)atof(argv[5 + i]);

    tinyhiponesweep::OnesweepSort onesweep(0);
    BVHGPUBuilder gpuBuilder("minimum_lbvh.cu", "./");
    Triangle* d_tris;
    CHECK(hipMalloc(&d_tris, sizeof(Triangle) * n));
    CHECK(hipMemcpy(d_tris, tris.data(), sizeof(Triangle) * n, hipMemcpyHostToDevice));
    gpuBuilder.build(d_tris, n, 0, onesweep, 0);

    RayGenerator rg;
    rg.lookat(make_float3(a[0], a[1], a[2]), make_float3(a[3], a[4], a[5]), make_float3(a[6], a[7], a[8]), a[9], w,
              h);
    uint32_t* d_px;
    CHECK(hipMalloc(&d_px, sizeof(uint32_t) * w * h));
    dim3 grid((w + 15) / 16, (h + 15) / 16), block(16, 16);
    normal<<<grid, block>>>(d_px, int2{w, h}, rg, gpuBuilder.m_rootNode, gpuBuilder.m_internals, d_tris);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());

    std::vector<InternalNode> nodes(n - 1);
    NodeIndex root;
    std::vector<uint32_t> px(w * h);
    CHECK(hipMemcpy(nodes.data(), gpuBuilder.m_internals, sizeof(InternalNode) * (n - 1), hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(&root, gpuBuilder.m_rootNode, sizeof(NodeIndex), hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(px.data(), d_px, sizeof(uint32_t) * w * h, hipMemcpyDeviceToHost));
    char path[1024];
    snprintf(path, sizeof(path), "%s_nodes.bin", argv[2]);
    if (!write_file(path, nodes.data(), sizeof(InternalNode) * (n - 1))) return 4;
    snprintf(path, sizeof(path), "%s_root.bin", argv[2]);
    if (!write_file(path, &root, sizeof(root))) return 4;
    snprintf(path, sizeof(path), "%s_pixels.bin", argv[2]);
    if (!write_file(path, px.data(), sizeof(uint32_t) * w * h)) return 4;

    // The reference's build kernel through the facade's build_lbvh, on the
    // library's sorted records and deltas (mlbvh_build_ex).
    {
        uint64_t* d_sorted;
        uint8_t* d_deltas;
        InternalNode *d_tmpNodes, *d_nodes2;
        NodeIndex *d_tmpRoot, *d_root2;
        CHECK(hipMalloc(&d_sorted, 8 * (size_t)n));
        CHECK(hipMalloc(&d_deltas, (size_t)n));
        CHECK(hipMalloc(&d_tmpNodes, sizeof(InternalNode) * (n - 1)));
        CHECK(hipMalloc(&d_nodes2, sizeof(InternalNode) * (n - 1)));
        CHECK(hipMalloc(&d_tmpRoot, sizeof(NodeIndex)));
        CHECK(hipMalloc(&d_root2, sizeof(NodeIndex)));
        if (mlbvh_build_ex(onesweep.context(), d_tris, n, 0, d_tmpNodes, d_tmpRoot, d_sorted, d_deltas, nullptr) !=
            MLBVH_OK) {
            fprintf(stderr, "mlbvh_build_ex: %s\n", mlbvh_last_error());
            return 3;
        }
        init_internals<<<(n - 1 + 255) / 256, 256>>>(d_nodes2, n - 1);
        build<<<(n + 255) / 256, 256>>>(d_root2, d_nodes2, d_tris, n, d_deltas, (const IndexedMorton*)d_sorted);
        CHECK(hipGetLastError());
        CHECK(hipDeviceSynchronize());
        std::vector<InternalNode> nodes2(n - 1);
        NodeIndex root2;
        CHECK(hipMemcpy(nodes2.data(), d_nodes2, sizeof(InternalNode) * (n - 1), hipMemcpyDeviceToHost));
        CHECK(hipMemcpy(&root2, d_root2, sizeof(NodeIndex), hipMemcpyDeviceToHost));
        snprintf(path, sizeof(path), "%s_leafbuild_nodes.bin", argv[2]);
        if (!write_file(path, nodes2.data(), sizeof(InternalNode) * (n - 1))) return 4;
        snprintf(path, sizeof(path), "%s_leafbuild_root.bin", argv[2]);
        if (!write_file(path, &root2, sizeof(root2))) return 4;
        CHECK(hipFree(d_sorted));
        CHECK(hipFree(d_deltas));
        CHECK(hipFree(d_tmpNodes));
        CHECK(hipFree(d_nodes2));
        CHECK(hipFree(d_tmpRoot));
        CHECK(hipFree(d_root2));
    }

    // The three traversals in a user kernel.
    {
        std::vector<float> rays;
        FILE* fr = fopen(argv[15], "rb");
        if (!fr) return 2;
        float v;
        while (fread(&v, sizeof(float), 1, fr) == 1) rays.push_back(v);
        fclose(fr);
        int nRays = (int)(rays.size() / 8);
        int depth = tree_depth(nodes, root);
        if (depth >= kStackDepth) {
            fprintf(stderr, "tree depth %d exceeds the stackfull kernel's %d-entry stack\n", depth, kStackDepth);
            return 5;
        }
        float* d_rays;
        Hit* d_hits;
        CHECK(hipMalloc(&d_rays, sizeof(float) * rays.size()));
        CHECK(hipMalloc(&d_hits, sizeof(Hit) * 4 * (size_t)nRays));
        CHECK(hipMemcpy(d_rays, rays.data(), sizeof(float) * rays.size(), hipMemcpyHostToDevice));
        trace_all<<<(nRays + 255) / 256, 256>>>(gpuBuilder.m_internals, d_tris, gpuBuilder.m_rootNode, d_rays, nRays,
                                                d_hits);
        CHECK(hipGetLastError());
        CHECK(hipDeviceSynchronize());
        std::vector<Hit> hits(4 * (size_t)nRays);
        CHECK(hipMemcpy(hits.data(), d_hits, sizeof(Hit) * hits.size(), hipMemcpyDeviceToHost));
        snprintf(path, sizeof(path), "%s_hits.bin", argv[2]);
        if (!write_file(path, hits.data(), sizeof(Hit) * hits.size())) return 4;
        printf("traced %d rays through 3 walks (tree depth %d)\n", nRays, depth);
        CHECK(hipFree(d_rays));
        CHECK(hipFree(d_hits));
    }

    // sort self-test of main_gpu.cpp:104-145: keys with indices, 64-bit
    const int N = (1 << 20) + 3;
    std::mt19937_64 rng(7);
    std::vector<uint64_t> xs(N);
    std::vector<uint32_t> idx(N);
    for (int i = 0; i < N; i++) {
        xs[i] = rng();
        idx[i] = i;
    }
    uint64_t *d_k, *d_kt;
    uint32_t *d_v, *d_vt;
    CHECK(hipMalloc(&d_k, 8 * N));
    CHECK(hipMalloc(&d_kt, 8 * N));
    CHECK(hipMalloc(&d_v, 4 * N));
    CHECK(hipMalloc(&d_vt, 4 * N));
    CHECK(hipMemcpy(d_k, xs.data(), 8 * N, hipMemcpyHostToDevice));
    CHECK(hipMemcpy(d_v, idx.data(), 4 * N, hipMemcpyHostToDevice));
    DeviceStopwatch sw(0);
    sw.start();
    onesweep.sort({d_k, d_v}, {d_kt, d_vt}, N, 0, 64, 0);
    sw.stop();
    float ms = sw.getElapsedMs();
    std::vector<uint64_t> sk(N);
    std::vector<uint32_t> si(N);
    CHECK(hipMemcpy(sk.data(), d_k, 8 * N, hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(si.data(), d_v, 4 * N, hipMemcpyDeviceToHost));
    int bad = 0;
    for (int i = 0; i < N; i++) {
        if (sk[i] != xs[si[i]]) bad++;
        if (i + 1 < N && sk[i] > sk[i + 1]) bad++;
    }
    printf("sort self-test N=%d: %s (%.3f ms)\n", N, bad ? "FAIL" : "OK", ms);
    printf("root %u leaf %u\n", (unsigned)root.m_index, (unsigned)root.m_isLeaf);
    CHECK(hipFree(d_k));
    CHECK(hipFree(d_kt));
    CHECK(hipFree(d_v));
    CHECK(hipFree(d_vt));
    CHECK(hipFree(d_px));
    CHECK(hipFree(d_tris));
    return bad ? 1 : 0;
}
