// Host-side use of the C++ facade (include/minimum_lbvh.h), no GPU needed:
// BVHCPUBuilder::build / validate, AABB::encodeMortonCode, hashPCG, and the
// three traversals (intersect_stackfree, intersect_stackfull,
// intersect_escape_link) called on the host.  tests/test_facade_cpu.py runs
// it and compares every output with the CPU oracle.
//
// usage: facade_cpu build <tris.bin> <rays.bin> <out_prefix> <buildOption>
//        facade_cpu corrupt <tris.bin>        (validate() must abort)
//        facade_cpu misc <points.bin> <out_prefix>
// rays.bin: 8 floats per ray (ox, oy, oz, tmax, dx, dy, dz, pad).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "minimum_lbvh.h"

using namespace minimum_lbvh;

template <class T>
static std::vector<T> read_all(const char* path)
{
    std::vector<T> v;
    FILE* f = fopen(path, "rb");
    if (!f) return v;
    T x;
    while (fread(&x, sizeof(T), 1, f) == 1) v.push_back(x);
    fclose(f);
    return v;
}

static bool write_file(const char* prefix, const char* suffix, const void* p, size_t bytes)
{
    char path[1024];
    snprintf(path, sizeof(path), "%s_%s", prefix, suffix);
    FILE* f = fopen(path, "wb");
    if (!f) return false;
    bool ok = fwrite(p, 1, bytes, f) == bytes;
    fclose(f);
    return ok;
}

struct Ray {
    float o[3], tmax, d[3], pad;
};

int main(int argc, char** argv)
{
    if (argc >= 3 && !strcmp(argv[1], "corrupt")) {
        std::vector<Triangle> tris = read_all<Triangle>(argv[2]);
        BVHCPUBuilder b;
        b.build(tris.data(), (int)tris.size(), 0);
        b.validate();
        // validate_lbvh requires delta(child) <= delta(parent): give an
        // internal child of the root a delta above the root's
        const InternalNode& r = b.m_internals[b.m_rootNode.m_index];
        NodeIndex child = r.children[0].m_isLeaf ? r.children[1] : r.children[0];
        if (child.m_isLeaf) return 5;
        b.m_deltas[child.m_index] = (uint8_t)(b.m_deltas[b.m_rootNode.m_index] + 1);
        b.validate();  // aborts
        printf("validate did not abort\n");
        return 0;
    }
    if (argc >= 4 && !strcmp(argv[1], "misc")) {
        std::vector<float3> pts = read_all<float3>(argv[2]);
        AABB box;
        box.setEmpty();
        for (const float3& p : pts) box.extend(p);
        std::vector<uint64_t> codes(pts.size());
        for (size_t i = 0; i < pts.size(); i++) codes[i] = box.encodeMortonCode(pts[i]);
        std::vector<uint32_t> hashes(4096);
        for (uint32_t i = 0; i < hashes.size(); i++) hashes[i] = hashPCG(i * 2654435761u);
        float boxOut[6] = {box.lower.x, box.lower.y, box.lower.z, box.upper.x, box.upper.y, box.upper.z};
        bool ok = write_file(argv[3], "box.bin", boxOut, sizeof(boxOut)) &&
                  write_file(argv[3], "codes.bin", codes.data(), codes.size() * 8) &&
                  write_file(argv[3], "hashes.bin", hashes.data(), hashes.size() * 4);
        return ok ? 0 : 4;
    }
    if (argc != 6 || strcmp(argv[1], "build")) {
        fprintf(stderr, "usage: %s build tris.bin rays.bin out_prefix buildOption | corrupt tris.bin | misc pts.bin out\n",
                argv[0]);
        return 2;
    }
    std::vector<Triangle> tris = read_all<Triangle>(argv[2]);
    std::vector<Ray> rays = read_all<Ray>(argv[3]);
    const char* out = argv[4];
    uint32_t opt = (uint32_t)atoi(argv[5]);
    int n = (int)tris.size();

    BVHCPUBuilder b;
    b.build(tris.data(), n, opt);
    if (b.empty()) return 3;
    b.validate();
    if (!write_file(out, "nodes.bin", b.m_internals.data(), b.m_internals.size() * sizeof(InternalNode)) ||
        !write_file(out, "root.bin", &b.m_rootNode, sizeof(NodeIndex)) ||
        !write_file(out, "deltas.bin", b.m_deltas.data(), b.m_deltas.size()))
        return 4;

    // hits: [stackfree closest, stackfree any, stackfull closest, escape link closest] per ray
    std::vector<Hit> hits(rays.size() * 4);
    std::vector<NodeIndex> stack((size_t)n + 1);  // a chain of equal keys is n - 1 deep
    for (size_t r = 0; r < rays.size(); r++) {
        float3 ro = make_float3(rays[r].o[0], rays[r].o[1], rays[r].o[2]);
        float3 rd = make_float3(rays[r].d[0], rays[r].d[1], rays[r].d[2]);
        float3 inv = invRd(rd);
        Hit* h = &hits[4 * r];
        for (int k = 0; k < 4; k++) h[k].t = rays[r].tmax;
        intersect_stackfree(&h[0], b.m_internals.data(), tris.data(), b.m_rootNode, ro, rd, inv, RAY_QUERY_CLOSEST);
        intersect_stackfree(&h[1], b.m_internals.data(), tris.data(), b.m_rootNode, ro, rd, inv, RAY_QUERY_ANY);
        intersect_stackfull(&h[2], b.m_internals.data(), tris.data(), b.m_rootNode, ro, rd, inv, stack.data());
        intersect_escape_link(&h[3], b.m_internals.data(), tris.data(), b.m_rootNode, ro, rd, inv,
                              RAY_QUERY_CLOSEST);
    }
    return write_file(out, "hits.bin", hits.data(), hits.size() * sizeof(Hit)) ? 0 : 4;
}
