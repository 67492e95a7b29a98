// Calibration probes for the roofline of the traversal kernels (built into
// _lib/libmlbvh_calib.so, used by the bench and the profiling workloads; not
// part of the library's C ABI).
//
// mlbvh_calib_gather: every lane fetches `iters` 64-byte nodes, each as four
// 16-byte loads like the traversal's node fetch, at pseudo-random node indices
// in a table of `nodes` nodes (a power of two).
//   dependent = 0: the next index is a hash of the lane and step only, so a
//                  lane keeps many fetches in flight: the gather throughput
//                  of the memory level the table lives in;
//   dependent = 1: the next index also hashes the fetched data, so each lane
//                  has one fetch in flight at a time: the latency-bound rate
//                  of dependent fetches at the launch's occupancy;
//   dependent = 2: as 0, but the 4 lanes of a quad load the 4 quarters of one
//                  node per instruction (quad-cooperative fetch).
// Each lane writes the xor of what it read, so no load is dead.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

__device__ inline uint32_t mix(uint32_t v)
{
    // PCG hash (Jarzynski and Olano 2020)
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28) + 4)) ^ state) * 277803737u;
    return (word >> 22) ^ word;
}

// kMode 0: each lane fetches its own node (next index from lane and step);
// 1: the same with the next index hashed from the fetched data (dependent);
// 2: quad-cooperative: per node fetch, 4 rounds in which the 4 lanes of a
//    quad load the 4 quarters of one quad member's node (same four 16-byte
//    loads per lane per fetch, but one node per quad and instruction).
template <int kMode>
__global__ void __launch_bounds__(256) k_gather(const uint4* __restrict__ table, uint32_t mask, uint32_t iters,
                                                uint32_t* __restrict__ out)
{
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, quadBase = lane & ~3, quarter = lane & 3;
    uint32_t idx = mix(tid) & mask;
    uint32_t acc = 0;
    for (uint32_t it = 0; it < iters; it++) {
        uint32_t v;
        if (kMode == 2) {
            uint4 q[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const uint32_t node = (uint32_t)__shfl((int)idx, quadBase + r, 64);
                q[r] = table[4 * (size_t)node + quarter];
            }
            v = q[0].x ^ q[1].y ^ q[2].z ^ q[3].w;
        } else {
            const uint4* p = table + 4 * (size_t)idx;
            const uint4 a = p[0], b = p[1], c = p[2], d = p[3];
            v = a.x ^ a.w ^ b.y ^ b.z ^ c.x ^ c.w ^ d.y ^ d.z;
        }
        acc ^= v;
        idx = kMode == 1 ? mix(idx ^ v ^ it) & mask : mix(tid * 0x9E3779B9u + it) & mask;
    }
    out[tid] = acc;
}

}  // namespace

// grid x 256 lanes; `out` holds grid * 256 words; table holds `nodes` x 64 B.
extern "C" __attribute__((visibility("default"))) int mlbvh_calib_gather(const void* table, uint32_t nodes,
                                                                        uint32_t iters, int dependent, uint32_t* out,
                                                                        int grid, void* stream)
{
    if (!table || !out || nodes == 0 || (nodes & (nodes - 1)) != 0 || grid < 1) return -1;
    if (dependent == 1)
        k_gather<1><<<grid, 256, 0, (hipStream_t)stream>>>((const uint4*)table, nodes - 1, iters, out);
    else if (dependent == 2)
        k_gather<2><<<grid, 256, 0, (hipStream_t)stream>>>((const uint4*)table, nodes - 1, iters, out);
    else
        k_gather<0><<<grid, 256, 0, (hipStream_t)stream>>>((const uint4*)table, nodes - 1, iters, out);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}
