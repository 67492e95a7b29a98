// tinyhiponesweep.h -- C++ facade of the gfx950 Onesweep radix sort.
//
// Keeps the reference class tinyhiponesweep::OnesweepSort
// (tinyhiponesweep.h:562-748 of Ushio/MinimumLBVH): the KeyValueSoA32/64
// argument structs and sort(elements, tmp, n, startBit, endBit, stream).  The
// kernels are compiled ahead of time into libmlbvh.so and launched through
// the C ABI (include/mlbvh.h); the constructor takes a HIP device ordinal
// instead of an oroDevice.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "mlbvh.h"

namespace tinyhiponesweep {

constexpr int N_RADIX = 8;
constexpr int BIN_SIZE = 1 << N_RADIX;

struct KeyValueSoA32 {
    uint32_t* key;
    uint32_t* value;
};
struct KeyValueSoA64 {
    uint64_t* key;
    uint32_t* value;
};

class OnesweepSort {
public:
    explicit OnesweepSort(int device = 0)
    {
        int rc = mlbvh_ctx_create(device, &m_ctx);
        if (rc != MLBVH_OK) fail("mlbvh_ctx_create");
    }
    ~OnesweepSort() { mlbvh_ctx_destroy(m_ctx); }
    OnesweepSort(const OnesweepSort&) = delete;
    void operator=(const OnesweepSort&) = delete;

    // Stable ascending sort by key bits [startBit, startBit + 8*ceil((endBit -
    // startBit)/8)); value == nullptr sorts keys only.  The result ends in
    // elementsToSort; tmp is clobbered.
    void sort(const KeyValueSoA32& elementsToSort, const KeyValueSoA32& tmp, uint32_t n, int startBit, int endBit,
              hipStream_t stream)
    {
        if (mlbvh_sort_u32(m_ctx, elementsToSort.key, elementsToSort.value, tmp.key, tmp.value, n, startBit, endBit,
                           stream) != MLBVH_OK)
            fail("mlbvh_sort_u32");
    }
    void sort(const KeyValueSoA64& elementsToSort, const KeyValueSoA64& tmp, uint32_t n, int startBit, int endBit,
              hipStream_t stream)
    {
        if (mlbvh_sort_u64(m_ctx, elementsToSort.key, elementsToSort.value, tmp.key, tmp.value, n, startBit, endBit,
                           stream) != MLBVH_OK)
            fail("mlbvh_sort_u64");
    }

    // The library context (shared with minimum_lbvh::BVHGPUBuilder::build).
    mlbvh_ctx* context() const { return m_ctx; }

private:
    static void fail(const char* what)
    {
        fprintf(stderr, "%s: %s\n", what, mlbvh_last_error());
        abort();
    }
    mlbvh_ctx* m_ctx = nullptr;
};

}  // namespace tinyhiponesweep
