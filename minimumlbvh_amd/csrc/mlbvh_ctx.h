// Per-device context: scratch owned by the sorter and the builder.
#pragma once

#include "mlbvh_device.h"

struct mlbvh_ctx {
    int device = 0;
    int cuCount = 256;

    // --- radix sort state -------------------------------------------------
    uint64_t* status = nullptr;  // decoupled look-back words, [tiles][256]
    size_t statusWords = 0;
    uint32_t* hist = nullptr;     // digit histograms, [8][256]
    uint32_t* tickets = nullptr;  // per-pass tile tickets, [16]
    uint32_t epoch = 0;           // pass tag stored in every look-back word

    // --- build scratch ------------------------------------------------------
    uint64_t* recA = nullptr;  // IndexedMorton records (ping)
    size_t recCap = 0;
    uint64_t* recB = nullptr;  // (pong)
    size_t recBCap = 0;
    uint8_t* deltas = nullptr;
    size_t deltasCap = 0;
    uint32_t* flags = nullptr;  // per-internal-node arrival word of the climb
    size_t flagsCap = 0;
    uint32_t* sceneBox = nullptr;  // 6 order-mapped uint32 (lo xyz, hi xyz)
};

namespace mlbvh {

// Grow a device buffer to hold at least `need` elements (contents dropped).
template <class T>
int ensure(T** p, size_t* cap, size_t need, const char* what)
{
    if (need <= *cap && *p) return MLBVH_OK;
    if (*p) {
        hipError_t e = hipFree(*p);
        *p = nullptr;
        *cap = 0;
        if (e != hipSuccess) return hip_status(e, what);
    }
    size_t n = need < 1 ? 1 : need;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return hip_status(e, what);
    }
    *cap = n;
    return MLBVH_OK;
}

int set_device(mlbvh_ctx* ctx);

// Radix sort passes shared by mlbvh_sort_* and the builder.  When histReady
// is true the caller has already filled ctx->hist[0..passes) and zeroed
// ctx->tickets (the builder fuses both into its Morton kernel).
int sort_records_u64(mlbvh_ctx* ctx, uint64_t* keys, uint64_t* tmp, uint32_t n, int startBit, int endBit,
                     bool histReady, hipStream_t stream);

}  // namespace mlbvh
