// Onesweep LSD radix sort for gfx950 (wave64).
//
// Semantics follow OnesweepSort::sort (tinyhiponesweep.h:592-668): a stable
// ascending sort by 8-bit digits starting at startBit, ceil((endBit -
// startBit) / 8) passes, result left in the caller's keys/values.
//
// Structure (one upfront histogram, then one kernel per digit pass):
//   k_sort_init   zero the per-pass digit histograms and tile tickets
//   k_hist        all digit histograms of the whole input in one read
//   k_pass        per tile: wave64 ballot multisplit ranking, per-digit
//                 decoupled look-back over a [tile][digit] status table,
//                 LDS-staged local scatter, coalesced writes of digit runs
//
// Tiles are handed out by an atomic ticket, so a block only waits on tiles
// that are already running; look-back words carry a pass tag (epoch), so the
// status table is never cleared between passes.
#include <stdarg.h>
#include <stdio.h>

#include "mlbvh_ctx.h"

namespace mlbvh {

constexpr int kSortBlock = 256;  // 4 waves
constexpr int kSortWaves = kSortBlock / 64;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortBlock * kSortItems;  // 4096 keys
constexpr int kHistBlock = 256;

constexpr uint64_t kFlagA = 1;  // tile aggregate published
constexpr uint64_t kFlagP = 2;  // inclusive prefix published

__device__ inline uint64_t lb_pack(uint32_t epoch, uint64_t flag, uint32_t value)
{
    return ((uint64_t)epoch << 34) | (flag << 32) | (uint64_t)value;
}

__device__ inline uint32_t digit_of(uint32_t k, int shift) { return (k >> shift) & 0xFFu; }
__device__ inline uint32_t digit_of(uint64_t k, int shift) { return (uint32_t)(k >> shift) & 0xFFu; }

// Exclusive scan of one value per thread over a 256-thread block.
__device__ inline uint32_t block_excl_scan256(uint32_t v, uint32_t* s_tmp, uint32_t* total)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_tmp[wave] = x;
    __syncthreads();
    uint32_t waveBase = 0, sum = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; w++) {
        uint32_t s = s_tmp[w];
        if (w < wave) waveBase += s;
        sum += s;
    }
    __syncthreads();
    if (total) *total = sum;
    return waveBase + x - v;
}

__global__ void k_sort_init(uint32_t* hist, int nHist, uint32_t* tickets)
{
    for (int i = threadIdx.x; i < nHist * 256; i += blockDim.x) hist[i] = 0;
    if (threadIdx.x < 16) tickets[threadIdx.x] = 0;
}

template <class K>
__global__ void __launch_bounds__(kHistBlock) k_hist(const K* __restrict__ keys, uint32_t n, int startBit,
                                                     int nPass, uint32_t* __restrict__ hist)
{
    __shared__ uint32_t s_h[8][256];
    for (int p = 0; p < 8; p++) s_h[p][threadIdx.x] = 0;
    __syncthreads();
    const int keyBits = 8 * (int)sizeof(K);
    for (uint32_t i = blockIdx.x * kHistBlock + threadIdx.x; i < n; i += gridDim.x * kHistBlock) {
        K k = keys[i];
        for (int p = 0; p < nPass; p++) {
            int sh = startBit + 8 * p;
            uint32_t d = sh < keyBits ? digit_of(k, sh) : 0u;
            atomicAdd(&s_h[p][d], 1u);
        }
    }
    __syncthreads();
    for (int p = 0; p < nPass; p++) {
        uint32_t c = s_h[p][threadIdx.x];
        if (c) atomicAdd(&hist[p * 256 + threadIdx.x], c);
    }
}

template <class K, bool kVals>
__global__ void __launch_bounds__(kSortBlock)
    k_pass(const K* __restrict__ kin, K* __restrict__ kout, const uint32_t* __restrict__ vin,
           uint32_t* __restrict__ vout, uint32_t n, int shift, const uint32_t* __restrict__ passHist,
           uint64_t* status, uint32_t* ticket, uint32_t epoch)
{
    __shared__ K s_keys[kSortTile];
    __shared__ uint32_t s_vals[kVals ? kSortTile : 1];
    __shared__ uint32_t s_wcnt[kSortWaves][256];
    __shared__ uint32_t s_dstart[256];
    __shared__ uint32_t s_gbase[256];
    __shared__ uint32_t s_tmp[kSortWaves];
    __shared__ uint32_t s_tile;

    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;

    if (t == 0) s_tile = atomicAdd(ticket, 1u);
    for (int w = 0; w < kSortWaves; w++) s_wcnt[w][t] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t tileBase = (uint64_t)tile * kSortTile;
    const uint64_t waveBase = tileBase + (uint64_t)wave * 64 * kSortItems;

    K key[kSortItems];
    uint32_t val[kVals ? kSortItems : 1];
    uint32_t off[kSortItems];
#pragma unroll
    for (int k = 0; k < kSortItems; k++) {
        uint64_t idx = waveBase + (uint64_t)k * 64 + lane;
        if (idx < n) {
            key[k] = kin[idx];
            if (kVals) val[k] = vin[idx];
        } else {
            key[k] = 0;
            if (kVals) val[k] = 0;
        }
    }

    // Wave64 multisplit: lanes whose 8-bit digits agree are found with eight
    // ballots; a lane's rank is the number of peers below it.
    const uint64_t ltMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int k = 0; k < kSortItems; k++) {
        uint64_t idx = waveBase + (uint64_t)k * 64 + lane;
        bool valid = idx < n;
        uint32_t d = digit_of(key[k], shift);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            bool bit = (d >> b) & 1u;
            uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t rank = (uint32_t)__popcll(peers & ltMask);
        uint32_t cnt = (uint32_t)__popcll(peers);
        uint32_t prior = valid ? s_wcnt[wave][d] : 0u;
        off[k] = prior + rank;
        if (valid && rank == 0) s_wcnt[wave][d] = prior + cnt;
    }
    __syncthreads();

    // Per digit (thread t owns digit t): wave offsets inside the tile.
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; w++) {
        uint32_t c = s_wcnt[w][t];
        s_wcnt[w][t] = tot;
        tot += c;
    }
    uint32_t dstart = block_excl_scan256(tot, s_tmp, nullptr);
    uint32_t gstart = block_excl_scan256(passHist[t], s_tmp, nullptr);
    s_dstart[t] = dstart;

    // Decoupled look-back for digit t.
    uint64_t* my = status + (uint64_t)tile * 256 + t;
    uint32_t excl = 0;
    if (tile == 0) {
        __hip_atomic_store(my, lb_pack(epoch, kFlagP, tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        __hip_atomic_store(my, lb_pack(epoch, kFlagA, tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int64_t j = (int64_t)tile - 1;
        while (true) {
            const uint64_t* src = status + (uint64_t)j * 256 + t;
            uint64_t w;
            while (true) {
                w = __hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((uint32_t)(w >> 34) == epoch && ((w >> 32) & 3u) != 0) break;
                __builtin_amdgcn_s_sleep(1);
            }
            excl += (uint32_t)w;
            if (((w >> 32) & 3u) == kFlagP) break;
            j--;
        }
        __hip_atomic_store(my, lb_pack(epoch, kFlagP, excl + tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    s_gbase[t] = gstart + excl - dstart;
    __syncthreads();

    // Stable local scatter into LDS (digit-major, original order within digit).
#pragma unroll
    for (int k = 0; k < kSortItems; k++) {
        uint64_t idx = waveBase + (uint64_t)k * 64 + lane;
        if (idx < n) {
            uint32_t d = digit_of(key[k], shift);
            uint32_t pos = s_dstart[d] + s_wcnt[wave][d] + off[k];
            s_keys[pos] = key[k];
            if (kVals) s_vals[pos] = val[k];
        }
    }
    __syncthreads();

    const uint32_t tileCount = (uint32_t)((n - tileBase) < (uint64_t)kSortTile ? (n - tileBase) : kSortTile);
    for (uint32_t i = t; i < tileCount; i += kSortBlock) {
        K k = s_keys[i];
        uint32_t d = digit_of(k, shift);
        uint32_t dst = s_gbase[d] + i;
        kout[dst] = k;
        if (kVals) vout[dst] = s_vals[i];
    }
}

static int hist_grid(mlbvh_ctx* ctx, uint32_t n)
{
    uint32_t want = (n + kHistBlock * 16 - 1) / (kHistBlock * 16);
    uint32_t cap = (uint32_t)ctx->cuCount * 4;
    if (want < 1) want = 1;
    return (int)(want < cap ? want : cap);
}

static int bump_epoch(mlbvh_ctx* ctx, hipStream_t stream)
{
    // 30-bit pass tag; on wrap, clear the table so no stale word can match.
    if (ctx->epoch >= (1u << 30) - 2) {
        hipError_t e = hipMemsetAsync(ctx->status, 0, ctx->statusWords * sizeof(uint64_t), stream);
        if (e != hipSuccess) return hip_status(e, "hipMemsetAsync(status)");
        ctx->epoch = 0;
    }
    ctx->epoch++;
    return MLBVH_OK;
}

static int ensure_status(mlbvh_ctx* ctx, uint32_t tiles)
{
    size_t need = (size_t)tiles * 256;
    if (need <= ctx->statusWords && ctx->status) return MLBVH_OK;
    int rc = ensure(&ctx->status, &ctx->statusWords, need, "hipMalloc(look-back status)");
    if (rc) return rc;
    hipError_t e = hipMemset(ctx->status, 0, ctx->statusWords * sizeof(uint64_t));
    if (e != hipSuccess) return hip_status(e, "hipMemset(status)");
    ctx->epoch = 0;
    return MLBVH_OK;
}

template <class K>
static int sort_impl(mlbvh_ctx* ctx, K* keys, uint32_t* vals, K* ktmp, uint32_t* vtmp, uint32_t n, int startBit,
                     int endBit, bool histReady, hipStream_t stream)
{
    const int keyBits = 8 * (int)sizeof(K);
    if (startBit < 0 || endBit <= startBit) {
        set_error("sort: invalid bit range [%d, %d)", startBit, endBit);
        return MLBVH_E_INVALID_ARG;
    }
    if (n > 0 && (!keys || !ktmp || (vals && !vtmp))) {
        set_error("sort: null buffer");
        return MLBVH_E_INVALID_ARG;
    }
    int passes = (endBit - startBit + 7) / 8;
    // Passes whose digit lies entirely past the key width see all-zero
    // digits; a stable pass over equal digits is the identity, so skip them.
    int live = 0;
    for (int p = 0; p < passes; p++)
        if (startBit + 8 * p < keyBits) live++;
    if (n <= 1 || live == 0) return MLBVH_OK;
    if (live > 8) live = 8;

    int rc = set_device(ctx);
    if (rc) return rc;
    uint32_t tiles = (uint32_t)((n + (uint64_t)kSortTile - 1) / kSortTile);
    if ((rc = ensure_status(ctx, tiles))) return rc;

    if (!histReady) {
        k_sort_init<<<1, 256, 0, stream>>>(ctx->hist, live, ctx->tickets);
        k_hist<K><<<hist_grid(ctx, n), kHistBlock, 0, stream>>>(keys, n, startBit, live, ctx->hist);
    }
    K* src = keys;
    K* dst = ktmp;
    uint32_t* vsrc = vals;
    uint32_t* vdst = vtmp;
    for (int p = 0; p < live; p++) {
        if ((rc = bump_epoch(ctx, stream))) return rc;
        int shift = startBit + 8 * p;
        if (vals)
            k_pass<K, true><<<tiles, kSortBlock, 0, stream>>>(src, dst, vsrc, vdst, n, shift, ctx->hist + 256 * p,
                                                             ctx->status, ctx->tickets + p, ctx->epoch);
        else
            k_pass<K, false><<<tiles, kSortBlock, 0, stream>>>(src, dst, nullptr, nullptr, n, shift,
                                                              ctx->hist + 256 * p, ctx->status, ctx->tickets + p,
                                                              ctx->epoch);
        K* tk = src;
        src = dst;
        dst = tk;
        uint32_t* tv = vsrc;
        vsrc = vdst;
        vdst = tv;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_status(e, "sort kernel launch");
    if (src != keys) {
        e = hipMemcpyAsync(keys, src, sizeof(K) * (size_t)n, hipMemcpyDeviceToDevice, stream);
        if (e != hipSuccess) return hip_status(e, "hipMemcpyAsync(keys)");
        if (vals) {
            e = hipMemcpyAsync(vals, vsrc, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, stream);
            if (e != hipSuccess) return hip_status(e, "hipMemcpyAsync(values)");
        }
    }
    return MLBVH_OK;
}

int sort_records_u64(mlbvh_ctx* ctx, uint64_t* keys, uint64_t* tmp, uint32_t n, int startBit, int endBit,
                     bool histReady, hipStream_t stream)
{
    return sort_impl<uint64_t>(ctx, keys, nullptr, tmp, nullptr, n, startBit, endBit, histReady, stream);
}

}  // namespace mlbvh

using namespace mlbvh;

extern "C" int mlbvh_sort_u32(mlbvh_ctx* ctx, uint32_t* keys, uint32_t* values, uint32_t* keys_tmp,
                              uint32_t* values_tmp, uint32_t n, int startBit, int endBit, void* stream)
{
    if (!ctx) {
        set_error("sort: null context");
        return MLBVH_E_INVALID_ARG;
    }
    return sort_impl<uint32_t>(ctx, keys, values, keys_tmp, values_tmp, n, startBit, endBit, false,
                               (hipStream_t)stream);
}

extern "C" int mlbvh_sort_u64(mlbvh_ctx* ctx, uint64_t* keys, uint32_t* values, uint64_t* keys_tmp,
                              uint32_t* values_tmp, uint32_t n, int startBit, int endBit, void* stream)
{
    if (!ctx) {
        set_error("sort: null context");
        return MLBVH_E_INVALID_ARG;
    }
    return sort_impl<uint64_t>(ctx, keys, values, keys_tmp, values_tmp, n, startBit, endBit, false,
                               (hipStream_t)stream);
}
