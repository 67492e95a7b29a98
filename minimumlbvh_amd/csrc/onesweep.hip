// Onesweep LSD radix sort for gfx950 (wave64).
//
// Semantics follow OnesweepSort::sort (tinyhiponesweep.h:592-668): a stable
// ascending sort by 8-bit digits starting at startBit, ceil((endBit -
// startBit) / 8) passes, result left in the caller's keys/values.
//
// Structure (one upfront histogram, then one kernel per digit pass):
//   k_hist         all digit histograms of the input in one read, written as
//                  per-block partials (no global atomics), and the first
//                  pass's per-tile digit counts; its block 0 hands every pass
//                  a fresh look-back tag and zeroes the tickets
//   k_scan_reduce  partials -> per-pass histograms, and the first pass's
//                  per-tile prefixes from the counts (its tiles do not look back)
//   k_pass         per tile: wave64 ballot multisplit ranking, per-digit
//                  decoupled look-back over a [tile][digit] status table,
//                  LDS-staged local scatter, writes of contiguous digit runs
//
// Tiles of the look-back passes are handed out by an atomic ticket, so a
// block only waits on tiles that are already running; look-back words carry a pass tag, so the status
// table is never cleared between passes.  The tile shape (threads x keys per
// thread) is one of kSortCfgs, chosen per context (mlbvh_set_tuning).
//
// Large standalone sorts take a hybrid route with the same result: the
// three highest digit passes run as above (the array is then stably ordered
// by a 24-bit prefix), and k_segment_finish puts each run of equal prefix in
// order -- for uniform keys such a run holds a few keys.  A run longer than
// the finish kernel's halo sets a device flag that turns on the full pass
// sequence (guarded launches) from the three-pass state.
#include <stdarg.h>
#include <stdio.h>

#include "mlbvh_ctx.h"

namespace mlbvh {

struct SortCfg {
    int block, items;
};
constexpr SortCfg kSortCfgs[kNumSortCfgs] = {{256, 16}, {256, 32}, {512, 16}, {512, 24}, {1024, 16}};
constexpr int kHistBlock = 256;

#ifndef MLBVH_LOOKBACK_WINDOW
#define MLBVH_LOOKBACK_WINDOW 4
#endif
constexpr int kLookWindow = MLBVH_LOOKBACK_WINDOW;  // predecessors read per look-back round
constexpr uint64_t kFlagA = 1;  // tile aggregate published
constexpr uint64_t kFlagP = 2;  // inclusive prefix published

__device__ inline uint64_t lb_pack(uint32_t epoch, uint64_t flag, uint32_t value)
{
    return ((uint64_t)epoch << 34) | (flag << 32) | (uint64_t)value;
}

__device__ inline uint32_t digit_of(uint32_t k, int shift) { return (k >> shift) & 0xFFu; }
__device__ inline uint32_t digit_of(uint64_t k, int shift) { return (uint32_t)(k >> shift) & 0xFFu; }

// The lanes of the wave (among those with `valid`) whose 8-bit digit equals
// this lane's, as the two 32-bit halves of a lane mask: for each bit, the
// ballot of the lanes that have it set, or its complement where this lane has
// it clear.  nb is this lane's bit inverted and sign-extended (0 when set, ~0
// when clear), so an XOR picks the ballot or its complement; the XOR and the
// AND are one three-input bit operation per half (kAndXor: its truth table for
// x & (y ^ z), from the operand patterns 0xF0, 0xCC and 0xAA).
constexpr unsigned kAndXor = 0xF0u & (0xCCu ^ 0xAAu);
__device__ inline void digit_peers(uint32_t d, bool valid, uint32_t& lo, uint32_t& hi)
{
    const uint64_t v = __ballot(valid);
    lo = (uint32_t)v;
    hi = (uint32_t)(v >> 32);
    const uint32_t nd = ~d;
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const uint32_t nb = (uint32_t)__builtin_amdgcn_sbfe((int)nd, (unsigned)b, 1u);
        const uint64_t m = __ballot(nb == 0u);
        lo = __builtin_amdgcn_bitop3_b32(lo, (uint32_t)m, nb, kAndXor);
        hi = __builtin_amdgcn_bitop3_b32(hi, (uint32_t)(m >> 32), nb, kAndXor);
    }
}

// One LDS histogram add per wave when its active lanes share the digit (as
// whole runs of equal keys do), else one per lane.
__device__ inline void lds_count(uint32_t* h, uint32_t d)
{
    const uint64_t active = __ballot(1);
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
    if (__ballot(d == d0) == active) {
        if ((uint32_t)__lane_id() == (uint32_t)__ffsll((unsigned long long)active) - 1u)
            atomicAdd(&h[d0], (uint32_t)__popcll(active));
    } else {
        atomicAdd(&h[d], 1u);
    }
}

// Exclusive scan of one value per thread over threads 0..255 (threads past
// 255 pass 0 and get an unspecified result); every thread of the block calls.
__device__ inline uint32_t scan256(uint32_t v, uint32_t* s_tmp)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (wave < 4 && lane == 63) s_tmp[wave] = x;
    __syncthreads();
    uint32_t base = 0;
#pragma unroll
    for (int w = 0; w < 4; w++)
        if (w < wave) base += s_tmp[w];
    __syncthreads();
    return base + x - v;
}

// With tileCounts, blocks take whole sort tiles of tileKeys keys (tile-
// strided) and also write each tile's counts of the first digit,
// tileCounts[digit * nTiles + tile], for k_scan_reduce.  With init, block 0
// first sets up the sort's state (sort_state_init: init->passes pass slots,
// init->hists histograms), which no kernel before the next one reads.
struct SortInit {
    uint32_t* ctrl;
    int passes, hists;
    uint32_t* hist;
    uint64_t* status;
    size_t statusWords;
};
// kShared: count through lds_count (keys whose digits are often shared by a
// whole wave, such as Morton codes of a mesh; on uniform keys the check costs
// more than it saves).
template <class K, bool kShared = false>
__global__ void __launch_bounds__(kHistBlock) k_hist(const K* __restrict__ keys, uint32_t n, int startBit,
                                                     int nPass, uint32_t* __restrict__ partials,
                                                     const uint32_t* __restrict__ guard = nullptr,
                                                     uint32_t tileKeys = 0, uint32_t nTiles = 0,
                                                     uint32_t* __restrict__ tileCounts = nullptr,
                                                     SortInit init = SortInit{})
{
    if (guard && *guard == 0) return;
    if (init.ctrl && blockIdx.x == 0)
        sort_state_init(init.ctrl, init.passes, init.hist, init.hists, init.status, init.statusWords);
    __shared__ uint32_t s_h[8][256];
    for (int p = 0; p < 8; p++) s_h[p][threadIdx.x] = 0;
    const int keyBits = 8 * (int)sizeof(K);
    if (tileCounts) {
        __shared__ uint32_t s_t[256];  // this tile's counts of the first digit
        s_t[threadIdx.x] = 0;
        __syncthreads();
        auto add = [&](uint32_t* h, uint32_t d) {
            if (kShared)
                lds_count(h, d);
            else
                atomicAdd(&h[d], 1u);
        };
        auto count = [&](K k) {
            add(s_t, startBit < keyBits ? digit_of(k, startBit) : 0u);
            for (int p = 1; p < nPass; p++) {
                int sh = startBit + 8 * p;
                add(s_h[p], sh < keyBits ? digit_of(k, sh) : 0u);
            }
        };
        for (uint32_t tile = blockIdx.x; tile < nTiles; tile += gridDim.x) {
            const uint64_t t0 = (uint64_t)tile * tileKeys;
            const uint64_t t1 = t0 + tileKeys < (uint64_t)n ? t0 + tileKeys : (uint64_t)n;
            uint64_t i = t0 + threadIdx.x;
            for (; i + 3 * kHistBlock < t1; i += 4 * kHistBlock) {
                K k[4];
#pragma unroll
                for (int u = 0; u < 4; u++) k[u] = keys[i + u * kHistBlock];
#pragma unroll
                for (int u = 0; u < 4; u++) count(k[u]);
            }
            for (; i < t1; i += kHistBlock) count(keys[i]);
            __syncthreads();
            const uint32_t c = s_t[threadIdx.x];
            tileCounts[(size_t)threadIdx.x * nTiles + tile] = c;
            s_h[0][threadIdx.x] += c;
            s_t[threadIdx.x] = 0;
            __syncthreads();
        }
        for (int p = 0; p < nPass; p++)
            partials[((size_t)blockIdx.x * nPass + p) * 256 + threadIdx.x] = s_h[p][threadIdx.x];
        return;
    }
    __syncthreads();
    const uint32_t stride = gridDim.x * kHistBlock;
    uint32_t i = blockIdx.x * kHistBlock + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        K k[4];
#pragma unroll
        for (int u = 0; u < 4; u++) k[u] = keys[i + u * stride];
#pragma unroll
        for (int u = 0; u < 4; u++)
            for (int p = 0; p < nPass; p++) {
                int sh = startBit + 8 * p;
                atomicAdd(&s_h[p][sh < keyBits ? digit_of(k[u], sh) : 0u], 1u);
            }
    }
    for (; i < n; i += stride) {
        K k = keys[i];
        for (int p = 0; p < nPass; p++) {
            int sh = startBit + 8 * p;
            atomicAdd(&s_h[p][sh < keyBits ? digit_of(k, sh) : 0u], 1u);
        }
    }
    __syncthreads();
    for (int p = 0; p < nPass; p++) partials[((size_t)blockIdx.x * nPass + p) * 256 + threadIdx.x] = s_h[p][threadIdx.x];
}

// hist[p][d] += sum over blocks of partials[block][p][d].  Grid (nPass,
// chunks): each block sums kReduceChunk partial blocks with independent,
// unrolled loads and adds its sum atomically (hist starts zeroed).
constexpr int kReduceChunk = 32;
__global__ void __launch_bounds__(256) k_hist_reduce(const uint32_t* __restrict__ partials, int nBlocks, int nPass,
                                                     uint32_t* __restrict__ hist, const uint32_t* __restrict__ guard)
{
    if (guard && *guard == 0) return;
    const int p = blockIdx.x;
    const int b0 = blockIdx.y * kReduceChunk;
    const int b1 = min(b0 + kReduceChunk, nBlocks);
    uint32_t s = 0;
#pragma unroll 8
    for (int b = b0; b < b1; b++) s += partials[((size_t)b * nPass + p) * 256 + threadIdx.x];
    if (s) atomicAdd(&hist[p * 256 + threadIdx.x], s);
}

// Two independent jobs in one launch (a launch costs about 4.5 us even when
// it does little):
//  - blocks 0 .. 255 (with tileCounts): the first digit pass's inclusive
//    prefixes filled in advance from per-tile counts: status[d * nTiles +
//    tile] = (that pass's tag, prefix flag, count of digit d in tiles
//    0..tile).  Digit-major, so that each lane writes whole lines; the other
//    passes' look-back words ([tile][digit]) carry other tags, so sharing the
//    table is safe.  One block per digit.
//  - the blocks after them: hist[p][d] += partials of nBlocks blocks, as
//    k_hist_reduce (block r: pass r % nPass, chunk r / nPass).
__global__ void __launch_bounds__(256) k_scan_reduce(const uint32_t* __restrict__ tileCounts, uint32_t nTiles,
                                                     uint64_t* __restrict__ status, const uint32_t* __restrict__ ctrl,
                                                     int pass, const uint32_t* __restrict__ partials, int nBlocks,
                                                     int nPass, uint32_t* __restrict__ hist)
{
    const uint32_t scanBlocks = tileCounts ? 256u : 0u;
    if (blockIdx.x >= scanBlocks) {
        const int r = (int)(blockIdx.x - scanBlocks);
        const int p = r % nPass;
        const int b0 = (r / nPass) * kReduceChunk;
        const int b1 = min(b0 + kReduceChunk, nBlocks);
        uint32_t s = 0;
#pragma unroll 8
        for (int b = b0; b < b1; b++) s += partials[((size_t)b * nPass + p) * 256 + threadIdx.x];
        if (s) atomicAdd(&hist[p * 256 + threadIdx.x], s);
        return;
    }
    __shared__ uint32_t s_tmp[4];
    const uint32_t d = blockIdx.x;
    const uint32_t epoch = ctrl[kCtrlEpochs + pass];
    const uint32_t per = (nTiles + 255) / 256;
    const uint32_t a = threadIdx.x * per < nTiles ? threadIdx.x * per : nTiles;
    const uint32_t b = a + per < nTiles ? a + per : nTiles;
    const uint32_t* col = tileCounts + (size_t)d * nTiles;
    uint32_t sum = 0;
    for (uint32_t j = a; j < b; j++) sum += col[j];
    uint32_t run = scan256(sum, s_tmp);
    uint64_t* out = status + (size_t)d * nTiles;
    for (uint32_t j = a; j < b; j++) {
        run += col[j];
        out[j] = lb_pack(epoch, kFlagP, run);
    }
}

// k_scan_reduce for the first pass's tile counts (or none) and histBlocks
// rows of nPass partial histograms (or none) into hist.
static void launch_scan_reduce(mlbvh_ctx* ctx, const uint32_t* tileCounts, uint32_t tiles, int histBlocks, int nPass,
                               uint32_t* hist, hipStream_t stream)
{
    const unsigned reduceBlocks = histBlocks > 0 ? (unsigned)(nPass * ((histBlocks + kReduceChunk - 1) / kReduceChunk)) : 0u;
    const unsigned grid = (tileCounts ? 256u : 0u) + reduceBlocks;
    if (grid)
        k_scan_reduce<<<grid, 256, 0, stream>>>(tileCounts, tiles, ctx->status, ctx->ctrl, 0, ctx->partials,
                                                histBlocks, nPass > 0 ? nPass : 1, hist);
}

template <class K, bool kVals, int kBlock, int kItems>
__global__ void __launch_bounds__(kBlock)
    k_pass(const K* __restrict__ kin, K* __restrict__ kout, const uint32_t* __restrict__ vin,
           uint32_t* __restrict__ vout, uint32_t n, int shift, const uint32_t* __restrict__ passHist,
           uint64_t* status, uint32_t* ctrl, int pass, const uint32_t* __restrict__ guard, int preScanned)
{
    if (guard && *guard == 0) return;  // a fallback pass that is not needed
    constexpr int kTile = kBlock * kItems;
    constexpr int kWaves = kBlock / 64;
    __shared__ K s_keys[kTile];  // keys, then (as uint32) values
    __shared__ uint32_t s_wcnt[kWaves][256];
    __shared__ uint32_t s_dstart[256];
    __shared__ uint32_t s_gbase[256];
    __shared__ uint32_t s_tmp[4];
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_hist[256];  // this tile's digit counts

    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    uint32_t* ticket = ctrl + kCtrlTickets + pass;
    const uint32_t epoch = ctrl[kCtrlEpochs + pass];

    // A pre-scanned pass never waits on another tile, so its tiles need no
    // ticket (whose device-scope atomic every tile would wait for before its
    // loads): tile = block index.
    if (t == 0 && !preScanned) s_tile = atomicAdd(ticket, 1u);
    for (int i = t; i < kWaves * 256; i += kBlock) (&s_wcnt[0][0])[i] = 0;
    if (t < 256) s_hist[t] = 0;
    __syncthreads();
    const uint32_t tile = preScanned ? blockIdx.x : s_tile;
    const uint64_t tileBase = (uint64_t)tile * kTile;
    const uint64_t waveBase = tileBase + (uint64_t)wave * 64 * kItems;

    K key[kItems];
    uint32_t val[kVals ? kItems : 1];
    uint32_t off[kItems];
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        uint64_t idx = waveBase + (uint64_t)k * 64 + lane;
        if (idx < n) {
            key[k] = kin[idx];
            if (kVals) val[k] = vin[idx];
        } else {
            key[k] = 0;
            if (kVals) val[k] = 0;
        }
    }

    // The tile's aggregate goes out as soon as its keys are in: counted with
    // LDS atomics before the ranking, so that successors looking back find it
    // one ranking earlier.  Tile 0 publishes its inclusive prefix directly.
    if (!preScanned) {
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            uint64_t idx = waveBase + (uint64_t)k * 64 + lane;
            if (idx < n) atomicAdd(&s_hist[digit_of(key[k], shift)], 1u);
        }
        __syncthreads();
        if (t < 256)
            __hip_atomic_store(status + (uint64_t)tile * 256 + t, lb_pack(epoch, tile == 0 ? kFlagP : kFlagA, s_hist[t]),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    // Wave64 multisplit: lanes whose 8-bit digits agree are found with eight
    // ballots (digit_peers); a lane's rank is the number of peers below it.
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        uint64_t idx = waveBase + (uint64_t)k * 64 + lane;
        bool valid = idx < n;
        uint32_t d = digit_of(key[k], shift);
        uint32_t plo, phi;
        digit_peers(d, valid, plo, phi);
        uint32_t rank = __builtin_amdgcn_mbcnt_hi(phi, __builtin_amdgcn_mbcnt_lo(plo, 0u));
        uint32_t cnt = (uint32_t)(__popc(plo) + __popc(phi));
        uint32_t prior = valid ? s_wcnt[wave][d] : 0u;
        off[k] = prior + rank;
        if (valid && rank == 0) s_wcnt[wave][d] = prior + cnt;
    }
    __syncthreads();

    // Per digit (threads 0..255 own digit t): wave offsets inside the tile.
    uint32_t tot = 0;
    if (t < 256) {
#pragma unroll
        for (int w = 0; w < kWaves; w++) {
            uint32_t c = s_wcnt[w][t];
            s_wcnt[w][t] = tot;
            tot += c;
        }
    }
    uint32_t dstart = scan256(tot, s_tmp);
    uint32_t gstart = scan256(t < 256 ? passHist[t] : 0u, s_tmp);

    if (t < 256) {
        s_dstart[t] = dstart;
        // Decoupled look-back for digit t.
        uint64_t* my = status + (uint64_t)tile * 256 + t;
        uint32_t excl = 0;
        if (preScanned) {  // k_scan_reduce has published every tile's inclusive prefix
            if (tile != 0) excl = (uint32_t)status[(uint64_t)t * gridDim.x + tile - 1];
        } else if (tile != 0) {  // the aggregate (tot) went out before the ranking
            // kLookWindow predecessors per round, loaded together: the words
            // are summed from the nearest up to the first published prefix;
            // an unpublished word ends the round, which resumes there.  The
            // walk is what costs: at 2^25 keys a tile's predecessors have
            // nearly always published their aggregates by the time it looks
            // back, and the nearest prefix is 18-28 tiles back, so a window
            // divides the dependent round trips (one word per round: 8 us per
            // tile; 4 or 8 words: 6 us, the rounds slower under the added
            // traffic; 16 words: slower overall).
            int64_t j = (int64_t)tile - 1;
            while (true) {
                uint64_t w[kLookWindow];
#pragma unroll
                for (int q = 0; q < kLookWindow; q++)
                    w[q] = j - q >= 0 ? __hip_atomic_load(status + (uint64_t)(j - q) * 256 + t, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT)
                                      : lb_pack(epoch, kFlagP, 0u);
                int used = 0;
                bool done = false;
#pragma unroll
                for (int q = 0; q < kLookWindow; q++) {
                    const bool pub = (uint32_t)(w[q] >> 34) == epoch && ((w[q] >> 32) & 3u) != 0;
                    if (!done && used == q && pub) {
                        excl += (uint32_t)w[q];
                        used = q + 1;
                        done = ((w[q] >> 32) & 3u) == kFlagP;
                    }
                }
                if (done) break;
                j -= used;
                if (used < kLookWindow) __builtin_amdgcn_s_sleep(1);  // word j is not published yet
            }
            __hip_atomic_store(my, lb_pack(epoch, kFlagP, excl + tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        s_gbase[t] = gstart + excl - dstart;
    }
    __syncthreads();

    // Stable local scatter into LDS (digit-major, input order within digit).
    uint32_t pos[kItems];
#pragma unroll
    for (int k = 0; k < kItems; k++) {
        uint64_t idx = waveBase + (uint64_t)k * 64 + lane;
        uint32_t d = digit_of(key[k], shift);
        pos[k] = s_dstart[d] + s_wcnt[wave][d] + off[k];
        if (idx < n) s_keys[pos[k]] = key[k];
    }
    __syncthreads();

    const uint32_t tileCount = (uint32_t)((n - tileBase) < (uint64_t)kTile ? (n - tileBase) : kTile);
    uint32_t dst[kVals ? kItems : 1];
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        uint32_t i = (uint32_t)(t + j * kBlock);
        if (i < tileCount) {
            K k = s_keys[i];
            uint32_t o = s_gbase[digit_of(k, shift)] + i;
            kout[o] = k;
            if (kVals) dst[j] = o;
        }
    }
    if (kVals) {
        uint32_t* s_vals = reinterpret_cast<uint32_t*>(s_keys);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kItems; k++) {
            uint64_t idx = waveBase + (uint64_t)k * 64 + lane;
            if (idx < n) s_vals[pos[k]] = val[k];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kItems; j++) {
            uint32_t i = (uint32_t)(t + j * kBlock);
            if (i < tileCount) vout[dst[j]] = s_vals[i];
        }
    }
}


// ---------------------------------------------------------------------------
// Hybrid finish: after the top-digit passes (three, or four) the array is
// stably ordered by a 24-bit (32-bit) prefix, and for large uniform inputs
// (or spatially spread 63-bit Morton codes) the runs of equal prefix
// ("segments") hold a few keys each.  Every key finds its segment by looking
// at its neighbours' prefixes and its place in it by counting the segment's
// keys that order before it (sort bits compared, ties by position): the
// stable order, written out of place.
// ---------------------------------------------------------------------------
constexpr int kFinishBlock = 256;
#ifndef MLBVH_FINISH_ITEMS
#define MLBVH_FINISH_ITEMS 6
#endif
// keys per thread: at 2^25 u64 the finish took 132 us with 6, 137 with 4,
// 146 with 8 (fewer blocks per CU), 152 with 3 and 171 with 2 (more halo)
constexpr int kFinishItems = MLBVH_FINISH_ITEMS;
constexpr int kFinishChunk = kFinishBlock * kFinishItems;
constexpr int kFinishHalo = 64;                        // longest segment handled
constexpr int kFinishNear = 8;                         // segments up to 9 keys take the fixed-count path
constexpr uint32_t kHybridMinKeys = 1u << 18;

// The prefix of the top-digit passes: bits [shift, shift + 8 * top digits),
// as masked by prefixMask.
template <class K>
__device__ inline uint32_t prefix_of(K k, int shift, uint32_t prefixMask)
{
    return shift < (int)(8 * sizeof(K)) ? (uint32_t)(k >> shift) & prefixMask : 0u;
}

// The order of a key among keys sharing its prefix: its sort bits below the
// prefix, [startBit, prefixShift), moved down to bit 0 (at most 8 * (8 - 3) =
// 40 bits), above its window position (12 bits).  The word then stays below
// bit 52, where the segment start goes.
template <class K>
__device__ inline uint64_t order_word(K k, K lowMask, int startBit, int pos)
{
    return ((uint64_t)((k & lowMask) >> startBit) << 12) | (uint64_t)(uint32_t)pos;
}

// sortMask selects the key bits [startBit, startBit + 8 * passes) that order
// keys.  Thread t loads window positions t + 256u; those in the middle chunk
// are its keys to place.  The window is cut into rows of 64 consecutive
// positions, one wave's lanes each, so the segment starts and ends within a
// row are two ballots; each row posts its last start and first end, which
// give the bounds [a, e) of a segment that crosses rows.  Every window
// position gets the word (top << 52 | order_word) in LDS, where top is the
// segment's start a when the segment holds at most kFinishNear + 1 keys and
// kFinishFar otherwise.  A key of a short segment is ranked by counting,
// among the kFinishNear + 1 words from a, those below its own: keys of later
// segments have a later start or kFinishFar, so only the segment's own keys
// that order before it are counted, one 64-bit comparison each.  A key of a
// longer segment compares the words of [a, e).
constexpr uint64_t kFinishFar = 0xFFFull;
template <class K, bool kVals>
__global__ void __launch_bounds__(kFinishBlock)
    k_segment_finish(const K* __restrict__ src, K* __restrict__ dst, const uint32_t* __restrict__ vsrc,
                     uint32_t* __restrict__ vdst, uint32_t n, int startBit, int prefixShift, uint32_t prefixMask,
                     K sortMask, uint32_t* ctrl)
{
    constexpr int kWin = kFinishChunk + 2 * kFinishHalo;
    constexpr int kRows = kWin / 64;
    constexpr int kWaves = kFinishBlock / 64;
    static_assert(kWin < (int)kFinishFar, "window positions must fit the 12 low bits of order_word, below kFinishFar");
    static_assert(kFinishHalo > kFinishNear, "the fixed-count reads stay inside the window");
    static_assert(kFinishHalo % 64 == 0 && kFinishBlock % 64 == 0, "rows are all keys or all halo");
    constexpr int kNone = 0x7FFFFFFF;
    // sort bits below the prefix (the prefix lies inside the key, below its top)
    const K lowMask = sortMask & (K)(((K)1 << prefixShift) - 1);
    __shared__ uint32_t s_pf[kWin];
    __shared__ uint64_t s_ow[kWin];
    __shared__ int s_rowLast[kRows];   // the row's last segment start, or -1
    __shared__ int s_rowFirst[kRows];  // one past the row's first segment end, or kNone
    __shared__ uint32_t s_overflow;    // a wave of the block has stored the fallback flag
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t lo = (int64_t)blockIdx.x * kFinishChunk - kFinishHalo;  // window start (may be < 0)
    // all loads in flight before the first LDS store
    constexpr int kLoads = (kWin + kFinishBlock - 1) / kFinishBlock;
    K ld[kLoads];
    uint32_t val[kVals ? kLoads : 1];
#pragma unroll
    for (int u = 0; u < kLoads; u++) {
        const int i = t + u * kFinishBlock;
        const int64_t pos = lo + i;
        const bool load = i < kWin && pos >= 0 && pos < (int64_t)n;
        ld[u] = load ? src[pos] : (K)0;
        if (kVals) val[u] = load && i >= kFinishHalo && i < kFinishHalo + kFinishChunk ? vsrc[pos] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kLoads; u++) {
        const int i = t + u * kFinishBlock;
        if (i < kWin) s_pf[i] = prefix_of(ld[u], prefixShift, prefixMask);
    }
    if (t == 0) s_overflow = 0;
    __syncthreads();
    const int wLo = (int)((lo < 0 ? 0 : lo) - lo);                                   // first valid window index
    const int wHi = (int)((lo + kWin < (int64_t)n ? lo + kWin : (int64_t)n) - lo);   // one past the last
    const uint64_t atOrBelow = lane == 63 ? ~0ull : ((2ull << lane) - 1);
    const uint64_t atOrAbove = ~((1ull << lane) - 1);
    // segment starts and ends of each row (wave-uniform masks)
    uint64_t startMask[kLoads], endMask[kLoads];
#pragma unroll
    for (int u = 0; u < kLoads; u++) {
        const int r = u * kWaves + wave, w0 = 64 * r, w = w0 + lane;
        startMask[u] = endMask[u] = 0;
        if (r >= kRows) continue;
        const bool valid = w >= wLo && w < wHi;
        const uint32_t pf = s_pf[w];
        startMask[u] = __ballot(valid && (w == wLo || s_pf[w - 1] != pf));
        endMask[u] = __ballot(valid && (w + 1 == wHi || s_pf[w + 1] != pf));
        if (lane == 0) {
            s_rowLast[r] = startMask[u] ? w0 + 63 - __clzll(startMask[u]) : -1;
            s_rowFirst[r] = endMask[u] ? w0 + __builtin_ctzll(endMask[u]) + 1 : kNone;
        }
    }
    __syncthreads();
    uint64_t mine[kLoads];
    int segA[kLoads], segE[kLoads];
#pragma unroll
    for (int u = 0; u < kLoads; u++) {
        const int r = u * kWaves + wave, w0 = 64 * r, w = w0 + lane;
        mine[u] = 0;
        segA[u] = segE[u] = 0;
        if (r >= kRows || w < wLo || w >= wHi) continue;
        const uint64_t below = startMask[u] & atOrBelow, above = endMask[u] & atOrAbove;
        int a, e;
        if (below) {
            a = w0 + 63 - __clzll(below);
        } else {  // the segment began in an earlier row
            int k = r - 1;
            while (k >= 0 && s_rowLast[k] < 0) k--;
            a = k >= 0 ? s_rowLast[k] : wLo;
        }
        if (above) {
            e = w0 + __builtin_ctzll(above) + 1;
        } else {  // the segment ends in a later row
            int k = r + 1;
            while (k < kRows && s_rowFirst[k] == kNone) k++;
            e = k < kRows ? s_rowFirst[k] : wHi;
        }
        const bool isShort = e - a <= kFinishNear + 1;
        // Inside a segment the prefix is common, so keys order by their sort
        // bits below the prefix, ties by window position: one comparison of
        // (low bits << 12 | position), window positions being < 4096.
        mine[u] = ((uint64_t)(isShort ? a : (int)kFinishFar) << 52) | order_word(ld[u], lowMask, startBit, w);
        s_ow[w] = mine[u];
        segA[u] = a;
        segE[u] = e;
    }
    // positions outside the array: words no key counts
#pragma unroll
    for (int u = 0; u < kLoads; u++) {
        const int i = t + u * kFinishBlock;
        if (i < kWin && (i < wLo || i >= wHi)) s_ow[i] = ~0ull;
    }
    __syncthreads();
    bool overflow = false;
#pragma unroll
    for (int u = 0; u < kLoads; u++) {
        const int w0 = u * kFinishBlock + (t - lane);
        const int w = w0 + lane;
        if (w0 < kFinishHalo || w0 >= kFinishHalo + kFinishChunk || w >= wHi) continue;
        const int a = segA[u], e = segE[u];
        // a segment reaching the window edge inside the array is longer than the halo
        if ((a == wLo && lo + wLo > 0) || (e == wHi && lo + wHi < (int64_t)n)) {
            overflow = true;
            continue;
        }
        int rank = 0;
        if (e - a <= kFinishNear + 1) {
#pragma unroll
            for (int j = 0; j <= kFinishNear; j++) rank += s_ow[a + j] < mine[u] ? 1 : 0;
        } else {
            for (int j = a; j < e; j++) rank += s_ow[j] < mine[u] ? 1 : 0;
        }
        dst[lo + a + rank] = ld[u];
        if (kVals) vdst[lo + a + rank] = val[u];
    }
    // one store per block, by the first wave that found a long run: with every
    // key of a long run storing, the stores to the one word queued for
    // milliseconds, and one per wave still for 2 ms at 2^25 keys (a block-wide
    // OR instead cost a barrier in every block of the common case)
    if (__ballot(overflow) != 0 && lane == 0 && atomicExch(&s_overflow, 1u) == 0u)
        __hip_atomic_store(&ctrl[kCtrlSortFallback], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// dst = src when the guard word is set (kWhenSet) or clear (device-side
// test): the finish's result, when it is not in the caller's buffers and the
// fallback did not run.
template <class K, bool kWhenSet = true>
__global__ void __launch_bounds__(256) k_copy_guarded(const K* __restrict__ src, K* __restrict__ dst,
                                                      const uint32_t* __restrict__ vsrc, uint32_t* __restrict__ vdst,
                                                      uint32_t n, const uint32_t* __restrict__ guard)
{
    if ((*guard != 0) != kWhenSet) return;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        dst[i] = src[i];
        if (vsrc) vdst[i] = vsrc[i];
    }
}

// ---------------------------------------------------------------------------
// The hybrid route's fallback in one launch.  When the finish found a run of
// equal prefix longer than its halo, every digit pass of the LSD sequence
// runs from the top-pass state here, inside one kernel: one 1024-thread
// block per CU (all resident), each a contiguous chunk of the keys, and three
// device-wide barriers per pass (chunk counts; per-digit scans over the
// blocks; the stable scatter of each chunk in 8192-key steps, ranked as in
// k_pass).  With the flag clear it returns at once: one launch of about 4 us
// in place of the eleven guarded ones (histograms, their reduction, every
// pass and a copy) that each cost as much while returning at once.
// ---------------------------------------------------------------------------
constexpr int kFbThreads = 1024;
constexpr int kFbWaves = kFbThreads / 64;
constexpr int kFbMaxBlocks = kFbThreads;  // the per-digit scan has one thread per block
constexpr int kFbItems = 8;                // keys per thread in one scatter step
constexpr uint32_t kFbStep = kFbThreads * kFbItems;
constexpr uint32_t kGridPollLimit = 1u << 22;

// Device-wide barrier of a launch whose blocks are all resident (one per
// CU).  ctrl[kCtrlGridArrive] only grows: every block adds 1 per barrier and
// waits until it reaches base + k * gridDim.x at its k-th barrier, base being
// ctrl[kCtrlGridBase] as read at the start of the launch (block 0 stores the
// final value there at the end, after every block has read it).  Every
// wave's stores have completed at the block barrier; thread 0 then releases
// at agent scope (L2 written back) before it arrives and acquires (caches
// invalidated) after the wait.  The wait gives up after 2^22 polls
// (seconds) and sets ctrl[kCtrlGridTimeout], so a launch whose blocks were
// not all resident ends (with wrong results) instead of hanging.
__device__ inline void grid_barrier(uint32_t* ctrl, uint32_t target)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(ctrl + kCtrlGridArrive, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t polls = 0;
        while ((int32_t)(__hip_atomic_load(ctrl + kCtrlGridArrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) -
                         target) < 0) {
            if (++polls == kGridPollLimit) {
                __hip_atomic_store(ctrl + kCtrlGridTimeout, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// Exclusive scan over the 1024 threads of the block (all of them call).
__device__ inline uint32_t scan1024(uint32_t v, uint32_t* s_tmp /* kFbWaves + 1 words */, uint32_t* total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_tmp[wave] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kFbWaves; w++) {
        const uint32_t c = s_tmp[w];
        before += w < wave ? c : 0u;
        all += c;
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

// counts: (2 * gridDim.x + 1) rows of 256 words: the blocks' chunk counts,
// their per-digit exclusive prefixes over the blocks, the digit totals.
template <class K, bool kVals>
__global__ void __launch_bounds__(kFbThreads)
    k_fallback_sort(K* state, K* other, uint32_t* vstate, uint32_t* vother, K* keys, uint32_t* vals, uint32_t n,
                    int startBit, int live, uint32_t* counts, uint32_t* ctrl)
{
    if (ctrl[kCtrlSortFallback] == 0) return;  // the finish placed every key
    __shared__ uint32_t s_cnt[256];
    __shared__ uint32_t s_off[256];
    __shared__ uint32_t s_wcnt[kFbWaves][256];
    __shared__ uint32_t s_tmp[kFbWaves + 1];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t G = gridDim.x, blk = blockIdx.x;
    const uint32_t base =
        t == 0 ? __hip_atomic_load(ctrl + kCtrlGridBase, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
    uint32_t barriers = 0;
    const uint32_t chunk = (n + G - 1) / G;
    const uint32_t lo = blk * chunk < n ? blk * chunk : n;
    const uint32_t hi = lo + chunk < n ? lo + chunk : n;
    const int keyBits = 8 * (int)sizeof(K);
    K* src = state;
    K* dst = other;
    uint32_t* vsrc = vstate;
    uint32_t* vdst = vother;
    uint32_t* rowPrefix = counts + (size_t)G * 256;
    uint32_t* rowTotal = counts + (size_t)2 * G * 256;
    for (int p = 0; p < live; p++) {
        const int shift = startBit + 8 * p;
        auto dig = [&](K k) -> uint32_t { return shift < keyBits ? digit_of(k, shift) : 0u; };
        // 1. this block's counts: one LDS atomic per distinct digit of a wave
        if (t < 256) s_cnt[t] = 0;
        __syncthreads();
        for (uint32_t b0 = lo; b0 < hi; b0 += kFbStep) {  // the same steps in every lane (ballots)
            K kk[kFbItems];
#pragma unroll
            for (int k = 0; k < kFbItems; k++) {  // all loads in flight first
                const uint32_t i = b0 + (uint32_t)(k * kFbThreads + t);
                kk[k] = i < hi ? src[i] : (K)0;
            }
#pragma unroll
            for (int k = 0; k < kFbItems; k++) {
                const bool valid = b0 + (uint32_t)(k * kFbThreads + t) < hi;
                const uint32_t d = dig(kk[k]);
                uint32_t plo, phi;
                digit_peers(d, valid, plo, phi);
                if (valid && __builtin_amdgcn_mbcnt_hi(phi, __builtin_amdgcn_mbcnt_lo(plo, 0u)) == 0)
                    atomicAdd(&s_cnt[d], (uint32_t)(__popc(plo) + __popc(phi)));
            }
        }
        __syncthreads();
        if (t < 256) counts[(size_t)blk * 256 + t] = s_cnt[t];
        grid_barrier(ctrl, base + (++barriers) * G);
        // 2. per digit: exclusive prefix over the blocks, and the total
        for (uint32_t d = blk; d < 256; d += G) {
            const uint32_t v = (uint32_t)t < G ? counts[(size_t)t * 256 + d] : 0u;
            uint32_t all;
            const uint32_t ex = scan1024(v, s_tmp, &all);
            if ((uint32_t)t < G) rowPrefix[(size_t)t * 256 + d] = ex;
            if (t == 0) rowTotal[d] = all;
        }
        grid_barrier(ctrl, base + (++barriers) * G);
        // 3. the block's first position for each digit, then its chunk in order
        {
            uint32_t all;
            const uint32_t tot = t < 256 ? rowTotal[t] : 0u;
            const uint32_t start = scan1024(tot, s_tmp, &all);
            if (t < 256) s_off[t] = start + rowPrefix[(size_t)blk * 256 + t];
        }
        for (int d = lane; d < 256; d += 64) s_wcnt[wave][d] = 0;
        __syncthreads();
        // kFbItems keys per thread per step; wave w's keys are the step's
        // positions [w * 64 * kFbItems, (w + 1) * 64 * kFbItems), item k
        // covering 64 consecutive ones, ranked as in k_pass
        for (uint32_t b0 = lo; b0 < hi; b0 += kFbStep) {
            K key[kFbItems];
            uint32_t val[kVals ? kFbItems : 1], off[kFbItems], d[kFbItems];
#pragma unroll
            for (int k = 0; k < kFbItems; k++) {
                const uint32_t i = b0 + (uint32_t)(wave * 64 * kFbItems + k * 64 + lane);
                const bool valid = i < hi;
                key[k] = valid ? src[i] : (K)0;
                if (kVals) val[k] = valid ? vsrc[i] : 0u;
            }
#pragma unroll
            for (int k = 0; k < kFbItems; k++) {
                const uint32_t i = b0 + (uint32_t)(wave * 64 * kFbItems + k * 64 + lane);
                const bool valid = i < hi;
                d[k] = dig(key[k]);
                uint32_t plo, phi;
                digit_peers(d[k], valid, plo, phi);
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi(phi, __builtin_amdgcn_mbcnt_lo(plo, 0u));
                const uint32_t cnt = (uint32_t)(__popc(plo) + __popc(phi));
                const uint32_t prior = valid ? s_wcnt[wave][d[k]] : 0u;
                off[k] = prior + rank;
                if (valid && rank == 0) s_wcnt[wave][d[k]] = prior + cnt;
            }
            __syncthreads();
            uint32_t tot = 0;
            if (t < 256) {  // digit t: the waves' offsets, and the step's total
#pragma unroll
                for (int w = 0; w < kFbWaves; w++) {
                    const uint32_t c = s_wcnt[w][t];
                    s_wcnt[w][t] = tot;
                    tot += c;
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < kFbItems; k++) {
                const uint32_t i = b0 + (uint32_t)(wave * 64 * kFbItems + k * 64 + lane);
                if (i < hi) {
                    const uint32_t pos = s_off[d[k]] + s_wcnt[wave][d[k]] + off[k];
                    dst[pos] = key[k];
                    if (kVals) vdst[pos] = val[k];
                }
            }
            __syncthreads();
            if (t < 256) {
                s_off[t] += tot;
#pragma unroll
                for (int w = 0; w < kFbWaves; w++) s_wcnt[w][t] = 0;
            }
            __syncthreads();
        }
        grid_barrier(ctrl, base + (++barriers) * G);
        K* tk = src;
        src = dst;
        dst = tk;
        uint32_t* tv = vsrc;
        vsrc = vdst;
        vdst = tv;
    }
    if (src != keys) {  // the result into the caller's buffers
        for (size_t i = (size_t)blk * kFbThreads + t; i < n; i += (size_t)G * kFbThreads) {
            keys[i] = src[i];
            if (kVals) vals[i] = vsrc[i];
        }
    }
    if (blk == 0 && t == 0)
        __hip_atomic_store(ctrl + kCtrlGridBase, base + barriers * G, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

static int hist_grid(mlbvh_ctx* ctx, uint32_t n)
{
    uint32_t want = (n + kHistBlock * 16 - 1) / (kHistBlock * 16);
    uint32_t cap = (uint32_t)ctx->cuCount * 4;
    if (want < 1) want = 1;
    return (int)(want < cap ? want : cap);
}

int sort_tile_keys(mlbvh_ctx* ctx) { return kSortCfgs[ctx->sortConfig].block * kSortCfgs[ctx->sortConfig].items; }

uint32_t sort_tiles(mlbvh_ctx* ctx, uint32_t n)
{
    uint64_t tile = (uint64_t)sort_tile_keys(ctx);
    return (uint32_t)((n + tile - 1) / tile);
}

// Look-back table for `tiles` tiles; a new table starts zeroed (tag 0 never
// matches a pass tag).
int ensure_status(mlbvh_ctx* ctx, uint32_t tiles)
{
    size_t need = (size_t)tiles * 256;
    if (need <= ctx->statusWords && ctx->status) return MLBVH_OK;
    int rc = ensure(ctx, &ctx->status, &ctx->statusWords, need, "hipMalloc(look-back status)");
    if (rc) return rc;
    hipError_t e = hipMemset(ctx->status, 0, ctx->statusWords * sizeof(uint64_t));
    if (e != hipSuccess) return hip_status(e, "hipMemset(status)");
    return MLBVH_OK;
}

int ensure_partials(mlbvh_ctx* ctx, int blocks, int nPass)
{
    return ensure(ctx, &ctx->partials, &ctx->partialsCap, (size_t)blocks * nPass * 256, "hipMalloc(histogram partials)");
}

template <class K, bool kVals, int B, int I>
static void launch_pass(hipStream_t s, uint32_t tiles, const K* src, K* dst, const uint32_t* vsrc, uint32_t* vdst,
                        uint32_t n, int shift, const uint32_t* hist, uint64_t* status, uint32_t* ctrl, int p,
                        const uint32_t* guard, int preScanned)
{
    k_pass<K, kVals, B, I><<<tiles, B, 0, s>>>(src, dst, vsrc, vdst, n, shift, hist, status, ctrl, p, guard,
                                               preScanned);
}

template <class K, bool kVals>
static void dispatch_pass(int cfg, hipStream_t s, uint32_t tiles, const K* src, K* dst, const uint32_t* vsrc,
                          uint32_t* vdst, uint32_t n, int shift, const uint32_t* hist, uint64_t* status,
                          uint32_t* ctrl, int p, const uint32_t* guard = nullptr, int preScanned = 0)
{
    switch (cfg) {
    case 0: launch_pass<K, kVals, 256, 16>(s, tiles, src, dst, vsrc, vdst, n, shift, hist, status, ctrl, p, guard, preScanned); break;
    case 1: launch_pass<K, kVals, 256, 32>(s, tiles, src, dst, vsrc, vdst, n, shift, hist, status, ctrl, p, guard, preScanned); break;
    case 2: launch_pass<K, kVals, 512, 16>(s, tiles, src, dst, vsrc, vdst, n, shift, hist, status, ctrl, p, guard, preScanned); break;
    case 3: launch_pass<K, kVals, 512, 24>(s, tiles, src, dst, vsrc, vdst, n, shift, hist, status, ctrl, p, guard, preScanned); break;
    default: launch_pass<K, kVals, 1024, 16>(s, tiles, src, dst, vsrc, vdst, n, shift, hist, status, ctrl, p, guard, preScanned); break;
    }
}

// `top` top-digit passes (3 or 4; keys -> tmp -> ...), the segment finish
// from the top-pass state into the other buffer, then the full pass
// sequence from that state in k_fallback_sort (launched always, returning at
// once when the finish did not set the fallback flag).  The result ends in
// the caller's buffers: with 4 top passes the state is in them, so the
// finish writes the temporaries and a copy guarded on the flag being clear
// brings it back; the fallback copies its result there itself.
template <class K>
static int sort_hybrid(mlbvh_ctx* ctx, K* keys, uint32_t* vals, K* ktmp, uint32_t* vtmp, uint32_t n, int startBit,
                       int live, uint32_t tiles, hipStream_t stream, int top, bool sharedDigits)
{
    int rc;
    const int keyBits = 8 * (int)sizeof(K);
    const int hg = hist_grid(ctx, n);
    // partials: hg blocks x `top` histograms, then the first pass's per-tile
    // counts (the fallback launch later reuses them: 2 rows per block and one of totals)
    const size_t topRows = (size_t)hg * top + tiles;
    const size_t fbRows = 2 * (size_t)(ctx->cuCount < kFbMaxBlocks ? ctx->cuCount : kFbMaxBlocks) + 1;
    const size_t partialRows = topRows > fbRows ? topRows : fbRows;
    if ((rc = ensure_partials(ctx, (int)partialRows, 1))) return rc;
    uint32_t* tileCounts = ctx->partials + (size_t)hg * top * 256;
    // pass slots: 0 .. top-1 = the top digits; top .. top+live-1 = the
    // fallback sequence (set up by k_hist's block 0).  The top digits'
    // histograms now (the rest only if the fallback runs), with the first
    // pass's per-tile counts: its look-back table is then filled in advance
    // and its tiles never wait on each other
    const SortInit init{ctx->ctrl, top + live, live, ctx->hist, ctx->status, ctx->statusWords};
    if (sharedDigits)
        k_hist<K, true><<<hg, kHistBlock, 0, stream>>>(keys, n, startBit + 8 * (live - top), top, ctx->partials,
                                                       nullptr, (uint32_t)sort_tile_keys(ctx), tiles, tileCounts, init);
    else
        k_hist<K><<<hg, kHistBlock, 0, stream>>>(keys, n, startBit + 8 * (live - top), top, ctx->partials, nullptr,
                                                 (uint32_t)sort_tile_keys(ctx), tiles, tileCounts, init);
    launch_scan_reduce(ctx, tileCounts, tiles, hg, top, ctx->hist + 256 * (live - top), stream);
    const uint32_t* guard = ctx->ctrl + kCtrlSortFallback;
    K* bufs[2] = {keys, ktmp};
    uint32_t* vbufs[2] = {vals, vtmp};
    for (int q = 0; q < top; q++) {
        const int p = live - top + q;
        const int shift = startBit + 8 * p;
        const int pre = q == 0;
        if (vals)
            dispatch_pass<K, true>(ctx->sortConfig, stream, tiles, bufs[q & 1], bufs[(q + 1) & 1], vbufs[q & 1],
                                   vbufs[(q + 1) & 1], n, shift, ctx->hist + 256 * p, ctx->status, ctx->ctrl, q,
                                   nullptr, pre);
        else
            dispatch_pass<K, false>(ctx->sortConfig, stream, tiles, bufs[q & 1], bufs[(q + 1) & 1], nullptr, nullptr, n,
                                    shift, ctx->hist + 256 * p, ctx->status, ctx->ctrl, q, nullptr, pre);
    }
    K* state = bufs[top & 1];
    K* other = bufs[(top + 1) & 1];
    uint32_t* vstate = vbufs[top & 1];
    uint32_t* vother = vbufs[(top + 1) & 1];
    // sort key = bits [startBit, min(startBit + 8 * live, keyBits))
    const int topBit = startBit + 8 * live < keyBits ? startBit + 8 * live : keyBits;
    const K high = topBit == keyBits ? ~(K)0 : (((K)1 << topBit) - 1);
    const K sortMask = high & ~(((K)1 << startBit) - 1);
    const int prefixShift = startBit + 8 * (live - top);
    const uint32_t prefixMask = top >= 4 ? 0xFFFFFFFFu : (1u << (8 * top)) - 1u;
    const unsigned fb = (unsigned)((n + kFinishChunk - 1) / kFinishChunk);
    if (vals)
        k_segment_finish<K, true><<<fb, kFinishBlock, 0, stream>>>(state, other, vstate, vother, n, startBit,
                                                                   prefixShift, prefixMask, sortMask, ctx->ctrl);
    else
        k_segment_finish<K, false><<<fb, kFinishBlock, 0, stream>>>(state, other, nullptr, nullptr, n, startBit,
                                                                    prefixShift, prefixMask, sortMask, ctx->ctrl);
    const unsigned gc = (unsigned)(((size_t)n + 255) / 256), capc = (unsigned)ctx->cuCount * 8;
    if (other != keys)  // the finish's result, unless the fallback runs (which keeps the state intact)
        k_copy_guarded<K, false><<<gc < capc ? gc : capc, 256, 0, stream>>>(other, keys, vother, vals, n, guard);
    // fallback: every pass from the top-pass state (which the finish does not
    // modify), in one launch that returns at once unless the finish set the flag
    const unsigned G = (unsigned)(ctx->cuCount < kFbMaxBlocks ? ctx->cuCount : kFbMaxBlocks);
    if (vals)
        k_fallback_sort<K, true><<<G, kFbThreads, 0, stream>>>(state, other, vstate, vother, keys, vals, n, startBit,
                                                               live, ctx->partials, ctx->ctrl);
    else
        k_fallback_sort<K, false><<<G, kFbThreads, 0, stream>>>(state, other, nullptr, nullptr, keys, nullptr, n,
                                                                startBit, live, ctx->partials, ctx->ctrl);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? MLBVH_OK : hip_status(e, "sort kernel launch");
}

template <class K>
static int sort_impl(mlbvh_ctx* ctx, K* keys, uint32_t* vals, K* ktmp, uint32_t* vtmp, uint32_t n, int startBit,
                     int endBit, bool histReady, hipStream_t stream, const uint32_t* readyTileCounts = nullptr,
                     int readyHistBlocks = 0, int hybridTop = 3, bool sharedDigits = false)
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
    uint32_t tiles = sort_tiles(ctx, n);
    if ((rc = ensure_status(ctx, tiles))) return rc;

    if (!histReady && n >= kHybridMinKeys && live > hybridTop)
        return sort_hybrid(ctx, keys, vals, ktmp, vtmp, n, startBit, live, tiles, stream, hybridTop, sharedDigits);

    // per-tile counts of the first digit (from the caller's histogram kernel
    // or from k_hist here) let the first pass skip its look-back; partial
    // histograms not reduced yet (histBlocks rows) are summed in the same launch
    const uint32_t* tileCounts = readyTileCounts;
    int histBlocks = readyHistBlocks;
    if (!histReady) {
        int hg = hist_grid(ctx, n);
        if ((rc = ensure_partials(ctx, hg * live + (int)tiles, 1))) return rc;
        uint32_t* counts = ctx->partials + (size_t)hg * live * 256;
        const SortInit init{ctx->ctrl, live, live, ctx->hist, ctx->status, ctx->statusWords};
        k_hist<K><<<hg, kHistBlock, 0, stream>>>(keys, n, startBit, live, ctx->partials, nullptr,
                                                 (uint32_t)sort_tile_keys(ctx), tiles, counts, init);
        tileCounts = counts;
        histBlocks = hg;
    }
    launch_scan_reduce(ctx, tileCounts, tiles, histBlocks, live, ctx->hist, stream);
    K* src = keys;
    K* dst = ktmp;
    uint32_t* vsrc = vals;
    uint32_t* vdst = vtmp;
    for (int p = 0; p < live; p++) {
        int shift = startBit + 8 * p;
        const int pre = p == 0 && tileCounts != nullptr;
        if (vals)
            dispatch_pass<K, true>(ctx->sortConfig, stream, tiles, src, dst, vsrc, vdst, n, shift,
                                   ctx->hist + 256 * p, ctx->status, ctx->ctrl, p, nullptr, pre);
        else
            dispatch_pass<K, false>(ctx->sortConfig, stream, tiles, src, dst, nullptr, nullptr, n, shift,
                                    ctx->hist + 256 * p, ctx->status, ctx->ctrl, p, nullptr, pre);
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
                     bool histReady, hipStream_t stream, const uint32_t* tileCounts, int histBlocks)
{
    return sort_impl<uint64_t>(ctx, keys, nullptr, tmp, nullptr, n, startBit, endBit, histReady, stream, tileCounts,
                               histBlocks);
}

int sort_pairs_u64(mlbvh_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t* ktmp, uint32_t* vtmp, uint32_t n,
                   int startBit, int endBit, hipStream_t stream, int hybridTop, bool sharedDigits)
{
    if (hybridTop < 3 || hybridTop > 4) {
        set_error("sort: %d top-digit passes (3 or 4)", hybridTop);
        return MLBVH_E_INVALID_ARG;
    }
    return sort_impl<uint64_t>(ctx, keys, vals, ktmp, vtmp, n, startBit, endBit, false, stream, nullptr, 0, hybridTop,
                               sharedDigits);
}

void launch_hist_reduce(mlbvh_ctx* ctx, int blocks, int nPass, hipStream_t stream, int firstHist,
                        const uint32_t* guard)
{
    dim3 grid((unsigned)nPass, (unsigned)((blocks + kReduceChunk - 1) / kReduceChunk), 1);
    k_hist_reduce<<<grid, 256, 0, stream>>>(ctx->partials, blocks, nPass, ctx->hist + 256 * firstHist, guard);
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
