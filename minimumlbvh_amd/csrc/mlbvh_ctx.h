// Per-device context: scratch owned by the sorter and the builder.
#pragma once

#include "mlbvh_device.h"

struct mlbvh_ctx {
    int device = 0;
    int cuCount = 256;

    // --- radix sort state -------------------------------------------------
    uint64_t* status = nullptr;  // decoupled look-back words, [tiles][256]
    size_t statusWords = 0;
    uint32_t* hist = nullptr;  // digit histograms, [8][256]
    uint32_t* partials = nullptr;  // per-block histogram partials, [blocks][passes][256]
    size_t partialsCap = 0;
    int sortConfig = 2;  // tile shape index into the sort's configuration table (2: 512 x 16)
    // Device control words (see kCtrl* below): the pass-tag counter, the
    // per-pass tile tickets and the per-pass tags.  They live in device memory
    // and are advanced by the init kernels, so a captured graph replays with
    // fresh tags.
    uint32_t* ctrl = nullptr;

    // --- build scratch ------------------------------------------------------
    uint64_t* recA = nullptr;  // IndexedMorton records (ping)
    size_t recCap = 0;
    uint64_t* recB = nullptr;  // (pong)
    size_t recBCap = 0;
    uint32_t* idxA = nullptr;  // triangle indices beside 63-bit keys (MORTON64 build), ping
    size_t idxCap = 0;
    uint32_t* idxB = nullptr;  // (pong)
    size_t idxBCap = 0;
    uint8_t* deltas = nullptr;
    size_t deltasCap = 0;
    uint32_t* flags = nullptr;  // per-internal-node arrival word of the climb
    size_t flagsCap = 0;
    uint32_t* frontier = nullptr;  // (node, lo, hi) frontier entries of the block-local levels, [n][3]
    size_t frontierCap = 0;
    uint32_t* superList = nullptr;  // the level-1 frontier in leaf order: the super-leaves, [n][3]
    size_t superCap = 0;
    uint32_t* blockCounts = nullptr;  // level-2 frontier counts, [level-2 blocks]
    size_t blockCountsCap = 0;
    uint64_t* l1Status = nullptr;  // look-back words: frontier compaction workgroups, then level-2 blocks
    size_t l1StatusCap = 0;
    uint32_t* topList = nullptr;   // level-2 frontier records for the top tree, [kTopMax][kTopWords]
    size_t topListCap = 0;
    uint64_t* sceneBox = nullptr;  // tagged copies of the scene box's 6 coordinates, kSceneBoxWords words

    // --- batched build / trace ------------------------------------------------
    uint32_t* meshBox = nullptr;  // [meshes][6] order-mapped bounds
    size_t meshBoxCap = 0;
    int traceBlocksPerCU = 8;  // persistent trace grid = cuCount * this
    int traceChunk = 64;       // rays a wave takes per work-queue pull (multiple of 64)

    // Set on a context view whose buffers are carved from fixed device
    // scratch (the allocation-free build): ensure() then never frees or
    // allocates and reports a buffer that is too small as an error.
    bool fixedScratch = false;
};

namespace mlbvh {

// Layout of mlbvh_ctx::ctrl (uint32 words).
constexpr int kCtrlEpochCounter = 0;
constexpr int kCtrlTickets = 16;  // [16]
constexpr int kCtrlEpochs = 32;   // [16]
constexpr int kCtrlWorkQueue = 48;  // ray counter of the persistent trace
constexpr int kCtrlFrontier = 49;   // entries of the climb's frontier list
constexpr int kCtrlSortFallback = 50;  // set when a prefix segment is too long for the LDS sort
constexpr int kCtrlSuperCount = 51;    // super-leaves of the build's second block-local level
constexpr int kCtrlAoTiles = 52;       // next pixel tile of the persistent AO kernel
constexpr int kCtrlPtTiles = 53;       // next pixel tile of the persistent path-tracing kernel
constexpr int kCtrlTopFallback = 54;   // set when the nodes above level 2 climb through device memory
// The words below are reset or advanced by level 1's block 0 at the start of each build or refit climb.
constexpr int kCtrlCompactTicket = 55;  // next workgroup of the level-1 frontier compaction
constexpr int kCtrlCompactTag = 56;     // the build's look-back words carry this + 1
constexpr int kCtrlL2Ticket = 57;       // next block of the build's second block-local level
constexpr int kCtrlTopCount = 58;       // entries of the top-tree list (level 2's frontier)
constexpr int kCtrlBuildTag = 59;       // the scene box slots carry this + 1 (advanced for the next build)
// Device-wide barriers of the sort's one-launch fallback (onesweep.hip, grid_barrier):
constexpr int kCtrlGridArrive = 60;   // arrivals, only ever incremented
constexpr int kCtrlGridBase = 61;     // kCtrlGridArrive's value when the next such launch starts
constexpr int kCtrlGridTimeout = 62;  // set when a barrier wait gave up (blocks not all resident)
constexpr int kCtrlWords = 64;
// The build's top-tree list (mlbvh_ctx::topList): at most kTopMax records of
// kTopWords words, one per level-2 frontier entry, in leaf order: node, last
// leaf, delta of the split after it (255: none), 0, then the entry's subtree
// box (lo xyz, hi xyz) and two zeros.
constexpr int kSceneBoxWords = 256;
constexpr uint32_t kTopMax = 4096;
constexpr int kTopWords = 12;
constexpr int kNumSortCfgs = 5;  // tile shapes of the radix sort (onesweep.hip)
constexpr uint32_t kEpochLimit = (1u << 30) - 1;  // look-back words keep 30 tag bits

// Sort-state set-up for one sort of nPass digit passes; run by a single
// 256-thread block on the sort's stream before the passes: zero the nHist
// digit histograms and the tile tickets, and hand each pass a fresh tag.
// When the tag counter would leave 30 bits the look-back table is cleared and
// the tags restart at 1 (tag 0 never matches, the table starts zeroed).
__device__ inline void sort_state_init(uint32_t* ctrl, int nPass, uint32_t* hist, int nHist, uint64_t* status,
                                       size_t statusWords)
{
    __shared__ uint32_t s_base;
    __shared__ int s_clear;
    if (threadIdx.x == 0) {
        uint32_t c = ctrl[kCtrlEpochCounter];
        s_clear = c + (uint32_t)nPass >= kEpochLimit;
        if (s_clear) c = 0;
        s_base = c;
        ctrl[kCtrlEpochCounter] = c + (uint32_t)nPass;
    }
    __syncthreads();
    if (s_clear)
        for (size_t i = threadIdx.x; i < statusWords; i += blockDim.x) status[i] = 0;
    for (int i = threadIdx.x; i < nHist * 256; i += blockDim.x) hist[i] = 0;
    if (threadIdx.x < 16) {
        ctrl[kCtrlTickets + threadIdx.x] = 0;
        ctrl[kCtrlEpochs + threadIdx.x] = s_base + 1 + threadIdx.x;
    }
    if (threadIdx.x == 0) ctrl[kCtrlSortFallback] = 0;
}

// A failed scratch allocation: under stream capture the message says what
// to do instead (a graph cannot record hipMalloc / hipFree).
inline int alloc_status(hipError_t e, const char* what)
{
    if (e == hipErrorStreamCaptureUnsupported || e == hipErrorStreamCaptureInvalidated ||
        e == hipErrorStreamCaptureImplicit || e == hipErrorCapturedEvent) {
        set_error("%s: the call needs more scratch than this context holds, and a stream is being captured (%s); "
                  "run the same call once outside capture at the largest size first, or use "
                  "mlbvh_build_ex_noalloc",
                  what, hipGetErrorString(e));
        return MLBVH_E_INVALID_ARG;
    }
    return hip_status(e, what);
}

// Grow a device buffer to hold at least `need` elements (contents dropped).
// Calls at a size the context has already served allocate nothing, so they
// may be captured in a graph.
template <class T>
int ensure(const mlbvh_ctx* ctx, T** p, size_t* cap, size_t need, const char* what)
{
    if (need <= *cap && *p) return MLBVH_OK;
    if (ctx->fixedScratch) {
        set_error("%s: %zu elements needed, the fixed scratch holds %zu (no allocation in this entry point)", what,
                  need, *cap);
        return MLBVH_E_INVALID_ARG;
    }
    if (*p) {
        hipError_t e = hipFree(*p);
        *p = nullptr;
        *cap = 0;
        if (e != hipSuccess) return alloc_status(e, what);
    }
    size_t n = need < 1 ? 1 : need;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return alloc_status(e, what);
    }
    *cap = n;
    return MLBVH_OK;
}

int set_device(mlbvh_ctx* ctx);

// Radix sort passes shared by mlbvh_sort_* and the builders.  When histReady
// is true the caller has already run sort_state_init and filled
// ctx->hist[0..passes) or the partials below (the builders fuse both into
// their own kernels) and has called ensure_status(ctx, sort_tiles(n)).
// With tileCounts (digit-major per-tile counts of the first digit, over
// sort_tiles(n) tiles of sort_tile_keys keys) the first pass reads prefixes
// instead of looking back.  With histBlocks > 0 the histograms are summed
// here from ctx->partials (histBlocks rows of [pass][256]) into the zeroed
// ctx->hist, in the same launch as the prefix scan.
int sort_records_u64(mlbvh_ctx* ctx, uint64_t* keys, uint64_t* tmp, uint32_t n, int startBit, int endBit,
                     bool histReady, hipStream_t stream, const uint32_t* tileCounts = nullptr, int histBlocks = 0);
// Key-value sort (u64 keys, u32 values) with its own histograms, as
// mlbvh_sort_u64.  hybridTop: the top-digit passes (3 or 4) of the route
// large sorts take before the segment finish; 4 (a 32-bit prefix) for keys
// whose 24-bit prefixes repeat often, such as the 63-bit Morton codes of a
// surface mesh.  sharedDigits: count the histograms with one LDS atomic per
// wave where its lanes share a digit (Morton codes), not per key.
int sort_pairs_u64(mlbvh_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t* ktmp, uint32_t* vtmp, uint32_t n,
                   int startBit, int endBit, hipStream_t stream, int hybridTop = 3, bool sharedDigits = false);
uint32_t sort_tiles(mlbvh_ctx* ctx, uint32_t n);
int sort_tile_keys(mlbvh_ctx* ctx);
int ensure_status(mlbvh_ctx* ctx, uint32_t tiles);
// Histogram partials of `blocks` blocks x nPass passes, and their reduction
// into ctx->hist[firstHist ..] (one launch on `stream`; with a guard word, the
// launch returns at once while the word is 0).
int ensure_partials(mlbvh_ctx* ctx, int blocks, int nPass);
void launch_hist_reduce(mlbvh_ctx* ctx, int blocks, int nPass, hipStream_t stream, int firstHist = 0,
                        const uint32_t* guard = nullptr);

// Resolve and zero the allocation-free build's device scratch (noalloc.hip);
// once per device, with that device current.
int noalloc_setup(int device);

// One plain build (the body of mlbvh_build_ex).
int build_one(mlbvh_ctx* ctx, const void* d_tris, int n, uint32_t opt, void* d_nodes, void* d_root,
              uint64_t* d_sorted, uint8_t* d_deltas, hipStream_t stream);

}  // namespace mlbvh
