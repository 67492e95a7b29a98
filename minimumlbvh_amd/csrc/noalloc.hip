// Allocation-free build: mlbvh_build_ex with every scratch buffer carved from
// one file-scope device array instead of the context's hipMalloc'd buffers,
// for callers that cannot allocate (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc).
//
// The array, `scratch`, holds a fixed layout sized for kNoAllocMaxTris
// triangles (any sort tile shape, any CU count): the control words and the
// two look-back tables (the sort's, and that of the build's frontier
// compaction) first, at fixed offsets, because the tags they carry stay
// valid from one call to the next; then the histograms, the records, the
// deltas, the climb's arrival words, its frontier, the frontier counts and
// the top-tree list (the super-leaf list reuses the sort's pong buffers).  Every kernel
// of the build clears or overwrites what it reads before reading it
// (histograms, sort tickets, arrival words, frontier counts), and the
// build's first block-local level returns the build's tickets to zero and
// advances the tags of its look-back words and scene-box slots, so a call
// does not depend on other contents a previous call left.  All calls on one GPU share
// that GPU's array: calls that overlap in time on one GPU race and give
// undefined results; order them on one stream.
//
// Set-up (mlbvh_ctx_create, once per device, with that device current)
// resolves the array's device address with hipGetSymbolAddress and zeroes
// its control words and look-back tables; host code uses only that address.
#include <mutex>

#include "mlbvh_ctx.h"

namespace mlbvh {

constexpr uint32_t kNoAllocMaxTris = 1u << 20;
constexpr size_t kAlign = 256;
constexpr size_t up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

// Bounds that hold for every sort configuration and CU count: the smallest
// sort tile is 256 x 16 keys; histogram grids are at most one block per
// 256 x 16 keys (onesweep.hip: hist_grid) and the Morton kernel's grid is at
// most one block per tile.
constexpr size_t kMinTileKeys = 256 * 16;
constexpr size_t kMaxTiles = (kNoAllocMaxTris + kMinTileKeys - 1) / kMinTileKeys;
constexpr size_t kMaxHistBlocks = kMaxTiles;
// partial-histogram rows: the 32-bit build (Morton blocks x 4 + tiles) or the
// 63-bit build's key-value sort (hist blocks x 8 passes, or x 3 + tiles)
constexpr size_t kPartialRows = kMaxHistBlocks * 8 + kMaxTiles;

// Byte offsets into `scratch`.
constexpr size_t kOffCtrl = 0;
constexpr size_t kOffSceneBox = kOffCtrl + up(kCtrlWords * 4);
constexpr size_t kOffStatus = kOffSceneBox + up(kSceneBoxWords * 8);
constexpr size_t kStatusWords = kMaxTiles * 256;
constexpr size_t kL1Blocks = (size_t)kNoAllocMaxTris / 256 + 1;  // level-1 blocks of at least 256 leaves
constexpr size_t kOffL1Status = kOffStatus + up(kStatusWords * 8);
constexpr size_t kOffHist = kOffL1Status + up(kL1Blocks * 8);
constexpr size_t kOffPartials = kOffHist + up(8 * 256 * 4);
constexpr size_t kOffRecA = kOffPartials + up(kPartialRows * 256 * 4);
constexpr size_t kOffRecB = kOffRecA + up((size_t)kNoAllocMaxTris * 8);
constexpr size_t kOffIdxA = kOffRecB + up((size_t)kNoAllocMaxTris * 8);
constexpr size_t kOffIdxB = kOffIdxA + up((size_t)kNoAllocMaxTris * 4);
constexpr size_t kOffDeltas = kOffIdxB + up((size_t)kNoAllocMaxTris * 4);
constexpr size_t kOffFlags = kOffDeltas + up((size_t)kNoAllocMaxTris);
constexpr size_t kOffFrontier = kOffFlags + up((size_t)kNoAllocMaxTris * 4);
constexpr size_t kOffCounts = kOffFrontier + up((size_t)kNoAllocMaxTris * 3 * 4);
constexpr size_t kCountWords = 2 * kL1Blocks;  // level-1 and level-2 frontier counts
constexpr size_t kOffTopList = kOffCounts + up(kCountWords * 4);
constexpr size_t kScratchBytes = kOffTopList + up((size_t)kTopMax * kTopWords * 4);
// The super-leaf list ([n][3] words) shares the sort's pong buffers
// (records B, indices A and B), which are dead once the sorted records are
// in records A, before the climb starts.
constexpr size_t kOffSuper = kOffRecB;
static_assert(kOffDeltas - kOffSuper >= (size_t)kNoAllocMaxTris * 3 * 4, "the super-leaf list fits the pong buffers");
// the zeroed prefix: control words, scene box, the look-back tables
constexpr size_t kZeroBytes = kOffHist;
static_assert(kScratchBytes < (48ull << 20), "scratch stays modest: under 48 MiB per GPU");

// The scratch array.  In host code `scratch` names a host-side stand-in; the
// device address is g_scratch[device], resolved at set-up.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static unsigned char* g_scratch[kMaxDevices];
static std::mutex g_scratchMutex;

__global__ void k_zero_words(uint64_t* p, size_t words)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Once per device, with `device` current: resolve the array's address and
// zero its control words and look-back tables (the tags start at 0, which
// no look-back word is written with).
int noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) {
        set_error("ctx_create: device %d is beyond the %d devices this library tracks", device, kMaxDevices);
        return MLBVH_E_INVALID_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (g_scratch[device]) return MLBVH_OK;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return hip_status(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    k_zero_words<<<64, 256, 0, nullptr>>>((uint64_t*)ptr, kZeroBytes / 8);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_status(e, "zeroing the allocation-free scratch");
    g_scratch[device] = (unsigned char*)ptr;
    return MLBVH_OK;
}

}  // namespace mlbvh

using namespace mlbvh;

// Declared in include/mlbvh.h.
extern "C" int mlbvh_build_ex_noalloc(mlbvh_ctx* ctx, const void* d_triangles, int n, uint32_t buildOption,
                                      void* d_nodes, void* d_root, uint64_t* d_sorted_records, uint8_t* d_deltas,
                                      void* stream)
{
    if (!ctx) {
        set_error("build_ex_noalloc: null context");
        return MLBVH_E_INVALID_ARG;
    }
    if (n > (int)kNoAllocMaxTris) {
        set_error("build_ex_noalloc: n = %d exceeds the allocation-free build's maximum of %u triangles "
                  "(use mlbvh_build_ex)", n, kNoAllocMaxTris);
        return MLBVH_E_INVALID_ARG;
    }
    unsigned char* base = ctx->device >= 0 && ctx->device < kMaxDevices ? g_scratch[ctx->device] : nullptr;
    if (!base) {
        set_error("build_ex_noalloc: the scratch of device %d was not set up (mlbvh_ctx_create)", ctx->device);
        return MLBVH_E_INVALID_ARG;
    }
    // A view of the context: its tuning and device, with every scratch buffer
    // in the fixed array (capacities are element counts).
    mlbvh_ctx view = *ctx;
    view.fixedScratch = true;
    view.ctrl = (uint32_t*)(base + kOffCtrl);
    view.sceneBox = (uint64_t*)(base + kOffSceneBox);
    view.status = (uint64_t*)(base + kOffStatus);
    view.statusWords = kStatusWords;
    view.hist = (uint32_t*)(base + kOffHist);
    view.partials = (uint32_t*)(base + kOffPartials);
    view.partialsCap = kPartialRows * 256;
    view.recA = (uint64_t*)(base + kOffRecA);
    view.recCap = kNoAllocMaxTris;
    view.recB = (uint64_t*)(base + kOffRecB);
    view.recBCap = kNoAllocMaxTris;
    view.idxA = (uint32_t*)(base + kOffIdxA);
    view.idxCap = kNoAllocMaxTris;
    view.idxB = (uint32_t*)(base + kOffIdxB);
    view.idxBCap = kNoAllocMaxTris;
    view.deltas = (uint8_t*)(base + kOffDeltas);
    view.deltasCap = kNoAllocMaxTris;
    view.flags = (uint32_t*)(base + kOffFlags);
    view.flagsCap = kNoAllocMaxTris;
    view.frontier = (uint32_t*)(base + kOffFrontier);
    view.frontierCap = (size_t)kNoAllocMaxTris * 3;
    view.superList = (uint32_t*)(base + kOffSuper);
    view.superCap = (size_t)kNoAllocMaxTris * 3;
    view.blockCounts = (uint32_t*)(base + kOffCounts);
    view.blockCountsCap = kCountWords;
    view.l1Status = (uint64_t*)(base + kOffL1Status);
    view.l1StatusCap = kL1Blocks;
    view.topList = (uint32_t*)(base + kOffTopList);
    view.topListCap = (size_t)kTopMax * kTopWords;
    view.meshBox = nullptr;
    view.meshBoxCap = 0;
    return build_one(&view, d_triangles, n, buildOption, d_nodes, d_root, d_sorted_records, d_deltas,
                     (hipStream_t)stream);
}
