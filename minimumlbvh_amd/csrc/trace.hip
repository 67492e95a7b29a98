// Ray traversal kernels: batch trace (the replacement for calling
// intersect_stackfree, minimum_lbvh.h:981-1061, per ray) and the normal / AO /
// path-tracing / pack kernels of main_gpu.cu.
//
// One ray per lane.  The image kernels give each wave64 an 8x8 pixel tile
// (a 256-thread block covers 16x16 pixels) so neighbouring lanes walk the
// same part of the tree.  The walks are in traverse.h: every kernel finds its
// hits with the per-lane LDS stack walk, which tests the leaves of the
// stack-free walk in the same order (so the Hit records are identical) but
// fetches each internal node once per visit; the stack-free and escape-link
// walks stay available through mlbvh_trace_stackfree /
// mlbvh_trace_escape_link.
#include "mlbvh_ctx.h"
#include "mlbvh_math.h"
#include "traverse.h"

namespace mlbvh {

constexpr int kTraceBlock = 256;
#ifndef MLBVH_AO_REFILL
#define MLBVH_AO_REFILL 8
#endif
constexpr int kAoRefill = MLBVH_AO_REFILL;  // lanes waiting before a wave starts new AO rays
// Per-lane stack entries of the AO kernel's walks: 12 entries (24 KB of LDS
// per block) let 6 blocks share a CU, against 5 with 16; the AO frame of the
// level-9 sphere took 5.27 ms against 5.56 ms, the more frequent stack-free
// restarts costing 0.1 % more node fetches.  The oracle's AO walks use the
// same depth (ORC_AO_STACK), so the kernel's counters can be compared.
constexpr uint32_t kAoStack = 12;
// Waves per SIMD the AO kernel is compiled for (the register budget): 6
// matches the 6 blocks per CU its LDS allows.
#ifndef MLBVH_AO_WAVES
#define MLBVH_AO_WAVES 6
#endif

// Walks of the plain trace.  kWalkStack: the stack walk (traverse_stack /
// stack_walk_step), the hits of intersect_stackfree with each internal node
// fetched once per visit (the default; k_trace_ww when each wave gets enough
// rays, else k_trace); kWalkStackFree: the parent-pointer walk itself
// (traverse); kWalkEscape: the escape-link walk (traverse_escape).  k_trace
// runs one ray per lane.
enum : int { kWalkStack = 0, kWalkStackFree = 1, kWalkEscape = 2 };

__device__ inline void write_hit(HitRec* hits, uint32_t i, const Hit& h)
{
    HitRec r;
    r.t = h.t;
    r.pad = 0;
    r.u = h.u;
    r.v = h.v;
    r.ngx = h.ng.x;
    r.ngy = h.ng.y;
    r.ngz = h.ng.z;
    r.tri = h.tri;
    hits[i] = r;
}

// One ray per lane.  TS is the triangle source (`const Tri*` or IndexedTris).
template <bool kAny, bool kCount, int kWalk, class TS>
__global__ void __launch_bounds__(kTraceBlock)
    k_trace(const Node* __restrict__ nodes, TS tris, const uint32_t* __restrict__ rootPtr,
            const float4* __restrict__ rays, HitRec* __restrict__ hits, int nRays, unsigned long long* counters)
{
    __shared__ uint32_t s_stack[kWalk == kWalkStack ? kLaneStack : 1][kTraceBlock];
    __shared__ float s_entry[kWalk == kWalkStack ? kLaneStack : 1][kTraceBlock];
    int i = blockIdx.x * kTraceBlock + threadIdx.x;
    uint32_t nFetch = 0, nTest = 0;
    if (i < nRays) {
        float4 a = rays[2 * i], b = rays[2 * i + 1];
        f3 ro = mk(a.x, a.y, a.z), rd = mk(b.x, b.y, b.z);
        Hit h;
        h.t = a.w;
        h.u = h.v = 0.0f;
        h.ng = mk(0, 0, 0);
        h.tri = 0xFFFFFFFFu;
        if (kWalk == kWalkEscape)
            traverse_escape<kAny, kCount>(nodes, tris, *rootPtr, ro, rd, inv_rd(rd), h, nFetch, nTest);
        else if (kWalk == kWalkStack)
            traverse_stack<kAny, kCount, kTraceBlock>(nodes, tris, *rootPtr, ro, rd, inv_rd(rd), h, nFetch, nTest,
                                                      &s_stack[0][threadIdx.x], &s_entry[0][threadIdx.x]);
        else
            traverse<kAny, kCount>(nodes, tris, *rootPtr, ro, rd, inv_rd(rd), h, nFetch, nTest);
        write_hit(hits, (uint32_t)i, h);
    }
    if (kCount) {
        wave_add_u64(&counters[0], nFetch);
        wave_add_u64(&counters[1], nTest);
    }
}

__global__ void k_queue_reset(uint32_t* counter) { *counter = 0; }

// "While-while" trace of the stack walk.  Each wave takes `chunk`
// consecutive rays at a time and hands them to its lanes one by one: a lane
// whose walk has ended takes the next ray of the wave's chunk once
// kTraceRefill lanes are idle (or the whole wave is), so lanes with short
// walks do not wait for the wave's longest one.  Each ray is walked by
// stack_walk_step exactly as traverse_stack walks it (same hits, same
// counters).  The chunks come from the work counter `queue` (one atomic per
// chunk), or, with queue null, wave w of the grid takes chunks w, w + W, ...
// kBatched: a ray names its mesh in the Ray pad word (u32) and is traced in
// that mesh's BVH (nodes at nodesAll + off[m] - m, triangles from off[m],
// root roots[m]); hits carry mesh-local triangle indices, and a ray naming no
// mesh gets a miss record.  Otherwise off is unused and roots[0] is the root.
// Against one ray per lane (traverse_stack to the end), 4M closest-hit rays
// on the level-9 sphere took 3.10 ms instead of 3.68 ms and the 4096-mesh
// batched trace 28.5 ms instead of 34.7 ms.  Refill at 8 idle lanes: 4 and 16
// were within 1 %, 2 and 32 slower by 4 % (batched).
#ifndef MLBVH_TRACE_REFILL
#define MLBVH_TRACE_REFILL 8
#endif
constexpr uint32_t kTraceRefill = MLBVH_TRACE_REFILL;
// Rays per wave of the plain trace, whose grid gives each wave one chunk
// (the hardware dispatcher balances the waves): 512 took 3.10 ms on 4M rays,
// 256 3.27, 1024 3.12, 2048 4.15.  Below 16 waves per CU at 512 rays each
// (2M rays on 256 CUs) one ray per lane is faster: on the level-9 sphere
// it took 0.73 against 0.84 ms at 512K rays (chunks of 128) and 1.23
// against 1.27 ms at 1M (chunks of 256), while at 2M rays it took 1.98
// against 1.81 ms with chunks of 512.
#ifndef MLBVH_PLAIN_CHUNK
#define MLBVH_PLAIN_CHUNK 512
#endif
constexpr uint32_t kPlainChunk = MLBVH_PLAIN_CHUNK;

template <bool kAny, bool kCount, bool kBatched, class TS>
__global__ void __launch_bounds__(kTraceBlock)
    k_trace_ww(const Node* __restrict__ nodesAll, TS trisAll, const uint32_t* __restrict__ off,
               const uint32_t* __restrict__ roots, int nMeshes, const float4* __restrict__ rays,
               HitRec* __restrict__ hits, uint32_t nRays, uint32_t chunk, uint32_t* queue,
               unsigned long long* counters)
{
    __shared__ uint32_t s_stack[kLaneStack][kTraceBlock];
    __shared__ float s_entry[kLaneStack][kTraceBlock];
    const int t = threadIdx.x, lane = t & 63;
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t nFetch = 0, nTest = 0;
    uint32_t cBase = 0, cEnd = 0;  // the wave's chunk of rays not yet handed out
    uint32_t nextChunk = blockIdx.x * (kTraceBlock / 64) + (t >> 6);
    bool queueDone = false, live = false;
    uint32_t ray = 0, nodeOff = 0, triOff = 0, root = kBatched ? 0u : roots[0];
    f3 ro = mk(0, 0, 0), rd = mk(0, 0, 1), inv = mk(0, 0, 1);
    Hit h;
    h.t = 0.0f;
    h.u = h.v = 0.0f;
    h.ng = mk(0, 0, 0);
    h.tri = 0xFFFFFFFFu;
    StackWalk w = stack_walk_start(root);
    uint32_t steps = 0;  // of the lane's ray; the cap only stops a malformed node array
    while (true) {
        const uint64_t liveMask = __ballot(live);
        if (liveMask == 0 && queueDone && cBase >= cEnd) break;
        const uint32_t nNeed = 64u - (uint32_t)__popcll(liveMask);
        if (liveMask == 0 || kTraceRefill <= nNeed) {
            if (cBase >= cEnd && !queueDone) {
                uint32_t base = 0;
                if (queue) {
                    if (lane == 0) base = atomicAdd(queue, chunk);
                    base = __shfl(base, 0, 64);
                } else {
                    // no work counter: wave w takes chunks w, w + W, w + 2W, ...
                    const uint64_t c = (uint64_t)nextChunk * chunk;
                    base = c < nRays ? (uint32_t)c : nRays;
                    nextChunk += gridDim.x * (kTraceBlock / 64);
                }
                if (base >= nRays) {
                    queueDone = true;
                } else {
                    cBase = base;
                    cEnd = nRays - base < chunk ? nRays : base + chunk;
                }
            }
            const uint32_t avail = cEnd - cBase;
            const uint32_t take = nNeed < avail ? nNeed : avail;
            if (!live) {
                const uint32_t rank = (uint32_t)__popcll(~liveMask & below);
                if (rank < take) {
                    ray = cBase + rank;
                    const float4 a = rays[2 * ray], b = rays[2 * ray + 1];
                    h.t = a.w;
                    h.u = h.v = 0.0f;
                    h.ng = mk(0, 0, 0);
                    h.tri = 0xFFFFFFFFu;
                    bool valid = true;
                    if (kBatched) {
                        const uint32_t m = __float_as_uint(b.w);
                        valid = m < (uint32_t)nMeshes;
                        if (valid) {
                            triOff = off[m];
                            nodeOff = triOff - m;
                            root = roots[m];
                        }
                    }
                    if (valid) {
                        ro = mk(a.x, a.y, a.z);
                        rd = mk(b.x, b.y, b.z);
                        inv = inv_rd(rd);
                        w = stack_walk_start(root);
                        steps = 0;
                        live = true;
                    } else {
                        write_hit(hits, ray, h);
                    }
                }
            }
            cBase += take;
        }
        if (!live) continue;
        bool done;
        if constexpr (kBatched)
            done = stack_walk_step<kAny, kCount, kTraceBlock, kLaneStack>(nodesAll + nodeOff, trisAll + triOff, root,
                                                                          ro, rd, inv, h, w, &s_stack[0][t],
                                                                          &s_entry[0][t], nFetch, nTest);
        else
            done = stack_walk_step<kAny, kCount, kTraceBlock, kLaneStack>(nodesAll, trisAll, root, ro, rd, inv, h, w,
                                                                          &s_stack[0][t], &s_entry[0][t], nFetch,
                                                                          nTest);
        if (done || ++steps == kMaxTraversalSteps) {
            write_hit(hits, ray, h);
            live = false;
        }
    }
    if (kCount) {
        wave_add_u64(&counters[0], nFetch);
        wave_add_u64(&counters[1], nTest);
    }
}

// Grid of the persistent batched trace: CUs x blocks per CU, fewer when the
// rays fill fewer waves.
static unsigned persistent_grid(const mlbvh_ctx* ctx, uint32_t nRays, uint32_t* chunkOut)
{
    const uint32_t chunk = (uint32_t)(ctx->traceChunk < 64 ? 64 : ctx->traceChunk / 64 * 64);
    unsigned grid = (unsigned)(ctx->cuCount * (ctx->traceBlocksPerCU < 1 ? 1 : ctx->traceBlocksPerCU));
    const uint64_t waves = ((uint64_t)nRays + chunk - 1) / chunk;
    if ((uint64_t)grid * (kTraceBlock / 64) > waves) grid = (unsigned)((waves + 3) / 4);
    *chunkOut = chunk;
    return grid;
}

// ---------------------------------------------------------------------------
// Shading helpers of main_gpu.cu:17-82
// ---------------------------------------------------------------------------
__device__ inline uint32_t pack_rgba(float r, float g, float b)
{
    int x = iclamp(f2i_trunc(r * 255.0f + 0.5f), 0, 255);
    int y = iclamp(f2i_trunc(g * 255.0f + 0.5f), 0, 255);
    int z = iclamp(f2i_trunc(b * 255.0f + 0.5f), 0, 255);
    return ((uint32_t)z << 16) | ((uint32_t)y << 8) | (uint32_t)x;
}

__device__ inline f3 hemisphere_cos(float xi0, float xi1)
{
    const float kPi = 3.14159265358979323846f;
    float phi = (xi0 * 2.0f) * kPi;
    float r = sqrtf(xi1);
    float sp, cp;
    fm::sincos(phi, &sp, &cp);  // in place of the reference's cosf / sinf
    float x = cp * r;
    float z = sp * r;
    float y = sqrtf(fmaxf(1.0f - r * r, 0.0f));
    return mk(x, y, z);
}

// Orthonormal basis of Duff et al. 2017.
__device__ inline void onb(f3 z, f3* xaxis, f3* yaxis)
{
    const float sign = copysignf(1.0f, z.z);
    const float a = -1.0f / (sign + z.z);
    const float b = (z.x * z.y) * a;
    *xaxis = mk(1.0f + ((sign * z.x) * z.x) * a, sign * b, -sign * z.x);
    *yaxis = mk(b, sign + (z.y * z.y) * a, -z.y);
}

__device__ inline uint32_t hash_pcg(uint32_t v)
{
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28) + 4)) ^ state) * 277803737u;
    return (word >> 22) ^ word;
}

struct Pcg {
    uint64_t state, inc;
    __device__ Pcg(uint64_t seed, uint64_t seq)
    {
        state = 0u;
        inc = (seq << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    __device__ uint32_t next()
    {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31));
    }
    __device__ float uniformf() { return __uint_as_float((next() >> 9) | 0x3f800000u) - 1.0f; }
};

// Owen-scrambled, shuffled 2D Sobol (Burley 2020) as in sobol.h:110-129.
__device__ inline uint32_t lk_perm(uint32_t x, uint32_t seed)
{
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}
__device__ inline uint32_t pascal(uint32_t v)
{
    v ^= v << 16;
    v ^= (v & 0x00FF00FFu) << 8;
    v ^= (v & 0x0F0F0F0Fu) << 4;
    v ^= (v & 0x33333333u) << 2;
    v ^= (v & 0x55555555u) << 1;
    return v;
}
__device__ inline float u24(uint32_t u) { return (float)(u >> 8) / 16777216.0f; }
__device__ inline void sobol2d(float* x, float* y, uint32_t index, uint32_t p0, uint32_t p1, uint32_t p2)
{
    uint32_t vx = p0 * 1664525u + 1013904223u;
    uint32_t vy = p1 * 1664525u + 1013904223u;
    uint32_t vz = p2 * 1664525u + 1013904223u;
    vx += vy * vz;
    vy += vz * vx;
    vz += vx * vy;
    vx ^= vx >> 16u;
    vy ^= vy >> 16u;
    vz ^= vz >> 16u;
    vx += vy * vz;
    vy += vz * vx;
    vz += vx * vy;
    uint32_t v = lk_perm(__brev(index), vx);
    *x = u24(__brev(lk_perm(__brev(v), vy)));
    *y = u24(__brev(lk_perm(__brev(pascal(v)), vz)));
}

__device__ inline void tile_pixel(int* xi, int* yi)
{
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    *xi = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    *yi = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
}

__global__ void __launch_bounds__(kTraceBlock)
    k_render_normal(uint32_t* pixels, int w, int h, mlbvh_raygen cam, const uint32_t* __restrict__ rootPtr,
                    const Node* __restrict__ nodes, const Tri* __restrict__ tris)
{
    int xi, yi;
    tile_pixel(&xi, &yi);
    if (xi >= w || yi >= h) return;
    f3 ro, rd;
    shoot(cam, &ro, &rd, (float)xi / (float)w, (float)yi / (float)h);
    Hit hit;
    hit.t = kFltMax;
    hit.ng = mk(0, 0, 0);
    hit.tri = 0xFFFFFFFFu;
    uint32_t a = 0, b = 0;
    __shared__ uint32_t s_stack[kLaneStack][kTraceBlock];
    __shared__ float s_entry[kLaneStack][kTraceBlock];
    traverse_stack<false, false, kTraceBlock>(nodes, tris, *rootPtr, ro, rd, inv_rd(rd), hit, a, b,
                                              &s_stack[0][threadIdx.x], &s_entry[0][threadIdx.x]);
    uint32_t px;
    if (hit.t != kFltMax) {
        f3 n = normalize(hit.ng);
        f3 c = (n + mk(1.0f, 1.0f, 1.0f)) * 0.5f;
        px = pack_rgba(c.x, c.y, c.z);
    } else {
        px = pack_rgba(0, 0, 0);
    }
    pixels[xi + yi * w] = px;
}

// 6 waves per SIMD (at most 80 VGPRs), matching the 6 blocks per CU that its
// LDS allows.
// One step of a ray of the AO kernels: the LDS stack walk with entry-distance
// re-test (this lane's stack entry k at stk[k * kTraceBlock]), stack-free
// from the root after a push past kDepth entries.  A primary ray (closest
// hit) narrows tmax and keeps the hit's geometric normal in n; an AO ray ends
// at its first hit (*blocked).  Returns true when the ray has ended.
template <bool kCount, uint32_t kDepth>
__device__ inline bool ao_walk_step(const Node* __restrict__ nodes, const Tri* __restrict__ tris, uint32_t root,
                                    bool primary, f3 o, f3 d, f3 iv, float& tmax, f3& n, uint32_t& curr,
                                    uint32_t& prev, uint32_t& pend, uint32_t& sp, bool& stackFree, uint32_t* stk,
                                    float* ent, bool& blocked, uint32_t& nFetch, uint32_t& nTest)
{
    bool done = false, pop = false;
    if (curr & kLeafBit) {
        const Tri tri = tris[curr & ~kLeafBit];
        if (kCount) nTest++;
        float tt, u, v;
        f3 ng;
        const bool hitTri = isect_tri(&tt, &u, &v, &ng, 0.0f, tmax, o, d, tri.v[0], tri.v[1], tri.v[2]);
        if (hitTri && !primary) {
            done = blocked = true;
        } else {
            if (hitTri) {  // primary: a closer hit (or an equal-t one visited later)
                tmax = tt;
                n = ng;
            }
            if (stackFree) {
                uint32_t tmp = curr;
                curr = prev;
                prev = tmp;
            } else {
                pop = true;
            }
        }
    } else {
        const NodeView nd = load_node(nodes, curr);
        if (kCount) nFetch++;
        const float2 rl = slabs(o, iv, nd.lo0, nd.hi0, tmax);
        const float2 rr = slabs(o, iv, nd.lo1, nd.hi1, tmax);
        const bool hitL = rl.x <= rl.y, hitR = rr.x <= rr.y, rFirst = rr.x < rl.x;
        if (stackFree) {
            const StepOut st = stackfree_decide(nd, prev, hitL, hitR, rFirst, pend);
            if (st.done || st.next == kInvalidNode) {
                done = true;
            } else {
                prev = curr;
                curr = st.next;
            }
        } else if (hitL && hitR) {
            if (sp == kDepth) {  // stack full: restart this ray stack-free
                stackFree = true;
                curr = root;
                prev = kInvalidNode;
                pend = 0;
            } else {
                stk[sp * kTraceBlock] = rFirst ? nd.c0 : nd.c1;
                ent[sp * kTraceBlock] = rFirst ? rl.x : rr.x;
                sp++;
                curr = rFirst ? nd.c1 : nd.c0;
            }
        } else if (hitL || hitR) {
            curr = hitL ? nd.c0 : nd.c1;
        } else {
            pop = true;
        }
    }
    if (pop) {  // the next far child still entered within tmax, if any
        done = true;
        while (sp != 0) {
            --sp;
            if (ent[sp * kTraceBlock] <= tmax) {
                curr = stk[sp * kTraceBlock];
                done = false;
                break;
            }
        }
    }
    return done;
}

// kSobol: AO directions from the Sobol stream (the reference's useSobol),
// else from the per-pixel PCG stream.
template <bool kCount, bool kSobol>
__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(MLBVH_AO_WAVES)))
    k_render_ao(uint32_t* pixels, float* aoOut, int w, int h, mlbvh_raygen cam, const uint32_t* __restrict__ rootPtr,
                const Node* __restrict__ nodes, const Tri* __restrict__ tris, int nSample, float radius, uint32_t seed,
                unsigned long long* counters)
{
    int xi, yi;
    tile_pixel(&xi, &yi);
    uint32_t nRays = 0, nFetch = 0, nTest = 0;
    const bool inImage = xi < w && yi < h;
    const uint32_t root = *rootPtr;
    const int pixel = xi + yi * w;

    // All of a lane's rays run in one "while-while" loop: first its primary ray
    // (closest hit), then its AO rays.  A lane whose ray ends takes its next
    // one without waiting for the rest of the wave, so a wave is not held up
    // by its slowest primary ray before any AO ray starts.
    //
    // Every ray walks with an explicit per-lane stack in LDS: near child
    // first, the far child pushed with its entry distance; a popped entry is
    // skipped unless its entry distance is still <= tmax.  Every internal node
    // is fetched once, where the stack-free walk fetches it again on the way
    // back up, and the leaves are tested in the stack-free walk's order, so
    // the primary hit (ties included) is the reference walk's: back at a node
    // from its near child, intersect_stackfree re-tests the far box against
    // the current hit distance, which is this entry-distance test.  An AO ray
    // only asks whether anything lies within `radius` (its tmax never
    // changes, so no entry is skipped).  A push past kAoStack entries
    // restarts that ray with the stack-free walk from the root, keeping the
    // primary ray's hit so far.  The oracle's orc_render_ao (walk 3) does the
    // same, counters included.  Node fetch, slab tests and triangle test are
    // shared by both walks; only the choice of the next node differs.
    __shared__ uint32_t s_stack[kAoStack][kTraceBlock];
    __shared__ float s_entry[kAoStack][kTraceBlock];
    const int t = threadIdx.x;
    // The lane's current ray; after the primary ray, o is the AO origin.
    f3 o = mk(0, 0, 0), d = mk(0, 0, 1), iv = mk(0, 0, 1);
    float tmax = kFltMax;  // primary: closest t so far; AO: radius
    // primary: the closest hit's geometric normal; then the shading frame's
    // normal (xaxis, zaxis complete the frame)
    f3 n = mk(0, 0, 0), xaxis = mk(1, 0, 0), zaxis = mk(0, 1, 0);
    bool primary = inImage;       // the primary ray is in flight
    bool live = inImage;          // a ray is in flight
    bool waiting = false;         // the next AO sample is due
    bool stackFree = false;       // the current ray walks stack-free (after a stack overflow)
    bool hasHit = false;
    uint32_t curr = kInvalidNode, prev = kInvalidNode, pend = 0, sp = 0;
    if (inImage) {
        shoot(cam, &o, &d, (float)xi / (float)w, (float)yi / (float)h);
        iv = inv_rd(d);
        curr = root;
    }
    Pcg rng(kSobol ? 0u : hash_pcg((uint32_t)pixel), 0);
    int sample = 0, occluded = 0;
    uint32_t guard = kMaxTraversalSteps;
    while (guard-- != 0) {
        const uint64_t liveMask = __ballot(live);
        const uint64_t waitMask = __ballot(waiting);
        if (waitMask == 0 && liveMask == 0) break;
        if (waitMask != 0 && (liveMask == 0 || kAoRefill <= __popcll(waitMask))) {
            if (waiting) {
                float r0, r1;
                if (kSobol) {
                    sobol2d(&r0, &r1, (uint32_t)sample, (uint32_t)xi, (uint32_t)yi, seed);
                } else {
                    r0 = rng.uniformf();
                    r1 = rng.uniformf();
                }
                f3 dl = hemisphere_cos(r0, r1);
                d = ((xaxis * dl.x) + (zaxis * dl.z)) + (n * dl.y);
                iv = inv_rd(d);
                tmax = radius;
                curr = root;
                prev = kInvalidNode;
                pend = 0;
                sp = 0;
                stackFree = false;
                live = true;
                waiting = false;
            }
        }
        if (!live) continue;
        bool blocked = false;
        const bool done = ao_walk_step<kCount, kAoStack>(nodes, tris, root, primary, o, d, iv, tmax, n, curr, prev, pend,
                                                         sp, stackFree, &s_stack[0][t], &s_entry[0][t], blocked,
                                                         nFetch, nTest);
        if (done) {
            nRays++;
            live = false;
            if (primary) {
                primary = false;
                hasHit = tmax != kFltMax;
                if (hasHit) {
                    n = normalize(n);
                    if (0.0f < dot(n, d)) n = -n;
                    onb(n, &xaxis, &zaxis);
                    o = (o + d * tmax) + n * 0.001f;
                    waiting = 0 < nSample;
                }
            } else {
                occluded += blocked ? 1 : 0;
                waiting = ++sample < nSample;
            }
        }
    }

    if (inImage) {
        if (!hasHit) {
            pixels[pixel] = pack_rgba(0, 0, 0);
            if (aoOut) aoOut[pixel] = -1.0f;
        } else {
            // ao += 1/N once per occluded sample, in float, as the reference does
            const float step = 1.0f / (float)nSample;
            float ao = 0.0f;
            for (int k = 0; k < occluded; k++) ao += step;
            float brightness = fm::pow_pos(1.0f - ao, 1.0f / 2.2f);
            pixels[pixel] = pack_rgba(brightness, brightness, brightness);
            if (aoOut) aoOut[pixel] = ao;
        }
    }
    if (kCount) {
        wave_add_u64(&counters[0], nRays);
        wave_add_u64(&counters[1], nFetch);
        wave_add_u64(&counters[2], nTest);
    }
}

// AO with a per-wave pool of AO rays (Sobol sampling).  The lanes of a wave
// still trace their own pixels' primary rays, but an AO ray is no longer
// bound to its pixel's lane: once a primary ray hits, its pixel's nSample AO
// rays join the wave's pool (the hit distance and normal go to LDS), and any
// lane whose ray has ended takes the next ray of the pool.  Lanes of a tile
// with short walks then take over the work of lanes with long ones, and a
// lane whose pixel missed still traces AO rays.  Each ray is traced exactly
// as in k_render_ao (same walk, same Sobol sample, same counters); a pixel's
// occluded samples are counted with LDS atomics and turned into its AO value
// once the wave's pool is empty.  The waves are persistent: one resident set
// of blocks, each wave taking 8x8 tiles from a device counter (reset by a
// one-thread kernel before the launch, so a captured graph replays it), so
// the CUs stay full until the last tiles (4.74 -> 4.57-4.67 ms per 1080p
// frame of the level-9 sphere against a grid of one block per 16x16 pixels).
#ifndef MLBVH_AO_POOL_STACK
#define MLBVH_AO_POOL_STACK 12
#endif
constexpr uint32_t kAoPoolStack = MLBVH_AO_POOL_STACK;
// Idle lanes a wave collects before it hands out pooled AO rays (measured on
// the level-9 1080p frame: 1 -> 6.3 ms, 4 -> 5.1, 8 -> 4.9, 16 -> 4.72, 32 -> 5.07).
#ifndef MLBVH_AO_POOL_REFILL
#define MLBVH_AO_POOL_REFILL 16
#endif
constexpr uint32_t kAoPoolRefill = MLBVH_AO_POOL_REFILL;
// Sobol AO with up to 255 samples runs on the pool kernel (0: the per-lane kernel).
#ifndef MLBVH_AO_POOL
#define MLBVH_AO_POOL 1
#endif
// Blocks per CU of the persistent AO grid: at least the 5 that fit (29.9 KB
// of LDS each); blocks beyond those start once the tiles are gone and exit.
constexpr unsigned kAoPoolBlocksPerCU = 8;

__global__ void k_set_word(uint32_t* p, uint32_t v)
{
    if (threadIdx.x == 0) *p = v;
}
constexpr bool kAoPool = MLBVH_AO_POOL != 0;

template <bool kCount>
__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(MLBVH_AO_WAVES)))
    k_render_ao_pool(uint32_t* pixels, float* aoOut, int w, int h, mlbvh_raygen cam,
                     const uint32_t* __restrict__ rootPtr, const Node* __restrict__ nodes,
                     const Tri* __restrict__ tris, int nSample, float radius, uint32_t seed,
                     unsigned long long* counters, uint32_t* tileCounter)
{
    constexpr int kWaves = kTraceBlock / 64;
    __shared__ uint32_t s_stack[kAoPoolStack][kTraceBlock];
    __shared__ float s_entry[kAoPoolStack][kTraceBlock];
    __shared__ float4 s_hit[kWaves][64];   // (t, shading normal) of each pixel's primary hit
    __shared__ uint32_t s_occ[kWaves][64];  // occluded samples per pixel
    __shared__ uint8_t s_ready[kWaves][64]; // pixels whose AO rays are in the pool, in arrival order
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t root = *rootPtr;
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t nRays = 0, nFetch = 0, nTest = 0;
    // Persistent waves: each takes the next 8x8 tile from the counter, in the
    // order of the 16x16 groups a block grid would take them, until none is
    // left; a slow tile then holds one wave, not its block's four slots.
    const uint32_t groupsX = (uint32_t)(w + 15) / 16, nTiles = groupsX * ((uint32_t)(h + 15) / 16) * 4;
    while (true) {
        uint32_t tile = 0;
        if (lane == 0) tile = atomicAdd(tileCounter, 1u);
        tile = __builtin_amdgcn_readfirstlane(tile);
        if (tile >= nTiles) break;
        const uint32_t grp = tile >> 2, sub = tile & 3;
        const int xi = (int)((grp % groupsX) * 16 + (sub & 1) * 8 + (uint32_t)(lane & 7));
        const int yi = (int)((grp / groupsX) * 16 + (sub >> 1) * 8 + (uint32_t)(lane >> 3));
    const int tileX = xi - (lane & 7), tileY = yi - (lane >> 3);
    const bool inImage = xi < w && yi < h;
    s_occ[wave][lane] = 0;

    f3 o = mk(0, 0, 0), d = mk(0, 0, 1), iv = mk(0, 0, 1);
    f3 n = mk(0, 0, 0);  // primary: the closest hit's geometric normal
    float tmax = kFltMax;
    bool primary = inImage, live = inImage, stackFree = false, hasHit = false, justHit = false;
    uint32_t curr = kInvalidNode, prev = kInvalidNode, pend = 0, sp = 0, rayPix = (uint32_t)lane;
    if (inImage) {
        shoot(cam, &o, &d, (float)xi / (float)w, (float)yi / (float)h);
        iv = inv_rd(d);
        curr = root;
    }
    // wave-uniform pool state: pixels in the pool, AO rays handed out
    uint32_t readyCount = 0, nextTicket = 0;
    uint32_t guard = kMaxTraversalSteps;
    while (guard-- != 0) {
        __builtin_amdgcn_wave_barrier();
        const uint64_t readyMask = __ballot(justHit);
        if (justHit) {
            s_ready[wave][readyCount + (uint32_t)__popcll(readyMask & below)] = (uint8_t)lane;
            justHit = false;
        }
        readyCount += (uint32_t)__popcll(readyMask);
        __builtin_amdgcn_wave_barrier();
        const uint64_t liveMask = __ballot(live);
        const uint32_t avail = (uint32_t)nSample * readyCount - nextTicket;
        if (liveMask == 0 && avail == 0) break;
        const uint64_t needMask = ~liveMask;
        const uint32_t nNeed = (uint32_t)__popcll(needMask);
        if (avail != 0 && (liveMask == 0 || kAoPoolRefill <= nNeed)) {
            const uint32_t take = nNeed < avail ? nNeed : avail;
            if (!live) {
                const uint32_t rank = (uint32_t)__popcll(needMask & below);
                if (rank < take) {
                    const uint32_t ticket = nextTicket + rank;
                    rayPix = s_ready[wave][ticket / (uint32_t)nSample];
                    const uint32_t sample = ticket % (uint32_t)nSample;
                    const float4 hv = s_hit[wave][rayPix];
                    const int px = tileX + (int)(rayPix & 7), py = tileY + (int)(rayPix >> 3);
                    f3 po, pd;
                    shoot(cam, &po, &pd, (float)px / (float)w, (float)py / (float)h);
                    const f3 nn = mk(hv.y, hv.z, hv.w);
                    f3 xaxis, zaxis;
                    onb(nn, &xaxis, &zaxis);
                    o = (po + pd * hv.x) + nn * 0.001f;
                    float r0, r1;
                    sobol2d(&r0, &r1, sample, (uint32_t)px, (uint32_t)py, seed);
                    const f3 dl = hemisphere_cos(r0, r1);
                    d = ((xaxis * dl.x) + (zaxis * dl.z)) + (nn * dl.y);
                    iv = inv_rd(d);
                    tmax = radius;
                    curr = root;
                    prev = kInvalidNode;
                    pend = 0;
                    sp = 0;
                    stackFree = false;
                    live = true;
                }
            }
            nextTicket += take;
        }
        if (!live) continue;
        bool blocked = false;
        const bool done = ao_walk_step<kCount, kAoPoolStack>(nodes, tris, root, primary, o, d, iv, tmax, n, curr, prev,
                                                             pend, sp, stackFree, &s_stack[0][t], &s_entry[0][t],
                                                             blocked, nFetch, nTest);
        if (done) {
            nRays++;
            live = false;
            if (primary) {
                primary = false;
                hasHit = tmax != kFltMax;
                if (hasHit) {
                    f3 nn = normalize(n);
                    if (0.0f < dot(nn, d)) nn = -nn;
                    s_hit[wave][lane] = make_float4(tmax, nn.x, nn.y, nn.z);
                    justHit = 0 < nSample;
                }
            } else if (blocked) {
                atomicAdd(&s_occ[wave][rayPix], 1u);
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    if (inImage) {
        const int pixel = xi + yi * w;
        if (!hasHit) {
            pixels[pixel] = pack_rgba(0, 0, 0);
            if (aoOut) aoOut[pixel] = -1.0f;
        } else {
            const uint32_t occluded = s_occ[wave][lane];
            // ao += 1/N once per occluded sample, in float, as the reference does
            const float step = 1.0f / (float)nSample;
            float ao = 0.0f;
            for (uint32_t k = 0; k < occluded; k++) ao += step;
            float brightness = fm::pow_pos(1.0f - ao, 1.0f / 2.2f);
            pixels[pixel] = pack_rgba(brightness, brightness, brightness);
            if (aoOut) aoOut[pixel] = ao;
        }
    }
    __builtin_amdgcn_wave_barrier();  // the next tile reuses the wave's LDS
    }
    if (kCount) {
        wave_add_u64(&counters[0], nRays);
        wave_add_u64(&counters[1], nFetch);
        wave_add_u64(&counters[2], nTest);
    }
}

struct Attrib {
    f3 sn[3];
    f3 reflectance;
    f3 emissive;
};
static_assert(sizeof(Attrib) == 60, "TriangleAttrib is 60 bytes");

__global__ void __launch_bounds__(kTraceBlock)
    k_render_pt(float4* acc, int w, int h, mlbvh_raygen cam, const uint32_t* __restrict__ rootPtr,
                const Node* __restrict__ nodes, const Tri* __restrict__ tris, const Attrib* __restrict__ attribs,
                int useSobol, int maxSPP, int maxDepth, uint32_t* tileCounter)
{
    __shared__ uint32_t s_stack[kLaneStack][kTraceBlock];
    __shared__ float s_entry[kLaneStack][kTraceBlock];
    uint32_t* const stk = &s_stack[0][threadIdx.x];
    float* const ent = &s_entry[0][threadIdx.x];
    const uint32_t root = *rootPtr;
    const int lane = threadIdx.x & 63;
    // Persistent waves over 8x8 pixel tiles from the counter, in the order of
    // the 16x16 groups a block grid would take them.
    const uint32_t groupsX = (uint32_t)(w + 15) / 16, nTiles = groupsX * ((uint32_t)(h + 15) / 16) * 4;
    while (true) {
    uint32_t tile = 0;
    if (lane == 0) tile = atomicAdd(tileCounter, 1u);
    tile = __builtin_amdgcn_readfirstlane(tile);
    if (tile >= nTiles) break;
    const uint32_t grp = tile >> 2, sub = tile & 3;
    const int xi = (int)((grp % groupsX) * 16 + (sub & 1) * 8 + (uint32_t)(lane & 7));
    const int yi = (int)((grp / groupsX) * 16 + (sub >> 1) * 8 + (uint32_t)(lane >> 3));
    if (xi >= w || yi >= h) continue;
    int pixel = xi + yi * w;
    float4 a = acc[pixel];
    int iteration = (int)a.w;
    Pcg rng(hash_pcg(hash_pcg((uint32_t)pixel) + (uint32_t)iteration), 0);
    uint32_t dim = 0;
    auto next2 = [&](float* r0, float* r1) {
        if (useSobol) {
            sobol2d(r0, r1, (uint32_t)iteration, (uint32_t)xi, (uint32_t)yi, dim++);
        } else {
            *r0 = rng.uniformf();
            *r1 = rng.uniformf();
        }
    };
    float r0, r1;
    next2(&r0, &r1);
    f3 ro, rd;
    shoot(cam, &ro, &rd, ((float)xi + r0) / (float)w, ((float)yi + r1) / (float)h);
    f3 color = mk(0, 0, 0);
    f3 tp = mk(1.0f, 1.0f, 1.0f);
    Hit hit;
    hit.t = kFltMax;
    hit.ng = mk(0, 0, 0);
    hit.tri = 0xFFFFFFFFu;
    uint32_t c0 = 0, c1 = 0;
    traverse_stack<false, false, kTraceBlock>(nodes, tris, root, ro, rd, inv_rd(rd), hit, c0, c1, stk, ent);
    if (hit.t != kFltMax) {
        for (int depth = 0; depth < maxDepth; depth++) {
            next2(&r0, &r1);
            f3 hp = ro + rd * hit.t;
            f3 hn = normalize(hit.ng);
            if (0.0f < dot(hn, rd)) hn = -hn;
            const Attrib at = attribs[hit.tri];
            color = color + tp * at.emissive;
            f3 xaxis, zaxis;
            onb(hn, &xaxis, &zaxis);
            f3 dl = hemisphere_cos(r0, r1);
            f3 nro = hp + hn * 0.0001f;
            f3 nrd = ((xaxis * dl.x) + (zaxis * dl.z)) + (hn * dl.y);
            tp = tp * at.reflectance;
            float maxComp = fmaxf(fmaxf(tp.x, tp.y), tp.z);
            float p = fminf(maxComp / 0.25f, 1.0f);
            float q0, q1;
            next2(&q0, &q1);
            if (p < q0) break;
            tp = tp / p;
            hit.t = kFltMax;
            hit.u = hit.v = 0.0f;
            hit.ng = mk(0, 0, 0);
            hit.tri = 0xFFFFFFFFu;
            traverse_stack<false, false, kTraceBlock>(nodes, tris, root, nro, nrd, inv_rd(nrd), hit, c0, c1, stk, ent);
            if (hit.t == kFltMax) break;
            ro = nro;
            rd = nrd;
        }
    }
    if (iteration < maxSPP) {
        a.x += color.x;
        a.y += color.y;
        a.z += color.z;
        a.w += 1.0f;
        acc[pixel] = a;
    }
    }
}

__global__ void __launch_bounds__(kTraceBlock) k_pack(uint32_t* pixels, const float4* __restrict__ acc, int n)
{
    int i = blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= n) return;
    float4 a = acc[i];
    pixels[i] = pack_rgba(fm::pow_pos(a.x / a.w, 1.0f / 2.2f), fm::pow_pos(a.y / a.w, 1.0f / 2.2f),
                          fm::pow_pos(a.z / a.w, 1.0f / 2.2f));
}

static int check_scene(mlbvh_ctx* ctx, const void* root, const void* nodes, const void* tris, const char* who)
{
    if (!ctx || !root || !nodes || !tris) {
        set_error("%s: null argument", who);
        return MLBVH_E_INVALID_ARG;
    }
    return set_device(ctx);
}

static dim3 image_grid(int w, int h) { return dim3((unsigned)((w + 15) / 16), (unsigned)((h + 15) / 16), 1); }

}  // namespace mlbvh

using namespace mlbvh;

template <int kWalk, class TS>
static int trace_impl(mlbvh_ctx* ctx, const void* d_nodes, TS tris, const void* d_root, const float* d_rays,
                      void* d_hits, int nRays, int rayQuery, unsigned long long* d_counters, void* stream,
                      const char* who)
{
    // the callers have checked the triangle source
    if (!ctx || !d_root || !d_nodes) {
        set_error("%s: null argument", who);
        return MLBVH_E_INVALID_ARG;
    }
    int rc = set_device(ctx);
    if (rc) return rc;
    if (nRays < 0 || (nRays > 0 && (!d_rays || !d_hits)) ||
        (rayQuery != MLBVH_RAY_QUERY_CLOSEST && rayQuery != MLBVH_RAY_QUERY_ANY)) {
        set_error("%s: invalid rays/hits/query", who);
        return MLBVH_E_INVALID_ARG;
    }
    if (nRays == 0) return MLBVH_OK;
    hipStream_t s = (hipStream_t)stream;
    const Node* nodes = (const Node*)d_nodes;
    const uint32_t* root = (const uint32_t*)d_root;
    const float4* rays = (const float4*)d_rays;
    HitRec* hits = (HitRec*)d_hits;
    const uint32_t n = (uint32_t)nRays;
    // stack walk: one chunk of kPlainChunk rays per wave once that still
    // makes 16 waves per CU; one ray per lane (k_trace) below that, and for
    // the other walks
    const uint32_t chunk = kPlainChunk;
    const bool ww = kWalk == kWalkStack && (uint64_t)n >= (uint64_t)ctx->cuCount * 16u * chunk;
    const uint32_t perBlock = ww ? chunk * (kTraceBlock / 64) : kTraceBlock;
    const unsigned g = (unsigned)(((uint64_t)n + perBlock - 1) / perBlock);
#define LAUNCH(A, C, CNT)                                                                                        \
    {                                                                                                            \
        if (ww)                                                                                                  \
            k_trace_ww<A, C, false><<<g, kTraceBlock, 0, s>>>(nodes, tris, nullptr, root, 1, rays, hits, n,      \
                                                              chunk, nullptr, CNT);                              \
        else                                                                                                     \
            k_trace<A, C, kWalk><<<g, kTraceBlock, 0, s>>>(nodes, tris, root, rays, hits, nRays, CNT);           \
    }
    const bool any = rayQuery == MLBVH_RAY_QUERY_ANY;
    if (d_counters) {
        if (any) LAUNCH(true, true, d_counters)
        else LAUNCH(false, true, d_counters)
    } else {
        if (any) LAUNCH(true, false, nullptr)
        else LAUNCH(false, false, nullptr)
    }
#undef LAUNCH
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? MLBVH_OK : hip_status(e, who);
}

extern "C" int mlbvh_trace(mlbvh_ctx* ctx, const void* d_nodes, const void* d_triangles, const void* d_root,
                           const float* d_rays, void* d_hits, int nRays, int rayQuery, unsigned long long* d_counters,
                           void* stream)
{
    if (!d_triangles) {
        set_error("trace: null argument");
        return MLBVH_E_INVALID_ARG;
    }
    return trace_impl<kWalkStack>(ctx, d_nodes, (const Tri*)d_triangles, d_root, d_rays, d_hits, nRays, rayQuery,
                                  d_counters, stream, "trace");
}

extern "C" int mlbvh_trace_stackfree(mlbvh_ctx* ctx, const void* d_nodes, const void* d_triangles,
                                     const void* d_root, const float* d_rays, void* d_hits, int nRays, int rayQuery,
                                     unsigned long long* d_counters, void* stream)
{
    if (!d_triangles) {
        set_error("trace_stackfree: null argument");
        return MLBVH_E_INVALID_ARG;
    }
    return trace_impl<kWalkStackFree>(ctx, d_nodes, (const Tri*)d_triangles, d_root, d_rays, d_hits, nRays,
                                      rayQuery, d_counters, stream, "trace_stackfree");
}

extern "C" int mlbvh_trace_escape_link(mlbvh_ctx* ctx, const void* d_nodes, const void* d_triangles,
                                       const void* d_root, const float* d_rays, void* d_hits, int nRays,
                                       int rayQuery, unsigned long long* d_counters, void* stream)
{
    if (!d_triangles) {
        set_error("trace_escape_link: null argument");
        return MLBVH_E_INVALID_ARG;
    }
    return trace_impl<kWalkEscape>(ctx, d_nodes, (const Tri*)d_triangles, d_root, d_rays, d_hits, nRays, rayQuery,
                                   d_counters, stream, "trace_escape_link");
}

extern "C" int mlbvh_trace_indexed(mlbvh_ctx* ctx, const void* d_nodes, const float* d_vertices, int nVertices,
                                   const uint32_t* d_indices, const void* d_root, const float* d_rays, void* d_hits,
                                   int nRays, int rayQuery, unsigned long long* d_counters, void* stream)
{
    if (!d_vertices || !d_indices || nVertices < 1) {
        set_error("trace_indexed: null vertex/index buffer or nVertices = %d", nVertices);
        return MLBVH_E_INVALID_ARG;
    }
    IndexedTris src{d_vertices, d_indices, (uint32_t)nVertices - 1u};
    return trace_impl<kWalkStack>(ctx, d_nodes, src, d_root, d_rays, d_hits, nRays, rayQuery, d_counters, stream,
                                  "trace_indexed");
}

extern "C" int mlbvh_trace_batched(mlbvh_ctx* ctx, const void* d_nodes, const void* d_triangles,
                                   const uint32_t* d_triOffsets, const uint32_t* d_roots, int nMeshes,
                                   const float* d_rays, void* d_hits, int nRays, int rayQuery,
                                   unsigned long long* d_counters, void* stream)
{
    int rc = check_scene(ctx, d_roots, d_nodes, d_triangles, "trace_batched");
    if (rc) return rc;
    if (!d_triOffsets || nMeshes < 1 || nRays < 0 || (nRays > 0 && (!d_rays || !d_hits)) ||
        (rayQuery != MLBVH_RAY_QUERY_CLOSEST && rayQuery != MLBVH_RAY_QUERY_ANY)) {
        set_error("trace_batched: invalid arguments");
        return MLBVH_E_INVALID_ARG;
    }
    if (nRays == 0) return MLBVH_OK;
    hipStream_t s = (hipStream_t)stream;
    uint32_t* queue = ctx->ctrl + kCtrlWorkQueue;
    uint32_t chunk;
    const unsigned grid = persistent_grid(ctx, (uint32_t)nRays, &chunk);
    k_queue_reset<<<1, 1, 0, s>>>(queue);
    const Node* nodes = (const Node*)d_nodes;
    const Tri* tris = (const Tri*)d_triangles;
    const float4* rays = (const float4*)d_rays;
    HitRec* hits = (HitRec*)d_hits;
    uint32_t n = (uint32_t)nRays;
    bool any = rayQuery == MLBVH_RAY_QUERY_ANY;
#define LAUNCH(A, C)                                                                                             \
    k_trace_ww<A, C, true><<<grid, kTraceBlock, 0, s>>>(nodes, tris, d_triOffsets, d_roots, nMeshes, rays, hits, n, \
                                                        chunk, queue, d_counters)
    if (d_counters) {
        if (any) LAUNCH(true, true);
        else LAUNCH(false, true);
    } else {
        if (any) LAUNCH(true, false);
        else LAUNCH(false, false);
    }
#undef LAUNCH
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? MLBVH_OK : hip_status(e, "trace_batched launch");
}

extern "C" int mlbvh_set_tuning(mlbvh_ctx* ctx, int sortConfig, int traceBlocksPerCU, int traceChunk)
{
    if (!ctx || sortConfig < 0 || sortConfig >= kNumSortCfgs || traceBlocksPerCU < 1 || traceBlocksPerCU > 64 ||
        traceChunk < 64 || traceChunk > 4096 || traceChunk % 64) {
        set_error("set_tuning: sortConfig in [0,%d), blocksPerCU in [1,64], chunk a multiple of 64 in [64,4096]",
                  kNumSortCfgs);
        return MLBVH_E_INVALID_ARG;
    }
    ctx->sortConfig = sortConfig;
    ctx->traceBlocksPerCU = traceBlocksPerCU;
    ctx->traceChunk = traceChunk;
    return MLBVH_OK;
}

extern "C" int mlbvh_get_tuning(mlbvh_ctx* ctx, int* sortConfig, int* traceBlocksPerCU, int* traceChunk)
{
    if (!ctx || !sortConfig || !traceBlocksPerCU || !traceChunk) {
        set_error("get_tuning: null argument");
        return MLBVH_E_INVALID_ARG;
    }
    *sortConfig = ctx->sortConfig;
    *traceBlocksPerCU = ctx->traceBlocksPerCU;
    *traceChunk = ctx->traceChunk;
    return MLBVH_OK;
}

extern "C" int mlbvh_render_normal(mlbvh_ctx* ctx, uint32_t* d_pixels, int width, int height,
                                   const mlbvh_raygen* camera, const void* d_root, const void* d_nodes,
                                   const void* d_triangles, void* stream)
{
    int rc = check_scene(ctx, d_root, d_nodes, d_triangles, "render_normal");
    if (rc) return rc;
    if (!d_pixels || !camera || width <= 0 || height <= 0) {
        set_error("render_normal: invalid image arguments");
        return MLBVH_E_INVALID_ARG;
    }
    k_render_normal<<<image_grid(width, height), kTraceBlock, 0, (hipStream_t)stream>>>(
        d_pixels, width, height, *camera, (const uint32_t*)d_root, (const Node*)d_nodes, (const Tri*)d_triangles);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? MLBVH_OK : hip_status(e, "render_normal launch");
}

extern "C" int mlbvh_render_ao(mlbvh_ctx* ctx, uint32_t* d_pixels, float* d_ao, int width, int height,
                               const mlbvh_raygen* camera, const void* d_root, const void* d_nodes,
                               const void* d_triangles, int useSobol, int nSample, float radius, uint32_t seed,
                               unsigned long long* d_counters, void* stream)
{
    int rc = check_scene(ctx, d_root, d_nodes, d_triangles, "render_ao");
    if (rc) return rc;
    if (!d_pixels || !camera || width <= 0 || height <= 0 || nSample < 0) {
        set_error("render_ao: invalid image arguments");
        return MLBVH_E_INVALID_ARG;
    }
    dim3 g = image_grid(width, height);
    hipStream_t s = (hipStream_t)stream;
    const uint32_t* root = (const uint32_t*)d_root;
    const Node* nodes = (const Node*)d_nodes;
    const Tri* tris = (const Tri*)d_triangles;
    if (kAoPool && useSobol && nSample > 0 && nSample <= 255) {
        uint32_t* tileCounter = ctx->ctrl + kCtrlAoTiles;
        k_set_word<<<1, 64, 0, s>>>(tileCounter, 0u);
        const unsigned resident = (unsigned)ctx->cuCount * kAoPoolBlocksPerCU, groups = g.x * g.y;
        const dim3 gp(resident < groups ? resident : groups, 1, 1);
        if (d_counters)
            k_render_ao_pool<true><<<gp, kTraceBlock, 0, s>>>(d_pixels, d_ao, width, height, *camera, root, nodes,
                                                              tris, nSample, radius, seed, d_counters, tileCounter);
        else
            k_render_ao_pool<false><<<gp, kTraceBlock, 0, s>>>(d_pixels, d_ao, width, height, *camera, root, nodes,
                                                               tris, nSample, radius, seed, nullptr, tileCounter);
    } else if (d_counters) {
        if (useSobol)
            k_render_ao<true, true><<<g, kTraceBlock, 0, s>>>(d_pixels, d_ao, width, height, *camera, root, nodes, tris,
                                                              nSample, radius, seed, d_counters);
        else
            k_render_ao<true, false><<<g, kTraceBlock, 0, s>>>(d_pixels, d_ao, width, height, *camera, root, nodes,
                                                               tris, nSample, radius, seed, d_counters);
    } else {
        if (useSobol)
            k_render_ao<false, true><<<g, kTraceBlock, 0, s>>>(d_pixels, d_ao, width, height, *camera, root, nodes,
                                                               tris, nSample, radius, seed, nullptr);
        else
            k_render_ao<false, false><<<g, kTraceBlock, 0, s>>>(d_pixels, d_ao, width, height, *camera, root, nodes,
                                                                tris, nSample, radius, seed, nullptr);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? MLBVH_OK : hip_status(e, "render_ao launch");
}

extern "C" int mlbvh_render_pt(mlbvh_ctx* ctx, float* d_accumulators, int width, int height,
                               const mlbvh_raygen* camera, const void* d_root, const void* d_nodes,
                               const void* d_triangles, const void* d_attribs, int useSobol, int maxSPP, int maxDepth,
                               void* stream)
{
    int rc = check_scene(ctx, d_root, d_nodes, d_triangles, "render_pt");
    if (rc) return rc;
    if (!d_accumulators || !d_attribs || !camera || width <= 0 || height <= 0 || maxDepth < 0) {
        set_error("render_pt: invalid arguments");
        return MLBVH_E_INVALID_ARG;
    }
    uint32_t* tileCounter = ctx->ctrl + kCtrlPtTiles;
    k_set_word<<<1, 64, 0, (hipStream_t)stream>>>(tileCounter, 0u);
    const dim3 g = image_grid(width, height);
    const unsigned resident = (unsigned)ctx->cuCount * kAoPoolBlocksPerCU, groups = g.x * g.y;
    k_render_pt<<<resident < groups ? resident : groups, kTraceBlock, 0, (hipStream_t)stream>>>(
        (float4*)d_accumulators, width, height, *camera, (const uint32_t*)d_root, (const Node*)d_nodes,
        (const Tri*)d_triangles, (const Attrib*)d_attribs, useSobol, maxSPP, maxDepth, tileCounter);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? MLBVH_OK : hip_status(e, "render_pt launch");
}

extern "C" int mlbvh_pack(mlbvh_ctx* ctx, uint32_t* d_pixels, const float* d_accumulators, int n, void* stream)
{
    if (!ctx || n < 0 || (n > 0 && (!d_pixels || !d_accumulators))) {
        set_error("pack: invalid arguments");
        return MLBVH_E_INVALID_ARG;
    }
    if (n == 0) return MLBVH_OK;
    int rc = set_device(ctx);
    if (rc) return rc;
    k_pack<<<(unsigned)((n + kTraceBlock - 1) / kTraceBlock), kTraceBlock, 0, (hipStream_t)stream>>>(
        d_pixels, (const float4*)d_accumulators, n);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? MLBVH_OK : hip_status(e, "pack launch");
}
