// Ray-BVH traversal core shared by the trace, render and differentiable
// kernels: node loads, slab and triangle tests, and the stack-free
// parent-pointer walk (intersect_stackfree, minimum_lbvh.h:981-1061) with an
// exact early exit.
//
// A node visit is four 16-byte loads of the 64-byte InternalNode; the slab
// test uses the hardware v_min/v_max (identical to the reference's host
// min/max for the NaN-free values a finite ray produces).
#pragma once

#include "mlbvh_device.h"

namespace mlbvh {

constexpr uint32_t kMaxTraversalSteps = 1u << 26;

struct NodeView {
    uint32_t c0, c1, parent;
    f3 lo0, hi0, lo1, hi1;
};

__device__ inline NodeView load_node(const Node* nodes, uint32_t i)
{
    const uint4* p = reinterpret_cast<const uint4*>(nodes + i);
    uint4 a = p[0], b = p[1], c = p[2], d = p[3];
    NodeView v;
    v.c0 = a.x;
    v.c1 = a.y;
    v.lo0 = mk(__uint_as_float(a.z), __uint_as_float(a.w), __uint_as_float(b.x));
    v.hi0 = mk(__uint_as_float(b.y), __uint_as_float(b.z), __uint_as_float(b.w));
    v.lo1 = mk(__uint_as_float(c.x), __uint_as_float(c.y), __uint_as_float(c.z));
    v.hi1 = mk(__uint_as_float(c.w), __uint_as_float(d.x), __uint_as_float(d.y));
    v.parent = d.z;
    return v;
}

// slabs(), minimum_lbvh.h:82-96
__device__ inline float2 slabs(f3 ro, f3 inv, f3 lo, f3 hi, float tHit)
{
    f3 t0 = (lo - ro) * inv;
    f3 t1 = (hi - ro) * inv;
    float tminx = fminf(t0.x, t1.x), tminy = fminf(t0.y, t1.y), tminz = fminf(t0.z, t1.z);
    float tmaxx = fmaxf(t0.x, t1.x), tmaxy = fmaxf(t0.y, t1.y), tmaxz = fmaxf(t0.z, t1.z);
    float rmin = fmaxf(fmaxf(tminx, tminy), tminz);
    float rmax = fminf(fminf(tmaxx, tmaxy), tmaxz) * 1.00000024f;
    rmin = fmaxf(rmin, 0.0f);
    rmax = fminf(rmax, tHit);
    return make_float2(rmin, rmax);
}

// invRd(), minimum_lbvh.h:898-901
__device__ inline f3 inv_rd(f3 rd)
{
    f3 r = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    r.x = fmaxf(-kFltMax, fminf(r.x, kFltMax));
    r.y = fmaxf(-kFltMax, fminf(r.y, kFltMax));
    r.z = fmaxf(-kFltMax, fminf(r.z, kFltMax));
    return r;
}

// intersectRayTriangle(), minimum_lbvh.h:363-413
__device__ inline bool isect_tri(float* tOut, float* uOut, float* vOut, f3* ngOut, float tmin, float tmax, f3 ro,
                                 f3 rd, f3 v0, f3 v1, f3 v2)
{
    f3 e0 = v1 - v0;
    f3 e1 = v2 - v1;
    f3 ng = cross(e0, e1);
    float t = dot(v0 - ro, ng) / dot(ng, rd);
    if (tmin <= t && t <= tmax) {
        f3 e2 = v0 - v2;
        f3 ro2 = ro * 2.0f;
        float u = dot(cross(e2, (v0 + v2) - ro2), rd);
        float v = dot(cross(e0, (v1 + v0) - ro2), rd);
        float w = dot(cross(e1, (v2 + v1) - ro2), rd);
        if (fminf(fminf(u, v), w) < 0.0f && 0.0f < fmaxf(fmaxf(u, v), w)) return false;
        float vol = (u + v) + w;
        *tOut = t;
        *uOut = u / vol;
        *vOut = v / vol;
        *ngOut = ng;
        return true;
    }
    return false;
}

struct Hit {
    float t, u, v;
    f3 ng;
    uint32_t tri;
};

// One step of the stack-free state machine at an internal node (the body of
// intersect_stackfree's loop, minimum_lbvh.h:1018-1059), plus a count of the
// far children still owed on the current path: `pend` rises when both
// children are entered-able on the way down and falls when the far child is
// taken on the way back.  Once the walk would climb with nothing owed, every
// remaining step would only climb to the root, so the caller can stop: the
// leaves tested, and so the hit, are exactly those of the full walk.
// hit.t only shrinks, so a far child taken on the way back was always
// counted on the way down.
struct StepOut {
    uint32_t next;
    bool done;
};
// The decision part of a step, given the two children's slab results
// (hitL / hitR: the child is entered; rFirst: the right child is entered
// first, which with both hit makes it the near child).
__device__ inline StepOut stackfree_decide(const NodeView& nd, uint32_t prev, bool hitL, bool hitR, bool rFirst,
                                           uint32_t& pend)
{
    uint32_t nearN = nd.c0, farN = nd.c1;
    int nHits = 0;
    if (hitL && hitR) {
        if (rFirst) {
            nearN = nd.c1;
            farN = nd.c0;
        }
        nHits = 2;
    } else if (hitL || hitR) {
        nHits = 1;
        nearN = hitR ? farN : nearN;
    }
    uint32_t next;
    if (prev == nd.parent) {
        next = 0 < nHits ? nearN : nd.parent;
        if (nHits == 2) pend++;
    } else if (prev == nearN && nHits == 2) {
        next = farN;
        pend--;
    } else {
        next = nd.parent;
    }
    return StepOut{next, next == nd.parent && pend == 0};
}

__device__ inline StepOut stackfree_step(const NodeView& nd, uint32_t prev, f3 ro, f3 inv, float tHit, uint32_t& pend)
{
    const float2 rl = slabs(ro, inv, nd.lo0, nd.hi0, tHit);
    const float2 rr = slabs(ro, inv, nd.lo1, nd.hi1, tHit);
    return stackfree_decide(nd, prev, rl.x <= rl.y, rr.x <= rr.y, rr.x < rl.x, pend);
}

// TS is the triangle source: `const Tri*` (a soup) or IndexedTris.
template <bool kAny, bool kCount, class TS>
__device__ inline void traverse(const Node* __restrict__ nodes, TS tris, uint32_t root, f3 ro, f3 rd, f3 inv,
                                Hit& hit, uint32_t& nFetch, uint32_t& nTest)
{
    uint32_t curr = root;
    uint32_t prev = kInvalidNode;
    uint32_t pend = 0;
    // A well-formed tree visits each node at most three times; the cap only
    // stops a malformed node array from looping forever.
    uint32_t guard = kMaxTraversalSteps;
    while (curr != kInvalidNode && guard-- != 0) {
        if (curr & kLeafBit) {
            const Tri tri = tris[curr & ~kLeafBit];
            if (kCount) nTest++;
            float t, u, v;
            f3 ng;
            if (isect_tri(&t, &u, &v, &ng, 0.0f, hit.t, ro, rd, tri.v[0], tri.v[1], tri.v[2])) {
                hit.t = t;
                hit.u = u;
                hit.v = v;
                hit.ng = ng;
                hit.tri = curr & ~kLeafBit;
                if (kAny) break;
            }
            uint32_t tmp = curr;
            curr = prev;
            prev = tmp;
            continue;
        }
        const NodeView nd = load_node(nodes, curr);
        if (kCount) nFetch++;
        StepOut s = stackfree_step(nd, prev, ro, inv, hit.t, pend);
        if (s.done) break;
        prev = curr;
        curr = s.next;
    }
}

// Closest- or any-hit walk with an explicit per-lane stack of kLaneStack
// entries in LDS (entry k of this lane at stackNode[k * kStride]): near child
// first, the far child pushed with its entry distance; a popped entry is
// skipped unless that distance is still <= hit.t.  It tests the leaves of the
// stack-free walk in the same order and ends on the same Hit, ties included:
// entry distances, which order the children, do not depend on hit.t, and back
// at a node from its near child the stack-free walk takes the far child iff
// its box is still entered, i.e. iff its entry distance is <= hit.t (its exit
// bound held when it was pushed).  Each internal node is fetched once per
// visit instead of again on the way back up.  A push past kLaneStack entries
// finishes the ray with the stack-free walk from the root, keeping the hit so
// far: every triangle at the final t is tested in both walks in the same
// depth-first order, and the last of them wins in both.
// 16 entries (the oracle's ORC_LANE_STACK, so the counters compare): with 12
// the 1080p normal render of the level-9 sphere took 2.95 ms against 0.98 ms
// and the 4M-ray trace 3.61 against 3.05 ms, with 20 the same times, with 24
// the batched trace 33.7 ms against 28.5 ms (fewer blocks per CU).
constexpr uint32_t kLaneStack = 16;
template <bool kAny, bool kCount, int kStride, class TS>
__device__ inline void traverse_stack(const Node* __restrict__ nodes, TS tris, uint32_t root, f3 ro, f3 rd, f3 inv,
                                      Hit& hit, uint32_t& nFetch, uint32_t& nTest, uint32_t* stackNode,
                                      float* stackEntry)
{
    uint32_t curr = root;
    uint32_t sp = 0;
    uint32_t guard = kMaxTraversalSteps;
    while (guard-- != 0) {
        bool pop = false;
        if (curr & kLeafBit) {
            const Tri tri = tris[curr & ~kLeafBit];
            if (kCount) nTest++;
            float t, u, v;
            f3 ng;
            if (isect_tri(&t, &u, &v, &ng, 0.0f, hit.t, ro, rd, tri.v[0], tri.v[1], tri.v[2])) {
                hit.t = t;
                hit.u = u;
                hit.v = v;
                hit.ng = ng;
                hit.tri = curr & ~kLeafBit;
                if (kAny) return;
            }
            pop = true;
        } else {
            const NodeView nd = load_node(nodes, curr);
            if (kCount) nFetch++;
            const float2 rl = slabs(ro, inv, nd.lo0, nd.hi0, hit.t);
            const float2 rr = slabs(ro, inv, nd.lo1, nd.hi1, hit.t);
            const bool hitL = rl.x <= rl.y, hitR = rr.x <= rr.y, rFirst = rr.x < rl.x;
            if (hitL && hitR) {
                if (sp == kLaneStack) {
                    traverse<kAny, kCount>(nodes, tris, root, ro, rd, inv, hit, nFetch, nTest);
                    return;
                }
                stackNode[sp * kStride] = rFirst ? nd.c0 : nd.c1;
                stackEntry[sp * kStride] = rFirst ? rl.x : rr.x;
                sp++;
                curr = rFirst ? nd.c1 : nd.c0;
            } else if (hitL || hitR) {
                curr = hitL ? nd.c0 : nd.c1;
            } else {
                pop = true;
            }
        }
        if (pop) {
            while (true) {
                if (sp == 0) return;
                --sp;
                if (stackEntry[sp * kStride] <= hit.t) break;
            }
            curr = stackNode[sp * kStride];
        }
    }
}

// traverse_stack as a single step, for kernels that keep a ray in flight per
// lane and give a lane its next ray as soon as its walk ends.  The state of a
// walk; a push past kDepth entries switches it to the stack-free walk from
// the root (prev, pend), as traverse_stack does.
struct StackWalk {
    uint32_t curr, prev, pend, sp;
    bool stackFree;
};
__device__ inline StackWalk stack_walk_start(uint32_t root)
{
    return StackWalk{root, kInvalidNode, 0u, 0u, false};
}
// One node or leaf of the walk; returns true when the ray's walk has ended.
template <bool kAny, bool kCount, int kStride, uint32_t kDepth, class TS>
__device__ inline bool stack_walk_step(const Node* __restrict__ nodes, TS tris, uint32_t root, f3 ro, f3 rd, f3 inv,
                                       Hit& hit, StackWalk& w, uint32_t* stackNode, float* stackEntry,
                                       uint32_t& nFetch, uint32_t& nTest)
{
    bool pop = false;
    if (w.curr & kLeafBit) {
        const Tri tri = tris[w.curr & ~kLeafBit];
        if (kCount) nTest++;
        float t, u, v;
        f3 ng;
        if (isect_tri(&t, &u, &v, &ng, 0.0f, hit.t, ro, rd, tri.v[0], tri.v[1], tri.v[2])) {
            hit.t = t;
            hit.u = u;
            hit.v = v;
            hit.ng = ng;
            hit.tri = w.curr & ~kLeafBit;
            if (kAny) return true;
        }
        if (w.stackFree) {
            const uint32_t tmp = w.curr;
            w.curr = w.prev;
            w.prev = tmp;
            return w.curr == kInvalidNode;
        }
        pop = true;
    } else {
        const NodeView nd = load_node(nodes, w.curr);
        if (kCount) nFetch++;
        const float2 rl = slabs(ro, inv, nd.lo0, nd.hi0, hit.t);
        const float2 rr = slabs(ro, inv, nd.lo1, nd.hi1, hit.t);
        const bool hitL = rl.x <= rl.y, hitR = rr.x <= rr.y, rFirst = rr.x < rl.x;
        if (w.stackFree) {
            const StepOut st = stackfree_decide(nd, w.prev, hitL, hitR, rFirst, w.pend);
            if (st.done || st.next == kInvalidNode) return true;
            w.prev = w.curr;
            w.curr = st.next;
            return false;
        }
        if (hitL && hitR) {
            if (w.sp == kDepth) {
                w.stackFree = true;
                w.curr = root;
                w.prev = kInvalidNode;
                w.pend = 0;
                return false;
            }
            stackNode[w.sp * kStride] = rFirst ? nd.c0 : nd.c1;
            stackEntry[w.sp * kStride] = rFirst ? rl.x : rr.x;
            w.sp++;
            w.curr = rFirst ? nd.c1 : nd.c0;
        } else if (hitL || hitR) {
            w.curr = hitL ? nd.c0 : nd.c1;
        } else {
            pop = true;
        }
    }
    if (pop) {
        while (w.sp != 0) {
            --w.sp;
            if (stackEntry[w.sp * kStride] <= hit.t) {
                w.curr = stackNode[w.sp * kStride];
                return false;
            }
        }
        return true;
    }
    return false;
}

// Escape-link walk (intersect_escape_link, minimum_lbvh.h:1063-1137).  An
// internal-node visit tests ONE child box: the left one when arriving from
// the parent, else the right one.  A missed left box re-enters the same node
// (which then tests its right box); a missed right box follows the node's
// `context` link, the node to resume at once its subtree is done (~0: none).
// After a leaf, the walk returns to the parent when the leaf was its left
// child and follows the parent's link otherwise.  Children are always taken
// left first, so closest-hit ties can resolve differently from the stack-free
// walk; hit.t is the same.  nFetch counts internal-node visits.
template <bool kAny, bool kCount, class TS>
__device__ inline void traverse_escape(const Node* __restrict__ nodes, TS tris, uint32_t root, f3 ro, f3 rd, f3 inv,
                                       Hit& hit, uint32_t& nFetch, uint32_t& nTest)
{
    uint32_t curr = root;
    uint32_t prev = kInvalidNode;
    uint32_t guard = kMaxTraversalSteps;
    while (curr != kInvalidNode && guard-- != 0) {
        if (curr & kLeafBit) {
            const Tri tri = tris[curr & ~kLeafBit];
            if (kCount) nTest++;
            float t, u, v;
            f3 ng;
            if (isect_tri(&t, &u, &v, &ng, 0.0f, hit.t, ro, rd, tri.v[0], tri.v[1], tri.v[2])) {
                hit.t = t;
                hit.u = u;
                hit.v = v;
                hit.ng = ng;
                hit.tri = curr & ~kLeafBit;
                if (kAny) break;
            }
            if (prev == kInvalidNode) break;  // a leaf root: nothing else to visit
            // prev is the leaf's parent: the walk only enters a leaf from it
            const uint32_t* pw = reinterpret_cast<const uint32_t*>(nodes + prev);
            if (pw[0] == curr) {
                const uint32_t tmp = curr;
                curr = prev;
                prev = tmp;
                continue;
            }
            const uint32_t link = pw[15];
            prev = curr;
            curr = link != 0xFFFFFFFFu ? link : kInvalidNode;
            continue;
        }
        const uint4* p = reinterpret_cast<const uint4*>(nodes + curr);
        const uint4 a = p[0], b = p[1], c = p[2], d = p[3];
        if (kCount) nFetch++;
        const bool descent = prev == d.z;
        f3 lo, hi;
        uint32_t hitChild, miss;
        if (descent) {
            lo = mk(__uint_as_float(a.z), __uint_as_float(a.w), __uint_as_float(b.x));
            hi = mk(__uint_as_float(b.y), __uint_as_float(b.z), __uint_as_float(b.w));
            hitChild = a.x;
            miss = curr;
        } else {
            lo = mk(__uint_as_float(c.x), __uint_as_float(c.y), __uint_as_float(c.z));
            hi = mk(__uint_as_float(c.w), __uint_as_float(d.x), __uint_as_float(d.y));
            hitChild = a.y;
            miss = d.w != 0xFFFFFFFFu ? d.w : kInvalidNode;
        }
        const float2 r = slabs(ro, inv, lo, hi, hit.t);
        prev = curr;
        curr = r.x <= r.y ? hitChild : miss;
    }
}

__device__ inline void wave_add_u64(unsigned long long* dst, uint64_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, (unsigned long long)v);
}

}  // namespace mlbvh
