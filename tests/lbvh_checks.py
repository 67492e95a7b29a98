"""Independent (numpy) characterisations of the LBVH used to pin the oracle
and to check GPU output at sizes where the oracle is not re-run.

expected_tree() derives the whole node array from sorted keys alone, as a
Cartesian tree over the deltas: node p (the boundary between sorted leaves p
and p+1) covers the leaves [L+1, R] where L is the nearest boundary to the
left with delta >= delta[p] and R the nearest to the right with delta >
delta[p]; its children split at p, and each child is the leftmost maximum-
delta boundary of its sub-range (or a leaf).  This is what the agglomerative
climb of minimum_lbvh.h:284-361 produces (ties in delta resolve toward the
right subtree first), derived here without simulating the climb.
"""
from __future__ import annotations

import numpy as np

LEAF = np.uint32(0x80000000)
INVALID = np.uint32(0x7FFFFFFF)
FLT_MAX = np.float32(3.402823466e38)


def numpy_morton_keys(tris: np.ndarray, use_normal: bool = False, full: bool = False) -> np.ndarray:
    """32-bit Morton keys restated with numpy float32 arithmetic; with
    ``full`` the whole 63-bit code (uint64) of the centroid instead."""
    t = np.ascontiguousarray(tris, np.float32).reshape(-1, 3, 3)
    lo = t.reshape(-1, 3).min(axis=0)
    hi = t.reshape(-1, 3).max(axis=0)
    c = ((t[:, 0] + t[:, 1]) + t[:, 2]) / np.float32(3.0)

    def quantise(p, blo, bhi):
        with np.errstate(all="ignore"):
            f = ((p - blo) / (bhi - blo)) * np.float32(2097152.0)
            ok = (f >= np.float32(-2147483648.0)) & (f < np.float32(2147483648.0))
            q = np.where(ok, np.trunc(np.where(ok, f, 0)).astype(np.int64), -(2**31))
        return np.clip(q, 0, 0x1FFFFF).astype(np.uint64)

    def code(q):
        out = np.zeros(q.shape[0], np.uint64)
        for i in range(21):
            for a in range(3):
                out |= ((q[:, a] >> np.uint64(i)) & np.uint64(1)) << np.uint64(3 * i + a)
        return out

    if full:
        return code(quantise(c, lo, hi))
    key = (code(quantise(c, lo, hi)) >> np.uint64(31)).astype(np.uint32)
    if use_normal:
        e1 = t[:, 1] - t[:, 0]
        e2 = t[:, 2] - t[:, 0]
        ng = np.stack([e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1], e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2],
                       e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]], axis=1).astype(np.float32)
        with np.errstate(all="ignore"):
            d = (ng[:, 0] * ng[:, 0] + ng[:, 1] * ng[:, 1]) + ng[:, 2] * ng[:, 2]
            inv = np.float32(1.0) / np.sqrt(d)
            ng = ng * inv[:, None]
        nk = (code(quantise(ng, np.float32(-1.0), np.float32(1.0))) >> np.uint64(31)).astype(np.uint32)
        key = (key & np.uint32(0xFFFFFFC0)) | (nk >> np.uint32(26))
    return key


def stable_order(keys: np.ndarray) -> np.ndarray:
    return np.argsort(keys, kind="stable").astype(np.uint32)


def deltas_of(sorted_keys: np.ndarray) -> np.ndarray:
    """Bit length of the xor of neighbouring keys (exact for 64-bit keys)."""
    x = (sorted_keys[:-1] ^ sorted_keys[1:]).astype(np.uint64)
    d = np.zeros(x.shape[0], np.uint8)
    for s in (32, 16, 8, 4, 2, 1):
        big = x >= (np.uint64(1) << np.uint64(s))
        d[big] += np.uint8(s)
        x = np.where(big, x >> np.uint64(s), x)
    d[x != 0] += np.uint8(1)
    return d


class _ArgMaxTable:
    """Sparse table answering 'leftmost argmax of d[a..b]'."""

    def __init__(self, d: np.ndarray):
        self.d = d.astype(np.int64)
        n = d.shape[0]
        self.levels = [np.arange(n, dtype=np.int64)]
        k = 1
        while (1 << k) <= n:
            prev = self.levels[-1]
            h = 1 << (k - 1)
            a = prev[: n - (1 << k) + 1]
            b = prev[h: h + n - (1 << k) + 1]
            self.levels.append(np.where(self.d[a] >= self.d[b], a, b))
            k += 1

    def query(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        ln = b - a + 1
        k = np.floor(np.log2(ln)).astype(np.int64)
        out = np.empty(a.shape[0], np.int64)
        for kk in np.unique(k):
            m = k == kk
            lv = self.levels[kk]
            i1 = lv[a[m]]
            i2 = lv[b[m] - (1 << kk) + 1]
            out[m] = np.where(self.d[i1] >= self.d[i2], i1, i2)
        return out


class _MinTable:
    def __init__(self, v: np.ndarray, op):
        self.op = op
        n = v.shape[0]
        self.levels = [v]
        k = 1
        while (1 << k) <= n:
            prev = self.levels[-1]
            h = 1 << (k - 1)
            self.levels.append(op(prev[: n - (1 << k) + 1], prev[h: h + n - (1 << k) + 1]))
            k += 1

    def query(self, a, b):
        ln = b - a + 1
        k = np.floor(np.log2(ln)).astype(np.int64)
        out = np.empty((a.shape[0],) + self.levels[0].shape[1:], self.levels[0].dtype)
        for kk in np.unique(k):
            m = k == kk
            lv = self.levels[kk]
            out[m] = self.op(lv[a[m]], lv[b[m] - (1 << kk) + 1])
        return out


def _ordered(f: np.ndarray) -> np.ndarray:
    """float32 -> uint32 keys whose integer order is the float order with
    -0 below +0 (NaN excluded)."""
    b = np.ascontiguousarray(f, np.float32).view(np.uint32)
    return np.where(b >> 31 == 1, ~b, b | np.uint32(0x80000000)).astype(np.uint32)


def _unordered(k: np.ndarray) -> np.ndarray:
    k = k.astype(np.uint32)
    return np.where(k >> 31 == 1, k & np.uint32(0x7FFFFFFF), ~k).astype(np.uint32).view(np.float32)


def ranges(d: np.ndarray):
    """lo/hi leaf range of every internal node via monotonic stacks."""
    n1 = d.shape[0]
    lo = np.zeros(n1, np.int64)
    hi = np.zeros(n1, np.int64)
    dl = d.tolist()
    st = []
    for p in range(n1):  # nearest left with d >= d[p]
        while st and dl[st[-1]] < dl[p]:
            st.pop()
        lo[p] = (st[-1] + 1) if st else 0
        st.append(p)
    st = []
    for p in range(n1 - 1, -1, -1):  # nearest right with d > d[p]
        while st and dl[st[-1]] <= dl[p]:
            st.pop()
        hi[p] = st[-1] if st else n1
        st.append(p)
    return lo, hi


def expected_tree(tris: np.ndarray, order: np.ndarray, d: np.ndarray):
    """Full expected node array (n-1, 16) uint32 and root from sorted order +
    deltas, computed without the agglomerative climb."""
    t = np.ascontiguousarray(tris, np.float32).reshape(-1, 3, 3)
    n = order.shape[0]
    n1 = n - 1
    lo, hi = ranges(d)
    am = _ArgMaxTable(d)
    p = np.arange(n1, dtype=np.int64)
    idx = order.astype(np.int64)

    c0 = np.where(lo == p, idx[p] | 0x80000000, 0).astype(np.int64)
    m = lo < p
    c0[m] = am.query(lo[m], p[m] - 1)
    c1 = np.where(p + 1 == hi, idx[np.minimum(p + 1, n - 1)] | 0x80000000, 0).astype(np.int64)
    m = p + 1 < hi
    c1[m] = am.query(p[m] + 1, hi[m] - 1)

    parent = np.full(n1, int(INVALID), np.int64)
    for c in (c0, c1):
        internal = c < 0x80000000
        parent[c[internal]] = p[internal]
    root = int(am.query(np.array([0]), np.array([n1 - 1]))[0]) if n1 > 0 else 0

    # leaf boxes: the host rule a<b?a:b over v0, v1, v2 from +-FLT_MAX
    # (AABB::extend(float3)); subtree unions: min / max in the order where
    # -0 < +0 (the merge's tie rule), taken on order-preserving integer keys
    lo_b = np.full(t.shape[:1] + (3,), np.float32(3.402823466e38), np.float32)
    hi_b = np.full(t.shape[:1] + (3,), np.float32(-3.402823466e38), np.float32)
    for k in range(3):
        v = t[:, k, :]
        lo_b = np.where(lo_b < v, lo_b, v)
        hi_b = np.where(hi_b > v, hi_b, v)
    mn = _MinTable(_ordered(lo_b[idx]), np.minimum)
    mx = _MinTable(_ordered(hi_b[idx]), np.maximum)
    box0 = np.concatenate([_unordered(mn.query(lo, p)), _unordered(mx.query(lo, p))], axis=1)
    box1 = np.concatenate([_unordered(mn.query(p + 1, hi)), _unordered(mx.query(p + 1, hi))], axis=1)

    nodes = np.zeros((n1, 16), np.uint32)
    nodes[:, 0] = c0.astype(np.uint32)
    nodes[:, 1] = c1.astype(np.uint32)
    nodes[:, 2:8] = box0.astype(np.float32).view(np.uint32)
    nodes[:, 8:14] = box1.astype(np.float32).view(np.uint32)
    nodes[:, 14] = parent.astype(np.uint32)
    nodes[:, 15] = np.where(hi < n1, hi, 0xFFFFFFFF).astype(np.uint32)
    return nodes, root


def structural_checks(nodes: np.ndarray, root: int, n: int):
    """Every leaf exactly once, every internal node reached once from the
    root, parent pointers consistent."""
    n1 = n - 1
    c = nodes[:, :2].astype(np.int64)
    leaf = c >= 0x80000000
    leaves = (c[leaf] - 0x80000000)
    assert leaves.shape[0] == n and np.array_equal(np.sort(leaves), np.arange(n)), "each leaf once"
    internal = c[~leaf]
    assert internal.shape[0] == n1 - 1, "n-2 internal children"
    assert np.array_equal(np.sort(np.concatenate([internal, [root]])), np.arange(n1)), "each node once"
    rows = np.repeat(np.arange(n1), 2).reshape(n1, 2)[~leaf]
    assert np.array_equal(nodes[internal, 14].astype(np.int64), rows), "parent pointers"
    assert nodes[root, 14] == INVALID
