"""Geometric nested dissection of a tensor grid, restated from the reference's definition
with grid geometry instead of a sparse graph (test helper; not product code).

The reference (src/partition.cpp:241-350, 514-587; src/tree.cpp:48-57) works on A^T A and
coordinate lists.  On a tensor grid with a 5-point (2D) or 7-point (3D) stencil the A^T A
neighbours of a cell are the cells within lattice distance |d|_1 <= 2, and the recursive
bisection acts on boxes of cells.  This module computes the same ordering from those two facts
with whole-grid numpy operations, as an independent check of the oracle's and the GPU
planner's graph-based implementations:

* bisection: a box is cut along its longest axis (the first on ties) at the value of the
  sorted coordinate at position N/2; cells <= that value go left (bissect_geo :241-279);
* separators, top level first: a left-part cell not yet in a separator whose |d|_1 <= 2
  neighbourhood holds an unclaimed cell of the sibling part (:306-333);
* interfaces (l, r): from each separator cell's A-row neighbours below its level, the side
  being the parity of the neighbour's ancestor just below the separator, keeping the
  lowest-level one (first in column order on ties); then, where a side is still empty, the
  highest common of that side over same-separator neighbours; then self (:514-587);
* ordering: stable sorts by the cluster ID, merged level by level (initPermutation).
"""
from __future__ import annotations

import itertools

import numpy as np


def _offsets(d: int):
    """Lattice offsets with 0 < |o|_1 <= 2 (A^T A neighbours of a 5/7-point stencil)."""
    out = []
    for o in itertools.product(range(-2, 3), repeat=d):
        if 0 < sum(abs(x) for x in o) <= 2:
            out.append(o)
    return out


def _shift(a: np.ndarray, o, fill):
    """b[x] = a[x + o] (fill outside the grid)."""
    b = np.full_like(a, fill)
    src = tuple(slice(max(0, oi), a.shape[i] + min(0, oi)) for i, oi in enumerate(o))
    dst = tuple(slice(max(0, -oi), a.shape[i] + min(0, -oi)) for i, oi in enumerate(o))
    b[dst] = a[src]
    return b


def leaf_parts(n: int, d: int, nlevels: int) -> np.ndarray:
    """Leaf part of every cell (grid-shaped), by recursive median bisection of boxes."""
    parts = np.zeros((n,) * d, dtype=np.int64)
    boxes = [tuple((0, n - 1) for _ in range(d))]
    for depth in range(nlevels - 1):
        nxt = []
        for k, box in enumerate(boxes):
            lens = [hi - lo for lo, hi in box]
            ax = int(np.argmax(lens))  # first axis with the largest range
            N = int(np.prod([hi - lo + 1 for lo, hi in box]))
            other = N // (box[ax][1] - box[ax][0] + 1)
            mid = box[ax][0] + (N // 2) // other  # value at sorted position N/2
            sl = tuple(slice(lo, hi + 1) for lo, hi in box)
            sub = parts[sl]
            coord = np.arange(box[ax][0], box[ax][1] + 1).reshape([-1 if i == ax else 1 for i in range(d)])
            parts[sl] = np.where(coord <= mid, 2 * k, 2 * k + 1) + 0 * sub
            left = list(box)
            right = list(box)
            left[ax] = (box[ax][0], mid)
            right[ax] = (mid + 1, box[ax][1])
            nxt += [tuple(left), tuple(right)]
        boxes = nxt
    return parts


def separators(n: int, d: int, nlevels: int):
    """(self_lvl, self_sep, section) per cell, grid-shaped (lvl 0 for interiors)."""
    parts = leaf_parts(n, d, nlevels)
    lvl = np.full(parts.shape, -1, dtype=np.int64)
    sep = np.zeros(parts.shape, dtype=np.int64)
    offs = _offsets(d)
    for depth in range(nlevels - 1):
        l = nlevels - depth - 1
        pi = parts >> (l - 1)
        free = lvl == -1
        cand = free & (pi % 2 == 0)
        hit = np.zeros(parts.shape, dtype=bool)
        for o in offs:
            nb_free = _shift(free, o, False)
            nb_pi = _shift(pi, o, -7)
            hit |= nb_free & (nb_pi == pi + 1)
        new = cand & hit
        lvl[new] = l
        sep[new] = pi[new] // 2
    leaf = lvl == -1
    lvl[leaf] = 0
    sep[leaf] = parts[leaf]
    return lvl, sep, parts


def _lowest(a, b):
    return a if a[0] < b[0] else b


def _highest_common(a, b):
    (l1, s1), (l2, s2) = a, b
    while l1 < l2:
        l1, s1 = l1 + 1, s1 // 2
    while l2 < l1:
        l2, s2 = l2 + 1, s2 // 2
    while s1 != s2:
        l1, l2, s1, s2 = l1 + 1, l2 + 1, s1 // 2, s2 // 2
    return (l1, s1)


def cluster_ids(A, n: int, d: int, nlevels: int):
    """ClusterID (self, l, r, section) of every unknown in natural order: list of tuples
    ((lvl, sep), (lvl, sep), (lvl, sep), section)."""
    lvl, sep, parts = separators(n, d, nlevels)
    lvl, sep, parts = lvl.ravel(), sep.ravel(), parts.ravel()
    N = lvl.size
    R = A.tocsr()
    R.sort_indices()
    NONE = (-1, 0)
    left = [NONE] * N
    right = [NONE] * N
    for i in range(N):
        if lvl[i] <= 0:
            continue
        sl = lvl[i]
        for j in R.indices[R.indptr[i]:R.indptr[i + 1]]:
            if lvl[j] < sl:
                nb = (int(lvl[j]), int(sep[j]))
                pj = nb[1] >> (sl - nb[0] - 1)
                if pj % 2 == 0:
                    left[i] = nb if left[i] == NONE else _lowest(nb, left[i])
                else:
                    right[i] = nb if right[i] == NONE else _lowest(nb, right[i])
    l1, r1 = list(left), list(right)
    for i in range(N):
        me = (int(lvl[i]), int(sep[i]))
        if r1[i] == NONE:
            for j in R.indices[R.indptr[i]:R.indptr[i + 1]]:
                if (int(lvl[j]), int(sep[j])) == me and r1[j] != NONE:
                    right[i] = r1[j] if right[i] == NONE else _highest_common(right[i], r1[j])
        elif l1[i] == NONE:
            for j in R.indices[R.indptr[i]:R.indptr[i + 1]]:
                if (int(lvl[j]), int(sep[j])) == me and l1[j] != NONE:
                    left[i] = l1[j] if left[i] == NONE else _highest_common(left[i], l1[j])
    ids = []
    for i in range(N):
        me = (int(lvl[i]), int(sep[i]))
        if lvl[i] == 0:
            ids.append((me, me, me, int(parts[i]), 2))
        else:
            ids.append((me, left[i] if left[i] != NONE else me, right[i] if right[i] != NONE else me,
                        int(parts[i]), 2))
    return ids


def _merge(cid, L):
    (s, l, r, sec, part) = cid
    up = lambda x: (x[0] + 1, x[1] // 2)
    return (s, up(l) if l[0] < L else l, up(r) if r[0] < L else r, sec // 2 if L > 0 else sec,
            2 if s[0] == L else part)


def _key(cid):
    (s, l, r, sec, part) = cid
    return (s, l, r, sec, part)


def ordering(A, n: int, d: int, nlevels: int):
    """cperm and the level-0 cluster list [(ClusterID, first, count)] (initPermutation)."""
    ids = cluster_ids(A, n, d, nlevels)
    pm = list(ids)
    perm = sorted(range(len(pm)), key=lambda i: _key(pm[i]))
    for L in range(1, nlevels):
        pm = [_merge(c, L) for c in pm]
        perm = sorted(perm, key=lambda i: _key(pm[i]))  # stable
    perm = np.asarray(perm, dtype=np.int64)
    leaves = []
    k = 0
    while k < len(perm):
        kn = k + 1
        while kn < len(perm) and ids[perm[kn]] == ids[perm[k]]:
            kn += 1
        leaves.append((ids[perm[k]], k, kn - k))
        k = kn
    return perm, leaves


def separator_sizes(n: int, d: int, nlevels: int) -> dict:
    """{level: sorted list of separator sizes} by counting cells (lvl > 0)."""
    lvl, sep, _ = separators(n, d, nlevels)
    out = {}
    for l in range(1, nlevels):
        m = lvl == l
        if m.any():
            out[l] = sorted(np.bincount(sep[m]).tolist())
    return out
