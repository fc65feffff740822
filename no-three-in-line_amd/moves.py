"""Host-side move bookkeeping for the 3-cycle neighbourhood: the legal
candidates of one slot (search.py:38-65) and applying a candidate
(search.py:68-81).  Scoring them is gpu_delta.evaluate_three_cycles."""
from __future__ import annotations

import numpy as np

from .state import PairPermutations, points_to_permutations


def _columns(p, q) -> np.ndarray:
    return np.column_stack((np.asarray(p, np.int32), np.asarray(q, np.int32))).reshape(-1)


def legal_three_cycle_candidates(p, q, selected_index: int) -> np.ndarray:
    """int32[C][4] rows (u, v, w, orientation) for u = selected_index and every
    v < w in rows other than u's and each other's, the three columns
    distinct, and no row left with two equal columns after the cycle -- the
    same rows, in the same order, as search.py:38-65."""
    cols = _columns(p, q)
    m = len(cols)
    u = int(selected_index)
    ru = u // 2
    slots = np.arange(m)
    rows = slots // 2
    other = cols[slots ^ 1]
    v, w = np.triu_indices(m, 1)
    keep = (rows[v] != ru) & (rows[w] != ru) & (rows[v] != rows[w])
    keep &= (cols[u] != cols[v]) & (cols[u] != cols[w]) & (cols[v] != cols[w])
    v, w = v[keep], w[keep]
    ou = cols[u ^ 1]
    ok0 = (cols[v] != ou) & (cols[w] != other[v]) & (cols[u] != other[w])   # u <- v, v <- w, w <- u
    ok1 = (cols[w] != ou) & (cols[v] != other[w]) & (cols[u] != other[v])   # u <- w, w <- v, v <- u
    rows_out = []
    order = np.argsort(v * m + w, kind="stable")  # (v, w) lexicographic, as the reference's double loop
    v, w, ok0, ok1 = v[order], w[order], ok0[order], ok1[order]
    cand = np.stack([np.full_like(v, u), v, w, np.zeros_like(v)], axis=1)
    c0 = cand[ok0]
    c1 = cand[ok1].copy()
    c1[:, 3] = 1
    # interleave per (v, w): orientation 0 first, then 1
    key0 = np.flatnonzero(ok0) * 2
    key1 = np.flatnonzero(ok1) * 2 + 1
    merged = np.concatenate([c0, c1])[np.argsort(np.concatenate([key0, key1]), kind="stable")]
    rows_out.append(merged)
    return np.ascontiguousarray(np.concatenate(rows_out).astype(np.int32).reshape(-1, 4))


def apply_three_cycle(p, q, candidate) -> PairPermutations:
    """search.py:68-81: move the three columns around and re-colour into P/Q."""
    n = len(p)
    u, v, w, orientation = (int(x) for x in candidate)
    cols = _columns(p, q)
    points = [(s // 2, int(cols[s])) for s in range(2 * n)]
    old = [(s // 2, int(cols[s])) for s in (u, v, w)]
    for point in old:
        points.remove(point)
    if orientation == 0:
        new_columns = (old[1][1], old[2][1], old[0][1])
    else:
        new_columns = (old[2][1], old[0][1], old[1][1])
    points.extend((old[i][0], new_columns[i]) for i in range(3))
    return points_to_permutations(points, n)
