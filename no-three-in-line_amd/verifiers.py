"""The two exact verifiers of a candidate (drop-in for verifier_a.py and
verifier_b.py): the same structural checks, messages and first witness as
the reference's, found by vectorised host searches.  Verifier A tests
integer determinants of index triples in lexicographic order; verifier B
looks, around each anchor in order, for the first repeated primitive
undirected direction."""
from __future__ import annotations

from math import comb, gcd

import numpy as np

from .state import Point


def _structure_a(points: list[Point], n: int) -> str | None:
    """verifier_a.py:19-34."""
    if len(points) != 2 * n:
        return f"expected {2 * n} points, got {len(points)}"
    if len(set(points)) != len(points):
        return "duplicate point"
    if any(not (0 <= r < n and 0 <= c < n) for r, c in points):
        return "point outside grid"
    rows, cols = np.bincount([r for r, _ in points], minlength=n), np.bincount([c for _, c in points], minlength=n)
    if np.any(rows != 2):
        return "row degree is not exactly two"
    if np.any(cols != 2):
        return "column degree is not exactly two"
    return None


def collinear_a(a: Point, b: Point, c: Point) -> bool:
    """verifier_a.py:8-9."""
    return (b[0] - a[0]) * (c[1] - a[1]) == (b[1] - a[1]) * (c[0] - a[0])


def first_collinear_triple_a(points: list[Point]) -> tuple[Point, Point, Point] | None:
    """verifier_a.py:12-16: the first index triple i < j < k (lexicographic)
    whose points are collinear."""
    pts = np.asarray(points, dtype=np.int64).reshape(-1, 2)
    k = len(pts)
    if k < 3:
        return None
    dr = pts[None, :, 0] - pts[:, None, 0]          # [i][j] = row_j - row_i
    dc = pts[None, :, 1] - pts[:, None, 1]
    for i in range(k - 2):
        # det(i, j, l) for j, l > i: dr[i, j] dc[i, l] - dc[i, j] dr[i, l]
        det = np.outer(dr[i, i + 1:], dc[i, i + 1:]) - np.outer(dc[i, i + 1:], dr[i, i + 1:])
        hits = np.argwhere(np.triu(det == 0, 1))
        if len(hits):
            j, l = hits[0]
            return points[i], points[i + 1 + j], points[i + 1 + l]
    return None


def verify_a(points: list[Point], n: int) -> tuple[bool, str]:
    """verifier_a.verify (verifier_a.py:19-38)."""
    problem = _structure_a(points, n)
    if problem:
        return False, problem
    witness = first_collinear_triple_a(points)
    if witness is not None:
        return False, f"collinear triple: {witness}"
    return True, f"checked {comb(len(points), 3)} integer determinants"


def primitive_undirected_direction(a: Point, b: Point) -> tuple[int, int]:
    """verifier_b.py:8-19."""
    dx, dy = b[0] - a[0], b[1] - a[1]
    g = gcd(abs(dx), abs(dy))
    if g == 0:
        raise ValueError("duplicate point has no direction")
    dx, dy = dx // g, dy // g
    if dx < 0 or (dx == 0 and dy < 0):
        dx, dy = -dx, -dy
    return dx, dy


def first_collinear_triple_b(points: list[Point]) -> tuple[Point, Point, Point] | None:
    """verifier_b.py:22-33: around each anchor in order, the first other
    point whose direction repeats an earlier one's."""
    for a_index, anchor in enumerate(points):
        seen: dict[tuple[int, int], Point] = {}
        for o_index, other in enumerate(points):
            if o_index == a_index:
                continue
            d = primitive_undirected_direction(anchor, other)
            previous = seen.get(d)
            if previous is not None:
                return anchor, previous, other
            seen[d] = other
    return None


def verify_b(points: list[Point], n: int) -> tuple[bool, str]:
    """verifier_b.verify (verifier_b.py:36-57)."""
    if len(points) != n + n:
        return False, f"cardinality mismatch: {len(points)} versus {2 * n}"
    if len(set(points)) != len(points):
        return False, "coordinates are not unique"
    rows, cols = [0] * n, [0] * n
    for point in points:
        r, c = point
        if not (0 <= r < n and 0 <= c < n):
            return False, f"out-of-range coordinate {point}"
        rows[r] += 1
        cols[c] += 1
    if set(rows) != {2}:
        return False, "not every row occurs twice"
    if set(cols) != {2}:
        return False, "not every column occurs twice"
    witness = first_collinear_triple_b(points)
    if witness:
        return False, f"repeated normalized direction at {witness[0]}: {witness}"
    return True, f"checked normalized directions around {len(points)} anchors"
