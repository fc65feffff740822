"""Conflict analysis of a candidate (drop-in for conflicts.py:1-83): the
violating lines with their triples, per-point badness C(k - 1, 2) summed
over the lines through the point, an optional line-weighted score and the
slope histogram; the full triple count and the recount delta of a
same-layer swap."""
from __future__ import annotations

from collections import Counter
from dataclasses import dataclass
from itertools import combinations
from math import comb

import numpy as np

from .lines import LineKey, ViolatingLine, line_key, primitive_slope, violating_lines
from .state import PairPermutations, Point


@dataclass(frozen=True)
class ConflictMetrics:
    """conflicts.py:21-33."""

    triples: int
    violating_lines: int
    max_line_occupancy: int
    max_point_badness: int
    weighted_score: int
    badness: tuple[int, ...]
    lines: tuple[ViolatingLine, ...]
    slope_histogram: dict[tuple[int, int], int]

    @property
    def lexicographic(self) -> tuple[int, int, int, int]:
        return self.triples, self.violating_lines, self.max_point_badness, self.weighted_score


def analyze_points(points: list[Point], line_weights: dict[LineKey, int] | None = None) -> ConflictMetrics:
    """conflicts.py:36-72 (points must be distinct)."""
    lines = violating_lines(points)
    index = {tuple(p): i for i, p in enumerate(points)}
    badness = [0] * len(points)
    slopes: Counter[tuple[int, int]] = Counter()
    triples = weighted = 0
    for line in lines:
        triples += line.triples
        weighted += (line_weights.get(line.key, 1) if line_weights else 1) * line.triples
        per_point = comb(len(line.points) - 1, 2)
        for point in line.points:
            badness[index[point]] += per_point
        slopes[primitive_slope(line.key)] += line.triples
    return ConflictMetrics(triples=triples, violating_lines=len(lines),
                           max_line_occupancy=max((len(line.points) for line in lines), default=2),
                           max_point_badness=max(badness, default=0), weighted_score=weighted,
                           badness=tuple(badness), lines=tuple(lines), slope_histogram=dict(slopes))


def full_triple_score(points: list[Point]) -> int:
    """conflicts.py:75-76: the number of index triples whose points are
    collinear (zero determinant), duplicates included."""
    pts = np.asarray(points, dtype=np.int64).reshape(-1, 2)
    if len(pts) < 3:
        return 0
    if len({tuple(p) for p in pts.tolist()}) == len(pts):
        return sum(line.triples for line in violating_lines([tuple(p) for p in pts.tolist()]))
    total = 0
    for a, b, c in combinations(range(len(pts)), 3):
        (ar, ac), (br, bc), (cr, cc) = pts[a], pts[b], pts[c]
        total += (br - ar) * (cc - ac) == (bc - ac) * (cr - ar)
    return int(total)


def reference_swap_delta(state: PairPermutations, layer: int, row_a: int, row_b: int) -> int:
    """conflicts.py:79-83: the triple-count change of swapping rows a and b
    of one layer, by recount."""
    return full_triple_score(state.swapped(layer, row_a, row_b).points()) - full_triple_score(state.points())


__all__ = ["ConflictMetrics", "analyze_points", "full_triple_score", "reference_swap_delta", "line_key"]
