"""Lines of a point set: the canonical integer line through two points and
the lines holding three or more points (geometry.py:14-30,
conflicts.py:36-72), as the breakout search weights them."""
from __future__ import annotations

from dataclasses import dataclass
from math import comb, gcd

import numpy as np

from .state import Point

LineKey = tuple[int, int, int]


def line_key(a: Point, b: Point) -> LineKey:
    """geometry.py:14-30: (A, B, C) with A row + B column + C = 0 through
    both points, divided by gcd(|A|, |B|, |C|), signed so that A > 0 or
    A = 0 < B."""
    if a == b:
        raise ValueError("two distinct points are required")
    coef_a = b[1] - a[1]
    coef_b = a[0] - b[0]
    coef_c = -(coef_a * a[0] + coef_b * a[1])
    g = gcd(gcd(abs(coef_a), abs(coef_b)), abs(coef_c))
    coef_a, coef_b, coef_c = coef_a // g, coef_b // g, coef_c // g
    if coef_a < 0 or (coef_a == 0 and coef_b < 0):
        coef_a, coef_b, coef_c = -coef_a, -coef_b, -coef_c
    return coef_a, coef_b, coef_c


@dataclass(frozen=True)
class ViolatingLine:
    """conflicts.py:14-18."""

    key: LineKey
    points: tuple[Point, ...]
    triples: int


def violating_lines(points: list[Point]) -> list[ViolatingLine]:
    """Every line holding >= 3 of the points with its C(k, 3) triples, in the
    order (-triples, key) of conflicts.py:66."""
    pts = np.asarray(points, dtype=np.int64).reshape(-1, 2)
    k = len(pts)
    if k < 3:
        return []
    i, j = np.triu_indices(k, 1)
    a = pts[j, 1] - pts[i, 1]
    b = pts[i, 0] - pts[j, 0]
    c = -(a * pts[i, 0] + b * pts[i, 1])
    g = np.gcd(np.gcd(np.abs(a), np.abs(b)), np.abs(c))
    if np.any(g == 0):
        raise ValueError("two distinct points are required")
    a, b, c = a // g, b // g, c // g
    flip = (a < 0) | ((a == 0) & (b < 0))
    a, b, c = np.where(flip, -a, a), np.where(flip, -b, b), np.where(flip, -c, c)
    members: dict[LineKey, set[int]] = {}
    for key_a, key_b, key_c, left, right in zip(a.tolist(), b.tolist(), c.tolist(), i.tolist(), j.tolist()):
        s = members.setdefault((key_a, key_b, key_c), set())
        s.add(left)
        s.add(right)
    out = [ViolatingLine(key, tuple(sorted(tuple(int(v) for v in pts[x]) for x in idx)), comb(len(idx), 3))
           for key, idx in members.items() if len(idx) >= 3]
    out.sort(key=lambda line: (-line.triples, line.key))
    return out


def primitive_slope(key: LineKey) -> tuple[int, int]:
    """geometry.py:33-40: the line's direction (dx, dy) = (B, -A) reduced,
    signed so that dx > 0 or dx = 0 < dy."""
    dx, dy = key[1], -key[0]
    g = gcd(abs(dx), abs(dy))
    dx, dy = dx // g, dy // g
    if dx < 0 or (dx == 0 and dy < 0):
        dx, dy = -dx, -dy
    return dx, dy
