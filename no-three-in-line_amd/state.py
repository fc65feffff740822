"""Host-side P/Q state model (the reference's state.py:9-95), used by the
tempering driver to validate and normalise states; no scoring happens here.
"""
from __future__ import annotations

from dataclasses import dataclass
from random import Random

import numpy as np

Point = tuple[int, int]


@dataclass(frozen=True)
class PairPermutations:
    """state.py:9-44: permutations P, Q with P[r] != Q[r]."""

    P: tuple[int, ...]
    Q: tuple[int, ...]

    @property
    def n(self) -> int:
        return len(self.P)

    def validate(self) -> None:
        n = self.n
        expected = set(range(n))
        if len(self.Q) != n:
            raise ValueError("P and Q have different lengths")
        if set(self.P) != expected:
            raise ValueError("P is not a permutation")
        if set(self.Q) != expected:
            raise ValueError("Q is not a permutation")
        collisions = [row for row in range(n) if self.P[row] == self.Q[row]]
        if collisions:
            raise ValueError(f"P/Q duplicate edges at rows {collisions}")

    def points(self) -> list[Point]:
        self.validate()
        return sorted([(r, self.P[r]) for r in range(self.n)] + [(r, self.Q[r]) for r in range(self.n)])

    def swapped(self, layer: int, row_a: int, row_b: int) -> "PairPermutations":
        if layer not in (0, 1):
            raise ValueError("layer must be 0 (P) or 1 (Q)")
        p, q = list(self.P), list(self.Q)
        target = p if layer == 0 else q
        target[row_a], target[row_b] = target[row_b], target[row_a]
        out = PairPermutations(tuple(p), tuple(q))
        out.validate()
        return out

    def columns(self) -> np.ndarray:
        """Interleaved slot columns [P0, Q0, P1, Q1, ...] (tempering.py:254)."""
        return np.column_stack((self.P, self.Q)).astype(np.int32).reshape(-1)


def points_to_permutations(points: list[Point], n: int) -> PairPermutations:
    """state.py:47-85: colour a row/column 2-factor into two perfect matchings
    with Kuhn's augmenting paths over sorted adjacency lists."""
    if len(points) != 2 * n or len(set(points)) != len(points):
        raise ValueError("a 2-factor needs exactly 2n distinct points")
    adjacency: list[list[int]] = [[] for _ in range(n)]
    column_degree = [0] * n
    for row, column in points:
        if not (0 <= row < n and 0 <= column < n):
            raise ValueError(f"point {(row, column)} outside grid")
        adjacency[row].append(column)
        column_degree[column] += 1
    if any(len(a) != 2 for a in adjacency) or any(d != 2 for d in column_degree):
        raise ValueError("points are not a row/column 2-factor")
    for a in adjacency:
        a.sort()
    owner = [-1] * n

    def augment(row: int, seen: list[bool]) -> bool:
        for column in adjacency[row]:
            if seen[column]:
                continue
            seen[column] = True
            previous = owner[column]
            if previous == -1 or augment(previous, seen):
                owner[column] = row
                return True
        return False

    for row in range(n):
        if not augment(row, [False] * n):
            raise ValueError("2-factor unexpectedly has no perfect matching")
    p = [-1] * n
    for column, row in enumerate(owner):
        p[row] = column
    q = [next(c for c in adjacency[r] if c != p[r]) for r in range(n)]
    state = PairPermutations(tuple(p), tuple(q))
    state.validate()
    return state


def random_state(n: int, rng: Random) -> PairPermutations:
    """state.py:88-95 (rejection sampling of Q)."""
    p = list(range(n))
    q = list(range(n))
    rng.shuffle(p)
    while True:
        rng.shuffle(q)
        if all(p[r] != q[r] for r in range(n)):
            return PairPermutations(tuple(p), tuple(q))


def check_columns(columns: np.ndarray) -> None:
    """Vectorised 2-factor check of interleaved states int32[R][2n]: every
    row's two columns differ and every column is used exactly twice (the
    conditions points_to_permutations enforces, tempering.py:132)."""
    r, m = columns.shape
    n = m // 2
    if columns.min(initial=0) < 0 or columns.max(initial=0) >= n:
        raise ValueError("points are outside the grid")
    if np.any(columns[:, 0::2] == columns[:, 1::2]):
        raise ValueError("a 2-factor needs exactly 2n distinct points")
    degree = np.zeros((r, n), dtype=np.int64)
    np.add.at(degree, (np.repeat(np.arange(r), m), columns.reshape(-1)), 1)
    if np.any(degree != 2):
        raise ValueError("points are not a row/column 2-factor")


def random_two_switches(columns: np.ndarray, moves: int, seed: int) -> np.ndarray:
    """Copies of interleaved states int32[R][2n] after `moves` random legal
    2-switches each (the tempering walk's move: swap the columns of two points
    in different rows so that every row keeps two distinct columns; the column
    counts are unchanged).  Seeded NumPy draws; illegal draws are redrawn."""
    cols = np.array(columns, dtype=np.int32, copy=True)
    if cols.ndim != 2 or cols.shape[1] % 2 or moves < 0:
        raise ValueError("columns must be int32[R][2n] and moves non-negative")
    check_columns(cols)
    r, m = cols.shape
    rng = np.random.default_rng(seed)
    rows = np.arange(r)
    for _ in range(moves):
        pending = rows.copy()
        while len(pending):
            u = rng.integers(0, m, len(pending))
            v = rng.integers(0, m, len(pending))
            cu, cv = cols[pending, u], cols[pending, v]
            ok = ((u >> 1) != (v >> 1)) & (cu != cv) & (cols[pending, u ^ 1] != cv) & (cols[pending, v ^ 1] != cu)
            done = pending[ok]
            cols[done, u[ok]] = cv[ok]
            cols[done, v[ok]] = cu[ok]
            pending = pending[~ok]
    return cols
