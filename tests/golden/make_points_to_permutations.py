"""Writes points_to_permutations.json: input point sets and the P/Q colouring
that the reference's `state.points_to_permutations` (state.py:47-85) returns
for them.

Which of the 2^(c-1) colourings of a c-component 2-factor comes back depends
on the order of its augmenting-path search, so the fixture records the
reference's answer.  This script imports the reference's pure-Python module
in the development container only (its source tree is not shipped); the
tests read the JSON it writes.  Run:  python tests/golden/make_points_to_permutations.py
"""
from __future__ import annotations

import json
import random
import sys
from pathlib import Path

REFERENCE_SRC = Path("/root/reference/src")
OUT = Path(__file__).resolve().parent / "points_to_permutations.json"


def random_two_factor(n: int, parts: int, rng: random.Random) -> list[tuple[int, int]]:
    """Points of a 2-regular row/column set with at least `parts` components:
    the rows and columns are cut into `parts` blocks (blocks of at least 2),
    each block gets two collision-free permutations, then rows and columns
    are relabelled at random."""
    cuts = sorted(rng.sample(range(2, n - 1), parts - 1)) if parts > 1 else []
    bounds, lo = [], 0
    for c in cuts:
        if c - lo >= 2 and n - c >= 2:
            bounds.append((lo, c))
            lo = c
    bounds.append((lo, n))
    points = []
    for lo, hi in bounds:
        size = hi - lo
        while True:
            p = list(range(size))
            q = list(range(size))
            rng.shuffle(p)
            rng.shuffle(q)
            if all(a != b for a, b in zip(p, q)):
                break
        for r in range(size):
            points.append((lo + r, lo + p[r]))
            points.append((lo + r, lo + q[r]))
    row_map = list(range(n))
    col_map = list(range(n))
    rng.shuffle(row_map)
    rng.shuffle(col_map)
    points = [(row_map[r], col_map[c]) for r, c in points]
    rng.shuffle(points)
    return points


def main() -> None:
    sys.path.insert(0, str(REFERENCE_SRC))
    from no_three_75.state import points_to_permutations  # the reference module

    rng = random.Random(20261204)
    cases = []
    for n in (5, 6, 7, 8, 10, 12, 16, 21, 32, 40, 52, 61, 75):
        for parts in (1, 2, 3, 5):
            if 2 * parts > n:
                continue
            for _ in range(3):
                points = random_two_factor(n, parts, rng)
                state = points_to_permutations(points, n)
                cases.append({"n": n, "points": [list(p) for p in points], "P": list(state.P), "Q": list(state.Q)})
    OUT.write_text(json.dumps({"source": "reference state.points_to_permutations (state.py:47-85)",
                               "cases": cases}, separators=(",", ":")) + "\n")
    print(f"wrote {len(cases)} cases to {OUT}")


if __name__ == "__main__":
    main()
