"""Breakout local search: steepest 2-switch descent on a line-weighted score,
raising the weights of the violating lines at every plateau (drop-in for
breakout.py:84-211).

The host loop and its random.Random(seed) calls follow the reference; the
full score and the raw 2-switch scores of the selected slot come from the
GPU (the `move_evaluator` hook and `scorer`).  Line weights change the
scores on the host: a weighted line of weight w with k points adds
(w - 1) C(k, 3) to the state's score.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from math import comb
from random import Random
from typing import Callable

import numpy as np

from .gpu import score_and_badness as gpu_score_and_badness
from .gpu_delta import evaluate_two_switches as gpu_evaluate_two_switches
from .lines import LineKey, violating_lines
from .search import LEGAL_BELOW, MoveEvaluator, Scorer, SearchSnapshot, switch_points
from .state import PairPermutations


@dataclass(frozen=True)
class BreakoutResult:
    """breakout.py:18-27."""

    state: PairPermutations
    score: int
    iterations: int
    elapsed_seconds: float
    moves_evaluated: int
    breakouts: int
    weighted_lines: int
    seed: int


def _slots(p: np.ndarray, q: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    rows = np.repeat(np.arange(len(p), dtype=np.int64), 2)
    columns = np.empty(2 * len(p), dtype=np.int64)
    columns[0::2] = p
    columns[1::2] = q
    return rows, columns


def _active(weights: dict[LineKey, int]) -> tuple[np.ndarray, np.ndarray] | None:
    """Coefficients [L, 3] and extra weights w - 1 [L] of the lines with w > 1."""
    items = [(key, w - 1) for key, w in weights.items() if w > 1]
    if not items:
        return None
    return (np.asarray([k for k, _ in items], dtype=np.int64).reshape(-1, 3),
            np.asarray([x for _, x in items], dtype=np.int64))


def weighted_candidate_scores(p: np.ndarray, q: np.ndarray, selected_index: int, raw_scores: np.ndarray,
                              weights: dict[LineKey, int]) -> np.ndarray:
    """breakout.py:39-68: raw 2-switch scores of slot u plus the weighted
    lines' extra (w - 1) C(k', 3) for the occupancy k' of each line after the
    switch (u to v's column in u's row, v to u's column in v's row)."""
    out = np.asarray(raw_scores).astype(np.int64, copy=True)
    active = _active(weights)
    if active is None:
        return out
    coef, extra = active
    rows, columns = _slots(p, q)
    a, b, c = coef[:, 0:1], coef[:, 1:2], coef[:, 2:3]
    on_line = a * rows[None, :] + b * columns[None, :] + c == 0            # [L, m]
    occupancy = on_line.sum(axis=1, keepdims=True)
    u = selected_index
    u_moved = a * rows[u] + b * columns[None, :] + c == 0                   # u at column of v
    v_moved = a * rows[None, :] + b * columns[u] + c == 0                   # v at column of u
    after = occupancy - on_line[:, u:u + 1] - on_line + u_moved + v_moved
    triples = np.where(after >= 3, after * (after - 1) * (after - 2) // 6, 0)
    return out + (extra[:, None] * triples).sum(axis=0)


def weighted_state_score(p: np.ndarray, q: np.ndarray, raw_score: int, weights: dict[LineKey, int]) -> int:
    """breakout.py:71-80."""
    rows, columns = _slots(p, q)
    value = int(raw_score)
    for (ka, kb, kc), w in weights.items():
        if w <= 1:
            continue
        k = int(np.count_nonzero(ka * rows + kb * columns + kc == 0))
        if k >= 3:
            value += (w - 1) * comb(k, 3)
    return value


def _weighted_badness(p: np.ndarray, q: np.ndarray, badness: np.ndarray, weights: dict[LineKey, int]) -> np.ndarray:
    """breakout.py:116-130: badness plus, for every weighted line through a
    slot, (w - 1) C(k - 1, 2)."""
    out = np.asarray(badness).astype(np.int64, copy=True)
    active = _active(weights)
    if active is None:
        return out
    coef, extra = active
    rows, columns = _slots(p, q)
    on_line = coef[:, 0:1] * rows[None, :] + coef[:, 1:2] * columns[None, :] + coef[:, 2:3] == 0
    k = on_line.sum(axis=1)
    per_point = extra * np.maximum(0, (k - 1) * (k - 2) // 2)
    return out + (per_point[:, None] * on_line).sum(axis=0)


def breakout_search(initial: PairPermutations, *, seed: int, iterations: int, plateau: int, weight_increment: int,
                    max_weighted_lines: int = 128, checkpoint: Callable[[SearchSnapshot], None] | None = None,
                    move_evaluator: MoveEvaluator = gpu_evaluate_two_switches,
                    scorer: Scorer = gpu_score_and_badness) -> BreakoutResult:
    """breakout.py:84-211."""
    initial.validate()
    rng = Random(seed)
    n = initial.n
    p = np.asarray(initial.P, dtype=np.int32).copy()
    q = np.asarray(initial.Q, dtype=np.int32).copy()
    raw, badness = scorer(p, q)
    best, best_p, best_q = int(raw), p.copy(), q.copy()
    weights: dict[LineKey, int] = {}
    weighted = int(raw)
    last_raw_improvement = 0
    moves_evaluated = 0
    breakouts = 0
    started = time.perf_counter()
    iteration = 0
    for iteration in range(1, iterations + 1):
        wbad = _weighted_badness(p, q, badness, weights)
        if rng.random() < 0.05 or int(wbad.max()) == 0:
            selected = rng.randrange(2 * n)
        else:
            top = int(wbad.max())
            selected = rng.choice([i for i in range(2 * n) if int(wbad[i]) == top])
        raw_scores = move_evaluator(p, q, selected, int(raw), badness)
        scores = weighted_candidate_scores(p, q, selected, raw_scores, weights)
        legal = np.flatnonzero(scores < LEGAL_BELOW)
        moves_evaluated += len(legal)
        if len(legal):
            minimum = int(scores[legal].min())
            chosen = rng.choice([int(i) for i in legal if int(scores[i]) == minimum])
            if minimum <= weighted:
                row_a, layer_a = divmod(selected, 2)
                row_b, layer_b = divmod(chosen, 2)
                moved = switch_points(p, q, selected, chosen)
                p, q = np.asarray(moved.P, dtype=np.int32), np.asarray(moved.Q, dtype=np.int32)
                raw, badness = scorer(p, q)
                weighted = weighted_state_score(p, q, int(raw), weights)
                if weighted != minimum:
                    raise AssertionError(f"weighted delta {minimum} != recomputed score {weighted}")
                if int(raw) < best:
                    best, best_p, best_q = int(raw), p.copy(), q.copy()
                    last_raw_improvement = iteration
                    if checkpoint is not None:
                        checkpoint(SearchSnapshot(iteration, time.perf_counter() - started, int(raw), best, layer_a,
                                                  row_a, row_b, layer_b,
                                                  PairPermutations(tuple(map(int, p)), tuple(map(int, q)))))
                    if best == 0:
                        break
        if iteration - last_raw_improvement >= plateau:
            lines = violating_lines([(r, int(p[r])) for r in range(n)] + [(r, int(q[r])) for r in range(n)])
            for line in lines:
                weights[line.key] = weights.get(line.key, 1) + weight_increment
            if breakouts and breakouts % 8 == 0:
                # decay: lines no longer violated lose one unit; weight 1 is dropped
                active_keys = {line.key for line in lines}
                weights = {key: w - (key not in active_keys) for key, w in weights.items()
                           if w - (key not in active_keys) > 1}
            if len(weights) > max_weighted_lines:
                weights = dict(sorted(weights.items(), key=lambda item: (-item[1], item[0]))[:max_weighted_lines])
            weighted = weighted_state_score(p, q, int(raw), weights)
            breakouts += 1
            last_raw_improvement = iteration
    return BreakoutResult(PairPermutations(tuple(map(int, best_p)), tuple(map(int, best_q))), best,
                          iteration if iterations else 0, time.perf_counter() - started, moves_evaluated, breakouts,
                          len(weights), seed)
