"""Tabu / annealing local search over 2-switches with 3-cycle plateau
escapes (drop-in for search.py:84-268).

The driver is the reference's host loop with the same random.Random(seed)
call sequence; its three evaluators default to the GPU kernels: the full
score and badness, the 2-switch scores of one selected slot (the
`move_evaluator` hook, search.py:94,124) and the 3-cycle scores of the
plateau escape.  Every evaluator returns the reference's exact integers, so
a run is the same trajectory whichever (exact) evaluators it is given.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
from random import Random
from typing import Callable

import numpy as np

from .gpu import score_and_badness as gpu_score_and_badness
from .gpu_delta import evaluate_three_cycles as gpu_evaluate_three_cycles
from .gpu_delta import evaluate_two_switches as gpu_evaluate_two_switches
from .moves import apply_three_cycle, legal_three_cycle_candidates
from .state import PairPermutations, points_to_permutations, random_state

LEGAL_BELOW = 1 << 29  # search.py:125: sentinel scores (1<<30 or 1<<60) are illegal moves

Scorer = Callable[[np.ndarray, np.ndarray], tuple[int, np.ndarray]]
MoveEvaluator = Callable[[np.ndarray, np.ndarray, int, int, np.ndarray], np.ndarray]
CycleEvaluator = Callable[[np.ndarray, np.ndarray, np.ndarray, int, np.ndarray], np.ndarray]


@dataclass(frozen=True)
class SearchSnapshot:
    """search.py:15-24."""

    iteration: int
    elapsed_seconds: float
    score: int
    best_score: int
    layer: int
    row_a: int
    row_b: int
    partner_layer: int
    state: PairPermutations


@dataclass(frozen=True)
class SearchResult:
    """search.py:27-34."""

    state: PairPermutations
    score: int
    iterations: int
    elapsed_seconds: float
    moves_evaluated: int
    seed: int


def _column(p: np.ndarray, q: np.ndarray, slot: int) -> int:
    row = slot >> 1
    return int(q[row] if slot & 1 else p[row])


def switch_points(p: np.ndarray, q: np.ndarray, a: int, b: int) -> PairPermutations:
    """The state after exchanging the columns of slots a and b, recoloured by
    points_to_permutations from the reference's point-list order
    (search.py:147-155)."""
    n = len(p)
    points = [(r, int(p[r])) for r in range(n)] + [(r, int(q[r])) for r in range(n)]
    point_a, point_b = (a >> 1, _column(p, q, a)), (b >> 1, _column(p, q, b))
    points.remove(point_a)
    points.remove(point_b)
    points.append((point_a[0], point_b[1]))
    points.append((point_b[0], point_a[1]))
    return points_to_permutations(points, n)


def _weighted_pick(rng: Random, badness: np.ndarray, exploration_probability: float) -> int:
    """search.py:112-122: a uniform slot with the exploration probability (or
    when every weight is zero), else a slot drawn in proportion to
    badness + 0.05 by a running double-precision sum."""
    weights = badness.astype(np.float64) + 0.05
    if rng.random() < exploration_probability or float(weights.sum()) == 0.0:
        return rng.randrange(len(weights))
    threshold = rng.random() * float(weights.sum())
    reached = np.flatnonzero(np.cumsum(weights) >= threshold)
    return int(reached[0]) if len(reached) else 0


def local_search(n: int, *, seed: int, iterations: int, initial: PairPermutations | None = None,
                 tabu_tenure: int = 11, plateau: int = 800, exploration_probability: float = 0.04,
                 checkpoint: Callable[[SearchSnapshot], None] | None = None,
                 move_evaluator: MoveEvaluator = gpu_evaluate_two_switches,
                 scorer: Scorer = gpu_score_and_badness,
                 cycle_evaluator: CycleEvaluator = gpu_evaluate_three_cycles) -> SearchResult:
    """search.py:84-268."""
    rng = Random(seed)
    state = initial if initial is not None else random_state(n, rng)
    state.validate()
    p = np.asarray(state.P, dtype=np.int32).copy()
    q = np.asarray(state.Q, dtype=np.int32).copy()
    current, badness = scorer(p, q)
    best_p, best_q, best = p.copy(), q.copy(), int(current)
    tabu_until: dict[tuple[int, int, int], int] = {}
    last_improvement = 0
    moves_evaluated = 0
    started = time.perf_counter()

    def snapshot(iteration, layer, row_a, row_b, partner_layer):
        if checkpoint is not None:
            checkpoint(SearchSnapshot(iteration, time.perf_counter() - started, int(current), best, layer, row_a,
                                      row_b, partner_layer, PairPermutations(tuple(map(int, p)), tuple(map(int, q)))))

    iteration = 0
    for iteration in range(1, iterations + 1):
        selected = _weighted_pick(rng, badness, exploration_probability)
        selected_row, layer = divmod(selected, 2)
        scores = move_evaluator(p, q, selected, int(current), badness)
        legal = [i for i in range(2 * n) if scores[i] < LEGAL_BELOW]
        moves_evaluated += len(legal)
        if not legal:
            continue
        legal.sort(key=lambda i: (int(scores[i]), rng.random()))
        chosen = next((i for i in legal
                       if tabu_until.get((min(selected, i), max(selected, i), 2), -1) <= iteration
                       or int(scores[i]) < best), legal[0])
        proposed = int(scores[chosen])
        temperature = max(0.05, 1.5 * (1.0 - min(1.0, (iteration - last_improvement) / max(1, plateau))))
        delta = proposed - int(current)
        if delta <= 0 or rng.random() < math.exp(-delta / temperature):
            partner_row, partner_layer = divmod(chosen, 2)
            moved = switch_points(p, q, selected, chosen)
            p, q = np.asarray(moved.P, dtype=np.int32), np.asarray(moved.Q, dtype=np.int32)
            current, badness = scorer(p, q)
            if int(current) != proposed:
                raise AssertionError(f"incremental score {proposed} != full score {current}")
            tabu_until[(min(selected, chosen), max(selected, chosen), 2)] = iteration + tabu_tenure
            if int(current) < best:
                best, best_p, best_q = int(current), p.copy(), q.copy()
                last_improvement = iteration
                snapshot(iteration, layer, selected_row, partner_row, partner_layer)
                if best == 0:
                    break

        if iteration - last_improvement >= plateau:
            # plateau escape: the best 3-cycle around a maximally bad slot
            worst = int(badness.max())
            anchors = [i for i in range(2 * n) if int(badness[i]) == worst]
            anchor = rng.choice(anchors) if anchors else rng.randrange(2 * n)
            candidates = legal_three_cycle_candidates(p, q, anchor)
            used_cycle = False
            if len(candidates):
                cycle_scores = cycle_evaluator(p, q, candidates, int(current), badness)
                moves_evaluated += len(candidates)
                pick = int(np.argmin(cycle_scores))
                cycle_score = int(cycle_scores[pick])
                if cycle_score <= int(current) + 3:
                    moved = apply_three_cycle(p, q, candidates[pick])
                    p, q = np.asarray(moved.P, dtype=np.int32), np.asarray(moved.Q, dtype=np.int32)
                    current, badness = scorer(p, q)
                    if int(current) != cycle_score:
                        raise AssertionError(f"three-cycle score {cycle_score} != full score {current}")
                    used_cycle = True
                    if int(current) < best:
                        best, best_p, best_q = int(current), p.copy(), q.copy()
                        cycle = candidates[pick]
                        snapshot(iteration, int(cycle[0]) % 2, int(cycle[0]) // 2, int(cycle[1]) // 2,
                                 int(cycle[1]) % 2)
            if used_cycle:
                last_improvement = iteration
                if best == 0:
                    break
                continue
            # otherwise 2-8 random legal-ish 2-switches (search.py:232-256)
            for _ in range(2 + rng.randrange(7)):
                for _attempt in range(100):
                    a, b = rng.randrange(2 * n), rng.randrange(2 * n)
                    if a >> 1 == b >> 1:
                        continue
                    if _column(p, q, b) == _column(p, q, a ^ 1) or _column(p, q, a) == _column(p, q, b ^ 1):
                        continue
                    moved = switch_points(p, q, a, b)
                    p, q = np.asarray(moved.P, dtype=np.int32), np.asarray(moved.Q, dtype=np.int32)
                    break
            current, badness = scorer(p, q)
            last_improvement = iteration

    return SearchResult(PairPermutations(tuple(map(int, best_p)), tuple(map(int, best_q))), best,
                        iteration if iterations else 0, time.perf_counter() - started, moves_evaluated, seed)
