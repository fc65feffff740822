"""Device GA search over the scorer (drop-in for gpu_device_search.py).

The mutation kernel, the scorer and the elite keys run on the GPU; selection
keeps the reference's host logic (stable score order, a quality share, a
shuffled energy window from random.Random(seed), lexicographic elite rank).
The mutation draws are Philox (child, generation, draw, "MUTA") instead of
the reference's xorshift64*, so children differ from the reference's for the
same seed while following the same law.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from random import Random
from typing import Callable

import numpy as np
import torch

from . import _lib
from .gpu import check_status, score_population_device
from .state import PairPermutations, points_to_permutations

DEFAULT_BAD_BIAS = 85  # gpu_device_search.py:238


@dataclass(frozen=True)
class GpuSearchSnapshot:
    """gpu_search.py:17-25."""

    generation: int
    elapsed_seconds: float
    evaluations: int
    generation_best: int
    best_score: int
    elite_scores: tuple[int, ...]
    elite_values: np.ndarray
    state: PairPermutations


@dataclass(frozen=True)
class GpuSearchResult:
    """gpu_search.py:28-35."""

    state: PairPermutations
    score: int
    generations: int
    evaluations: int
    elapsed_seconds: float
    seed: int


def normalize_raw(p: np.ndarray, q: np.ndarray) -> PairPermutations:
    """gpu_device_search.py:179-182: the canonical P/Q colouring of a raw
    (possibly layer-mixed) 2-factor."""
    n = int(p.shape[0])
    points = [(row, int(p[row])) for row in range(n)] + [(row, int(q[row])) for row in range(n)]
    return points_to_permutations(points, n)


def elite_key(score: int, badness: np.ndarray) -> tuple[int, int, int]:
    """gpu_device_search.py:200-202."""
    values = np.asarray(badness).astype(np.int64)
    return int(score), int(values.max()), int(np.dot(values, values))


def rank_elites(scores: np.ndarray, badness: np.ndarray) -> np.ndarray:
    """gpu_device_search.py:185-197: order by score, then max point badness,
    then squared badness mass; ties keep index order."""
    scores = np.asarray(scores).astype(np.int64)
    bad = np.asarray(badness).astype(np.int64)
    if len(scores) == 0:
        return np.empty(0, dtype=np.int32)
    return np.lexsort((np.einsum("ij,ij->i", bad, bad), bad.max(axis=1), scores)).astype(np.int32)


def elite_keys_device(badness: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """Max badness (int32) and squared badness mass (int64) per row of a
    device tensor int32[P][2n]."""
    count, m = int(badness.shape[0]), int(badness.shape[1])
    if badness.dtype != torch.int32 or not badness.is_cuda or not badness.is_contiguous() or m % 2:
        raise ValueError("badness must be a contiguous int32 device tensor of shape (P, 2n)")
    dev = _lib.device(badness.device.index)
    mx = torch.empty(count, dtype=torch.int32, device=dev)
    sq = torch.empty(count, dtype=torch.int64, device=dev)
    _lib.check(_lib.lib().n3_elite_keys(badness.data_ptr(), count, m // 2, mx.data_ptr(), sq.data_ptr(),
                                        _lib.stream_handle(dev)))
    return mx, sq


def rank_elites_device(scores: torch.Tensor, badness: torch.Tensor) -> torch.Tensor:
    """rank_elites for device tensors: the keys come from the elite-keys
    kernel, the order from three stable sorts (least significant key first),
    which is the lexicographic order with ties in index order."""
    mx, sq = elite_keys_device(badness)
    order = torch.sort(sq, stable=True).indices
    order = order[torch.sort(mx[order], stable=True).indices]
    order = order[torch.sort(scores.to(torch.int64)[order], stable=True).indices]
    return order.to(torch.int32)


def mutate_population(parent_p: torch.Tensor, parent_q: torch.Tensor, parent_badness: torch.Tensor, *,
                      quality_count: int, population: int, active_strength: int, max_strength: int, bad_bias: int,
                      seed: int, generation: int, out: tuple[torch.Tensor, torch.Tensor] | None = None
                      ) -> tuple[torch.Tensor, torch.Tensor]:
    """The mutation kernel on device tensors: parents P/Q int32[E][n] with
    badness int32[E][2n] -> children P/Q int32[population][n] (enqueued)."""
    elite_count, n = int(parent_p.shape[0]), int(parent_p.shape[1])
    for t, shape in ((parent_p, (elite_count, n)), (parent_q, (elite_count, n)), (parent_badness, (elite_count, 2 * n))):
        if t.shape != shape or t.dtype != torch.int32 or not t.is_cuda or not t.is_contiguous():
            raise ValueError("parents must be contiguous int32 device tensors P/Q (E, n) and badness (E, 2n)")
    dev = _lib.device(parent_p.device.index)
    if out is None:
        out = (torch.empty((population, n), dtype=torch.int32, device=dev),
               torch.empty((population, n), dtype=torch.int32, device=dev))
    child_p, child_q = out
    if child_p.shape != (population, n) or child_q.shape != (population, n):
        raise ValueError("output tensors must have shape (population, n)")
    _lib.check(_lib.lib().n3_mutate_population(
        parent_p.data_ptr(), parent_q.data_ptr(), parent_badness.data_ptr(), elite_count, int(quality_count), n,
        int(population), int(active_strength), int(max_strength), int(bad_bias), int(seed) & (2**64 - 1),
        int(generation), child_p.data_ptr(), child_q.data_ptr(), _lib.stream_handle(dev)))
    return child_p, child_q


def mutate_population_device(parent_p: np.ndarray, parent_q: np.ndarray, *, population: int, active_strength: int,
                             max_strength: int, seed: int, generation: int,
                             bad_bias: int = DEFAULT_BAD_BIAS) -> tuple[np.ndarray, np.ndarray]:
    """gpu_device_search.py:205-251: host wrapper of the mutation kernel
    (parent badness from the scorer, quality pool = a quarter of the
    parents)."""
    parent_p = np.ascontiguousarray(parent_p, dtype=np.int32)
    parent_q = np.ascontiguousarray(parent_q, dtype=np.int32)
    if parent_p.ndim != 2 or parent_q.shape != parent_p.shape:
        raise ValueError("P/Q batches must match")
    dev = _lib.device()
    p = torch.from_numpy(parent_p).to(dev)
    q = torch.from_numpy(parent_q).to(dev)
    bad = torch.empty((p.shape[0], 2 * p.shape[1]), dtype=torch.int32, device=dev)
    score_population_device(p, q, badness=bad)
    quality_count = max(1, min(p.shape[0], p.shape[0] // 4))
    cp, cq = mutate_population(p, q, bad, quality_count=quality_count, population=population,
                               active_strength=active_strength, max_strength=max_strength, bad_bias=bad_bias,
                               seed=seed, generation=generation)
    return _lib.to_host(cp, cq)


def verify_zero(state: PairPermutations) -> None:
    """The reference verifies an E=0 find with two independent verifiers
    (gpu_device_search.py:413-418); here both host verifiers must accept it,
    and the determinant kernel and the direction-histogram kernel must both
    count zero triples."""
    from .verifiers import verify_a, verify_b

    ok_a, detail_a = verify_a(state.points(), state.n)
    ok_b, detail_b = verify_b(state.points(), state.n)
    if not (ok_a and ok_b):
        raise AssertionError(f"E=0 failed verification: A={detail_a}; B={detail_b}")
    dev = _lib.device()
    p = torch.tensor([state.P], dtype=torch.int32, device=dev)
    q = torch.tensor([state.Q], dtype=torch.int32, device=dev)
    det = int(score_population_device(p, q, method=_lib.SCORE_DETERMINANT)[0])
    slope = int(score_population_device(p, q, method=_lib.SCORE_AUTO)[0])
    if det != 0 or slope != 0:
        raise AssertionError(f"E=0 failed verification: determinant count {det}, direction count {slope}")


def _base_elites(initial: PairPermutations, initial_elites, elite_size: int) -> np.ndarray:
    base = np.column_stack((initial.P, initial.Q)).astype(np.int32)[None, :, :]
    if initial_elites is not None:
        supplied = np.ascontiguousarray(initial_elites, dtype=np.int32)
        if supplied.ndim != 3 or supplied.shape[1:] != (initial.n, 2):
            raise ValueError("initial_elites must have shape (count,n,2)")
        for values in supplied:
            normalize_raw(values[:, 0], values[:, 1]).validate()
        base = np.concatenate((base, supplied), axis=0)
    return base[:elite_size]


def _score_rank(p: torch.Tensor, q: torch.Tensor, n: int):
    """Scores and badness of device states, ranked: returns the reordered
    (p, q, badness, host scores, host badness)."""
    dev = p.device
    bad = torch.empty((p.shape[0], 2 * n), dtype=torch.int32, device=dev)
    scores = score_population_device(p, q, badness=bad)
    s_host, b_host = scores.cpu().numpy(), bad.cpu().numpy()
    order = rank_elites(s_host, b_host)
    idx = torch.from_numpy(order.astype(np.int64)).to(dev)
    return p[idx].contiguous(), q[idx].contiguous(), bad[idx].contiguous(), s_host[order], b_host[order]


def _fill_selection(selected: list[int], candidates, elite_size: int) -> None:
    seen = set(selected)
    for raw in candidates:
        index = int(raw)
        if index not in seen:
            selected.append(index)
            seen.add(index)
            if len(selected) == elite_size:
                return


def select_elites(scores: np.ndarray, best_score: int, energy_window: int, elite_size: int, rng: Random) -> list[int]:
    """gpu_device_search.py:346-366 on host scores: the elite_size // 4 best
    children (stable order), then the children within energy_window of the
    best score in rng-shuffled order, then the best remaining ones."""
    order = np.argsort(scores, kind="stable")
    selected = list(map(int, order[:max(1, elite_size // 4)]))
    diverse = list(map(int, np.flatnonzero(scores <= best_score + energy_window)))
    rng.shuffle(diverse)
    _fill_selection(selected, diverse, elite_size)
    if len(selected) < elite_size:
        _fill_selection(selected, order, elite_size)
    return selected


def select_elites_device(scores: torch.Tensor, best_score: int, energy_window: int, elite_size: int,
                         rng: Random) -> list[int]:
    """select_elites with the sort and the window search on the device: only
    the quality share, the window's indices and (when those do not fill the
    beam) the full order cross to the host.  Same picks and the same rng
    calls as the host form."""
    order = torch.sort(scores, stable=True).indices
    selected = order[:max(1, elite_size // 4)].cpu().tolist()
    diverse = torch.nonzero(scores <= best_score + energy_window).flatten().cpu().tolist()
    rng.shuffle(diverse)
    _fill_selection(selected, diverse, elite_size)
    if len(selected) < elite_size:
        _fill_selection(selected, order.cpu().tolist(), elite_size)
    return selected


def gpu_device_beam_search(initial: PairPermutations, *, seed: int, generations: int, population: int,
                           elite_size: int, max_strength: int, energy_window: int,
                           initial_elites: np.ndarray | None = None, report_every: int = 5,
                           bad_bias: int = DEFAULT_BAD_BIAS,
                           callback: Callable[[GpuSearchSnapshot], None] | None = None) -> GpuSearchResult:
    """gpu_device_search.py:254-432: mutate the beam on the device, score the
    children, keep a quality share plus a shuffled energy window, re-rank by
    the elite key; stops at E=0 after verifying it twice."""
    initial.validate()
    if not (1 <= elite_size <= population):
        raise ValueError("elite_size must be between one and population")
    if max_strength < 1:
        raise ValueError("max_strength must be positive")
    if not 0 <= bad_bias <= 100:
        raise ValueError("bad_bias must be between zero and 100")
    n = initial.n
    dev = _lib.device()
    rng = Random(seed)
    base = _base_elites(initial, initial_elites, elite_size)
    elite_p = torch.from_numpy(np.ascontiguousarray(base[:, :, 0])).to(dev)
    elite_q = torch.from_numpy(np.ascontiguousarray(base[:, :, 1])).to(dev)
    elite_p, elite_q, elite_bad, elite_scores, elite_bad_host = _score_rank(elite_p, elite_q, n)
    best_score = int(elite_scores[0])
    best_key = elite_key(best_score, elite_bad_host[0])
    best_p, best_q = elite_p[0].cpu().numpy(), elite_q[0].cpu().numpy()
    evaluations = len(base)
    last_improvement = 0
    started = time.perf_counter()
    child_p = torch.empty((population, n), dtype=torch.int32, device=dev)
    child_q = torch.empty_like(child_p)
    child_scores = torch.empty(population, dtype=torch.int32, device=dev)
    completed = 0
    for generation in range(1, generations + 1):
        completed = generation
        active_strength = min(max_strength, 2 + (generation - last_improvement) // 10)
        quality_count = max(1, min(len(elite_scores), elite_size // 4))
        mutate_population(elite_p, elite_q, elite_bad, quality_count=quality_count, population=population,
                          active_strength=active_strength, max_strength=max_strength, bad_bias=bad_bias, seed=seed,
                          generation=generation, out=(child_p, child_q))
        score_population_device(child_p, child_q, child_scores, check=False)
        check_status(dev)
        evaluations += population
        selected = select_elites_device(child_scores, best_score, energy_window, elite_size, rng)
        sel = torch.from_numpy(np.asarray(selected, dtype=np.int64)).to(dev)
        elite_p, elite_q, elite_bad, elite_scores, elite_bad_host = _score_rank(
            child_p[sel].contiguous(), child_q[sel].contiguous(), n)
        generation_best = int(elite_scores[0])
        generation_key = elite_key(generation_best, elite_bad_host[0])
        improved = generation_key < best_key
        if improved:
            best_score, best_key = generation_best, generation_key
            best_p, best_q = elite_p[0].cpu().numpy(), elite_q[0].cpu().numpy()
            last_improvement = generation
        if callback is not None and (improved or generation % max(1, report_every) == 0):
            raw_p, raw_q = elite_p.cpu().numpy(), elite_q.cpu().numpy()
            values = []
            for p, q in zip(raw_p, raw_q):
                state = normalize_raw(p, q)
                values.append(np.column_stack((state.P, state.Q)))
            callback(GpuSearchSnapshot(generation, time.perf_counter() - started, evaluations, generation_best,
                                       best_score, tuple(map(int, elite_scores)), np.asarray(values, dtype=np.int32),
                                       normalize_raw(best_p, best_q)))
        if best_score == 0:
            verify_zero(normalize_raw(best_p, best_q))
            break
    return GpuSearchResult(normalize_raw(best_p, best_q), best_score, completed, evaluations,
                           time.perf_counter() - started, seed)
