"""Exhaustive 2-switch beam over the scorer (drop-in for
gpu_exhaustive_beam.py).

Each generation scores every 2-switch that touches a bad slot of an elite
(every 2-switch, every `global_every` generations).  The reference
materialises all children and recounts each one from scratch; here the
child scores come from the device 2-switch table of each elite (O(m) per
child instead of O(m^3)), and only the prospects that selection looks at are
materialised by the enumeration kernel.  Scores, legality and therefore the
selection are the same as the reference's (tested).
"""
from __future__ import annotations

import time
from random import Random
from typing import Callable

import numpy as np
import torch

from . import _lib
from .gpu import score_population_device
from .gpu_delta import delta_table_device
from .gpu_device_search import (GpuSearchResult, GpuSearchSnapshot, _base_elites, _score_rank, elite_key,
                                normalize_raw, rank_elites, verify_zero)
from .state import PairPermutations

ILLEGAL_SCORE = 1 << 29  # gpu_exhaustive_beam.py:179
SENTINEL32 = 1 << 30     # delta table entries of illegal switches


def candidate_descriptors(badness: np.ndarray, *, global_anchors: bool = False) -> np.ndarray:
    """gpu_exhaustive_beam.py:67-82: (parent, u, v) for every unordered
    switch touching a bad slot (or every switch), in the reference's order:
    parents, then anchors u ascending, then v ascending; v skips u's row and
    the anchors below u."""
    badness = np.asarray(badness)
    parent_count, edge_count = badness.shape
    slots = np.arange(edge_count)
    out = []
    for parent in range(parent_count):
        anchors = slots if global_anchors else np.flatnonzero(badness[parent])
        if len(anchors) == 0:
            continue
        is_anchor = np.zeros(edge_count, dtype=bool)
        is_anchor[anchors] = True
        u = np.repeat(anchors, edge_count)
        v = np.tile(slots, len(anchors))
        keep = (v != u) & ((v >> 1) != (u >> 1)) & ~(is_anchor[v] & (v < u))
        rows = np.empty((int(keep.sum()), 3), dtype=np.int32)
        rows[:, 0] = parent
        rows[:, 1] = u[keep]
        rows[:, 2] = v[keep]
        out.append(rows)
    if not out:
        return np.empty((0, 3), dtype=np.int32)
    return np.concatenate(out, axis=0)


def enumerate_switches(parent_p: torch.Tensor, parent_q: torch.Tensor, descriptors: torch.Tensor
                       ) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """The enumeration kernel on device tensors: children P/Q and validity."""
    parents, n = int(parent_p.shape[0]), int(parent_p.shape[1])
    count = int(descriptors.shape[0])
    for t, shape in ((parent_p, (parents, n)), (parent_q, (parents, n)), (descriptors, (count, 3))):
        if t.shape != shape or t.dtype != torch.int32 or not t.is_cuda or not t.is_contiguous():
            raise ValueError("parents P/Q (E, n) and descriptors (count, 3) must be contiguous int32 device tensors")
    dev = _lib.device(parent_p.device.index)
    child_p = torch.empty((count, n), dtype=torch.int32, device=dev)
    child_q = torch.empty_like(child_p)
    valid = torch.empty(count, dtype=torch.uint8, device=dev)
    _lib.check(_lib.lib().n3_enumerate_switches(parent_p.data_ptr(), parent_q.data_ptr(), parents,
                                                descriptors.data_ptr(), count, n, child_p.data_ptr(),
                                                child_q.data_ptr(), valid.data_ptr(), _lib.stream_handle(dev)))
    return child_p, child_q, valid


def candidate_scores(elite_p: torch.Tensor, elite_q: torch.Tensor, elite_scores: torch.Tensor,
                     elite_badness: torch.Tensor, descriptors: torch.Tensor) -> tuple[np.ndarray, np.ndarray]:
    """Exact score and legality of every described child, read from the
    device 2-switch table of its parent.  The table also scores a switch of
    two points in one column (accelerated.py:146-150 allow it: a no-op);
    enumerate_switches rejects it (gpu_exhaustive_beam.py:46-47), so it is
    masked out here."""
    table = delta_table_device(elite_p, elite_q, elite_scores, elite_badness)
    d = descriptors.long()
    cols = torch.stack((elite_p, elite_q), dim=2).reshape(elite_p.shape[0], -1)
    distinct = cols[d[:, 0], d[:, 1]] != cols[d[:, 0], d[:, 2]]
    picked = table[d[:, 0], d[:, 1], d[:, 2]]
    valid = ((picked != SENTINEL32) & distinct).cpu().numpy()
    return np.where(valid, picked.cpu().numpy(), ILLEGAL_SCORE).astype(np.int32), valid


def _signature(p: np.ndarray, q: np.ndarray) -> bytes:
    """gpu_exhaustive_beam.py:85-88: the state's rows as sorted column
    pairs, int16 bytes (the same point set gives the same signature
    whichever of P and Q holds which point)."""
    return _signatures(p[None, :], q[None, :])[0]


def _signatures(p: np.ndarray, q: np.ndarray) -> list[bytes]:
    """_signature of every row of P/Q arrays [count][n], in one sort."""
    pairs = np.sort(np.stack((p, q), axis=2), axis=2).astype(np.int16)
    return [row.tobytes() for row in pairs.reshape(len(pairs), -1)]


class _BeamPick:
    """The beam being chosen from materialised prospects: members in the
    order they were taken, no two with one signature, and (when asked) none
    whose signature an earlier generation already kept."""

    def __init__(self, signature_of: dict[int, bytes], seen: set[bytes]):
        self.signature_of = signature_of  # prospect index -> signature
        self.seen = seen
        self.members: list[int] = []
        self.signatures: set[bytes] = set()

    def fill(self, candidates, until: int, *, fresh_only: bool) -> None:
        """Offers the candidates in order; stops right after the one that
        brings the beam to `until` members.  Candidates that were never
        materialised are passed over."""
        for candidate in candidates:
            sig = self.signature_of.get(candidate)
            if sig is not None and sig not in self.signatures and not (fresh_only and sig in self.seen):
                self.members.append(candidate)
                self.signatures.add(sig)
            if len(self.members) == until:
                return


def gpu_exhaustive_beam_search(initial: PairPermutations, *, seed: int, generations: int, elite_size: int,
                               energy_window: int, initial_elites: np.ndarray | None = None, global_every: int = 25,
                               report_every: int = 5,
                               callback: Callable[[GpuSearchSnapshot], None] | None = None) -> GpuSearchResult:
    """gpu_exhaustive_beam.py:91-325: walk a diverse beam through complete
    bad-slot 2-switch neighbourhoods."""
    initial.validate()
    if elite_size < 1 or generations < 1:
        raise ValueError("elite_size and generations must be positive")
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
    host_p, host_q = elite_p.cpu().numpy(), elite_q.cpu().numpy()
    seen = {_signature(host_p[i], host_q[i]) for i in range(len(base))}
    evaluations = len(base)
    started = time.perf_counter()
    completed = 0
    for generation in range(1, generations + 1):
        completed = generation
        descriptors = candidate_descriptors(elite_bad_host,
                                            global_anchors=bool(global_every and generation % global_every == 0))
        if not len(descriptors):
            break
        desc_dev = torch.from_numpy(descriptors).to(dev)
        scores_dev = torch.from_numpy(np.ascontiguousarray(elite_scores, dtype=np.int32)).to(dev)
        scores, valid = candidate_scores(elite_p, elite_q, scores_dev, elite_bad, desc_dev)
        evaluations += int(valid.sum())
        legal_order = np.flatnonzero(valid)
        legal_order = legal_order[np.argsort(scores[legal_order], kind="stable")]
        if not len(legal_order):
            break

        # prospects: the best 24 x beam legal children by score, then the
        # first 24 x beam of the shuffled energy window (one rng call per
        # generation, as gpu_exhaustive_beam.py:186-196 makes it)
        reach = elite_size * 24
        window = [int(i) for i in np.flatnonzero(valid & (scores <= best_score + energy_window))]
        rng.shuffle(window)
        prospects = list(dict.fromkeys([int(i) for i in legal_order[:reach]] + window[:reach]))
        prospect_desc = torch.from_numpy(np.ascontiguousarray(descriptors[prospects])).to(dev)
        prospect_p_dev, prospect_q_dev, _ = enumerate_switches(elite_p, elite_q, prospect_desc)
        # re-rank the prospects by the full elite key (score, max, sum b^2)
        prospect_bad = torch.empty((len(prospects), 2 * n), dtype=torch.int32, device=dev)
        score_population_device(prospect_p_dev, prospect_q_dev, badness=prospect_bad)
        ranked = [prospects[int(i)] for i in
                  rank_elites(scores[np.asarray(prospects, dtype=np.int64)], prospect_bad.cpu().numpy())]
        signatures = _signatures(prospect_p_dev.cpu().numpy(), prospect_q_dev.cpu().numpy())
        pick = _BeamPick(dict(zip(prospects, signatures)), seen)
        pick.fill(ranked, max(1, elite_size // 4), fresh_only=True)
        pick.fill(window, elite_size, fresh_only=True)
        if len(pick.members) < elite_size:
            pick.fill(ranked + window, elite_size, fresh_only=False)
        if not pick.members:
            # every prospect was seen before: forget the history, keep the best
            seen.clear()
            pick = _BeamPick(pick.signature_of, seen)
            pick.members = ranked[:elite_size]
            pick.signatures = {pick.signature_of[c] for c in pick.members}
        selected = pick.members
        seen.update(pick.signatures)
        if len(seen) > 500000:
            seen = set(pick.signatures)
        lookup = {candidate: position for position, candidate in enumerate(prospects)}

        positions = torch.from_numpy(np.asarray([lookup[c] for c in selected], dtype=np.int64)).to(dev)
        elite_p, elite_q, elite_bad, elite_scores, elite_bad_host = _score_rank(
            prospect_p_dev[positions].contiguous(), prospect_q_dev[positions].contiguous(), n)
        generation_best = int(elite_scores[0])
        generation_key = elite_key(generation_best, elite_bad_host[0])
        improved = generation_key < best_key
        if improved:
            best_score, best_key = generation_best, generation_key
            best_p, best_q = elite_p[0].cpu().numpy(), elite_q[0].cpu().numpy()
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
