"""Replica-exchange tempering on the GPU: a drop-in for the reference's
TemperingChains (tempering.py:119-225) and its driver (tempering.py:228-319).

Differences from the reference, all deliberate:
  - random numbers are Philox4x32-10 words keyed by the seed with counter
    (step, sweep, slot) for proposals and (pair, round, ladder) for exchanges,
    so that any trajectory can be replayed on the CPU from (seed, sweep,
    round); there is no per-replica RNG state to carry (checkpoints still hold
    the reference's `rngs`, `round` and `host_rng` keys, so either engine can
    resume the other's files);
  - acceptance compares a 32-bit word with the precomputed integer threshold
    floor(2^32 exp(-x)) (n3_metropolis_thresholds / n3_exchange_thresholds)
    instead of log(U) < -x in double;
  - exchange runs on the device (no host round trip) and, like the
    reference's, swaps the replicas' states between fixed-temperature slots;
  - umbrella potentials (strengths != 0, --umbrella-max) run on the device
    walk with the reference's energy delta + s (2 (removed - center) dd +
    dd^2), accepted against floor(2^32 exp(-E / T)) computed by a fixed
    sequence of fma operations (so the CPU replay reproduces it); the batched
    and learned walks do not take them (NotImplementedError).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import time
from pathlib import Path

import numpy as np
import torch

from . import _lib
from .gpu import score_population_device
from .io import load_candidate, write_candidate
from .state import check_columns, points_to_permutations

MAX_EXCHANGE_TABLE = 2048


def move_table_size(temperatures: np.ndarray) -> int:
    """Entries per temperature such that the last threshold is 0:
    2^32 exp(-d/T) < 1 for d > 32 ln 2 T."""
    return max(8, int(math.ceil(float(np.max(temperatures)) * 32.0 * math.log(2.0))) + 2)


def exchange_table_size(temperatures: np.ndarray, ladder: int) -> int:
    t = temperatures.astype(np.float64)
    c = np.abs(1.0 / t[:-1] - 1.0 / t[1:])
    same_ladder = (np.arange(len(t) - 1) % ladder) != ladder - 1
    c = c[same_ladder & (c > 0)]
    if c.size == 0:
        return 8
    return int(min(MAX_EXCHANGE_TABLE, max(8, math.ceil(32.0 * math.log(2.0) / float(c.min())) + 2)))


def metropolis_thresholds(temperatures: np.ndarray, k: int) -> np.ndarray:
    t = np.ascontiguousarray(temperatures, dtype=np.float32)
    out = np.empty((len(t), k), dtype=np.uint32)
    _lib.check(_lib.lib().n3_metropolis_thresholds(t.ctypes.data, len(t), k, out.ctypes.data))
    return out


def exchange_thresholds(temperatures: np.ndarray, ladder: int, k: int) -> tuple[np.ndarray, np.ndarray]:
    t = np.ascontiguousarray(temperatures, dtype=np.float32)
    out = np.empty((len(t), k), dtype=np.uint32)
    sign = np.empty(len(t), dtype=np.int8)
    _lib.check(_lib.lib().n3_exchange_thresholds(t.ctypes.data, len(t), ladder, k, out.ctypes.data, sign.ctypes.data))
    return out, sign


class TemperingChains:
    """Replicas of one size n, slot-ordered; every array is a device tensor."""

    def __init__(self, columns, temperatures, seed: int, *, device: torch.device | None = None):
        columns = np.ascontiguousarray(columns, dtype=np.int32)
        if columns.ndim != 2 or columns.shape[1] % 2 or not 4 <= columns.shape[1] <= 200:
            raise ValueError("columns must have shape (replicas,2n), 2<=n<=100")
        self.n = columns.shape[1] // 2
        temperatures = np.asarray(temperatures, dtype=np.float32)
        if temperatures.shape != (len(columns),) or np.any(~np.isfinite(temperatures)) or np.any(temperatures <= 0):
            raise ValueError("one positive finite temperature is required per replica")
        check_columns(columns)
        dev = _lib.device(None if device is None else device.index)
        self.device = dev
        self.seed = int(seed) & (2**64 - 1)
        self.columns = torch.from_numpy(columns).to(dev)
        self.badness = torch.empty_like(self.columns)
        self.scores = torch.empty(len(columns), dtype=torch.int32, device=dev)
        self._rescore(self.columns, self.scores, self.badness)
        self.best_columns = self.columns.clone()
        self.best_scores = self.scores.clone()
        self.temperatures_host = temperatures
        self.temperatures = torch.from_numpy(temperatures).to(dev)
        self.accepted = torch.zeros(len(columns), dtype=torch.int32, device=dev)
        self.reference = self.columns[0].clone()
        reference = set((i // 2, int(c)) for i, c in enumerate(columns[0]))
        self.removed_points = torch.tensor(
            [len(reference - set((i // 2, int(c)) for i, c in enumerate(row))) for row in columns],
            dtype=torch.int32, device=dev)
        self.centers = torch.zeros(len(columns), dtype=torch.float32, device=dev)
        self.strengths = torch.zeros(len(columns), dtype=torch.float32, device=dev)
        # 1 / T in double for the umbrella acceptance (host division, as in the replay)
        self.inverse_temperatures = torch.from_numpy(1.0 / temperatures.astype(np.float64)).to(dev)
        self.k_move = move_table_size(temperatures)
        self.move_thresholds = torch.from_numpy(metropolis_thresholds(temperatures, self.k_move)).to(dev)
        self._exchange_tables: dict[int, tuple[torch.Tensor, torch.Tensor, int]] = {}
        # sweep, round, exchanges (64-bit); valid proposals, accepted moves
        self.counters = torch.zeros(4, dtype=torch.int32, device=dev)
        self.proposals = torch.zeros(2, dtype=torch.int64, device=dev)
        # replicas (wavefronts) per workgroup of the walk; 0 lets the library
        # choose (see autotune.py for the measured choice per n)
        self.block_waves = 0
        # optional device log of uniform-walk proposals (learned.ProposalLog)
        self.log = None

    # -- helpers -----------------------------------------------------------
    def _rescore(self, columns: torch.Tensor, scores: torch.Tensor, badness: torch.Tensor,
                 method: int = _lib.SCORE_AUTO) -> None:
        p = columns[:, 0::2].contiguous()
        q = columns[:, 1::2].contiguous()
        score_population_device(p, q, scores, badness, method=method)

    def _stream(self) -> int:
        return _lib.stream_handle(self.device)

    def _exchange_table(self, ladder: int):
        if ladder not in self._exchange_tables:
            k = exchange_table_size(self.temperatures_host, ladder)
            thr, sign = exchange_thresholds(self.temperatures_host, ladder, k)
            self._exchange_tables[ladder] = (torch.from_numpy(thr).to(self.device),
                                             torch.from_numpy(sign).to(self.device), k)
        return self._exchange_tables[ladder]

    def _chains(self, ladder: int | None = None) -> _lib.Chains:
        c = _lib.Chains()
        c.n = self.n
        c.replicas = len(self.temperatures_host)
        c.columns = self.columns.data_ptr()
        c.badness = self.badness.data_ptr()
        c.scores = self.scores.data_ptr()
        c.best_columns = self.best_columns.data_ptr()
        c.best_scores = self.best_scores.data_ptr()
        c.accepted = self.accepted.data_ptr()
        c.reference = self.reference.data_ptr()
        c.removed_points = self.removed_points.data_ptr()
        c.move_thresholds = self.move_thresholds.data_ptr()
        c.k_move = self.k_move
        if ladder is not None:
            thr, sign, k = self._exchange_table(ladder)
            c.exchange_thresholds = thr.data_ptr()
            c.exchange_sign = sign.data_ptr()
            c.k_exchange = k
        c.counters = self.counters.data_ptr()
        c.proposals = self.proposals.data_ptr()
        c.block_waves = int(self.block_waves)
        if self.log is not None:
            c.log = self.log.entries.data_ptr()
            c.log_capacity = self.log.capacity
            c.log_count = self.log.count.data_ptr()
        if self._umbrella():
            c.centers = self.centers.data_ptr()
            c.strengths = self.strengths.data_ptr()
            c.inverse_temperatures = self.inverse_temperatures.data_ptr()
        return c

    def _umbrella(self) -> bool:
        """Whether any slot has an umbrella potential; the arrays are made
        float32, contiguous and resident on this device (in place, so that
        captured graphs keep valid pointers)."""
        for name in ("centers", "strengths"):
            t = getattr(self, name)
            if t.dtype != torch.float32 or not t.is_contiguous() or t.device != self.device or \
                    t.shape != (len(self.temperatures_host),):
                setattr(self, name, torch.as_tensor(t, dtype=torch.float32, device=self.device)
                        .reshape(len(self.temperatures_host)).contiguous())
        return bool(torch.any(self.strengths != 0).item())

    def _check_umbrella(self) -> None:
        """The batched and learned walks have no umbrella term."""
        if self._umbrella():
            raise NotImplementedError("umbrella potentials are supported by advance() and run() only")

    @property
    def sweep(self) -> int:
        return int(self.counters[0].item())

    @property
    def round(self) -> int:
        return int(self.counters[1].item())

    # -- reference API -------------------------------------------------------
    def to_state(self, row: np.ndarray):
        return points_to_permutations([(i // 2, int(c)) for i, c in enumerate(row)], self.n)

    def advance(self, steps: int, bad_bias: int = 0) -> None:
        """tempering.py:158-167: `steps` proposals per replica (synchronous)."""
        if steps < 1 or not 0 <= bad_bias <= 100:
            raise ValueError("steps must be positive and bad_bias in [0,100]")
        chains = self._chains()
        _lib.check(_lib.lib().n3_temper_advance(ctypes.byref(chains), int(steps), int(bad_bias), self.seed,
                                                self._stream()))
        torch.cuda.synchronize(self.device)

    def exchange(self, ladder_size: int) -> int:
        """tempering.py:169-191: neighbouring-temperature swaps inside each ladder."""
        if ladder_size < 2 or len(self.temperatures_host) % ladder_size:
            raise ValueError("ladder_size must divide replica count")
        if self._umbrella():
            raise ValueError("temperature-only exchange is disabled with umbrella potentials")
        before = self.exchanges_total()
        chains = self._chains(ladder_size)
        _lib.check(_lib.lib().n3_temper_exchange(ctypes.byref(chains), int(ladder_size), self.seed, self._stream()))
        return self.exchanges_total() - before

    def run(self, sweeps: int, steps: int, ladder_size: int = 0, bad_bias: int = 0) -> None:
        """`sweeps` x (advance(steps); exchange(ladder_size)) from a captured
        hipGraph, enqueued on the current stream without synchronising;
        ladder_size 0 or 1 runs without exchanges."""
        if steps < 1 or not 0 <= bad_bias <= 100 or sweeps < 0:
            raise ValueError("steps must be positive, sweeps non-negative and bad_bias in [0,100]")
        ladder = int(ladder_size)
        if ladder > 1 and len(self.temperatures_host) % ladder:
            raise ValueError("ladder_size must divide replica count")
        if ladder > 1 and self._umbrella():
            raise ValueError("temperature-only exchange is disabled with umbrella potentials")
        chains = self._chains(ladder if ladder > 1 else None)
        _lib.check(_lib.lib().n3_temper_run(ctypes.byref(chains), int(sweeps), int(steps), ladder if ladder > 1 else 0,
                                            int(bad_bias), self.seed, self._stream()))

    def exchanges_total(self) -> int:
        c = self.counters[2:4].cpu().numpy().astype(np.uint32)
        return int(c[0]) | (int(c[1]) << 32)

    def validate(self, indices=None) -> None:
        """tempering.py:193-204: every selected replica's incremental score,
        badness and removed count against a full recount (the determinant
        kernel, independent of the incremental direction-histogram updates)."""
        total = len(self.temperatures_host)
        idx = np.arange(total) if indices is None else np.asarray(list(indices), dtype=np.int64) % total
        if idx.size == 0:
            return
        sel = torch.from_numpy(idx).to(self.device)
        cols = self.columns.index_select(0, sel)
        scores = torch.empty(len(idx), dtype=torch.int32, device=self.device)
        bad = torch.empty_like(cols)
        self._rescore(cols, scores, bad, method=_lib.SCORE_DETERMINANT)
        cols_h = cols.cpu().numpy()
        check_columns(cols_h)
        if not torch.equal(scores, self.scores.index_select(0, sel)) or \
                not torch.equal(bad, self.badness.index_select(0, sel)):
            raise AssertionError("incremental score/badness differs from full determinant check")
        reference = set((j // 2, int(c)) for j, c in enumerate(self.reference.cpu().numpy()))
        removed = self.removed_points.index_select(0, sel).cpu().numpy()
        for row, value in zip(cols_h, removed):
            if len(reference - set((j // 2, int(c)) for j, c in enumerate(row))) != int(value):
                raise AssertionError("incremental distance mismatch")

    # checkpoint keys of the reference (tempering.py:206-225); `rngs`,
    # `round` and `host_rng` are its RNG state, which this engine does not
    # use but writes so that the reference can resume from these files
    REFERENCE_KEYS = ("columns", "badness", "scores", "rngs", "best_columns", "best_scores", "accepted",
                      "reference", "removed_points", "centers", "strengths")

    def save(self, path) -> None:
        """tempering.py:206-211: np.savez_compressed with the reference's keys
        (plus `seed`, `counters` and `proposals`, which resume this engine's
        Philox streams exactly)."""
        arrays = {name: getattr(self, name).cpu().numpy() for name in (
            "columns", "badness", "scores", "best_columns", "best_scores", "accepted", "reference",
            "removed_points", "centers", "strengths", "counters", "proposals")}
        host = np.random.default_rng([self.seed & (2**63 - 1), self.round])
        rngs = host.integers(1, 2**64 - 1, size=len(self.temperatures_host), dtype=np.uint64)
        np.savez_compressed(path, **arrays, rngs=rngs, temperatures=self.temperatures_host,
                            round=np.asarray(self.round), host_rng=np.asarray(json.dumps(host.bit_generator.state)),
                            seed=np.asarray(self.seed, np.uint64))

    @classmethod
    def restore(cls, path):
        """tempering.py:213-225: a checkpoint of this engine, or of the
        reference (which carries no seed or sweep counter: the seed is then
        taken from its host_rng state and both counters from its round, so a
        resumed run still draws fresh Philox counters)."""
        with np.load(path, allow_pickle=False) as saved:
            if "seed" in saved:
                seed = int(saved["seed"])
            else:
                state = json.loads(str(saved["host_rng"]))
                seed = int(state["state"]["state"]) & (2**64 - 1)
            obj = cls(saved["columns"], saved["temperatures"], seed)
            for name in ("columns", "badness", "scores", "best_columns", "best_scores", "accepted", "reference",
                         "removed_points", "centers", "strengths", "counters", "proposals"):
                if name in saved:
                    setattr(obj, name, torch.from_numpy(np.ascontiguousarray(saved[name])).to(obj.device))
            if "counters" not in saved and "round" in saved:
                r = int(saved["round"])
                obj.counters = torch.tensor([r, r, 0, 0], dtype=torch.int32, device=obj.device)
        obj.validate()
        return obj


def main(argv=None) -> int:
    """tempering.py:228-319, umbrella sampling included (--umbrella-max: one
    center per ladder, linspace(0, max), strength --umbrella-strength, no
    exchanges, and a frontier of the best state per distance bin of 8)."""
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--initial", nargs="+", required=True)
    parser.add_argument("--output", required=True)
    parser.add_argument("--seed", type=int, required=True)
    parser.add_argument("--replicas", type=int, default=128)
    parser.add_argument("--ladder-size", type=int, default=16)
    parser.add_argument("--tmin", type=float, default=0.2)
    parser.add_argument("--tmax", type=float, default=4)
    parser.add_argument("--rounds", type=int, default=200)
    parser.add_argument("--steps", type=int, default=100)
    parser.add_argument("--bad-bias", type=int, default=0)
    parser.add_argument("--resume")
    parser.add_argument("--umbrella-max", type=float, default=0)
    parser.add_argument("--umbrella-strength", type=float, default=0.2)
    args = parser.parse_args(argv)
    output = Path(args.output)
    output.mkdir(parents=True, exist_ok=True)
    states, n = [], None
    for path in args.initial:
        size, points, _ = load_candidate(path)
        if n is not None and n != size:
            raise ValueError("all initial sizes must agree")
        n = size
        states.append(points_to_permutations(points, n).columns())
    if args.replicas % args.ladder_size or args.ladder_size < 2:
        raise ValueError("replicas must be a multiple of ladder-size >=2")
    if args.resume:
        chains = TemperingChains.restore(Path(args.resume))
        if len(chains.temperatures_host) != args.replicas or chains.n != n:
            raise ValueError("checkpoint dimensions differ from arguments")
    else:
        columns = np.asarray([states[(i // args.ladder_size) % len(states)] for i in range(args.replicas)],
                             dtype=np.int32)
        temperatures = np.tile(np.geomspace(args.tmin, args.tmax, args.ladder_size), args.replicas // args.ladder_size)
        chains = TemperingChains(columns, temperatures, args.seed)
    if args.umbrella_max:
        # tempering.py:265-267
        centers = np.repeat(np.linspace(0, args.umbrella_max, args.replicas // args.ladder_size), args.ladder_size)
        chains.centers = torch.from_numpy(centers.astype(np.float32)).to(chains.device)
        chains.strengths = torch.full((args.replicas,), np.float32(args.umbrella_strength), dtype=torch.float32,
                                      device=chains.device)
    reference = set((i // 2, int(c)) for i, c in enumerate(states[0]))
    archive = {}
    (output / "parameters.json").write_text(json.dumps(vars(args), indent=2) + "\n", encoding="utf-8")
    start = time.perf_counter()
    best, proposals, exchanges = 10**9, 0, 0
    for epoch in range(1, args.rounds + 1):
        chains.advance(args.steps, args.bad_bias)
        proposals += args.steps * args.replicas
        exchanges = chains.exchange(args.ladder_size) if not args.umbrella_max else 0
        best_scores = chains.best_scores.cpu().numpy()
        index = int(np.argmin(best_scores))
        if int(best_scores[index]) < best:
            best = int(best_scores[index])
            row = chains.best_columns[index].cpu().numpy()
            state = chains.to_state(row)
            exact = _recount(chains, row)
            if exact != best:
                raise AssertionError("record differs from independent full score")
            write_candidate(output / f"best-E{best}.json", n, state.points(), conflicts=best, seed=args.seed,
                            epoch=epoch, proposals=proposals, method="gfx950 replica exchange")
            print(f"epoch={epoch} best={best} proposals={proposals} seconds={time.perf_counter()-start:.2f}",
                  flush=True)
            if best == 0:
                from .verifiers import verify_a, verify_b

                assert verify_a(state.points(), n)[0] and verify_b(state.points(), n)[0]
                break
        if epoch % 10 == 0 or epoch == args.rounds:
            chains.validate([epoch % args.replicas])
            columns = chains.columns.cpu().numpy()
            scores = chains.scores.cpu().numpy()
            distances = [len(reference.symmetric_difference((i // 2, int(c)) for i, c in enumerate(row)))
                         for row in columns]
            if args.umbrella_max:
                # tempering.py:296-303: best state per distance bin
                for i, (score, distance) in enumerate(zip(scores, distances)):
                    bucket = distance // 8
                    if int(score) < archive.get(bucket, 10**9):
                        archive[bucket] = int(score)
                        candidate = chains.to_state(columns[i])
                        write_candidate(output / f"frontier-bin{bucket:02d}.json", n, candidate.points(),
                                        conflicts=int(score), symmetric_difference=distance, epoch=epoch,
                                        seed=args.seed)
            payload = {"epoch": epoch, "proposals": proposals, "elapsed_seconds": time.perf_counter() - start,
                       "best": best, "scores": scores.tolist(), "distances": distances,
                       "exchanges": exchanges, "accepted": int(chains.accepted.cpu().numpy().astype(np.int64).sum())}
            with (output / "history.jsonl").open("a", encoding="utf-8") as stream:
                stream.write(json.dumps(payload) + "\n")
        if epoch % 50 == 0:
            chains.save(output / "checkpoint.npz")
    chains.validate()
    chains.save(output / "checkpoint.npz")
    result = {"best": best, "proposals": proposals, "elapsed_seconds": time.perf_counter() - start,
              "accepted": int(chains.accepted.cpu().numpy().astype(np.int64).sum()), "distance_frontier": archive}
    (output / "summary.json").write_text(json.dumps(result, indent=2) + "\n", encoding="utf-8")
    print(json.dumps(result), flush=True)
    return 0


def _recount(chains: TemperingChains, row: np.ndarray) -> int:
    cols = torch.from_numpy(np.ascontiguousarray(row[None, :], dtype=np.int32)).to(chains.device)
    scores = torch.empty(1, dtype=torch.int32, device=chains.device)
    bad = torch.empty_like(cols)
    chains._rescore(cols, scores, bad, method=_lib.SCORE_DETERMINANT)
    return int(scores.item())
