"""Launch-shape tuner for the tempering walk.

The walk keeps one replica per wavefront; the tunable is how many replicas
(wavefronts) share a workgroup -- they share one LDS copy of the direction
table, so more waves per block means fewer table copies but coarser blocks
and, at large n, LDS pressure (the library rejects shapes that do not fit).

- `tune` times each n alone: a short graph-captured run of every candidate,
  keeping the fastest.  Alone, a small population fills the GPU only with
  small blocks.
- `tune_mixed` tunes the shapes of a batched run of several sizes on that
  run itself: starting from the most waves that fit in LDS for every n, it
  tries each candidate for one n at a time with the others fixed (largest n
  first) and keeps the fastest.  Mixed, the other sizes fill the compute
  units, so the best shapes are larger than each size's own best.

Results are cached as JSON per device.
"""
from __future__ import annotations

import json
import os
import time
from pathlib import Path

import torch

from . import _lib
from .batched import seeded_chains

CANDIDATES = (1, 2, 4, 6, 8, 12, 16)


def default_cache_path() -> Path:
    base = Path(os.environ.get("NO3_CACHE", Path.home() / ".cache" / "no3"))
    return base / "autotune.json"


def tune_size(n: int, replicas: int, *, ladder: int = 64, sweeps: int = 4, candidates=CANDIDATES,
              seed: int = 0, persistent: bool = False) -> dict:
    """Times `sweeps` sweeps of 2n proposals for every candidate block shape:
    per-sweep walk launches, or (persistent) the one-launch batched run of
    this size alone."""
    from .batched import BatchGroup, run_batched

    dev = _lib.device()
    ladder = min(ladder, replicas)
    chains = seeded_chains(n, replicas, ladder, seed=seed, device=dev)

    def run(count):
        if persistent:
            run_batched([BatchGroup(chains, ladder, 2 * n)], count, persistent=True)
        else:
            chains.run(count, 2 * n, ladder)

    timings = {}
    for w in candidates:
        chains.block_waves = w
        try:
            run(1)
        except ValueError:
            continue  # does not fit in LDS
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        run(sweeps)
        torch.cuda.synchronize(dev)
        elapsed = time.perf_counter() - t0
        timings[w] = replicas * sweeps * 2 * n / elapsed
    _lib.lib().n3_temper_release()
    best = max(timings, key=timings.get)
    return {"n": n, "replicas": replicas, "block_waves": best, "persistent": persistent,
            "proposals_per_s": {str(k): v for k, v in timings.items()}}


def tune(sizes: dict[int, int], *, cache: Path | None = None, refresh: bool = False, **kw) -> dict[int, int]:
    """block_waves per n for the given {n: replicas}; cached per device name."""
    cache = default_cache_path() if cache is None else Path(cache)
    name = torch.cuda.get_device_name(_lib.device())
    table = {}
    if cache.exists() and not refresh:
        try:
            table = json.loads(cache.read_text()).get(name, {})
        except (ValueError, OSError):
            table = {}
    out = {}
    persistent = bool(kw.get("persistent", False))
    for n, replicas in sizes.items():
        key = f"{n}:{replicas}" + (":persistent" if persistent else "")
        if key not in table:
            table[key] = tune_size(n, replicas, **kw)
        out[n] = int(table[key]["block_waves"])
    try:
        cache.parent.mkdir(parents=True, exist_ok=True)
        data = json.loads(cache.read_text()) if cache.exists() else {}
        data[name] = table
        cache.write_text(json.dumps(data, indent=1))
    except (ValueError, OSError):
        pass
    return out


def _fits(n: int, replicas: int, ladder: int, w: int, seed: int) -> bool:
    """Whether a block of `w` replicas of size n fits in LDS (one sweep)."""
    from .batched import BatchGroup, run_batched

    chains = seeded_chains(n, min(replicas, 64 * w), min(ladder, replicas), seed=seed, device=_lib.device())
    chains.block_waves = w
    try:
        run_batched([BatchGroup(chains, min(ladder, replicas), 2 * n)], 1)
    except ValueError:
        return False
    return True


def tune_mixed_shapes(sizes: dict[int, int], *, ladder: int = 64, sweeps: int = 8, candidates=(4, 6, 8, 12, 16),
                      seed: int = 0, persistent: bool = True) -> dict:
    """Coordinate descent over the block shapes of the batched run of all
    `sizes` ({n: replicas}) together, timing `sweeps` sweeps per trial."""
    from .batched import BatchGroup, run_batched

    dev = _lib.device()
    fits = {n: [w for w in sorted(candidates) if _fits(n, r, ladder, w, seed)] for n, r in sizes.items()}
    shapes = {n: ws[-1] for n, ws in fits.items()}
    chains = {n: seeded_chains(n, r, min(ladder, r), seed=seed + i, device=dev)
              for i, (n, r) in enumerate(sizes.items())}
    proposals = sum(r * 2 * n for n, r in sizes.items()) * sweeps

    def rate(trial: dict) -> float:
        groups = []
        for n, c in chains.items():
            c.block_waves = trial[n]
            groups.append(BatchGroup(c, min(ladder, len(c.temperatures_host)), 2 * n))
        run_batched(groups, 1, persistent=persistent)
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        run_batched(groups, sweeps, persistent=persistent)
        torch.cuda.synchronize(dev)
        return proposals / (time.perf_counter() - t0)

    trials = {}
    best_rate = rate(shapes)
    trials[json.dumps(shapes)] = best_rate
    for n in sorted(sizes, reverse=True):
        for w in fits[n]:
            if w == shapes[n]:
                continue
            trial = dict(shapes)
            trial[n] = w
            r = rate(trial)
            trials[json.dumps(trial)] = r
            if r > best_rate:
                best_rate, shapes = r, trial
    _lib.lib().n3_temper_release()
    return {"block_waves": {str(n): w for n, w in shapes.items()}, "proposals_per_s": best_rate,
            "persistent": persistent, "trials": len(trials)}


def tune_mixed(sizes: dict[int, int], *, cache: Path | None = None, refresh: bool = False, **kw) -> dict[int, int]:
    """block_waves per n for a batched run of all `sizes` together; cached
    per device name and size mix."""
    cache = default_cache_path() if cache is None else Path(cache)
    name = torch.cuda.get_device_name(_lib.device())
    persistent = bool(kw.get("persistent", True))
    key = "mixed:" + ",".join(f"{n}:{r}" for n, r in sorted(sizes.items())) + (":persistent" if persistent else "")
    table = {}
    if cache.exists() and not refresh:
        try:
            table = json.loads(cache.read_text()).get(name, {})
        except (ValueError, OSError):
            table = {}
    if key not in table:
        table[key] = tune_mixed_shapes(sizes, **kw)
    try:
        cache.parent.mkdir(parents=True, exist_ok=True)
        data = json.loads(cache.read_text()) if cache.exists() else {}
        data.setdefault(name, {}).update(table)
        cache.write_text(json.dumps(data, indent=1))
    except (ValueError, OSError):
        pass
    return {int(n): int(w) for n, w in table[key]["block_waves"].items()}
