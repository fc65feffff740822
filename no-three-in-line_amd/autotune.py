"""Per-size launch-shape tuner for the tempering walk.

The walk keeps one replica per wavefront; the tunable is how many replicas
(wavefronts) share a workgroup -- they share one LDS copy of the direction
table, so more waves per block means fewer table copies but coarser blocks
and, at large n, LDS pressure (the library rejects shapes that do not fit).
For each n the tuner times a short graph-captured run of every candidate on
the device and keeps the fastest; results are cached as JSON per device.
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
