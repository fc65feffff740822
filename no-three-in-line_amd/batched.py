"""Several populations of different grid sizes tempered together in one
walk launch per sweep, graph-captured (include/no3.h, n3_temper_batched_run).

Each group keeps its own TemperingChains (device arrays, temperatures,
thresholds, seed and counters); a batched run leaves every group exactly where
`chains.run(sweeps, steps, ladder)` on that group alone would.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .tempering import TemperingChains


@dataclass
class BatchGroup:
    chains: TemperingChains
    ladder: int
    steps: int
    bad_bias: int = 0


def run_batched(groups: list[BatchGroup], sweeps: int, persistent: bool = False) -> None:
    """Enqueues `sweeps` sweeps of every group on the current stream: one
    walk launch per sweep replayed from captured graphs, or (persistent) all
    sweeps in one cooperative launch; both give the same trajectories."""
    if not 1 <= len(groups) <= _lib.MAX_GROUPS:
        raise ValueError(f"between 1 and {_lib.MAX_GROUPS} groups are supported")
    if sweeps < 0:
        raise ValueError("negative sweeps")
    devices = {g.chains.device for g in groups}
    if len(devices) != 1:
        raise ValueError("all groups must live on one device")
    arr = (_lib.Group * len(groups))()
    for i, g in enumerate(groups):
        if g.steps < 1 or not 0 <= g.bad_bias <= 100:
            raise ValueError("steps must be positive and bad_bias in [0,100]")
        ladder = int(g.ladder)
        if ladder > 1 and len(g.chains.temperatures_host) % ladder:
            raise ValueError("ladder_size must divide replica count")
        g.chains._check_umbrella()
        arr[i].chains = g.chains._chains(ladder if ladder > 1 else None)
        arr[i].ladder = ladder if ladder > 1 else 0
        arr[i].steps = int(g.steps)
        arr[i].bad_bias = int(g.bad_bias)
        arr[i].seed = g.chains.seed
    dev = devices.pop()
    entry = _lib.lib().n3_temper_batched_persistent if persistent else _lib.lib().n3_temper_batched_run
    _lib.check(entry(arr, len(groups), int(sweeps), _lib.stream_handle(dev)))


def seeded_chains(n: int, replicas: int, ladder: int, *, seed: int, tmin: float = 0.05, tmax: float = 2.0,
                  state_seed: int | None = None, device: torch.device | None = None) -> TemperingChains:
    """Chains over seeded generator states with a geometric ladder repeated
    replicas / ladder times (the population layout of the benchmark runs)."""
    from .gpu import generate_states

    dev = _lib.device(None if device is None else device.index)
    p, q = generate_states(seed if state_seed is None else state_seed, n, replicas, device=dev)
    cols = torch.stack((p, q), dim=2).reshape(replicas, 2 * n).cpu().numpy()
    temps = np.tile(np.geomspace(tmin, tmax, ladder), replicas // ladder).astype(np.float32)
    return TemperingChains(cols, temps, seed, device=dev)
