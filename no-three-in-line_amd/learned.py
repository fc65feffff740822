"""Learned move proposals for the tempering walk.

The proposal network (8 features -> 128 hidden -> 1 logit, bf16 parameters,
fp32 accumulation; MFMA kernels in csrc/mlp.hip) scores a candidate moved
point from its collinear-pair count, the replica's temperature and the
state's energy.  A move's score is the sum of its two points' scores and the
network is trained online, with the hand-written backward kernel, on the
(move, accepted, delta) tuples that plain uniform-proposal runs log on the
device.  For sampling, the network is tabulated over (temperature, energy,
count) into integer weights; the walk draws the first point from
0.2 uniform + 0.8 weights, the second uniformly, and corrects with the
Hastings ratio, so the exact integer delta stays the judge of acceptance
(csrc/learned.hip).
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .tempering import TemperingChains

MAX_ENERGY = 1023
PARAMS = 8 * 128 + 128 + 128 + 1


def to_bf16(x: np.ndarray) -> np.ndarray:
    """float32 -> bf16 bits, round to nearest even (finite values)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bf16(h: np.ndarray) -> np.ndarray:
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def point_features(counts, log_t, energy) -> np.ndarray:
    """float32 features [N, 8] of candidate points (the formula of
    point_features in csrc/mlp.hip)."""
    b = np.asarray(counts, dtype=np.float32)
    lt = np.broadcast_to(np.asarray(log_t, dtype=np.float32), b.shape)
    e = np.broadcast_to(np.asarray(energy, dtype=np.float32), b.shape)
    f = np.empty(b.shape + (8,), dtype=np.float32)
    f[..., 0] = b * np.float32(0.125)
    f[..., 1] = np.sqrt(b) * np.float32(1.0 / 3.0)
    f[..., 2] = np.log1p(b) * np.float32(1.0 / 3.0)
    f[..., 3] = (b > 0).astype(np.float32)
    f[..., 4] = lt
    f[..., 5] = e * np.float32(1.0 / 150.0)
    f[..., 6] = np.sqrt(e) * np.float32(1.0 / 12.0)
    f[..., 7] = b * np.float32(0.125) * lt
    return f


class ProposalNetwork:
    """Parameters on the device: bf16 copy used by the kernels, fp32 master
    copy and Adam moments."""

    def __init__(self, seed: int = 0, device: torch.device | None = None):
        dev = _lib.device(None if device is None else device.index)
        self.device = dev
        assert _lib.lib().n3_mlp_params() == PARAMS
        rng = np.random.default_rng(seed)
        w1 = rng.normal(0.0, math.sqrt(2.0 / 8.0), size=(128, 8))
        b1 = np.full(128, 0.01)
        w2 = rng.normal(0.0, math.sqrt(1.0 / 128.0), size=128)
        master = np.concatenate([w1.reshape(-1), b1, w2, [0.0]]).astype(np.float32)
        self.master = torch.from_numpy(master).to(dev)
        self.params = torch.from_numpy(to_bf16(master).view(np.int16)).to(dev)
        self.m = torch.zeros(PARAMS, dtype=torch.float32, device=dev)
        self.v = torch.zeros(PARAMS, dtype=torch.float32, device=dev)
        self.step = 0

    def _stream(self) -> int:
        return _lib.stream_handle(self.device)

    def forward(self, x_bf16: torch.Tensor) -> torch.Tensor:
        """Logits of feature rows (int16 tensor holding bf16 bits, [N, 8])."""
        rows = int(x_bf16.shape[0])
        out = torch.empty(rows, dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().n3_mlp_forward(self.params.data_ptr(), x_bf16.data_ptr(), rows, out.data_ptr(),
                                             self._stream()))
        return out

    def gradient(self, xu: torch.Tensor, xv: torch.Tensor, y: torch.Tensor) -> tuple[torch.Tensor, float]:
        grad = torch.zeros(PARAMS, dtype=torch.float32, device=self.device)
        loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().n3_mlp_grad(self.params.data_ptr(), xu.data_ptr(), xv.data_ptr(), y.data_ptr(),
                                          int(y.shape[0]), grad.data_ptr(), loss.data_ptr(), self._stream()))
        return grad, float(loss.item()) / int(y.shape[0])

    def adam(self, grad: torch.Tensor, lr: float = 3e-3) -> None:
        self.step += 1
        _lib.check(_lib.lib().n3_mlp_adam(self.master.data_ptr(), self.params.data_ptr(), grad.data_ptr(),
                                          self.m.data_ptr(), self.v.data_ptr(), self.step, lr, 0.9, 0.999, 1e-8,
                                          self._stream()))

    def train(self, xu: torch.Tensor, xv: torch.Tensor, y: torch.Tensor, *, epochs: int = 2, batch: int = 32768,
              lr: float = 3e-3, seed: int = 0) -> list[float]:
        """Mini-batch Adam on (xu, xv, y); returns the mean loss per epoch."""
        gen = torch.Generator(device="cpu").manual_seed(seed)
        losses = []
        total = int(y.shape[0])
        for _ in range(epochs):
            order = torch.randperm(total, generator=gen).to(self.device)
            epoch_loss = 0.0
            for start in range(0, total, batch):
                idx = order[start:start + batch]
                grad, loss = self.gradient(xu[idx].contiguous(), xv[idx].contiguous(), y[idx].contiguous())
                self.adam(grad, lr)
                epoch_loss += loss * len(idx)
            losses.append(epoch_loss / total)
        return losses

    def weight_table(self, log_temps: np.ndarray, max_energy: int = MAX_ENERGY) -> torch.Tensor:
        lt = torch.from_numpy(np.ascontiguousarray(log_temps, dtype=np.float32)).to(self.device)
        table = torch.empty((len(log_temps), max_energy + 1, 64), dtype=torch.int32, device=self.device)
        _lib.check(_lib.lib().n3_mlp_weight_table(self.params.data_ptr(), lt.data_ptr(), len(log_temps), max_energy,
                                                  table.data_ptr(), self._stream()))
        return table


class ProposalLog:
    """Device ring of logged proposals (bad_u, bad_v, energy, slot, delta,
    accepted), filled by uniform-proposal walks while attached to chains."""

    def __init__(self, capacity: int, device: torch.device):
        self.capacity = int(capacity)
        self.entries = torch.zeros((self.capacity, 6), dtype=torch.int32, device=device)
        self.count = torch.zeros(1, dtype=torch.int64, device=device)

    def attach(self, chains: TemperingChains) -> None:
        chains.log = self

    def rows(self) -> torch.Tensor:
        filled = min(int(self.count.item()), self.capacity)
        return self.entries[:filled]


def training_tensors(entries: torch.Tensor, temperatures: np.ndarray, device) -> tuple:
    """Feature rows of the two moved points and the accepted labels."""
    e = entries.cpu().numpy()
    log_t = np.log(temperatures.astype(np.float32))[e[:, 3]]
    xu = to_bf16(point_features(e[:, 0], log_t, e[:, 2])).view(np.int16)
    xv = to_bf16(point_features(e[:, 1], log_t, e[:, 2])).view(np.int16)
    y = e[:, 5].astype(np.float32)
    return (torch.from_numpy(np.ascontiguousarray(xu)).to(device),
            torch.from_numpy(np.ascontiguousarray(xv)).to(device), torch.from_numpy(y).to(device))


@dataclass
class LearnedProposals:
    """Tables the learned walk reads for one set of chains."""

    weights: torch.Tensor          # int32 [temps][MAX_ENERGY + 1][64]
    temp_index: torch.Tensor       # int32 [R]
    thr64: torch.Tensor            # int64 [R][2 k64]
    k64: int
    max_energy: int = MAX_ENERGY


def temperature_index(temperatures: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Distinct temperatures (as log T) and the index of every slot's."""
    values, index = np.unique(temperatures.astype(np.float32), return_inverse=True)
    return np.log(values).astype(np.float32), index.astype(np.int32)


def thresholds64(temperatures: np.ndarray, k: int) -> np.ndarray:
    t = np.ascontiguousarray(temperatures, dtype=np.float32)
    out = np.empty((len(t), 2 * k), dtype=np.uint64)
    _lib.check(_lib.lib().n3_metropolis_thresholds64(t.ctypes.data, len(t), k, out.ctypes.data))
    return out


def prepare(chains: TemperingChains, network: ProposalNetwork) -> LearnedProposals:
    log_temps, index = temperature_index(chains.temperatures_host)
    table = network.weight_table(log_temps)
    thr = thresholds64(chains.temperatures_host, chains.k_move)
    return LearnedProposals(table, torch.from_numpy(index).to(chains.device),
                            torch.from_numpy(thr.view(np.int64)).to(chains.device), chains.k_move)


def run_learned(chains: TemperingChains, proposals: LearnedProposals, sweeps: int, steps: int,
                ladder_size: int = 0) -> None:
    """`sweeps` x (learned-proposal walk; exchange), graph-captured, enqueued."""
    if steps < 1 or sweeps < 0:
        raise ValueError("steps must be positive and sweeps non-negative")
    ladder = int(ladder_size)
    if ladder > 1 and len(chains.temperatures_host) % ladder:
        raise ValueError("ladder_size must divide replica count")
    chains._check_umbrella()
    c = chains._chains(ladder if ladder > 1 else None)
    _lib.check(_lib.lib().n3_temper_run_learned(
        ctypes.byref(c), proposals.weights.data_ptr(), proposals.temp_index.data_ptr(), proposals.max_energy,
        proposals.thr64.data_ptr(), proposals.k64, int(sweeps), int(steps), ladder if ladder > 1 else 0,
        chains.seed, chains._stream()))


def threshold_times(trace: list[tuple[float, int]], thresholds) -> dict:
    """First time (seconds) at which the best energy reached each threshold."""
    out = {}
    for thr in thresholds:
        out[str(thr)] = next((t for t, e in trace if e <= thr), None)
    return out


def ab_run(n: int, replicas: int, ladder: int, *, seconds: float, seed: int, state_seed: int,
           train_sweeps: int = 200, log_capacity: int = 1 << 21, chunk: int = 50,
           thresholds=(80, 70, 60, 55, 50, 45, 40, 12), tmin: float = 0.05, tmax: float = 2.0,
           device: torch.device | None = None) -> dict:
    """Uniform vs learned proposals on the same seeds, replicas and GPU.

    Both start from the same seeded states.  The uniform run proceeds in
    graph-captured chunks; the learned run first runs `train_sweeps` plain
    sweeps that log proposals, trains the network on them (timed), then
    continues with learned proposals.  The trace records (elapsed seconds,
    best energy) after every chunk; training time counts against the learned
    run."""
    from .batched import seeded_chains

    dev = _lib.device(None if device is None else device.index)
    steps = 2 * n
    result = {}
    # uniform
    uni = seeded_chains(n, replicas, ladder, seed=seed, state_seed=state_seed, tmin=tmin, tmax=tmax, device=dev)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    trace_u = []
    while time.perf_counter() - t0 < seconds:
        uni.run(chunk, steps, ladder)
        torch.cuda.synchronize(dev)
        trace_u.append((time.perf_counter() - t0, int(uni.best_scores.min().item())))
    uni.validate(range(0, replicas, max(1, replicas // 64)))
    result["uniform"] = {"trace": trace_u, "sweeps": chunk * len(trace_u),
                         "threshold_seconds": threshold_times(trace_u, thresholds)}
    # learned
    lrn = seeded_chains(n, replicas, ladder, seed=seed, state_seed=state_seed, tmin=tmin, tmax=tmax, device=dev)
    net = ProposalNetwork(seed=seed, device=dev)
    log = ProposalLog(log_capacity, dev)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    trace_l = []
    lrn.log = log
    done = 0
    while done < train_sweeps:
        lrn.run(min(chunk, train_sweeps - done), steps, ladder)
        done += min(chunk, train_sweeps - done)
        torch.cuda.synchronize(dev)
        trace_l.append((time.perf_counter() - t0, int(lrn.best_scores.min().item())))
    lrn.log = None
    xu, xv, y = training_tensors(log.rows(), lrn.temperatures_host, dev)
    losses = net.train(xu, xv, y, epochs=3, seed=seed)
    proposals = prepare(lrn, net)
    torch.cuda.synchronize(dev)
    train_seconds = time.perf_counter() - t0
    while time.perf_counter() - t0 < seconds:
        run_learned(lrn, proposals, chunk, steps, ladder)
        torch.cuda.synchronize(dev)
        trace_l.append((time.perf_counter() - t0, int(lrn.best_scores.min().item())))
    lrn.validate(range(0, replicas, max(1, replicas // 64)))
    result["learned"] = {"trace": trace_l, "sweeps": train_sweeps + chunk * (len(trace_l) - (train_sweeps + chunk - 1)
                                                                             // chunk),
                         "training_tuples": int(y.shape[0]), "training_losses": losses,
                         "seconds_until_learned_proposals": train_seconds,
                         "threshold_seconds": threshold_times(trace_l, thresholds)}
    result["config"] = {"n": n, "replicas": replicas, "ladder": ladder, "seconds": seconds, "seed": seed,
                        "state_seed": state_seed, "tmin": tmin, "tmax": tmax, "train_sweeps": train_sweeps}
    return result
