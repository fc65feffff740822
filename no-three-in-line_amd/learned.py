"""Learned move proposals for the tempering walk.

The proposal network (8 features -> 128 hidden -> 1 logit, bf16 parameters,
fp32 accumulation; MFMA kernels in csrc/mlp.hip) scores a candidate moved
point from its collinear-pair count, the replica's temperature and the
state's energy.  A move's score is the sum of its two points' scores and the
network is trained online, with the hand-written backward kernel, on the
(move, accepted, delta) tuples that plain uniform-proposal runs log on the
device.  For sampling, the network is tabulated over (temperature, energy,
count) into integer weights w.  With probability 0.2 the walk proposes a
uniform pair of slots; otherwise it draws both slots in proportion to w (by
the inverse CDF), which samples the softmax over all pairs exactly, since
the pair score is additive.  The Hastings ratio of the forward and reverse
proposal probabilities corrects the acceptance test, and the exact integer
delta still decides acceptance (csrc/learned.hip).
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


GRID_SCALE, GRID_MAX = 128, 255
W2_OFFSET = 8 * 128 + 128  # W1 and b1 come first in the parameter vector


def network_bf16(master: np.ndarray) -> np.ndarray:
    """bf16 parameters from fp32 master weights: W1 and b1 rounded onto the
    grid of multiples of 1/128 in (-2, 2) (what the Adam kernel writes, and
    what the point network's exact inference needs), w2 and b2 to bf16."""
    master = np.ascontiguousarray(master, dtype=np.float32)
    out = to_bf16(master)
    k = np.clip(np.rint(master[:W2_OFFSET] * np.float32(GRID_SCALE)), -GRID_MAX, GRID_MAX).astype(np.float32)
    out[:W2_OFFSET] = to_bf16(k * np.float32(1.0 / GRID_SCALE))
    return out


def temperature_codes(temperatures) -> np.ndarray:
    """lt_q = round(32 ln T) per replica (the point network's temperature
    feature is lt_q / 128)."""
    t = np.asarray(temperatures, dtype=np.float32).astype(np.float64)
    return np.clip(np.rint(32.0 * np.log(t)), -GRID_MAX, GRID_MAX).astype(np.int32)


def point_codes(counts, rows, cols, lt_q, energy) -> np.ndarray:
    """Feature codes [N, 8] of candidate points for the point network
    (csrc/mlp_common.h point_codes); the features are codes / 128."""
    b = np.asarray(counts, dtype=np.int64)
    shape = b.shape
    e = np.broadcast_to(np.asarray(energy, dtype=np.int64), shape)
    out = np.empty(shape + (8,), dtype=np.int64)
    out[..., 0] = 4 * np.minimum(b, 63)
    out[..., 1] = 32 * np.minimum(b, 7)
    out[..., 2] = np.where(b > 0, GRID_SCALE, 0)
    out[..., 3] = np.broadcast_to(np.asarray(rows, dtype=np.int64), shape)
    out[..., 4] = np.broadcast_to(np.asarray(cols, dtype=np.int64), shape)
    out[..., 5] = np.broadcast_to(np.asarray(lt_q, dtype=np.int64), shape)
    out[..., 6] = np.minimum(e, 510) >> 1
    out[..., 7] = 2 * np.minimum(e, 63)
    return out


def point_codes_device(counts: torch.Tensor, rows: torch.Tensor, cols: torch.Tensor, lt_q: torch.Tensor,
                       energy: torch.Tensor) -> torch.Tensor:
    """point_codes on device tensors (int [N] each) -> bf16 bits [N, 8]
    (int16), exact."""
    b = counts.to(torch.int32)
    e = energy.to(torch.int32)
    codes = torch.stack((4 * torch.clamp(b, max=63), 32 * torch.clamp(b, max=7), (b > 0).to(torch.int32) * GRID_SCALE,
                         rows.to(torch.int32), cols.to(torch.int32), lt_q.to(torch.int32),
                         torch.clamp(e, max=510) // 2, 2 * torch.clamp(e, max=63)), dim=1)
    return (codes.to(torch.float32) / GRID_SCALE).to(torch.bfloat16).view(torch.int16).contiguous()


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


def point_features_device(counts: torch.Tensor, log_t: torch.Tensor, energy: torch.Tensor) -> torch.Tensor:
    """point_features on device tensors (float32 [N] each) -> float32 [N, 8]."""
    b = counts.to(torch.float32)
    e = energy.to(torch.float32)
    lt = log_t.to(torch.float32)
    third = torch.tensor(1.0 / 3.0, dtype=torch.float32, device=b.device)
    return torch.stack((b * 0.125, torch.sqrt(b) * third, torch.log1p(b) * third, (b > 0).to(torch.float32), lt,
                        e * torch.tensor(1.0 / 150.0, dtype=torch.float32, device=b.device),
                        torch.sqrt(e) * torch.tensor(1.0 / 12.0, dtype=torch.float32, device=b.device),
                        b * 0.125 * lt), dim=1)


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
        self.params = torch.from_numpy(network_bf16(master).view(np.int16)).to(dev)
        self.m = torch.zeros(PARAMS, dtype=torch.float32, device=dev)
        self.v = torch.zeros(PARAMS, dtype=torch.float32, device=dev)
        self.step = 0

    def _stream(self) -> int:
        return _lib.stream_handle(self.device)

    def clone(self) -> "ProposalNetwork":
        """An independent copy (parameters, master weights, Adam state)."""
        other = ProposalNetwork.__new__(ProposalNetwork)
        other.device = self.device
        for name in ("master", "params", "m", "v"):
            setattr(other, name, getattr(self, name).clone())
        other.step = self.step
        return other

    def forward(self, x_bf16: torch.Tensor) -> torch.Tensor:
        """Logits of feature rows (int16 tensor holding bf16 bits, [N, 8])."""
        rows = int(x_bf16.shape[0])
        out = torch.empty(rows, dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().n3_mlp_forward(self.params.data_ptr(), x_bf16.data_ptr(), rows, out.data_ptr(),
                                             self._stream()))
        return out

    def gradient(self, xu: torch.Tensor, xv: torch.Tensor, y: torch.Tensor,
                 loss: torch.Tensor | None = None) -> tuple[torch.Tensor, float | None]:
        """Gradient of the mean pair loss.  With `loss` given, the summed loss
        is added to it on the device (no synchronisation) and None returned
        in place of the mean loss."""
        grad = torch.zeros(PARAMS, dtype=torch.float32, device=self.device)
        acc = torch.zeros(1, dtype=torch.float32, device=self.device) if loss is None else loss
        _lib.check(_lib.lib().n3_mlp_grad(self.params.data_ptr(), xu.data_ptr(), xv.data_ptr(), y.data_ptr(),
                                          int(y.shape[0]), grad.data_ptr(), acc.data_ptr(), self._stream()))
        if loss is not None:
            return grad, None
        return grad, float(acc.item()) / int(y.shape[0])

    def adam(self, grad: torch.Tensor, lr: float = 1e-2) -> None:
        self.step += 1
        _lib.check(_lib.lib().n3_mlp_adam(self.master.data_ptr(), self.params.data_ptr(), grad.data_ptr(),
                                          self.m.data_ptr(), self.v.data_ptr(), self.step, lr, 0.9, 0.999, 1e-8,
                                          self._stream()))

    def train(self, xu: torch.Tensor, xv: torch.Tensor, y: torch.Tensor, *, steps: int | None = None,
              epochs: int = 2, batch: int = 32768, lr: float = 1e-2, seed: int = 0, reports: int = 4) -> list[float]:
        """Mini-batch Adam on (xu, xv, y).  With `steps`, that many batches
        drawn with replacement on the device; otherwise `epochs` passes over a
        permutation.  Returns the mean loss over each of `reports` equal
        parts of the run (one synchronisation at the end)."""
        total = int(y.shape[0])
        if total == 0:
            raise ValueError("no training tuples")
        gen = torch.Generator(device=self.device).manual_seed(seed)
        batches = []
        if steps is not None:
            for _ in range(int(steps)):
                batches.append(torch.randint(0, total, (min(batch, total),), generator=gen, device=self.device))
        else:
            for _ in range(int(epochs)):
                order = torch.randperm(total, generator=gen, device=self.device)
                batches.extend(order[start:start + batch] for start in range(0, total, batch))
        parts = max(1, min(int(reports), len(batches)))
        sums = torch.zeros(parts, dtype=torch.float32, device=self.device)
        counts = [0] * parts
        for i, idx in enumerate(batches):
            part = i * parts // len(batches)
            grad, _ = self.gradient(xu[idx], xv[idx], y[idx], loss=sums[part:part + 1])
            counts[part] += len(idx)
            self.adam(grad, lr)
        return [float(v) / c for v, c in zip(sums.cpu().tolist(), counts)]

    def weight_table(self, log_temps: np.ndarray, max_energy: int = MAX_ENERGY) -> torch.Tensor:
        lt = torch.from_numpy(np.ascontiguousarray(log_temps, dtype=np.float32)).to(self.device)
        table = torch.empty((len(log_temps), max_energy + 1, 64), dtype=torch.int32, device=self.device)
        _lib.check(_lib.lib().n3_mlp_weight_table(self.params.data_ptr(), lt.data_ptr(), len(log_temps), max_energy,
                                                  table.data_ptr(), self._stream()))
        return table


LOG_FIELDS = 8


class ProposalLog:
    """Device log of proposals (bad_u, bad_v, energy, slot, delta, accepted,
    u | v << 16, column u | column v << 16), filled by uniform-proposal walks
    while attached to chains.  Every replica
    reserves one entry per proposal of a sweep; entries of invalid proposals
    keep the fill value -1, and sweeps whose reservation does not fit are not
    logged."""

    def __init__(self, capacity: int, device: torch.device):
        self.capacity = int(capacity)
        self.entries = torch.full((self.capacity, LOG_FIELDS), -1, dtype=torch.int32, device=device)
        self.count = torch.zeros(1, dtype=torch.int64, device=device)

    def attach(self, chains: TemperingChains) -> None:
        chains.log = self

    def reset(self) -> None:
        self.entries.fill_(-1)
        self.count.zero_()

    def rows(self) -> torch.Tensor:
        """The logged proposals in slot order (stable): a replica's entries
        come in sweep and step order whatever order the replicas reserved
        theirs in, so a log that holds every reservation yields the same rows,
        in the same order, on every run of the same walk."""
        filled = min(int(self.count.item()), self.capacity)
        head = self.entries[:filled]
        head = head[head[:, 3] >= 0]
        return head[torch.argsort(head[:, 3], stable=True)]


def training_tensors(entries: torch.Tensor, temperatures: np.ndarray, device, *, label: str = "metropolis",
                     temperature_floor: float = 0.25, features: str = "table") -> tuple:
    """Feature rows (bf16 bits) of the two moved points and their labels,
    computed on the device.

    label "accepted" is the logged outcome (0/1).  label "metropolis" is the
    acceptance probability min(1, exp(-delta / max(T, temperature_floor)))
    of the logged exact delta: the expectation of the outcome (at T >= the
    floor) without its sampling noise, and, below the floor, the acceptance
    at the floor, since the outcomes there are almost all rejections and
    carry little signal about which points are worth moving."""
    e = entries.to(device)
    temps = torch.from_numpy(np.ascontiguousarray(temperatures, dtype=np.float32)).to(device)
    slot = e[:, 3].long()
    if features == "table":
        lt = torch.log(temps)[slot]
        xu = point_features_device(e[:, 0], lt, e[:, 2]).to(torch.bfloat16).view(torch.int16).contiguous()
        xv = point_features_device(e[:, 1], lt, e[:, 2]).to(torch.bfloat16).view(torch.int16).contiguous()
    elif features == "point":
        # the moved slots and their columns before the move (log fields 6, 7)
        lt_q = torch.from_numpy(temperature_codes(temperatures)).to(device)[slot]
        u, v = e[:, 6] & 0xFFFF, e[:, 6] >> 16
        cu, cv = e[:, 7] & 0xFFFF, e[:, 7] >> 16
        xu = point_codes_device(e[:, 0], u >> 1, cu, lt_q, e[:, 2])
        xv = point_codes_device(e[:, 1], v >> 1, cv, lt_q, e[:, 2])
    else:
        raise ValueError(f"unknown features {features!r}")
    if label == "accepted":
        y = e[:, 5].to(torch.float32)
    elif label == "metropolis":
        t_eff = torch.clamp(temps[slot], min=float(temperature_floor))
        y = torch.exp(-torch.clamp(e[:, 4].to(torch.float32), min=0.0) / t_eff)
    else:
        raise ValueError(f"unknown label {label!r}")
    return xu, xv, y.contiguous()


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


def refresh(proposals: LearnedProposals, chains: TemperingChains, network: ProposalNetwork) -> LearnedProposals:
    """Rewrites the weight table in place from the network's current
    parameters (the tensors keep their addresses, so captured learned-walk
    graphs stay valid and are reused)."""
    log_temps, _ = temperature_index(chains.temperatures_host)
    proposals.weights.copy_(network.weight_table(log_temps, proposals.max_energy))
    return proposals


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


@dataclass
class PointProposals:
    """Tables the point-network walk reads for one set of chains and one
    network (whose bf16 parameters it reads in place)."""

    params: torch.Tensor           # the network's bf16 parameters (int16 bits)
    consts: torch.Tensor           # float32 [260]: b1, w2 in MFMA-lane order, b2
    expw: torch.Tensor             # int32 [737]
    lt_q: torch.Tensor             # int32 [R]
    thr64: torch.Tensor            # int64 [R][2 k64]
    k64: int


_EXPW: np.ndarray | None = None


def point_exp_table() -> np.ndarray:
    global _EXPW
    if _EXPW is None:
        out = np.empty(737, dtype=np.uint32)
        _lib.check(_lib.lib().n3_point_exp_table(out.ctypes.data))
        _EXPW = out
    return _EXPW


def point_constants(params_bf16: np.ndarray) -> np.ndarray:
    """The walk's float constants of a bf16 parameter vector (ValueError if
    W1 or b1 is off the 1/128 grid)."""
    p = np.ascontiguousarray(params_bf16).view(np.uint16)
    out = np.empty(260, dtype=np.float32)
    _lib.check(_lib.lib().n3_point_constants(p.ctypes.data, out.ctypes.data))
    return out


def prepare_point(chains: TemperingChains, network: ProposalNetwork) -> PointProposals:
    dev = chains.device
    consts = torch.from_numpy(point_constants(network.params.cpu().numpy())).to(dev)
    thr = thresholds64(chains.temperatures_host, chains.k_move)
    return PointProposals(network.params, consts, torch.from_numpy(point_exp_table().view(np.int32)).to(dev),
                          torch.from_numpy(temperature_codes(chains.temperatures_host)).to(dev),
                          torch.from_numpy(thr.view(np.int64)).to(dev), chains.k_move)


def refresh_point(proposals: PointProposals, network: ProposalNetwork) -> PointProposals:
    """Rewrites the constants in place from the network's current parameters
    (captured graphs stay valid: every address is unchanged)."""
    if network.params.data_ptr() != proposals.params.data_ptr():
        raise ValueError("the proposals were prepared for another network")
    proposals.consts.copy_(torch.from_numpy(point_constants(network.params.cpu().numpy())))
    return proposals


def point_weights_device(proposals: PointProposals, columns: torch.Tensor, badness: torch.Tensor,
                         scores: torch.Tensor, lt_q: torch.Tensor) -> torch.Tensor:
    """The point weight of every slot of the given states (int32 [count][2n])."""
    count, m = int(columns.shape[0]), int(columns.shape[1])
    out = torch.empty((count, m), dtype=torch.int32, device=columns.device)
    _lib.check(_lib.lib().n3_point_weights(columns.data_ptr(), badness.data_ptr(), scores.data_ptr(),
                                           lt_q.data_ptr(), m // 2, count, proposals.params.data_ptr(),
                                           proposals.consts.data_ptr(), proposals.expw.data_ptr(), out.data_ptr(),
                                           _lib.stream_handle(columns.device)))
    return out


def run_learned_point(chains: TemperingChains, proposals: PointProposals, sweeps: int, steps: int,
                      ladder_size: int = 0) -> None:
    """`sweeps` x (point-proposal walk; exchange), graph-captured, enqueued."""
    if steps < 1 or sweeps < 0:
        raise ValueError("steps must be positive and sweeps non-negative")
    ladder = int(ladder_size)
    if ladder > 1 and len(chains.temperatures_host) % ladder:
        raise ValueError("ladder_size must divide replica count")
    chains._check_umbrella()
    c = chains._chains(ladder if ladder > 1 else None)
    _lib.check(_lib.lib().n3_temper_run_learned_point(
        ctypes.byref(c), proposals.params.data_ptr(), proposals.consts.data_ptr(), proposals.expw.data_ptr(),
        proposals.lt_q.data_ptr(), proposals.thr64.data_ptr(), proposals.k64, int(sweeps), int(steps),
        ladder if ladder > 1 else 0, chains.seed, chains._stream()))


def threshold_times(trace: list[tuple[float, int]], thresholds) -> dict:
    """First time (seconds) at which the best energy reached each threshold."""
    out = {}
    for thr in thresholds:
        out[str(thr)] = next((t for t, e in trace if e <= thr), None)
    return out


def ab_run(n: int, replicas: int, ladder: int, *, seconds: float, seed: int, state_seed: int,
           train_sweeps: int = 200, log_sweeps: int = 2, train_steps: int = 400, retrain_seconds: float = 2.0,
           retrain_steps: int = 100, log_capacity: int = 1 << 21,
           chunk: int = 50, thresholds=(80, 70, 60, 55, 50, 45, 40, 12), tmin: float = 0.05, tmax: float = 2.0,
           start: np.ndarray | None = None, network: "ProposalNetwork | None" = None, modes=("table",),
           device: torch.device | None = None) -> dict:
    """Uniform vs learned proposals on the same seeds, replicas and GPU.

    Every side starts from the same states: seeded generator states, or the
    rows of `start` (int32[replicas][2n]) when given, and gets the same
    wall-clock budget.  The uniform run proceeds in graph-captured chunks.
    Each learned side (one per entry of `modes`: "table" for the tabulated
    (temperature, energy, count) weights, "point" for the point network with
    row and column features evaluated inside the walk) first runs
    `train_sweeps` plain sweeps, the last `log_sweeps` of them logged, trains
    a fresh network on that log (`train_steps` Adam steps) and continues
    with learned proposals; every `retrain_seconds` it logs `log_sweeps` more
    plain sweeps and takes `retrain_steps` more steps on them (online
    training).  All of that counts against its budget.  With `network` given
    (trained beforehand), a clone of it starts the learned side at once and
    only the online retraining happens inside the budget.  The traces record
    (elapsed seconds, best energy) after every chunk."""
    from .batched import seeded_chains

    dev = _lib.device(None if device is None else device.index)
    steps = 2 * n
    if not 0 < log_sweeps <= train_sweeps:
        raise ValueError("need 0 < log_sweeps <= train_sweeps")
    if log_sweeps * steps * replicas > log_capacity:
        raise ValueError("log_capacity is too small for log_sweeps")
    for mode in modes:
        if mode not in ("table", "point"):
            raise ValueError(f"unknown mode {mode!r}")
    result = {}

    def chains():
        if start is None:
            return seeded_chains(n, replicas, ladder, seed=seed, state_seed=state_seed, tmin=tmin, tmax=tmax,
                                 device=dev)
        cols = np.ascontiguousarray(start, dtype=np.int32)
        if cols.shape != (replicas, 2 * n):
            raise ValueError("start must have shape (replicas, 2n)")
        temps = np.tile(np.geomspace(tmin, tmax, ladder), replicas // ladder).astype(np.float32)
        return TemperingChains(cols, temps, seed, device=dev)

    # uniform
    uni = chains()
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
    for mode in modes:
        result["learned" if mode == "table" else "learned_point"] = _learned_side(
            chains(), mode, seconds=seconds, seed=seed, steps=steps, ladder=ladder, chunk=chunk,
            train_sweeps=train_sweeps, log_sweeps=log_sweeps, train_steps=train_steps,
            retrain_seconds=retrain_seconds, retrain_steps=retrain_steps, log_capacity=log_capacity,
            thresholds=thresholds, network=None if network is None else network.clone(), dev=dev)
    result["config"] = {"n": n, "replicas": replicas, "ladder": ladder, "seconds": seconds, "seed": seed,
                        "state_seed": state_seed if start is None else None, "tmin": tmin, "tmax": tmax,
                        "train_sweeps": train_sweeps, "log_sweeps": log_sweeps, "train_steps": train_steps,
                        "retrain_seconds": retrain_seconds, "retrain_steps": retrain_steps,
                        "pretrained": network is not None, "modes": list(modes)}
    return result


def _learned_side(lrn: TemperingChains, mode: str, *, seconds, seed, steps, ladder, chunk, train_sweeps, log_sweeps,
                  train_steps, retrain_seconds, retrain_steps, log_capacity, thresholds, network, dev) -> dict:
    """One learned side of ab_run (see there) on the chains `lrn`."""
    replicas = len(lrn.temperatures_host)
    net = ProposalNetwork(seed=seed, device=dev) if network is None else network
    log = ProposalLog(log_capacity, dev)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    trace = []
    sweeps = 0

    def plain(count, logged):
        nonlocal sweeps
        lrn.log = log if logged else None
        lrn.run(count, steps, ladder)
        lrn.log = None
        sweeps += count

    tuples, losses = 0, []
    if network is None:
        done = 0
        while done < train_sweeps - log_sweeps:
            step = min(chunk, train_sweeps - log_sweeps - done)
            plain(step, False)
            done += step
            torch.cuda.synchronize(dev)
            trace.append((time.perf_counter() - t0, int(lrn.best_scores.min().item())))
        plain(log_sweeps, True)
        xu, xv, y = training_tensors(log.rows(), lrn.temperatures_host, dev, features=mode)
        tuples = int(y.shape[0])
        losses = net.train(xu, xv, y, steps=train_steps, seed=seed)
    if mode == "table":
        proposals = prepare(lrn, net)
        walk = run_learned
    else:
        proposals = prepare_point(lrn, net)
        walk = run_learned_point
    torch.cuda.synchronize(dev)
    train_seconds = time.perf_counter() - t0
    last_train = train_seconds
    retrains = 0
    while time.perf_counter() - t0 < seconds:
        if retrain_seconds > 0 and time.perf_counter() - t0 - last_train >= retrain_seconds:
            log.reset()
            plain(log_sweeps, True)
            xu, xv, y = training_tensors(log.rows(), lrn.temperatures_host, dev, features=mode)
            tuples += int(y.shape[0])
            losses += net.train(xu, xv, y, steps=retrain_steps, seed=seed + retrains + 1)
            if mode == "table":
                proposals = refresh(proposals, lrn, net)
            else:
                proposals = refresh_point(proposals, net)
            torch.cuda.synchronize(dev)
            last_train = time.perf_counter() - t0
            retrains += 1
        walk(lrn, proposals, chunk, steps, ladder)
        sweeps += chunk
        torch.cuda.synchronize(dev)
        trace.append((time.perf_counter() - t0, int(lrn.best_scores.min().item())))
    lrn.validate(range(0, replicas, max(1, replicas // 64)))
    return {"trace": trace, "sweeps": sweeps, "training_tuples": tuples, "training_losses": losses,
            "retrains": retrains, "seconds_until_learned_proposals": train_seconds,
            "threshold_seconds": threshold_times(trace, thresholds), "mode": mode}


def pretrain(n: int, replicas: int, ladder: int, *, seed: int, start: np.ndarray | None = None, state_seed: int = 2,
             sweeps: int = 300, log_sweeps: int = 2, steps: int = 400, tmin: float = 0.05, tmax: float = 2.0,
             features: str = "table", device: torch.device | None = None) -> tuple["ProposalNetwork", dict]:
    """A network trained on the logs of a separate plain calibration run
    (`sweeps` uniform sweeps from `start` or seeded states, the last
    `log_sweeps` logged, `steps` Adam steps): returns it with the run's
    timing."""
    from .batched import seeded_chains

    dev = _lib.device(None if device is None else device.index)
    t0 = time.perf_counter()
    if start is None:
        chains = seeded_chains(n, replicas, ladder, seed=seed, state_seed=state_seed, tmin=tmin, tmax=tmax, device=dev)
    else:
        temps = np.tile(np.geomspace(tmin, tmax, ladder), replicas // ladder).astype(np.float32)
        chains = TemperingChains(np.ascontiguousarray(start, dtype=np.int32), temps, seed, device=dev)
    chains.run(sweeps - log_sweeps, 2 * n, ladder)
    log = ProposalLog(log_sweeps * 2 * n * replicas, dev)
    chains.log = log
    chains.run(log_sweeps, 2 * n, ladder)
    chains.log = None
    xu, xv, y = training_tensors(log.rows(), chains.temperatures_host, dev, features=features)
    net = ProposalNetwork(seed=seed, device=dev)
    losses = net.train(xu, xv, y, steps=steps, seed=seed)
    torch.cuda.synchronize(dev)
    return net, {"seconds": time.perf_counter() - t0, "tuples": int(y.shape[0]), "losses": losses,
                 "sweeps": sweeps, "steps": steps}
