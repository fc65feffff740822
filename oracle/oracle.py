"""ctypes front-end of the CPU oracle (``oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``, always as the checker or as the
timed CPU baseline, never by the product package.

Each function mirrors the reference routine cited in ``oracle.c``; the array
conventions follow the reference (``p``/``q`` int32 permutations of length n,
badness int64 on the CPU side as in accelerated.py:25).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
SENTINEL64 = 1 << 60
SENTINEL32 = 1 << 30

_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


def build(force: bool = False) -> Path:
    """Compile liboracle.so with the Makefile next to this file."""
    if force or not LIB_PATH.exists() or LIB_PATH.stat().st_mtime < (HERE / "oracle.c").stat().st_mtime:
        subprocess.run(["make", "-B" if force else "-s", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        L.n3o_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.n3o_generate_states.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, _i32p, _i32p]
        L.n3o_score_and_badness.argtypes = [_i32p, _i32p, ctypes.c_int, _i64p]
        L.n3o_score_and_badness.restype = ctypes.c_int64
        L.n3o_score_slope.argtypes = [_i32p, _i32p, ctypes.c_int, _i64p]
        L.n3o_score_slope.restype = ctypes.c_int64
        L.n3o_score_population.argtypes = [_i32p, _i32p, ctypes.c_int, ctypes.c_int64, _i32p, ctypes.c_void_p, ctypes.c_int]
        L.n3o_evaluate_transpositions.argtypes = [_i32p, _i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64, _i64p, _i64p]
        L.n3o_evaluate_two_switches.argtypes = [_i32p, _i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int64, _i64p, _i64p]
        L.n3o_delta_table.argtypes = [_i32p, _i32p, ctypes.c_int, ctypes.c_int64, _i32p]
        L.n3o_evaluate_three_cycles.argtypes = [_i32p, _i32p, ctypes.c_int, _i32p, ctypes.c_int64, ctypes.c_int64,
                                                _i64p, _i64p]
        L.n3o_metropolis_thresholds.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, _u32p]
        L.n3o_exchange_thresholds.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _u32p, _i8p]
        L.n3o_advance.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32,
            _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _u32p, ctypes.c_int, ctypes.c_int,
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
        ]
        L.n3o_advance.restype = ctypes.c_int64
        L.n3o_exp_word.argtypes = [ctypes.c_double]
        L.n3o_exp_word.restype = ctypes.c_uint32
        L.n3o_exchange.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32,
            _i32p, _i32p, _i32p, _i32p, _u32p, _i8p, ctypes.c_int,
        ]
        L.n3o_exchange.restype = ctypes.c_int
        _u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
        L.n3o_metropolis_thresholds64.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, _u64p]
        L.n3o_advance_learned.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32,
            _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _u32p, _i32p, ctypes.c_int, _u64p,
            ctypes.c_int, ctypes.c_int,
        ]
        L.n3o_advance_learned.restype = ctypes.c_int64
        _u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
        L.n3o_point_weights.argtypes = [ctypes.c_int, ctypes.c_int64, _i32p, _i32p, _i32p, _i32p, _u16p, _u32p, _u32p]
        L.n3o_point_weights.restype = None
        L.n3o_advance_point.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32,
            _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _u16p, _u32p, _i32p, _u64p,
            ctypes.c_int, ctypes.c_int,
        ]
        L.n3o_advance_point.restype = ctypes.c_int64
        L.n3o_set_learned_variant.argtypes = [ctypes.c_int]
        L.n3o_set_learned_variant.restype = None
        _u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
        L.n3o_mutate_population.argtypes = [
            _i32p, _i32p, _i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
            ctypes.c_int, ctypes.c_uint64, ctypes.c_int, _i32p, _i32p,
        ]
        L.n3o_enumerate_switches.argtypes = [_i32p, _i32p, _i32p, ctypes.c_int64, ctypes.c_int, _i32p, _i32p, _u8p]
        _lib = L
    return _lib


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def philox4x32_10(counter, key) -> np.ndarray:
    out = np.zeros(4, dtype=np.uint32)
    lib().n3o_philox4x32_10(np.asarray(counter, dtype=np.uint32), np.asarray(key, dtype=np.uint32), out)
    return out


def generate_states(seed: int, n: int, count: int, first_index: int = 0) -> tuple[np.ndarray, np.ndarray]:
    p = np.empty((count, n), dtype=np.int32)
    q = np.empty((count, n), dtype=np.int32)
    lib().n3o_generate_states(seed, n, count, first_index, p, q)
    return p, q


def score_and_badness(p, q) -> tuple[int, np.ndarray]:
    """accelerated.py:14-37."""
    p, q = _i32(p), _i32(q)
    bad = np.zeros(2 * len(p), dtype=np.int64)
    score = lib().n3o_score_and_badness(p, q, len(p), bad)
    return int(score), bad


def score_slope(p, q) -> tuple[int, np.ndarray]:
    """Normalized-direction counting (verifier_b.py:8-33); raises on coincident points."""
    p, q = _i32(p), _i32(q)
    bad = np.zeros(2 * len(p), dtype=np.int64)
    score = lib().n3o_score_slope(p, q, len(p), bad)
    if score < 0:
        raise ValueError("duplicate point has no direction")
    return int(score), bad


def score_population(p_batch, q_batch, *, badness: bool = True, threads: int = 0):
    p, q = _i32(p_batch), _i32(q_batch)
    count, n = p.shape
    scores = np.empty(count, dtype=np.int32)
    bad = np.empty((count, 2 * n), dtype=np.int32) if badness else None
    lib().n3o_score_population(p, q, n, count, scores, None if bad is None else bad.ctypes.data, threads)
    return scores, bad


def evaluate_transpositions(p, q, layer: int, selected_row: int, current_score: int, badness) -> np.ndarray:
    """accelerated.py:40-119 (int64, sentinel 1<<60)."""
    p, q = _i32(p), _i32(q)
    out = np.empty(len(p), dtype=np.int64)
    lib().n3o_evaluate_transpositions(p, q, len(p), layer, selected_row, current_score,
                                      np.ascontiguousarray(badness, dtype=np.int64), out)
    return out


def evaluate_two_switches(p, q, selected_index: int, current_score: int, badness) -> np.ndarray:
    """accelerated.py:122-195 (int64, sentinel 1<<60)."""
    p, q = _i32(p), _i32(q)
    out = np.empty(2 * len(p), dtype=np.int64)
    lib().n3o_evaluate_two_switches(p, q, len(p), selected_index, current_score,
                                    np.ascontiguousarray(badness, dtype=np.int64), out)
    return out


def evaluate_three_cycles(p, q, candidates, current_score: int, badness) -> np.ndarray:
    """accelerated.py:198-276 (int64 per candidate row (u, v, w, orientation))."""
    p, q = _i32(p), _i32(q)
    cand = np.ascontiguousarray(candidates, dtype=np.int32).reshape(-1, 4)
    out = np.empty(len(cand), dtype=np.int64)
    lib().n3o_evaluate_three_cycles(p, q, len(p), cand, len(cand), current_score,
                                    np.ascontiguousarray(badness, dtype=np.int64), out)
    return out


def mutate_population(parent_p, parent_q, parent_badness, *, quality_count: int, population: int,
                      active_strength: int, max_strength: int, bad_bias: int, seed: int, generation: int):
    """gpu_device_search.py:27-113 with the Philox draws (child, generation,
    draw, "MUTA"): children P/Q int32[population][n]."""
    p, q = _i32(parent_p), _i32(parent_q)
    elite_count, n = p.shape
    bad = _i32(parent_badness).reshape(elite_count, 2 * n)
    cp = np.empty((population, n), dtype=np.int32)
    cq = np.empty((population, n), dtype=np.int32)
    lib().n3o_mutate_population(p, q, bad, elite_count, quality_count, n, population, active_strength, max_strength,
                                bad_bias, seed & (2**64 - 1), generation, cp, cq)
    return cp, cq


def enumerate_switches(parent_p, parent_q, descriptors):
    """gpu_exhaustive_beam.py:17-57: children and validity of the
    descriptors (parent, u, v)."""
    p, q = _i32(parent_p), _i32(parent_q)
    n = p.shape[1]
    d = _i32(descriptors).reshape(-1, 3)
    cp = np.empty((len(d), n), dtype=np.int32)
    cq = np.empty((len(d), n), dtype=np.int32)
    valid = np.empty(len(d), dtype=np.uint8)
    lib().n3o_enumerate_switches(p, q, d, len(d), n, cp, cq, valid)
    return cp, cq, valid.astype(bool)


def delta_table(p_batch, q_batch) -> np.ndarray:
    """Full m x m 2-switch table per state, int32 with sentinel 1<<30."""
    p, q = _i32(p_batch), _i32(q_batch)
    count, n = p.shape
    table = np.empty((count, 2 * n, 2 * n), dtype=np.int32)
    lib().n3o_delta_table(p, q, n, count, table)
    return table


def exp_word(x: float) -> int:
    """floor(2^32 exp(-x)) as the engine's umbrella acceptance computes it
    (n3o_exp_word)."""
    return int(lib().n3o_exp_word(float(x)))


def metropolis_thresholds(temperatures, k: int) -> np.ndarray:
    t = np.ascontiguousarray(temperatures, dtype=np.float32)
    thr = np.empty((len(t), k), dtype=np.uint32)
    lib().n3o_metropolis_thresholds(t, len(t), k, thr)
    return thr


def exchange_thresholds(temperatures, ladder: int, k: int) -> tuple[np.ndarray, np.ndarray]:
    t = np.ascontiguousarray(temperatures, dtype=np.float32)
    thr = np.empty((len(t), k), dtype=np.uint32)
    sign = np.empty(len(t), dtype=np.int8)
    lib().n3o_exchange_thresholds(t, len(t), ladder, k, thr, sign)
    return thr, sign


class TemperingReplay:
    """CPU replay of the tempering state machine (tempering.py:119-191) on the
    Philox stream; arrays are slot-ordered like TemperingChains'."""

    def __init__(self, columns, temperatures, seed: int, *, reference=None, k_move: int, k_exchange: int):
        self.columns = _i32(columns).copy()
        self.replicas, m = self.columns.shape
        self.n = m // 2
        self.temperatures = np.ascontiguousarray(temperatures, dtype=np.float32)
        self.seed = int(seed)
        scores, bad = [], []
        for row in self.columns:
            s, b = score_and_badness(row[0::2], row[1::2])
            scores.append(s)
            bad.append(b)
        self.scores = np.asarray(scores, dtype=np.int32)
        self.badness = np.asarray(bad, dtype=np.int32)
        self.best_columns = self.columns.copy()
        self.best_scores = self.scores.copy()
        self.accepted = np.zeros(self.replicas, dtype=np.int32)
        self.reference = _i32(self.columns[0] if reference is None else reference).copy()
        ref = set((i // 2, int(c)) for i, c in enumerate(self.reference))
        self.removed_points = np.asarray(
            [len(ref - set((i // 2, int(c)) for i, c in enumerate(row))) for row in self.columns], dtype=np.int32)
        self.k_move = k_move
        self.k_exchange = k_exchange
        self.centers = np.zeros(self.replicas, dtype=np.float32)
        self.strengths = np.zeros(self.replicas, dtype=np.float32)
        self.thr = metropolis_thresholds(self.temperatures, k_move)
        self.sweep = 0
        self.round = 0

    def advance(self, steps: int, bad_bias: int = 0, threads: int = 0) -> int:
        """One sweep of `steps` proposals per slot, with the umbrella term of
        tempering.py:92-95 when any of `strengths` is nonzero."""
        umbrella = None
        if np.any(self.strengths != 0):
            centers = np.ascontiguousarray(self.centers, dtype=np.float32)
            strengths = np.ascontiguousarray(self.strengths, dtype=np.float32)
            inverse_t = 1.0 / self.temperatures.astype(np.float64)
            umbrella = (centers, strengths, inverse_t)
        ptrs = [None, None, None] if umbrella is None else [a.ctypes.data for a in umbrella]
        valid = lib().n3o_advance(
            self.n, self.replicas, steps, bad_bias, self.seed, self.sweep, self.columns, self.badness, self.scores,
            self.best_columns, self.best_scores, self.accepted, self.reference, self.removed_points, self.thr,
            self.k_move, threads, *ptrs)
        self.sweep += 1
        return int(valid)

    def advance_learned(self, steps: int, weights, temp_index, max_energy: int, thr64, threads: int = 0) -> int:
        """The learned-proposal walk (see n3o_advance_learned) on the same state."""
        weights = np.ascontiguousarray(weights, dtype=np.uint32)
        thr64 = np.ascontiguousarray(thr64, dtype=np.uint64)
        k64 = thr64.shape[1] // 2
        valid = lib().n3o_advance_learned(
            self.n, self.replicas, steps, self.seed, self.sweep, self.columns, self.badness, self.scores,
            self.best_columns, self.best_scores, self.accepted, self.reference, self.removed_points,
            weights.reshape(-1), np.ascontiguousarray(temp_index, dtype=np.int32), max_energy, thr64, k64, threads)
        self.sweep += 1
        return int(valid)

    def advance_point(self, steps: int, params, expw, lt_q, thr64, threads: int = 0) -> int:
        """The walk with point-network proposals and the two-stage
        Metropolis-Hastings test (see n3o_advance_point) on the same state."""
        thr64 = np.ascontiguousarray(thr64, dtype=np.uint64)
        valid = lib().n3o_advance_point(
            self.n, self.replicas, steps, self.seed, self.sweep, self.columns, self.badness, self.scores,
            self.best_columns, self.best_scores, self.accepted, self.reference, self.removed_points,
            np.ascontiguousarray(params, dtype=np.uint16), np.ascontiguousarray(expw, dtype=np.uint32),
            np.ascontiguousarray(lt_q, dtype=np.int32), thr64, thr64.shape[1] // 2, threads)
        self.sweep += 1
        return int(valid)

    def exchange(self, ladder_size: int) -> int:
        if np.any(self.strengths != 0):
            raise ValueError("temperature-only exchange is disabled with umbrella potentials")
        thr, sign = exchange_thresholds(self.temperatures, ladder_size, self.k_exchange)
        count = lib().n3o_exchange(self.n, self.replicas, ladder_size, self.seed, self.round, self.columns,
                                   self.badness, self.scores, self.removed_points, thr, sign, self.k_exchange)
        self.round += 1
        return int(count)


def point_weights(columns, badness, scores, lt_q, params, expw) -> np.ndarray:
    """The point network's integer weight of every slot of every state
    (n3o_point_weights): columns and badness int32[count][2n]."""
    columns = _i32(columns)
    count, m = columns.shape
    out = np.empty((count, m), dtype=np.uint32)
    lib().n3o_point_weights(m // 2, count, columns, _i32(badness), _i32(scores), _i32(lt_q),
                            np.ascontiguousarray(params, dtype=np.uint16), np.ascontiguousarray(expw, dtype=np.uint32),
                            out)
    return out


def set_learned_variant(variant: int) -> None:
    """Plants a mistake in the oracle's learned walk (0 = none; see
    n3o_set_learned_variant), for tests that show their check catches it."""
    lib().n3o_set_learned_variant(int(variant))


def cpu_count() -> int:
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)


def metropolis_thresholds64(temperatures, k: int) -> np.ndarray:
    t = np.ascontiguousarray(temperatures, dtype=np.float32)
    thr = np.empty((len(t), 2 * k), dtype=np.uint64)
    lib().n3o_metropolis_thresholds64(t, len(t), k, thr)
    return thr


# -- proposal network in float64 (the check of the MFMA kernels) ---------------

MLP_IN, MLP_HIDDEN = 8, 128


def bf16_to_f64(h) -> np.ndarray:
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def mlp_unpack(params_bf16):
    p = bf16_to_f64(params_bf16)
    w1 = p[: MLP_HIDDEN * MLP_IN].reshape(MLP_HIDDEN, MLP_IN)
    b1 = p[MLP_HIDDEN * MLP_IN: MLP_HIDDEN * MLP_IN + MLP_HIDDEN]
    w2 = p[MLP_HIDDEN * MLP_IN + MLP_HIDDEN: MLP_HIDDEN * MLP_IN + 2 * MLP_HIDDEN]
    b2 = p[-1]
    return w1, b1, w2, b2


def mlp_forward(params_bf16, x_bf16) -> np.ndarray:
    """g(x) = w2 . relu(W1 x + b1) + b2 in float64 on bf16-valued inputs."""
    w1, b1, w2, b2 = mlp_unpack(params_bf16)
    x = bf16_to_f64(x_bf16).reshape(-1, MLP_IN)
    return np.maximum(x @ w1.T + b1, 0.0) @ w2 + b2


def mlp_pair_grad(params_bf16, xu_bf16, xv_bf16, y) -> tuple[np.ndarray, float]:
    """Gradient (float64, packed like the parameters) and summed loss of the
    mean BCE(sigmoid(g(xu) + g(xv) - b2), y) over the pairs."""
    w1, b1, w2, b2 = mlp_unpack(params_bf16)
    xu = bf16_to_f64(xu_bf16).reshape(-1, MLP_IN)
    xv = bf16_to_f64(xv_bf16).reshape(-1, MLP_IN)
    y = np.asarray(y, dtype=np.float64)
    pre_u, pre_v = xu @ w1.T + b1, xv @ w1.T + b1
    h_u, h_v = np.maximum(pre_u, 0), np.maximum(pre_v, 0)
    z = h_u @ w2 + h_v @ w2 + b2
    loss = float(np.sum(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z)))))
    dz = (1.0 / (1.0 + np.exp(-z)) - y) / len(y)
    dh_u = np.outer(dz, w2) * (pre_u > 0)
    dh_v = np.outer(dz, w2) * (pre_v > 0)
    g_w1 = dh_u.T @ xu + dh_v.T @ xv
    g_b1 = dh_u.sum(0) + dh_v.sum(0)
    g_w2 = h_u.T @ dz + h_v.T @ dz
    g_b2 = dz.sum()
    return np.concatenate([g_w1.reshape(-1), g_b1, g_w2, [g_b2]]), loss
