"""2-switch and 3-cycle move scores on the GPU.

score_two_switches_gpu is the drop-in for gpu_delta.py:86-116 (int32 with
the sentinel 1<<30); evaluate_two_switches / evaluate_transpositions keep the
int64 signatures and 1<<60 sentinels of accelerated.py:40-195; delta_table
produces the full m x m table of a population.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib


def _state_arrays(p, q, badness, dtype):
    p_host = np.ascontiguousarray(p, dtype=np.int32)
    q_host = np.ascontiguousarray(q, dtype=np.int32)
    bad_host = np.ascontiguousarray(badness, dtype=dtype)
    if p_host.ndim != 1 or q_host.shape != p_host.shape or bad_host.shape != (2 * len(p_host),):
        raise ValueError("invalid P/Q/badness shapes")
    return p_host, q_host, bad_host


def score_two_switches_gpu(p, q, selected_index: int, current_score: int, badness) -> np.ndarray:
    """gpu_delta.py:86-116: scores after switching slot `selected_index` with every slot."""
    p_host, q_host, bad_host = _state_arrays(p, q, badness, np.int32)
    n = len(p_host)
    dev = _lib.device()
    with _lib.STAGING_LOCK:
        # staged words: P, Q, badness, the status word (zeroed here), the
        # scores: one copy in, one copy out (status and scores), one sync
        st = _lib.staging(dev, 6 * n + 1)
        h = st.host_np
        h[:n], h[n:2 * n], h[2 * n:4 * n], h[4 * n] = p_host, q_host, bad_host, 0
        st.dev[:4 * n + 1].copy_(st.host[:4 * n + 1], non_blocking=True)
        base = st.dev.data_ptr()
        _lib.check(_lib.lib().n3_two_switch_scores_status(base, base + 4 * n, base + 8 * n, n, int(selected_index),
                                                          int(current_score), base + 16 * n + 4, base + 16 * n,
                                                          _lib.stream_handle(dev)))
        st.host[4 * n:6 * n + 1].copy_(st.dev[4 * n:6 * n + 1], non_blocking=True)
        torch.cuda.current_stream(dev).synchronize()
        if h[4 * n] & 1:
            raise ValueError("a column lies outside [0, n)")
        return h[4 * n + 1:6 * n + 1].copy()


def evaluate_two_switches(p, q, selected_index: int, current_score: int, badness) -> np.ndarray:
    """accelerated.py:122-195 computed on the GPU (int64, sentinel 1<<60)."""
    p_host, q_host, bad_host = _state_arrays(p, q, badness, np.int64)
    n = len(p_host)
    dev = _lib.device()
    tp, tq, tb = (torch.from_numpy(a).to(dev) for a in (p_host, q_host, bad_host))
    out = torch.empty(2 * n, dtype=torch.int64, device=dev)
    _lib.check(_lib.lib().n3_evaluate_two_switches(tp.data_ptr(), tq.data_ptr(), tb.data_ptr(), n,
                                                   int(selected_index), int(current_score), out.data_ptr(),
                                                   _lib.stream_handle(dev)))
    return out.cpu().numpy()


def evaluate_transpositions(p, q, layer: int, selected_row: int, current_score: int, badness) -> np.ndarray:
    """accelerated.py:40-119 computed on the GPU (int64[n], sentinel 1<<60)."""
    p_host, q_host, bad_host = _state_arrays(p, q, badness, np.int64)
    n = len(p_host)
    dev = _lib.device()
    tp, tq, tb = (torch.from_numpy(a).to(dev) for a in (p_host, q_host, bad_host))
    out = torch.empty(n, dtype=torch.int64, device=dev)
    _lib.check(_lib.lib().n3_evaluate_transpositions(tp.data_ptr(), tq.data_ptr(), tb.data_ptr(), n, int(layer),
                                                     int(selected_row), int(current_score), out.data_ptr(),
                                                     _lib.stream_handle(dev)))
    return out.cpu().numpy()


def evaluate_three_cycles(p, q, candidates, current_score: int, badness) -> np.ndarray:
    """accelerated.py:198-276 computed on the GPU: int64 score per candidate
    row (u, v, w, orientation)."""
    p_host, q_host, bad_host = _state_arrays(p, q, badness, np.int64)
    cand = np.ascontiguousarray(candidates, dtype=np.int32).reshape(-1, 4)
    n = len(p_host)
    if len(cand) == 0:
        return np.empty(0, dtype=np.int64)
    dev = _lib.device()
    tp, tq, tb, tc = (torch.from_numpy(a).to(dev) for a in (p_host, q_host, bad_host, cand))
    out = torch.empty(len(cand), dtype=torch.int64, device=dev)
    _lib.check(_lib.lib().n3_evaluate_three_cycles(tp.data_ptr(), tq.data_ptr(), tb.data_ptr(), n, tc.data_ptr(),
                                                   len(cand), int(current_score), out.data_ptr(),
                                                   _lib.stream_handle(dev)))
    return out.cpu().numpy()


def delta_table_device(p: torch.Tensor, q: torch.Tensor, scores: torch.Tensor, badness: torch.Tensor,
                       out: torch.Tensor | None = None) -> torch.Tensor:
    """Full 2-switch table int32[P][2n][2n] (sentinel 1<<30) of device batches."""
    count, n = int(p.shape[0]), int(p.shape[1])
    dev = _lib.device(p.device.index)
    for t, shape in ((q, (count, n)), (scores, (count,)), (badness, (count, 2 * n))):
        if t.shape != shape or t.dtype != torch.int32 or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("invalid P/Q/scores/badness tensors")
    if out is None:
        out = torch.empty((count, 2 * n, 2 * n), dtype=torch.int32, device=dev)
    _lib.check(_lib.lib().n3_delta_table(p.data_ptr(), q.data_ptr(), scores.data_ptr(), badness.data_ptr(), n, count,
                                         out.data_ptr(), _lib.stream_handle(dev)))
    return out


def delta_table(p_batch, q_batch) -> np.ndarray:
    """Host form of delta_table_device; scores and badness are computed first."""
    from .gpu import _validate_batches, score_population_device

    p_host, q_host = _validate_batches(p_batch, q_batch)
    dev = _lib.device()
    tp = torch.from_numpy(p_host).to(dev)
    tq = torch.from_numpy(q_host).to(dev)
    bad = torch.empty((p_host.shape[0], 2 * p_host.shape[1]), dtype=torch.int32, device=dev)
    scores = score_population_device(tp, tq, badness=bad)
    return _lib.to_host(delta_table_device(tp, tq, scores, bad))[0]
