"""Population scoring on the GPU (drop-in for the reference's gpu.py and the
score_population_badness kernel of gpu_device_search.py).

Host-array entry points copy to the device and back like the reference's
(gpu.py:69-77), the outputs into page-locked buffers; device entry points
take torch tensors that are already resident in HBM and enqueue on the
current stream.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib


def _validate_batches(p_batch, q_batch) -> tuple[np.ndarray, np.ndarray]:
    """gpu.py:61-66."""
    p = np.ascontiguousarray(p_batch, dtype=np.int32)
    q = np.ascontiguousarray(q_batch, dtype=np.int32)
    if p.ndim != 2 or q.shape != p.shape or p.shape[1] < 2:
        raise ValueError("P and Q must have the same two-dimensional shape")
    return p, q


def _check_device_batch(p: torch.Tensor, q: torch.Tensor) -> tuple[int, int]:
    if p.dim() != 2 or q.shape != p.shape or p.shape[1] < 2:
        raise ValueError("P and Q must have the same two-dimensional shape")
    if p.dtype != torch.int32 or q.dtype != torch.int32:
        raise ValueError("P and Q must be int32")
    if not (p.is_cuda and q.is_cuda and p.is_contiguous() and q.is_contiguous()):
        raise ValueError("P and Q must be contiguous device tensors")
    return int(p.shape[0]), int(p.shape[1])


def score_population_device(p: torch.Tensor, q: torch.Tensor, scores: torch.Tensor | None = None,
                            badness: torch.Tensor | None = None, *, method: int = _lib.SCORE_AUTO,
                            check: bool = True) -> torch.Tensor:
    """Exact triple counts (and badness, when a tensor is given) for device
    batches p, q int32[P][n] -- the score_population / score_population_badness
    kernels (gpu.py:13-42, gpu_device_search.py:118-161).  With check=False the
    call only enqueues (no synchronisation); call check_status() later."""
    count, n = _check_device_batch(p, q)
    dev = _lib.device(p.device.index)
    if scores is None:
        scores = torch.empty(count, dtype=torch.int32, device=dev)
    if scores.shape != (count,) or scores.dtype != torch.int32 or not scores.is_contiguous():
        raise ValueError("scores must be a contiguous int32 tensor of length P")
    if badness is not None and (badness.shape != (count, 2 * n) or badness.dtype != torch.int32
                                or not badness.is_contiguous()):
        raise ValueError("badness must be a contiguous int32 tensor of shape (P, 2n)")
    _lib.check(_lib.lib().n3_score_population(p.data_ptr(), q.data_ptr(), n, count, scores.data_ptr(),
                                              _lib.ptr(badness), int(method), _lib.stream_handle(dev)))
    if check:
        check_status(dev)
    return scores


def check_status(dev: torch.device | None = None) -> None:
    """Synchronises and raises ValueError if a scored column was out of range."""
    dev = _lib.device(None if dev is None else dev.index)
    _lib.check(_lib.lib().n3_score_status(_lib.stream_handle(dev)))


def _score_host(p: np.ndarray, q: np.ndarray, with_badness: bool):
    """Scores (and badness) of host batches: the inputs go in with the
    runtime's pageable copies (copying them in chunks, on side streams
    behind the scoring, measured slower: the pageable copies do not overlap
    the kernels); the outputs come back into page-locked host buffers, which
    the returned arrays view (39.3 MB of badness at n=75, P=65,536: 4x faster
    than a copy into pageable memory).  Raises like score_population_device
    for out-of-range columns."""
    count, n = p.shape
    dev = _lib.device()
    tp = torch.from_numpy(p).to(dev)
    tq = torch.from_numpy(q).to(dev)
    bad = torch.empty((count, 2 * n), dtype=torch.int32, device=dev) if with_badness else None
    scores = score_population_device(tp, tq, badness=bad)
    if bad is None:
        return _lib.to_host(scores)[0], None
    return _lib.to_host(scores, bad)


def score_population_gpu(p_batch, q_batch) -> np.ndarray:
    """Return exact integer triple counts for a population of P/Q states (gpu.py:69-77)."""
    p, q = _validate_batches(p_batch, q_batch)
    return _score_host(p, q, False)[0]


def score_population_badness(p_batch, q_batch) -> tuple[np.ndarray, np.ndarray]:
    """Scores and per-slot badness int32[P][2n] of host batches
    (the outputs of gpu_device_search.py:118-161)."""
    p, q = _validate_batches(p_batch, q_batch)
    return _score_host(p, q, True)


def generate_states(seed: int, n: int, count: int, first_index: int = 0,
                    device: torch.device | None = None) -> tuple[torch.Tensor, torch.Tensor]:
    """Seeded two-permutation population on the device (include/no3.h,
    n3_generate_states): P, Q int32[count][n] with P[r] != Q[r]."""
    dev = _lib.device(None if device is None else device.index)
    p = torch.empty((count, n), dtype=torch.int32, device=dev)
    q = torch.empty((count, n), dtype=torch.int32, device=dev)
    _lib.check(_lib.lib().n3_generate_states(int(seed) & (2**64 - 1), n, count, first_index, p.data_ptr(),
                                             q.data_ptr(), _lib.stream_handle(dev)))
    return p, q


def gpu_device_name() -> str:
    dev = _lib.device()
    return torch.cuda.get_device_name(dev)


def score_and_badness(p, q) -> tuple[int, np.ndarray]:
    """accelerated.py:14-37 on the GPU for one state: (score, badness int64[2n]).
    The state goes through the device's pinned staging buffer: one copy in,
    one copy out."""
    p_host, q_host = _validate_batches(np.asarray(p)[None, :], np.asarray(q)[None, :])
    n = p_host.shape[1]
    dev = _lib.device()
    with _lib.STAGING_LOCK:
        st = _lib.staging(dev, 4 * n + 1)
        h = st.host_np
        h[:n], h[n:2 * n] = p_host[0], q_host[0]
        st.dev[:2 * n].copy_(st.host[:2 * n], non_blocking=True)
        base = st.dev.data_ptr()
        stream = _lib.stream_handle(dev)
        _lib.check(_lib.lib().n3_score_population(base, base + 4 * n, n, 1, base + 16 * n, base + 8 * n,
                                                  _lib.SCORE_AUTO, stream))
        st.host[2 * n:4 * n + 1].copy_(st.dev[2 * n:4 * n + 1], non_blocking=True)
        _lib.check(_lib.lib().n3_score_status(stream))  # synchronises the stream
        return int(h[4 * n]), h[2 * n:4 * n].astype(np.int64)
