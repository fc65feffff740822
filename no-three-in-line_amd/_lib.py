"""ctypes binding of libno3.so (the C ABI declared in include/no3.h).

torch is imported first on purpose: it loads the HIP runtime
(libamdhip64.so.7) that libno3.so then shares, so device pointers and streams
from torch are valid in the library.  There is no CPU fallback: every entry
point raises if the library or a gfx950 device is missing.
"""
from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import torch  # noqa: F401  (shares its HIP runtime with libno3.so)

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libno3.so"

N3_OK = 0
N3_EINVAL = -22
N3_ENODEV = -19
N3_ENOMEM = -12
N3_EHIP = -1000

SCORE_AUTO = 0
SCORE_DETERMINANT = 1
SCORE_SLOPE = 2

SENTINEL32 = 1 << 30
SENTINEL64 = 1 << 60


class Chains(ctypes.Structure):
    """n3_chains (include/no3.h)."""

    _fields_ = [
        ("n", ctypes.c_int),
        ("replicas", ctypes.c_int),
        ("columns", ctypes.c_void_p),
        ("badness", ctypes.c_void_p),
        ("scores", ctypes.c_void_p),
        ("best_columns", ctypes.c_void_p),
        ("best_scores", ctypes.c_void_p),
        ("accepted", ctypes.c_void_p),
        ("reference", ctypes.c_void_p),
        ("removed_points", ctypes.c_void_p),
        ("move_thresholds", ctypes.c_void_p),
        ("k_move", ctypes.c_int),
        ("exchange_thresholds", ctypes.c_void_p),
        ("exchange_sign", ctypes.c_void_p),
        ("k_exchange", ctypes.c_int),
        ("counters", ctypes.c_void_p),
        ("proposals", ctypes.c_void_p),
        ("block_waves", ctypes.c_int),
        ("log", ctypes.c_void_p),
        ("log_capacity", ctypes.c_int64),
        ("log_count", ctypes.c_void_p),
        ("centers", ctypes.c_void_p),
        ("strengths", ctypes.c_void_p),
        ("inverse_temperatures", ctypes.c_void_p),
    ]


class Group(ctypes.Structure):
    """n3_group (include/no3.h)."""

    _fields_ = [
        ("chains", Chains),
        ("ladder", ctypes.c_int),
        ("steps", ctypes.c_int),
        ("bad_bias", ctypes.c_int),
        ("seed", ctypes.c_uint64),
    ]


MAX_GROUPS = 8


_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64

# name -> (restype, argtypes); every symbol of include/no3.h
SIGNATURES = {
    "n3_version": (_i, []),
    "n3_last_error": (ctypes.c_char_p, []),
    "n3_struct_size": (_i64, [_i]),
    "n3_chains_offsets": (_i, [_vp, _i]),
    "n3_device_init": (_i, [_i]),
    "n3_dirtable": (_i, [_i, _vp, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "n3_generate_states": (_i, [_u64, _i, _i64, _i64, _vp, _vp, _vp]),
    "n3_score_population": (_i, [_vp, _vp, _i, _i64, _vp, _vp, _i, _vp]),
    "n3_score_status": (_i, [_vp]),
    "n3_two_switch_scores": (_i, [_vp, _vp, _vp, _i, _i, _i, _vp, _vp]),
    "n3_two_switch_scores_status": (_i, [_vp, _vp, _vp, _i, _i, _i, _vp, _vp, _vp]),
    "n3_evaluate_two_switches": (_i, [_vp, _vp, _vp, _i, _i, _i64, _vp, _vp]),
    "n3_evaluate_transpositions": (_i, [_vp, _vp, _vp, _i, _i, _i, _i64, _vp, _vp]),
    "n3_delta_table": (_i, [_vp, _vp, _vp, _vp, _i, _i64, _vp, _vp]),
    "n3_evaluate_three_cycles": (_i, [_vp, _vp, _vp, _i, _vp, _i64, _i64, _vp, _vp]),
    "n3_mutate_population": (_i, [_vp, _vp, _vp, _i, _i, _i, _i64, _i, _i, _i, _u64, _i, _vp, _vp, _vp]),
    "n3_enumerate_switches": (_i, [_vp, _vp, _i, _vp, _i64, _i, _vp, _vp, _vp, _vp]),
    "n3_elite_keys": (_i, [_vp, _i64, _i, _vp, _vp, _vp]),
    "n3_temper_advance": (_i, [ctypes.POINTER(Chains), _i, _i, _u64, _vp]),
    "n3_temper_exchange": (_i, [ctypes.POINTER(Chains), _i, _u64, _vp]),
    "n3_temper_run": (_i, [ctypes.POINTER(Chains), _i, _i, _i, _i, _u64, _vp]),
    "n3_temper_batched_run": (_i, [ctypes.POINTER(Group), _i, _i, _vp]),
    "n3_temper_batched_persistent": (_i, [ctypes.POINTER(Group), _i, _i, _vp]),
    "n3_temper_release": (None, []),
    "n3_metropolis_thresholds": (_i, [_vp, _i, _i, _vp]),
    "n3_mlp_params": (_i, []),
    "n3_mlp_features": (_i, [_i, ctypes.c_float, _i, _vp]),
    "n3_mlp_forward": (_i, [_vp, _vp, _i64, _vp, _vp]),
    "n3_mlp_grad": (_i, [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp]),
    "n3_mlp_grad_noalloc": (_i, [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp]),
    "n3_mlp_adam": (_i, [_vp, _vp, _vp, _vp, _vp, _i, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                         ctypes.c_float, _vp]),
    "n3_mlp_weight_table": (_i, [_vp, _vp, _i, _i, _vp, _vp]),
    "n3_metropolis_thresholds64": (_i, [_vp, _i, _i, _vp]),
    "n3_temper_run_learned": (_i, [ctypes.POINTER(Chains), _vp, _vp, _i, _vp, _i, _i, _i, _i, _u64, _vp]),
    "n3_learned_release": (None, []),
    "n3_point_exp_table": (_i, [_vp]),
    "n3_point_constants": (_i, [_vp, _vp]),
    "n3_point_weights": (_i, [_vp, _vp, _vp, _vp, _i, _i64, _vp, _vp, _vp, _vp, _vp]),
    "n3_temper_run_learned_point": (_i, [ctypes.POINTER(Chains), _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _u64,
                                         _vp]),
    "n3_exchange_thresholds": (_i, [_vp, _i, _i, _i, _vp, _vp]),
}

_lib = None
_lock = threading.Lock()
_initialised_devices: set[int] = set()


def lib() -> ctypes.CDLL:
    """The loaded library; raises RuntimeError if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
            L = ctypes.CDLL(str(LIB_PATH))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def check(status: int) -> None:
    """Maps an n3 status to the reference's exception types."""
    if status == N3_OK:
        return
    message = lib().n3_last_error().decode(errors="replace")
    if status == N3_EINVAL:
        raise ValueError(message)
    raise RuntimeError(f"no3 status {status}: {message}")


def device(index: int | None = None) -> torch.device:
    """The current gfx950 device; RuntimeError if there is none (gpu.py:50-51)."""
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device is available")
    idx = torch.cuda.current_device() if index is None else int(index)
    if idx not in _initialised_devices:
        with torch.cuda.device(idx):
            check(lib().n3_device_init(idx))
        _initialised_devices.add(idx)
    return torch.device("cuda", idx)


class Staging:
    """A pinned host buffer and a device buffer of int32 words, reused by the
    single-state calls of one device: one H2D of the packed inputs and one
    D2H of the outputs per call instead of a pageable copy per array (such
    calls are latency-bound).  Callers hold STAGING_LOCK while they use it."""

    def __init__(self, dev: torch.device, words: int):
        self.host = torch.empty(words, dtype=torch.int32, pin_memory=True)
        self.host_np = self.host.numpy()
        self.dev = torch.empty(words, dtype=torch.int32, device=dev)


_STAGING: dict = {}
STAGING_LOCK = threading.Lock()


def staging(dev: torch.device, words: int) -> Staging:
    """The device's staging buffers, grown to at least `words` words."""
    st = _STAGING.get(dev.index)
    if st is None or st.host.numel() < words:
        st = Staging(dev, max(words, 1024))
        _STAGING[dev.index] = st
    return st


def to_host(*tensors: torch.Tensor):
    """Copies device tensors into page-locked host buffers on the current
    stream, synchronises it, and returns NumPy views of the buffers (a D2H
    into pageable memory runs at a fraction of the link's rate)."""
    out = []
    for t in tensors:
        host = torch.empty(tuple(t.shape), dtype=t.dtype, pin_memory=True)
        host.copy_(t, non_blocking=True)
        out.append(host)
    if tensors:
        torch.cuda.current_stream(tensors[0].device).synchronize()
    return tuple(h.numpy() for h in out)


def stream_handle(dev: torch.device) -> int:
    return torch.cuda.current_stream(dev).cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()
