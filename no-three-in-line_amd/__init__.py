"""gfx950 (MI355X) engine for the no-three-in-line hot path.

Exact collinear-triple scoring of P/Q permutation populations, 2-switch move
scores and replica-exchange tempering, computed by hand-written HIP kernels
in libno3.so (C ABI: include/no3.h).  The module names mirror the reference
package's (gpu, gpu_delta, tempering, state, io), so callers can switch by
changing the import.
"""
from . import _lib
from ._lib import SCORE_AUTO, SCORE_DETERMINANT, SCORE_SLOPE, SENTINEL32, SENTINEL64
from .gpu import (
    check_status,
    generate_states,
    score_and_badness,
    score_population_badness,
    score_population_device,
    score_population_gpu,
)
from .gpu_delta import (
    delta_table,
    delta_table_device,
    evaluate_three_cycles,
    evaluate_transpositions,
    evaluate_two_switches,
    score_two_switches_gpu,
)
from .moves import apply_three_cycle, legal_three_cycle_candidates
from .state import PairPermutations, points_to_permutations, random_state, random_two_switches
from .tempering import TemperingChains
from .batched import BatchGroup, run_batched, seeded_chains
from . import autotune, breakout, conflicts, gpu_device_search, gpu_exhaustive_beam, io, learned, lines, search, verifiers
from .breakout import BreakoutResult, breakout_search
from .search import SearchResult, SearchSnapshot, local_search
from .gpu_device_search import (GpuSearchResult, gpu_device_beam_search, mutate_population_device, rank_elites,
                                rank_elites_device)
from .gpu_exhaustive_beam import candidate_descriptors, gpu_exhaustive_beam_search

__all__ = [
    "SCORE_AUTO", "SCORE_DETERMINANT", "SCORE_SLOPE", "SENTINEL32", "SENTINEL64",
    "check_status", "generate_states", "score_and_badness", "score_population_badness", "score_population_device", "score_population_gpu",
    "delta_table", "delta_table_device", "evaluate_three_cycles", "evaluate_transpositions",
    "apply_three_cycle", "legal_three_cycle_candidates", "evaluate_two_switches", "score_two_switches_gpu",
    "PairPermutations", "points_to_permutations", "random_state", "random_two_switches", "TemperingChains", "_lib",
    "BatchGroup", "run_batched", "seeded_chains", "autotune", "io", "learned",
    "gpu_device_search", "gpu_exhaustive_beam", "GpuSearchResult", "gpu_device_beam_search", "mutate_population_device",
    "rank_elites", "rank_elites_device", "candidate_descriptors", "gpu_exhaustive_beam_search",
    "breakout", "conflicts", "lines", "search", "verifiers", "BreakoutResult", "breakout_search", "SearchResult", "SearchSnapshot", "local_search",
]
