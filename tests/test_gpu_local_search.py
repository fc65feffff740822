"""The local-search consumers on the GPU evaluators: the same trajectory as
with the oracle's evaluators (every evaluator returns exact integers, and
the host loop's random calls are the same)."""
from random import Random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run_local(pkg, oracle, n, seed, gpu):
    snaps = []
    hooks = {} if gpu else dict(move_evaluator=oracle.evaluate_two_switches, scorer=oracle.score_and_badness,
                                cycle_evaluator=oracle.evaluate_three_cycles)
    result = pkg.local_search(n, seed=seed, iterations=500, plateau=30, checkpoint=snaps.append, **hooks)
    return result, [(s.iteration, s.score, s.best_score, s.layer, s.row_a, s.row_b, s.partner_layer, s.state)
                    for s in snaps]


@pytest.mark.parametrize("n,seed", [(10, 1), (20, 2), (33, 5)])
def test_local_search_trajectory_matches_oracle(pkg, oracle, dev, n, seed):
    got, got_snaps = _run_local(pkg, oracle, n, seed, gpu=True)
    ref, ref_snaps = _run_local(pkg, oracle, n, seed, gpu=False)
    assert got_snaps == ref_snaps
    assert (got.state, got.score, got.iterations, got.moves_evaluated) == \
        (ref.state, ref.score, ref.iterations, ref.moves_evaluated)


@pytest.mark.parametrize("n,seed", [(12, 3), (24, 4)])
def test_breakout_trajectory_matches_oracle(pkg, oracle, dev, n, seed):
    state = pkg.random_state(n, Random(seed))
    kwargs = dict(seed=seed, iterations=400, plateau=12, weight_increment=2)
    got = pkg.breakout_search(state, **kwargs)
    ref = pkg.breakout_search(state, move_evaluator=oracle.evaluate_two_switches, scorer=oracle.score_and_badness,
                              **kwargs)
    assert (got.state, got.score, got.iterations, got.moves_evaluated, got.breakouts, got.weighted_lines) == \
        (ref.state, ref.score, ref.iterations, ref.moves_evaluated, ref.breakouts, ref.weighted_lines)


def test_local_search_n75_improves(pkg, oracle, dev):
    state = pkg.random_state(75, Random(1))
    start = oracle.score_and_badness(state.P, state.Q)[0]
    result = pkg.local_search(75, seed=1, iterations=300, initial=state)
    assert result.score == oracle.score_and_badness(result.state.P, result.state.Q)[0] < start
    # the reference's cuda-delta backend (int32 table, sentinel 1<<30) as the hook
    alt = pkg.local_search(75, seed=1, iterations=300, initial=state, move_evaluator=pkg.score_two_switches_gpu)
    assert alt.state == result.state and alt.score == result.score
