"""2-switch and transposition scores on the GPU against the oracle."""
from random import Random

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_delta_table_config1(pkg, oracle, dev):
    """n=32, P=256 seeded states: the full 64x64 2-switch table of every
    state, bit-exact against the oracle (accelerated.evaluate_two_switches
    for every anchor)."""
    n, count = 32, 256
    hp, hq = oracle.generate_states(0, n, count)
    table = pkg.delta_table(hp, hq)
    ref = oracle.delta_table(hp, hq)
    assert table.shape == (count, 2 * n, 2 * n)
    assert np.array_equal(table, ref)
    # every legal entry equals a full recount of the moved state
    scores, _ = oracle.score_population(hp, hq)
    rng = np.random.default_rng(5)
    checked = 0
    for s in rng.choice(count, 16, replace=False):
        for u, v in zip(rng.integers(0, 2 * n, 40), rng.integers(0, 2 * n, 40)):
            if table[s, u, v] >= (1 << 29):
                continue
            cols = np.column_stack((hp[s], hq[s])).reshape(-1).copy()
            cols[u], cols[v] = cols[v], cols[u]
            assert table[s, u, v] == oracle.score_and_badness(cols[0::2], cols[1::2])[0]
            checked += 1
    assert checked > 100


@pytest.mark.parametrize("n", [4, 7, 12, 31, 32, 33, 50, 75, 100])
def test_single_anchor_entry_points(pkg, oracle, dev, n):
    hp, hq = oracle.generate_states(77 + n, n, 6)
    for s in range(len(hp)):
        p, q = hp[s], hq[s]
        score, bad = oracle.score_and_badness(p, q)
        for u in (0, 1, 2 * n - 1, (7 * s + 3) % (2 * n)):
            ref = oracle.evaluate_two_switches(p, q, u, score, bad)
            assert np.array_equal(pkg.evaluate_two_switches(p, q, u, score, bad), ref)
            got32 = pkg.score_two_switches_gpu(p, q, u, score, bad.astype(np.int32))
            sent = ref >= (1 << 59)
            assert np.array_equal(got32 >= (1 << 29), sent)
            assert np.array_equal(got32[~sent], ref[~sent].astype(np.int32))
        for layer in (0, 1):
            row = (5 * s + layer) % n
            ref = oracle.evaluate_transpositions(p, q, layer, row, score, bad)
            assert np.array_equal(pkg.evaluate_transpositions(p, q, layer, row, score, bad), ref)


def test_reference_delta_test_states(pkg, oracle, dev):
    """The states of the reference's tests/test_delta.py (Random(9817) and
    Random(443) through state.random_state) evaluated on the GPU."""
    rng = Random(9817)
    comparisons = 0
    for n in range(4, 13):
        for _ in range(8):
            state = pkg.random_state(n, rng)
            p = np.asarray(state.P, dtype=np.int32)
            q = np.asarray(state.Q, dtype=np.int32)
            score, bad = oracle.score_and_badness(p, q)
            for _ in range(4):
                layer = rng.randrange(2)
                row = rng.randrange(n)
                got = pkg.evaluate_transpositions(p, q, layer, row, score, bad)
                assert np.array_equal(got, oracle.evaluate_transpositions(p, q, layer, row, score, bad))
                comparisons += int((got < (1 << 59)).sum())
    assert comparisons >= 1000
    rng = Random(443)
    for n in range(4, 11):
        for _ in range(5):
            state = pkg.random_state(n, rng)
            p = np.asarray(state.P, dtype=np.int32)
            q = np.asarray(state.Q, dtype=np.int32)
            score, bad = oracle.score_and_badness(p, q)
            selected = rng.randrange(2 * n)
            got = pkg.evaluate_two_switches(p, q, selected, score, bad)
            assert np.array_equal(got, oracle.evaluate_two_switches(p, q, selected, score, bad))


def test_inconsistent_badness_follows_formula(pkg, oracle, dev):
    """The incremental formula uses the caller's score and badness as given
    (accelerated.py:159,194): arbitrary inputs give the same arbitrary output."""
    hp, hq = oracle.generate_states(4, 20, 2)
    rng = np.random.default_rng(0)
    bad = rng.integers(0, 50, size=40).astype(np.int64)
    for u in range(40):
        ref = oracle.evaluate_two_switches(hp[0], hq[0], u, 1234, bad)
        assert np.array_equal(pkg.evaluate_two_switches(hp[0], hq[0], u, 1234, bad), ref)


def test_coincident_points_delta(pkg, oracle, dev):
    """States with P[r] == Q[r] (outside the model) still follow the reference."""
    rng = np.random.default_rng(3)
    n = 15
    p = rng.permutation(n).astype(np.int32)
    q = rng.permutation(n).astype(np.int32)
    q[3] = p[3]
    q[8] = p[8]
    score, bad = oracle.score_and_badness(p, q)
    for u in range(2 * n):
        assert np.array_equal(pkg.evaluate_two_switches(p, q, u, score, bad),
                              oracle.evaluate_two_switches(p, q, u, score, bad))


def test_delta_shape_errors(pkg, dev):
    with pytest.raises(ValueError):
        pkg.score_two_switches_gpu(np.arange(5), np.arange(5), 0, 0, np.zeros(9))
    # a column outside [0, n) is reported through the staged status word, and
    # the next call on the same staging buffer starts from a clear status
    p, q = np.arange(5), np.roll(np.arange(5), 1)
    with pytest.raises(ValueError):
        pkg.score_two_switches_gpu(p, np.where(q == 4, 7, q), 0, 0, np.zeros(10))
    assert pkg.score_two_switches_gpu(p, q, 0, 0, np.zeros(10)).shape == (10,)
    with pytest.raises(ValueError):
        pkg.evaluate_two_switches(np.arange(5), np.roll(np.arange(5), 1), 10, 0, np.zeros(10))


@pytest.mark.parametrize("n", [4, 9, 20, 32, 33, 75, 100])
def test_three_cycles(pkg, oracle, dev, n):
    hp, hq = oracle.generate_states(500 + n, n, 3)
    rng = np.random.default_rng(n)
    for s in range(3):
        p, q = hp[s], hq[s]
        score, bad = oracle.score_and_badness(p, q)
        u = int(rng.integers(0, 2 * n))
        cand = pkg.legal_three_cycle_candidates(p, q, u)
        if len(cand) > 300:
            cand = cand[rng.choice(len(cand), 300, replace=False)]
        assert np.array_equal(pkg.evaluate_three_cycles(p, q, cand, score, bad),
                              oracle.evaluate_three_cycles(p, q, cand, score, bad))


def test_three_cycles_reference_states(pkg, oracle, dev):
    """The states of the reference's test_three_cycle_scores_match_full_recompute."""
    from random import Random

    rng = Random(991)
    for n in range(4, 10):
        state = pkg.random_state(n, rng)
        p = np.asarray(state.P, dtype=np.int32)
        q = np.asarray(state.Q, dtype=np.int32)
        score, bad = oracle.score_and_badness(p, q)
        selected = rng.randrange(2 * n)
        cand = pkg.legal_three_cycle_candidates(p, q, selected)
        if len(cand) == 0:
            continue
        sample = cand[: min(30, len(cand))]
        assert np.array_equal(pkg.evaluate_three_cycles(p, q, sample, score, bad),
                              oracle.evaluate_three_cycles(p, q, sample, score, bad))
