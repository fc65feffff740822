"""Host local-search consumers (search.py, breakout.py) and the line
geometry they weight, run with the oracle's evaluators (no GPU)."""
from itertools import combinations
from math import comb
from random import Random

import numpy as np


def _oracle_hooks(oracle):
    return dict(move_evaluator=oracle.evaluate_two_switches, scorer=oracle.score_and_badness)


def test_line_key_normalization(pkg):
    """tests/test_geometry.py:11-14 and the normalisation rules."""
    key = pkg.lines.line_key
    expected = key((1, 2), (4, 8))
    assert key((4, 8), (1, 2)) == expected and key((7, 14), (1, 2)) == expected
    assert key((0, 0), (0, 5)) == key((0, 3), (0, 1)) == (1, 0, 0)  # row 0: A = 1
    assert key((2, 3), (5, 3))[0] == 0 and key((2, 3), (5, 3))[1] > 0  # column 3: A = 0 < B
    a, b, c = key((1, 1), (3, 4))
    assert a * 1 + b * 1 + c == 0 and a * 3 + b * 4 + c == 0


def test_violating_lines_against_brute_force(pkg, oracle):
    rng = np.random.default_rng(6)
    for n in (4, 7, 11):
        for _ in range(4):
            p, q = rng.permutation(n), rng.permutation(n)
            q = np.where(p == q, np.roll(q, 1), q)
            points = sorted(set([(r, int(p[r])) for r in range(n)] + [(r, int(q[r])) for r in range(n)]))
            lines = pkg.lines.violating_lines(points)
            brute = {}
            for x, y in combinations(points, 2):
                members = frozenset(z for z in points if (y[0] - x[0]) * (z[1] - x[1]) == (y[1] - x[1]) * (z[0] - x[0]))
                if len(members) >= 3:
                    brute[members] = comb(len(members), 3)
            assert sorted(map(frozenset, (l.points for l in lines)), key=sorted) == sorted(brute, key=sorted)
            assert sum(l.triples for l in lines) == sum(brute.values())
            assert [(-l.triples, l.key) for l in lines] == sorted((-l.triples, l.key) for l in lines)


def test_weighted_scores_equal_recounts(pkg, oracle):
    """weighted_candidate_scores of every 2-switch of a slot equals the
    weighted score of the switched state recounted from scratch."""
    br = pkg.breakout
    rng = Random(4)
    for n in (8, 13):
        state = pkg.random_state(n, rng)
        p, q = np.asarray(state.P, np.int32), np.asarray(state.Q, np.int32)
        points = state.points()
        lines = pkg.lines.violating_lines(points)
        keys = [l.key for l in lines] + [pkg.lines.line_key(points[0], points[1])]
        weights = {k: 1 + rng.randrange(5) for k in keys}
        score, bad = oracle.score_and_badness(p, q)
        for u in range(2 * n):
            raw = oracle.evaluate_two_switches(p, q, u, score, bad)
            weighted = br.weighted_candidate_scores(p, q, u, raw, weights)
            for v in range(2 * n):
                if raw[v] >= 1 << 29:
                    assert weighted[v] >= 1 << 29
                    continue
                moved = pkg.search.switch_points(p, q, u, v)
                mp, mq = np.asarray(moved.P, np.int32), np.asarray(moved.Q, np.int32)
                assert weighted[v] == br.weighted_state_score(mp, mq, oracle.score_and_badness(mp, mq)[0], weights)


def test_local_search_with_oracle_evaluators(pkg, oracle):
    snapshots = []
    result = pkg.local_search(12, seed=3, iterations=600, plateau=40, checkpoint=snapshots.append,
                              cycle_evaluator=oracle.evaluate_three_cycles, **_oracle_hooks(oracle))
    result.state.validate()
    assert result.score == oracle.score_and_badness(result.state.P, result.state.Q)[0]
    assert result.iterations <= 600 and result.moves_evaluated > 0
    assert [s.best_score for s in snapshots] == sorted((s.best_score for s in snapshots), reverse=True)
    assert snapshots and snapshots[-1].best_score == result.score


def test_breakout_with_oracle_evaluators(pkg, oracle):
    state = pkg.random_state(12, Random(9))
    result = pkg.breakout_search(state, seed=2, iterations=400, plateau=15, weight_increment=2,
                                 **_oracle_hooks(oracle))
    result.state.validate()
    assert result.score == oracle.score_and_badness(result.state.P, result.state.Q)[0]
    assert result.breakouts > 0 and result.weighted_lines <= 128


def test_conflict_metrics(pkg, oracle):
    """conflicts.analyze_points against the oracle's score and badness, a
    hand-made configuration and the determinant recount."""
    c = pkg.conflicts
    rng = Random(11)
    for n in (6, 12, 30):
        state = pkg.random_state(n, rng)
        points = [(r, state.P[r]) for r in range(n)] + [(r, state.Q[r]) for r in range(n)]
        metrics = c.analyze_points(points)
        score, bad = oracle.score_and_badness(np.array(state.P), np.array(state.Q))
        assert metrics.triples == score == c.full_triple_score(points)
        # badness is reported in the input order: P points then Q points
        slot_order = [2 * r for r in range(n)] + [2 * r + 1 for r in range(n)]
        assert list(metrics.badness) == [int(bad[s]) for s in slot_order]
        assert sum(metrics.slope_histogram.values()) == score
        assert metrics.lexicographic[0] == score
    # four points on the diagonal and one off it: one line, C(4,3) triples
    pts = [(0, 0), (1, 1), (2, 2), (3, 3), (0, 3)]
    m = c.analyze_points(pts, {pkg.lines.line_key((0, 0), (1, 1)): 5})
    assert (m.triples, m.violating_lines, m.max_line_occupancy, m.max_point_badness) == (4, 1, 4, 3)
    assert m.weighted_score == 20 and m.badness == (3, 3, 3, 3, 0) and m.slope_histogram == {(1, 1): 4}
    assert c.full_triple_score([(0, 0), (0, 0), (1, 1)]) == 1  # duplicates: zero determinant
    state = pkg.random_state(10, Random(3))
    for layer in (0, 1):
        delta = c.reference_swap_delta(state, layer, 2, 7)
        moved = state.swapped(layer, 2, 7)
        after = oracle.score_and_badness(np.array(moved.P), np.array(moved.Q))[0]
        before = oracle.score_and_badness(np.array(state.P), np.array(state.Q))[0]
        assert delta == after - before
