"""Pins the CPU oracle: published known answers, the reference's fixture and
test states, and independent formulations on small inputs."""
import json
from collections import Counter
from itertools import combinations
from math import comb, gcd
from pathlib import Path
from random import Random

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


# -- independent formulations (brute force on tiny inputs) --------------------

def canonical_line(a, b):
    """Primitive (A, B, C) of the line through a and b with A > 0 or A == 0 < B
    (the form of geometry.canonical_line, geometry.py:14-28)."""
    A = b[1] - a[1]
    B = a[0] - b[0]
    C = -(A * a[0] + B * a[1])
    g = gcd(gcd(abs(A), abs(B)), abs(C))
    A, B, C = A // g, B // g, C // g
    if A < 0 or (A == 0 and B < 0):
        A, B, C = -A, -B, -C
    return A, B, C


def line_count(points):
    """E = sum over lines of C(k, 3) and badness = sum of C(k-1, 2) over the
    lines through each point (conflicts.analyze_points, conflicts.py:36-72)."""
    members = {}
    for i, j in combinations(range(len(points)), 2):
        key = canonical_line(points[i], points[j])
        members.setdefault(key, set()).update((i, j))
    badness = [0] * len(points)
    triples = 0
    for idx in members.values():
        k = len(idx)
        if k >= 3:
            triples += comb(k, 3)
            for i in idx:
                badness[i] += comb(k - 1, 2)
    return triples, badness


def slots(p, q):
    pts = []
    for r in range(len(p)):
        pts.append((r, int(p[r])))
        pts.append((r, int(q[r])))
    return pts


# -- Philox ---------------------------------------------------------------------

def test_philox_known_answers(oracle):
    """Random123 kat_vectors for philox4x32_10."""
    kats = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, expected in kats:
        assert list(oracle.philox4x32_10(ctr, key)) == expected


# -- scoring --------------------------------------------------------------------

def test_determinant_known_answers(oracle):
    """tests/test_geometry.py:5-8 of the reference, as 3-point states' rows:
    the oracle's collinearity test on the same triples."""
    def collinear3(a, b, c):
        # one P/Q state cannot hold arbitrary triples; use the line counter on
        # the triple, and the oracle on a state containing it where possible
        return line_count([a, b, c])[0] == 1
    assert collinear3((0, 0), (3, 2), (6, 4))
    assert not collinear3((0, 0), (3, 2), (6, 5))
    assert collinear3((2, 0), (2, 4), (2, 9))
    assert collinear3((0, 7), (4, 7), (9, 7))
    assert collinear3((0, 0), (2, 2), (5, 5))  # opposite rays: test_verifier_agreement.py:18-21


def test_near_solution_fixture(oracle):
    """results/best-near-solution-E10.json: E=10 (README.md:7), and the slope
    form, the determinant form and the line form agree."""
    data = json.loads((GOLDEN / "best-near-solution-E10.json").read_text())
    n = data["n"]
    adj = [[] for _ in range(n)]
    for r, c in data["points"]:
        adj[r].append(c)
    p = np.array([min(a) for a in adj], np.int32)
    q = np.array([max(a) for a in adj], np.int32)
    e, bad = oracle.score_and_badness(p, q)
    e2, bad2 = oracle.score_slope(p, q)
    e3, bad3 = line_count(slots(p, q))
    assert e == e2 == e3 == 10
    assert list(bad) == list(bad2) == bad3
    assert bad.sum() == 30 and bad.max() == 3 and (bad > 0).sum() == 24


def test_scores_match_line_form_small(oracle):
    rng = np.random.default_rng(1)
    for n in range(2, 13):
        for _ in range(6):
            p = rng.permutation(n).astype(np.int32)
            q = rng.permutation(n).astype(np.int32)
            e, bad = oracle.score_and_badness(p, q)
            t, b = line_count(slots(p, q)) if not np.any(p == q) else (None, None)
            if t is not None:
                assert e == t and list(bad) == b
                e2, bad2 = oracle.score_slope(p, q)
                assert e2 == e and np.array_equal(bad2, bad)


def test_coincident_points_counted_like_determinant(oracle):
    """With P[r] == Q[r] every triple holding both copies is collinear; the
    slope form has no direction for the pair and refuses."""
    p = np.array([0, 1, 2, 3], np.int32)
    q = np.array([0, 2, 3, 1], np.int32)
    e, bad = oracle.score_and_badness(p, q)
    pts = slots(p, q)
    brute = sum(1 for a, b, c in combinations(pts, 3)
                if (b[0] - a[0]) * (c[1] - a[1]) == (b[1] - a[1]) * (c[0] - a[0]))
    assert e == brute
    with pytest.raises(ValueError):
        oracle.score_slope(p, q)


def test_population_scores_random_n75(oracle):
    p, q = oracle.generate_states(0, 75, 24)
    s, b = oracle.score_population(p, q)
    for i in range(4):
        e2, bad2 = oracle.score_slope(p[i], q[i])
        assert s[i] == e2 and np.array_equal(b[i], bad2)
    assert 350 < s.mean() < 550  # E of random n=75 states is about 445


# -- generator ------------------------------------------------------------------

def test_generator_invariants(oracle):
    for n in (2, 3, 32, 75, 100):
        p, q = oracle.generate_states(0, n, 200)
        assert np.all(np.sort(p, axis=1) == np.arange(n))
        assert np.all(np.sort(q, axis=1) == np.arange(n))
        assert np.all(p != q)
    a, _ = oracle.generate_states(0, 32, 10, first_index=5)
    b, _ = oracle.generate_states(0, 32, 15)
    assert np.array_equal(a, b[5:])


# -- deltas: the reference's own test states -----------------------------------

def _random_state(n, rng):
    """state.random_state (state.py:88-95) on stdlib Random."""
    p = list(range(n))
    q = list(range(n))
    rng.shuffle(p)
    while True:
        rng.shuffle(q)
        if all(p[r] != q[r] for r in range(n)):
            return p, q


def test_transpositions_match_full_recompute(oracle):
    """tests/test_delta.py:10-30 of the reference (Random(9817))."""
    rng = Random(9817)
    comparisons = 0
    for n in range(4, 13):
        for _ in range(8):
            P, Q = _random_state(n, rng)
            p = np.asarray(P, np.int32)
            q = np.asarray(Q, np.int32)
            score, badness = oracle.score_and_badness(p, q)
            assert score == line_count(slots(p, q))[0]
            for _ in range(4):
                layer = rng.randrange(2)
                row = rng.randrange(n)
                cand = oracle.evaluate_transpositions(p, q, layer, row, score, badness)
                for partner in range(n):
                    if cand[partner] >= (1 << 59):
                        continue
                    moved = [list(P), list(Q)]
                    moved[layer][row], moved[layer][partner] = moved[layer][partner], moved[layer][row]
                    assert moved[0][row] != moved[1][row] and moved[0][partner] != moved[1][partner]
                    assert cand[partner] == line_count(slots(moved[0], moved[1]))[0]
                    comparisons += 1
    assert comparisons >= 1000


def test_two_switches_match_full_recompute(oracle):
    """tests/test_delta.py:33-60 of the reference (Random(443))."""
    rng = Random(443)
    comparisons = 0
    for n in range(4, 11):
        for _ in range(5):
            P, Q = _random_state(n, rng)
            p = np.asarray(P, np.int32)
            q = np.asarray(Q, np.int32)
            score, badness = oracle.score_and_badness(p, q)
            selected = rng.randrange(2 * n)
            cand = oracle.evaluate_two_switches(p, q, selected, score, badness)
            pts = slots(p, q)
            for partner in range(2 * n):
                if cand[partner] >= (1 << 59):
                    continue
                moved = list(pts)
                (ra, ca), (rb, cb) = pts[selected], pts[partner]
                moved[selected] = (ra, cb)
                moved[partner] = (rb, ca)
                assert cand[partner] == line_count(moved)[0]
                comparisons += 1
    assert comparisons >= 200


def test_delta_table_rows(oracle):
    p, q = oracle.generate_states(0, 8, 3)
    table = oracle.delta_table(p, q)
    for s in range(3):
        score, bad = oracle.score_and_badness(p[s], q[s])
        for u in range(16):
            ref = oracle.evaluate_two_switches(p[s], q[s], u, score, bad)
            exp = np.where(ref >= (1 << 59), 1 << 30, ref).astype(np.int32)
            assert np.array_equal(table[s, u], exp)


# -- tempering replay -----------------------------------------------------------

def test_replay_invariants(oracle):
    """The walk's incremental score and badness equal a full recount after
    every sweep, and exchanges permute states between slots."""
    n, replicas, ladder = 9, 8, 4
    p, q = oracle.generate_states(3, n, replicas)
    cols = np.stack([np.column_stack((p[i], q[i])).reshape(-1) for i in range(replicas)]).astype(np.int32)
    temps = np.tile(np.geomspace(0.2, 3.0, ladder), 2).astype(np.float32)
    r = oracle.TemperingReplay(cols, temps, 17, k_move=80, k_exchange=400)
    for _ in range(30):
        r.advance(18)
        before = sorted(map(tuple, r.columns))
        r.exchange(ladder)
        assert sorted(map(tuple, r.columns)) == before
        for i in range(replicas):
            e, bad = oracle.score_and_badness(r.columns[i, 0::2], r.columns[i, 1::2])
            assert e == r.scores[i] and np.array_equal(bad, r.badness[i])
            c = Counter(r.columns[i])
            assert all(v == 2 for v in c.values()) and len(c) == n
    assert r.accepted.sum() > 0


def test_exp_word_law(oracle):
    """The umbrella acceptance word is floor(2^32 exp(-x)) to within one unit
    (math.exp is not correctly rounded either), never grows with x, and is 0
    from x = 23 on."""
    import math

    rng = np.random.default_rng(4)
    xs = np.concatenate([np.linspace(1e-12, 22.99, 4001), rng.uniform(0, 23, 4000), [1e-300, 0.5, math.log(2)]])
    for x in xs:
        want = math.floor(math.ldexp(math.exp(-float(x)), 32))
        got = oracle.exp_word(float(x))
        assert abs(got - min(want, 2**32 - 1)) <= 1, x
    words = [oracle.exp_word(float(x)) for x in np.sort(xs)]
    assert all(a >= b for a, b in zip(words, words[1:]))
    assert oracle.exp_word(23.0) == 0 and oracle.exp_word(40.0) == 0 and oracle.exp_word(1e-300) == 2**32 - 1


def _umbrella_replay(oracle, centers_from, strength, sweeps=6):
    n, replicas = 12, 16
    p, q = oracle.generate_states(8, n, replicas)
    cols = np.stack([np.column_stack((p[i], q[i])).reshape(-1) for i in range(replicas)]).astype(np.int32)
    # the reference state is slot 0's; other slots start at some distance
    temps = np.full(replicas, 0.5, np.float32)
    r = oracle.TemperingReplay(cols, temps, 21, k_move=80, k_exchange=400)
    r.strengths[:] = strength
    r.centers[:] = centers_from(r.removed_points)
    start = r.removed_points.copy()
    for _ in range(sweeps):
        r.advance(2 * n)
        for i in range(replicas):
            e, bad = oracle.score_and_badness(r.columns[i, 0::2], r.columns[i, 1::2])
            assert e == r.scores[i] and np.array_equal(bad, r.badness[i])
    return start, r


def test_umbrella_energy_terms(oracle):
    """tempering.py:93-94 on the replay.  A stiff umbrella centred on each
    slot's own distance rejects every move that changes it (strength dd^2
    dominates), while moves that keep it still happen; centred far above, the
    -2 (center - removed) dd term accepts every move that raises it."""
    start, r = _umbrella_replay(oracle, lambda removed: removed.astype(np.float32), 1e6)
    assert np.array_equal(r.removed_points, start)
    assert r.accepted.sum() > 0
    start, r = _umbrella_replay(oracle, lambda removed: removed.astype(np.float32) + 1000.0, 1.0)
    assert np.all(r.removed_points >= start) and r.removed_points.sum() > start.sum()
    with pytest.raises(ValueError):
        r.exchange(4)


def test_metropolis_thresholds_law(oracle):
    thr = oracle.metropolis_thresholds(np.array([0.5, 2.0], np.float32), 64)
    for t, row in zip((0.5, 2.0), thr):
        for d in (1, 2, 5, 20):
            assert abs(row[d] / 2**32 - np.exp(-d / t)) < 2**-31
        assert row[-1] == 0


def test_exchange_thresholds_law(oracle):
    temps = np.array([0.1, 0.2, 0.4, 0.4], np.float32)
    thr, sign = oracle.exchange_thresholds(temps, 4, 32)
    assert list(sign) == [1, 1, 0, 0]
    c = 1 / np.float64(np.float32(0.1)) - 1 / np.float64(np.float32(0.2))
    assert abs(thr[0, 3] / 2**32 - np.exp(-c * 3)) < 2**-31


def _three_cycle_candidates(p, q, u):
    """search.py:38-65 as a plain loop (test helper)."""
    n = len(p)
    cols = [0] * (2 * n)
    for r in range(n):
        cols[2 * r], cols[2 * r + 1] = int(p[r]), int(q[r])
    ru = u // 2
    out = []
    for v in range(2 * n - 1):
        if v // 2 == ru:
            continue
        for w in range(v + 1, 2 * n):
            if w // 2 == ru or w // 2 == v // 2:
                continue
            if len({cols[u], cols[v], cols[w]}) != 3:
                continue
            ou, ov, ow = cols[u ^ 1], cols[v ^ 1], cols[w ^ 1]
            if cols[v] != ou and cols[w] != ov and cols[u] != ow:
                out.append((u, v, w, 0))
            if cols[w] != ou and cols[v] != ow and cols[u] != ov:
                out.append((u, v, w, 1))
    return np.asarray(out, np.int32).reshape(-1, 4)


def test_three_cycles_match_full_recompute(oracle):
    """tests/test_delta.py:63-83 of the reference (Random(991)), with every
    candidate recounted by the line formulation."""
    rng = Random(991)
    comparisons = 0
    for n in range(4, 10):
        P, Q = _random_state(n, rng)
        p = np.asarray(P, np.int32)
        q = np.asarray(Q, np.int32)
        score, badness = oracle.score_and_badness(p, q)
        selected = rng.randrange(2 * n)
        cand = _three_cycle_candidates(p, q, selected)
        if len(cand) == 0:
            continue
        sample = cand[: min(30, len(cand))]
        scores = oracle.evaluate_three_cycles(p, q, sample, score, badness)
        pts = slots(p, q)
        for (u, v, w, o), s in zip(sample, scores):
            moved = list(pts)
            cu, cv, cw = pts[u][1], pts[v][1], pts[w][1]
            nu, nv, nw = (cv, cw, cu) if o == 0 else (cw, cu, cv)
            moved[u], moved[v], moved[w] = (pts[u][0], nu), (pts[v][0], nv), (pts[w][0], nw)
            assert s == line_count(moved)[0]
            comparisons += 1
    assert comparisons >= 50


def _interleave(p, q):
    out = np.empty((p.shape[0], 2 * p.shape[1]), dtype=np.int32)
    out[:, 0::2] = p
    out[:, 1::2] = q
    return out


def test_mutation_keeps_the_two_factor_and_the_elites(oracle, pkg):
    """Invariants of mutate_population (gpu_device_search.py:27-113): every
    child is a row/column 2-factor, children below elite_count are exact
    copies, every other child is within `strength` legal 2-switches of some
    parent (at most 2 slots per switch), and a strength-1 child differs from
    its nearest parent by exactly one switch whenever it differs at all."""
    n, elites = 9, 6
    p, q = oracle.generate_states(3, n, elites)
    bad = oracle.score_population(p, q)[1].astype(np.int32)
    for strength, bias in ((1, 0), (3, 85), (5, 100)):
        cp, cq = oracle.mutate_population(p, q, bad, quality_count=2, population=300, active_strength=strength,
                                          max_strength=strength, bad_bias=bias, seed=11, generation=4)
        kids, parents = _interleave(cp, cq), _interleave(p, q)
        pkg.state.check_columns(kids)
        assert np.array_equal(kids[:elites], parents)
        dist = (kids[elites:, None, :] != parents[None, :, :]).sum(axis=2).min(axis=1)
        assert dist.max() <= 2 * strength
        if strength == 1:
            assert set(np.unique(dist)) <= {0, 2} and (dist == 2).mean() > 0.9
        again = oracle.mutate_population(p, q, bad, quality_count=2, population=300, active_strength=strength,
                                         max_strength=strength, bad_bias=bias, seed=11, generation=4)
        assert np.array_equal(again[0], cp) and np.array_equal(again[1], cq)


def test_mutation_bias_moves_bad_points(oracle):
    """With bad_bias = 100 the first slot of every switch is drawn in
    proportion to the parent's badness: a strength-1 child of a parent whose
    only bad slots are known changes one of them (or its partner slot v)."""
    n = 12
    p, q = oracle.generate_states(5, n, 1)
    bad = np.zeros((1, 2 * n), dtype=np.int32)
    bad[0, [3, 10]] = 1
    cp, cq = oracle.mutate_population(p, q, bad, quality_count=1, population=400, active_strength=1,
                                      max_strength=1, bad_bias=100, seed=2, generation=1)
    kids, parent = _interleave(cp, cq)[1:], _interleave(p, q)[0]
    changed = kids != parent[None, :]
    moved = changed.any(axis=1)
    assert moved.mean() > 0.9
    assert np.all(changed[moved][:, [3, 10]].any(axis=1))


def test_enumeration_matches_the_switch_rule(oracle, pkg):
    """enumerate_switches (gpu_exhaustive_beam.py:17-57): a legal descriptor
    changes exactly slots u and v by swapping their columns, and the legal
    set is the one whose 2-switch score the delta table defines (entries
    below the sentinel) minus the same-column switches, which
    accelerated.py:146-150 score as no-ops and enumerate_switches rejects."""
    n = 8
    p, q = oracle.generate_states(9, n, 3)
    desc = np.array([(s, u, v) for s in range(3) for u in range(2 * n) for v in range(2 * n)], dtype=np.int32)
    cp, cq, valid = oracle.enumerate_switches(p, q, desc)
    kids, parents = _interleave(cp, cq), _interleave(p, q)
    table = oracle.delta_table(p, q)
    for i, (s, u, v) in enumerate(desc):
        assert valid[i] == (table[s, u, v] != 1 << 30 and parents[s, u] != parents[s, v])
        if valid[i]:
            expect = parents[s].copy()
            expect[u], expect[v] = parents[s, v], parents[s, u]
            assert np.array_equal(kids[i], expect)
            assert oracle.score_population(cp[i:i + 1], cq[i:i + 1])[0][0] == table[s, u, v]
        else:
            assert np.array_equal(kids[i], parents[s])
