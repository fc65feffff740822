"""The two exact verifiers (verifier_a.py, verifier_b.py): the reference's
agreement tests (tests/test_verifier_agreement.py), first-witness order
against a plain restatement of the loops, structural messages, and the
saved E=10 near-solution, which both must reject."""
from itertools import combinations
from pathlib import Path
from random import Random

import numpy as np

ROOT = Path(__file__).resolve().parents[1]


def _first_det_witness(points):
    for a, b, c in combinations(points, 3):
        if (b[0] - a[0]) * (c[1] - a[1]) == (b[1] - a[1]) * (c[0] - a[0]):
            return a, b, c
    return None


def test_independent_verifiers_agree_on_random_small_sets(pkg):
    v = pkg.verifiers
    rng = Random(20260904)
    for _ in range(3000):
        n = rng.randint(2, 9)
        grid = [(r, c) for r in range(n) for c in range(n)]
        points = rng.sample(grid, rng.randint(0, min(14, len(grid))))
        wa, wb = v.first_collinear_triple_a(points), v.first_collinear_triple_b(points)
        assert (wa is None) == (wb is None)
        assert wa == _first_det_witness(points)


def test_opposite_rays_are_treated_as_same_line(pkg):
    points = [(0, 0), (2, 2), (5, 5)]
    assert pkg.verifiers.first_collinear_triple_a(points) is not None
    assert pkg.verifiers.first_collinear_triple_b(points) is not None


def test_verify_messages_and_fixture(pkg, oracle):
    v = pkg.verifiers
    square = [(0, 0), (0, 1), (1, 0), (1, 1)]
    assert v.verify_a(square, 2) == (True, "checked 4 integer determinants")
    assert v.verify_b(square, 2) == (True, "checked normalized directions around 4 anchors")
    assert v.verify_a(square[:3], 2) == (False, "expected 4 points, got 3")
    assert v.verify_b(square[:3], 2) == (False, "cardinality mismatch: 3 versus 4")
    assert v.verify_a(square[:3] + [(0, 0)], 2) == (False, "duplicate point")
    assert v.verify_b(square[:3] + [(0, 0)], 2) == (False, "coordinates are not unique")
    assert v.verify_a([(0, 0), (0, 1), (1, 0), (1, 2)], 2) == (False, "point outside grid")
    assert v.verify_b([(0, 0), (0, 1), (1, 0), (1, 2)], 2) == (False, "out-of-range coordinate (1, 2)")
    assert v.verify_a([(0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (2, 2)], 3) == (False, "row degree is not exactly two")
    assert v.verify_b([(0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (2, 2)], 3) == (False, "not every row occurs twice")
    _, points, _ = pkg.io.load_candidate(ROOT / "tests" / "golden" / "best-near-solution-E10.json", 75)
    ok_a, detail_a = v.verify_a(points, 75)
    ok_b, detail_b = v.verify_b(points, 75)
    assert not ok_a and detail_a.startswith("collinear triple")
    assert not ok_b and detail_b.startswith("repeated normalized direction")
    # a random 2-factor: the verifiers' verdicts agree with the oracle's count
    for n in (8, 20):
        state = pkg.random_state(n, Random(n))
        score = oracle.score_and_badness(np.array(state.P), np.array(state.Q))[0]
        assert v.verify_a(state.points(), n)[0] == v.verify_b(state.points(), n)[0] == (score == 0)
