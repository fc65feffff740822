"""The C-ABI library without a GPU: it loads, exports every symbol declared in
include/no3.h, and its host-only functions agree with the oracle."""
import re
from math import gcd
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "no3.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(n3_\w+)\s*\(", text, re.M)))


def test_library_exports_header(pkg):
    lib = pkg._lib.lib()
    names = declared_symbols()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) <= set(pkg._lib.SIGNATURES)
    assert lib.n3_version() >= 10000


def test_struct_layouts_match_the_header(pkg):
    """The ctypes mirrors of n3_chains and n3_group have the C layout."""
    import ctypes

    import numpy as np

    lib = pkg._lib.lib()
    assert lib.n3_struct_size(0) == ctypes.sizeof(pkg._lib.Chains)
    assert lib.n3_struct_size(1) == ctypes.sizeof(pkg._lib.Group)
    fields = [name for name, _ in pkg._lib.Chains._fields_]
    out = np.zeros(64, dtype=np.int64)
    count = lib.n3_chains_offsets(out.ctypes.data, len(out))
    assert count == len(fields)
    assert list(out[:count]) == [getattr(pkg._lib.Chains, f).offset for f in fields]


def test_direction_table(pkg):
    import ctypes

    lib = pkg._lib.lib()
    for n in (2, 5, 12, 75):
        w = 2 * n - 1
        table = np.empty(w * w, np.uint16)
        d = ctypes.c_int()
        hw = ctypes.c_int()
        assert lib.n3_dirtable(n, table.ctypes.data, ctypes.byref(d), ctypes.byref(hw)) == 0
        table = table.reshape(w, w)
        ids = {}
        for dr in range(-(n - 1), n):
            for dc in range(-(n - 1), n):
                v = int(table[dr + n - 1, dc + n - 1])
                if dr == 0 and dc == 0:
                    assert v == d.value
                    continue
                g = gcd(abs(dr), abs(dc))
                a, b = dr // g, dc // g
                if a < 0 or (a == 0 and b < 0):
                    a, b = -a, -b
                assert ids.setdefault((a, b), v) == v
        assert len(set(ids.values())) == len(ids) == d.value
        assert hw.value * 4 >= d.value + 1 + 4
    # the count of canonical directions at n=75 quoted for the reference grid
    assert lib.n3_dirtable(75, None, ctypes.byref(d), None) == 0 and d.value == 6784
    assert lib.n3_dirtable(101, None, None, None) == pkg._lib.N3_EINVAL


def test_thresholds_match_oracle(pkg, oracle):
    from no_three_in_line_amd import tempering

    temps = np.tile(np.geomspace(0.05, 2.0, 64), 2).astype(np.float32)
    k = tempering.move_table_size(temps)
    assert np.array_equal(tempering.metropolis_thresholds(temps, k), oracle.metropolis_thresholds(temps, k))
    kx = tempering.exchange_table_size(temps, 64)
    thr, sign = tempering.exchange_thresholds(temps, 64, kx)
    othr, osign = oracle.exchange_thresholds(temps, 64, kx)
    assert np.array_equal(thr, othr) and np.array_equal(sign, osign)
    assert np.all(tempering.metropolis_thresholds(temps, k)[:, -1] == 0)


def test_host_validation_errors(pkg):
    import pytest

    with pytest.raises(ValueError):
        pkg.TemperingChains(np.zeros((4, 7), np.int32), np.ones(4, np.float32), 0)
    with pytest.raises(ValueError):
        pkg.TemperingChains(np.zeros((4, 8), np.int32), np.ones(3, np.float32), 0)
    cols = np.tile(np.array([0, 1, 1, 2, 2, 3, 3, 0], np.int32), (2, 1))
    bad = cols.copy()
    bad[0, 1] = 0  # row 0 uses column 0 twice
    with pytest.raises(ValueError):
        pkg.TemperingChains(bad, np.ones(2, np.float32), 0)
    with pytest.raises(ValueError):
        pkg.score_population_gpu(np.zeros((3, 4)), np.zeros((3, 5)))


def test_points_to_permutations(pkg):
    from random import Random

    rng = Random(123)
    state = pkg.random_state(75, rng)
    state.validate()
    again = pkg.points_to_permutations(state.points(), 75)
    assert again.points() == state.points()
    # matching colourings keep the point set (tests/test_state.py:34-38)
    s = pkg.PairPermutations((0, 1, 2, 3, 4, 5), (1, 0, 3, 2, 5, 4))
    assert pkg.points_to_permutations(s.points(), 6).points() == s.points()


def _components(n, points):
    """Connected components of the row/column graph of a point set."""
    parent = list(range(2 * n))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for r, c in points:
        parent[find(r)] = find(n + c)
    return len({find(a) for a in range(2 * n)})


def test_points_to_permutations_returns_the_reference_colouring(pkg):
    """The colouring (which of the 2^(c-1) choices) equals the reference's
    on 141 random 2-factors, n = 5..75 with 1 to 5+ components
    (tests/golden/points_to_permutations.json, written by the reference's
    state.points_to_permutations)."""
    import json

    data = json.loads((ROOT / "tests" / "golden" / "points_to_permutations.json").read_text())
    multi = 0
    for case in data["cases"]:
        n = case["n"]
        points = [tuple(p) for p in case["points"]]
        state = pkg.points_to_permutations(points, n)
        assert list(state.P) == case["P"] and list(state.Q) == case["Q"], n
        multi += _components(n, points) > 1
    assert len(data["cases"]) == 141 and multi >= 90


def test_three_cycle_candidates_match_reference_loop(pkg):
    import importlib

    helper = importlib.import_module("test_oracle")._three_cycle_candidates
    rng = np.random.default_rng(8)
    for n in (3, 4, 7, 12, 30):
        for _ in range(3):
            p = rng.permutation(n).astype(np.int32)
            q = rng.permutation(n).astype(np.int32)
            for u in (0, 2 * n - 1, int(rng.integers(0, 2 * n))):
                assert np.array_equal(pkg.legal_three_cycle_candidates(p, q, u), helper(p, q, u))


def test_apply_three_cycle(pkg):
    state = pkg.random_state(9, __import__("random").Random(3))
    p, q = np.asarray(state.P), np.asarray(state.Q)
    cand = pkg.legal_three_cycle_candidates(p, q, 4)
    assert len(cand) > 0
    moved = pkg.apply_three_cycle(p, q, cand[0])
    moved.validate()
    before, after = set(state.points()), set(moved.points())
    assert len(before - after) == 3 and len(after - before) == 3


def test_random_two_switches_keep_the_two_factor(pkg):
    P, Q = np.arange(9), np.roll(np.arange(9), 1)
    cols = np.repeat(np.stack((P, Q), 1).reshape(1, 18), 5, axis=0).astype(np.int32)
    out = pkg.random_two_switches(cols, 7, seed=3)
    pkg.state.check_columns(out)
    assert out.shape == cols.shape and not np.array_equal(out, cols)
    # every move changes the columns of exactly two slots, so at most 14 differ
    assert np.all((out != cols).sum(1) <= 14)
    assert np.array_equal(out, pkg.random_two_switches(cols, 7, seed=3))
    assert np.array_equal(pkg.random_two_switches(cols, 0, seed=3), cols)


def _reference_descriptors(badness, global_anchors=False):
    """The nested loops of gpu_exhaustive_beam.py:67-82, restated."""
    out = []
    for parent in range(badness.shape[0]):
        anchors = list(range(badness.shape[1])) if global_anchors else list(map(int, np.flatnonzero(badness[parent])))
        anchor_set = set(anchors)
        for u in anchors:
            for v in range(badness.shape[1]):
                if v == u or (v >> 1) == (u >> 1):
                    continue
                if v in anchor_set and v < u:
                    continue
                out.append((parent, u, v))
    return np.asarray(out, dtype=np.int32).reshape(-1, 3)


def test_candidate_descriptors_follow_the_reference_order(pkg):
    rng = np.random.default_rng(4)
    for m in (4, 8, 30):
        bad = np.where(rng.random((5, m)) < 0.3, rng.integers(1, 5, (5, m)), 0)
        bad[2] = 0
        for global_anchors in (False, True):
            got = pkg.candidate_descriptors(bad, global_anchors=global_anchors)
            assert np.array_equal(got, _reference_descriptors(bad, global_anchors))
    # every unordered pair of different rows once, per parent
    full = pkg.candidate_descriptors(np.zeros((1, 10)), global_anchors=True)
    assert len(full) == 10 * 9 // 2 - 5
    assert len(pkg.candidate_descriptors(np.zeros((3, 10)))) == 0


def test_rank_elites_is_the_lexicographic_key_order(pkg):
    rng = np.random.default_rng(8)
    scores = rng.integers(0, 4, 200)
    bad = rng.integers(0, 3, (200, 12))
    order = pkg.rank_elites(scores, bad)
    keys = [pkg.gpu_device_search.elite_key(int(scores[i]), bad[i]) for i in range(200)]
    expected = sorted(range(200), key=lambda i: keys[i])  # stable: ties keep index order
    assert order.dtype == np.int32 and list(order) == expected
    assert len(pkg.rank_elites(np.empty(0), np.empty((0, 4)))) == 0
