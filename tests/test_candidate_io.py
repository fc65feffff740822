"""Candidate formats (io.py:1-132): the strict JSON contract the reference's
tests/test_strict_candidate_input.py pins, the compact codec and the
permissive text form."""
import json
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SQUARE = [[0, 0], [0, 1], [1, 0], [1, 1]]


@pytest.fixture(scope="module")
def io(pkg):
    return pkg.io


def save(tmp_path, payload, name="synthetic.json"):
    path = tmp_path / name
    path.write_text(json.dumps(payload), encoding="utf-8")
    return path


@pytest.mark.parametrize("points", [
    [[bool(r), bool(c)] for r, c in SQUARE],
    [[float(r), float(c)] for r, c in SQUARE],
    [[r + .25, c + .75] for r, c in SQUARE],
    [[str(r), str(c)] for r, c in SQUARE],
    ["00", "01", "10", "11"],
    [[0, 0, 99], [0, 1], [1, 0], [1, 1]],
    [[0], [0, 1], [1, 0], [1, 1]],
    [None, [0, 1], [1, 0], [1, 1]],
])
def test_json_coordinates_are_never_coerced(io, tmp_path, points):
    with pytest.raises(ValueError, match="two integers"):
        io.load_candidate(save(tmp_path, {"n": 2, "points": points}))


@pytest.mark.parametrize("points", [None, True, 3, "00", {"0": [0, 0]}])
def test_points_container_is_a_list(io, tmp_path, points):
    with pytest.raises(ValueError):
        io.load_candidate(save(tmp_path, {"n": 2, "points": points}))


@pytest.mark.parametrize("bad_n", [True, False, 2.0, 2.5, "2", None, 0, -1])
def test_declared_json_n_is_strict_even_with_override(io, tmp_path, bad_n):
    path = save(tmp_path, {"n": bad_n, "points": SQUARE})
    for explicit in (None, 2):
        with pytest.raises(ValueError, match="positive integer"):
            io.load_candidate(path, n=explicit)


@pytest.mark.parametrize("bad_n", [True, False, 2.0, 2.5, "2", 0, -1])
def test_api_n_is_strict(io, tmp_path, bad_n):
    with pytest.raises(ValueError, match="positive integer"):
        io.load_candidate(save(tmp_path, {"n": 2, "points": SQUARE}), n=bad_n)
    with pytest.raises(ValueError, match="positive integer"):
        io.decode_compact(".0101", n=bad_n)


@pytest.mark.parametrize("payload", [SQUARE, {"points": SQUARE}, {"n": 2, "points": SQUARE, "label": "synthetic"},
                                     {"n": 2, "P": [0, 1], "Q": [1, 0]}])
def test_valid_integers_remain_zero_based_and_unchanged(io, tmp_path, payload):
    n, points, metadata = io.load_candidate(save(tmp_path, payload))
    assert n == 2 and sorted(points) == [tuple(p) for p in SQUARE]
    assert all(type(x) is int for point in points for x in point)
    assert metadata["source_path"].endswith("synthetic.json")
    if isinstance(payload, dict) and "label" in payload:
        assert metadata["label"] == "synthetic"


def test_override_compact_and_text_contracts(io, tmp_path):
    path = save(tmp_path, {"n": 2, "points": SQUARE})
    assert io.load_candidate(path, n=3)[0] == 3
    assert io.decode_compact(".0101")[2] == [tuple(p) for p in SQUARE]
    text = tmp_path / "synthetic.txt"
    text.write_text("0 0\n0 1\n1 0\n1 1\n", encoding="utf-8")
    assert io.load_candidate(text)[:2] == (2, [tuple(p) for p in SQUARE])
    one_based = [[r + 1, c + 1] for r, c in SQUARE]
    n, points, _ = io.load_candidate(save(tmp_path, {"n": 2, "points": one_based}))
    assert points == [tuple(p) for p in one_based]  # never shifted
    odd = tmp_path / "odd.txt"
    odd.write_text("points: 1, 2 and 3\n", encoding="utf-8")
    with pytest.raises(ValueError, match="odd"):
        io.load_candidate(odd)


@pytest.mark.parametrize("bad", [True, 0.0, "0", None, [0, 1]])
def test_pq_coordinates_are_strict(io, tmp_path, bad):
    with pytest.raises(ValueError):
        io.load_candidate(save(tmp_path, {"n": 2, "P": [bad, 1], "Q": [1, 0]}))


@pytest.mark.parametrize("bad", ["01", {}, None, True])
def test_pq_containers_are_lists(io, tmp_path, bad):
    with pytest.raises(ValueError):
        io.load_candidate(save(tmp_path, {"n": 2, "P": bad, "Q": [1, 0]}))


def test_compact_codec(io, pkg, tmp_path):
    from random import Random

    for n in (2, 9, 75, 90):
        state = pkg.random_state(n, Random(n))
        for symmetry in ("iden", "rot2", "full", ":"):
            code = io.encode_compact(state.points(), n, symmetry)
            size, name, points = io.decode_compact(code)
            assert size == n and sorted(points) == state.points()
            assert name == io.SYMMETRY_NAMES[code[0]]
        path = tmp_path / f"c{n}.txt"
        path.write_text(code + "\n", encoding="utf-8")
        size, points, metadata = io.load_candidate(path)
        assert size == n and metadata["compact"] == code and sorted(points) == state.points()
    assert len(io.ALPHABET) == len(set(io.ALPHABET)) == 90
    for bad in ("", "Q0101", ".012", ".0!01"):
        with pytest.raises(ValueError):
            io.decode_compact(bad)
    with pytest.raises(ValueError):
        io.decode_compact(".0102", n=2)  # column 2 outside n=2
    with pytest.raises(ValueError):
        io.encode_compact([(0, 0), (0, 1), (1, 0)], 2)
    with pytest.raises(ValueError):
        io.encode_compact(SQUARE, 2, "bogus")


def test_write_candidate_and_fixture(io, tmp_path):
    path = tmp_path / "out.json"
    io.write_candidate(path, 2, [(1, 1), (0, 1), (1, 0), (0, 0)], conflicts=0, method="x")
    text = path.read_text(encoding="utf-8")
    assert text.endswith("\n")
    data = json.loads(text)
    assert list(data) == ["n", "points", "conflicts", "method"] and data["points"] == SQUARE
    n, points, metadata = io.load_candidate(ROOT / "tests" / "golden" / "best-near-solution-E10.json")
    assert n == 75 and len(points) == 150 and len(set(points)) == 150
