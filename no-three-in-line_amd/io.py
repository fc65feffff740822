"""Candidate files: read and write the three on-disk forms of a point set.

Format specification (what files written by the reference's tools look like,
reference io.py:1-132; this module is written from this specification):

1. Grid size.  A grid size is a Python int >= 1; bool, float and str values
   are refused ("... must be a positive integer").  The size is taken from,
   in this order of precedence: the caller's explicit ``n``; the size the
   file declares (JSON key "n"); one more than the largest row among the
   points.  An explicit ``n`` and a declared "n" are both checked even when
   the explicit one wins.
2. Which form.  A file is JSON when its suffix is ``.json`` (any case) or
   its stripped text starts with "{" or "[".  Otherwise it is a compact
   code when it is a single whitespace-free token starting with a symmetry
   prefix; anything else is plain text.
3. JSON.  Either a bare list of points, or an object with "points", or an
   object with list-valued "P" and "Q" (row r holds the points (r, P[r])
   and (r, Q[r]); the P points come first).  A point is a list of exactly
   two genuine ints; nothing is coerced, rounded or shifted.  Every key of
   an object except "points" is returned as metadata.
4. Compact code.  One prefix character naming the symmetry class, then two
   characters per row (row 0 first, each row's columns in increasing
   order) from a fixed 90-character alphabet, so that slot i of the payload
   is a point of row i // 2.  The size is half the payload length unless
   given.
5. Plain text.  Every (possibly negative) decimal integer in the text, in
   order, taken two at a time as (row, column); an odd count is an error.
6. Writing.  JSON with keys "n", "points" (sorted), then the caller's
   metadata, indented by 2, with a trailing newline.

The alphabet, the prefixes and their names are part of the file format and
must be the reference's characters (io.py:9-21) for codes to be
interchangeable.
"""
from __future__ import annotations

import json
import re
from collections import Counter
from pathlib import Path
from typing import Iterable

from .state import Point

ALPHABET = "0123456789ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz#$%&@?!()[]<>{}=*+|-/~^_:;,."
SYMMETRY_PREFIXES = ".:/-ocx+*"
SYMMETRY_NAMES = dict(zip(SYMMETRY_PREFIXES, ("iden", "rot2", "dia1", "ort1", "rot4", "rct4", "dia2", "ort2", "full")))
_PREFIX_OF = {name: prefix for prefix, name in SYMMETRY_NAMES.items()}
_VALUE_OF = {ch: i for i, ch in enumerate(ALPHABET)}
_INTEGER = re.compile(r"-?\d+")


def _positive_int(value: object, what: str = "n") -> int:
    """Rule 1: exactly an int (no bool or other subclass) of at least 1."""
    if type(value) is not int or value < 1:
        raise ValueError(f"{what} must be a positive integer")
    return value


def _strict_points(items: object) -> list[Point]:
    """Rule 3: a list whose items are lists of exactly two genuine ints."""
    if not isinstance(items, list):
        raise ValueError("JSON points must be a list")

    bad = next((i for i, item in enumerate(items)
                if not (isinstance(item, list) and len(item) == 2 and all(type(x) is int for x in item))), None)
    if bad is not None:
        raise ValueError(f"point {bad} must be a list of exactly two integers")
    return [(item[0], item[1]) for item in items]


def _size(explicit: int | None, declared: int | None, points: list[Point]) -> int:
    """Rule 1's precedence; the result is checked again."""
    if explicit is not None:
        return _positive_int(explicit)
    if declared is not None:
        return _positive_int(declared)
    return _positive_int(1 + max((row for row, _ in points), default=-1))


def _is_json(path: Path, text: str) -> bool:
    return path.suffix.lower() == ".json" or text[:1] in ("{", "[")


def _read_json(text: str) -> tuple[int | None, list[Point], dict]:
    """(declared n, points, metadata) of a JSON candidate (rule 3)."""
    document = json.loads(text)
    if isinstance(document, list):
        return None, _strict_points(document), {}
    if not isinstance(document, dict):
        raise ValueError("JSON candidate must be an object or a points list")
    declared = _positive_int(document["n"], "JSON n") if "n" in document else None
    items = document.get("points")
    if items is None:
        if "P" not in document or "Q" not in document:
            raise ValueError("JSON candidate has neither points nor P/Q")
        layers = document["P"], document["Q"]
        if not all(isinstance(layer, list) for layer in layers):
            raise ValueError("JSON P and Q must be lists")
        items = [[row, column] for layer in layers for row, column in enumerate(layer)]
    extra = {key: value for key, value in document.items() if key != "points"}
    return declared, _strict_points(items), extra


def _read_text(text: str) -> list[Point]:
    """Rule 5."""
    values = [int(token) for token in _INTEGER.findall(text)]
    if len(values) % 2:
        raise ValueError("text candidate contains an odd number of integers")
    it = iter(values)
    return list(zip(it, it))


def decode_compact(code: str, n: int | None = None) -> tuple[int, str, list[Point]]:
    """Rule 4: (n, symmetry name, points) of a compact code."""
    if n is not None:
        _positive_int(n)
    code = code.strip()
    prefix, payload = code[:1], code[1:]
    if prefix == "" or prefix not in SYMMETRY_NAMES:
        raise ValueError("compact code has no recognized symmetry prefix")
    if n is None:
        if len(payload) & 1:
            raise ValueError("compact payload length is not even")
        n = _positive_int(len(payload) // 2)
    if len(payload) != 2 * n:
        raise ValueError(f"compact payload has {len(payload)} columns, expected {2 * n}")
    points: list[Point] = []
    for slot, ch in enumerate(payload):
        if ch not in _VALUE_OF:
            raise ValueError(f"unknown compact character {ch!r}")
        if _VALUE_OF[ch] >= n:
            raise ValueError(f"column {_VALUE_OF[ch]} outside n={n}")
        points.append((slot >> 1, _VALUE_OF[ch]))
    return n, SYMMETRY_NAMES[prefix], points


def encode_compact(points: Iterable[Point], n: int, symmetry: str = "iden") -> str:
    """Rule 4 in the other direction; `symmetry` is a name or a prefix."""
    prefix = _PREFIX_OF.get(symmetry) or (symmetry if symmetry in SYMMETRY_NAMES else None)
    if prefix is None:
        raise ValueError(f"unknown symmetry {symmetry!r}")
    ordered = sorted(points)
    for row, column in ordered:
        if not (0 <= row < n and 0 <= column < n):
            raise ValueError(f"point {(row, column)} outside n={n}")
    per_row = Counter(row for row, _ in ordered)
    if len(ordered) != 2 * n or any(per_row[row] != 2 for row in range(n)):
        raise ValueError("compact format requires exactly two points in every row")
    # sorted (row, column) pairs are already row by row, columns increasing
    return prefix + "".join(ALPHABET[column] for _, column in ordered)


def load_candidate(path, n: int | None = None) -> tuple[int, list[Point], dict]:
    """(n, points, metadata) of a candidate file in any of the three forms;
    metadata always holds "source_path"."""
    if n is not None:
        _positive_int(n)
    source = Path(path)
    text = source.read_text(encoding="utf-8").strip()
    metadata: dict = {"source_path": str(source)}
    if _is_json(source, text):
        declared, points, extra = _read_json(text)
        metadata.update(extra)
        return _size(n, declared, points), points, metadata
    tokens = text.split()
    if len(tokens) == 1 and tokens[0][0] in SYMMETRY_NAMES:
        size, symmetry, points = decode_compact(tokens[0], n=n)
        metadata["compact"] = tokens[0]
        metadata["symmetry"] = symmetry
        return size, points, metadata
    points = _read_text(text)
    return _size(n, None, points), points, metadata


def write_candidate(path, n: int, points: Iterable[Point], **metadata: object) -> None:
    """Rule 6."""
    document = {"n": int(n), "points": [[r, c] for r, c in sorted((int(r), int(c)) for r, c in points)]}
    document.update(metadata)
    Path(path).write_text(json.dumps(document, indent=2) + "\n", encoding="utf-8")
