"""Candidate files in the reference's formats (io.py:1-132).

- JSON (strict): an object {"n": n, "points": [[row, column], ...]} or
  {"n": n, "P": [...], "Q": [...]} with free-form metadata keys, or a bare
  list of points.  Coordinates must be genuine integers (no bool, float or
  string aliases) and n a positive integer; nothing is coerced or shifted.
- Compact code: one token, a symmetry prefix character then two column
  characters per row from the reference's 90-character alphabet.
- Plain text (permissive): every integer in the text, taken in pairs.
"""
from __future__ import annotations

import json
import re
from pathlib import Path
from typing import Iterable

from .state import Point

# The compact alphabet and symmetry prefixes are a file format: they must be
# the reference's characters (io.py:9-21) for codes to be interchangeable.
ALPHABET = "0123456789ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz#$%&@?!()[]<>{}=*+|-/~^_:;,."
SYMMETRY_PREFIXES = ".:/-ocx+*"
SYMMETRY_NAMES = {".": "iden", ":": "rot2", "/": "dia1", "-": "ort1", "o": "rot4", "c": "rct4", "x": "dia2",
                  "+": "ort2", "*": "full"}
_COLUMN_OF = {ch: value for value, ch in enumerate(ALPHABET)}


def _grid_n(value: object, label: str = "n") -> int:
    """A positive int, and only an int (io.py:25-28)."""
    if type(value) is not int or value < 1:
        raise ValueError(f"{label} must be a positive integer")
    return value


def _pairs(value: object) -> list[Point]:
    """io.py:31-40: a list of exactly-two-integer lists."""
    if not isinstance(value, list):
        raise ValueError("JSON points must be a list")
    points: list[Point] = []
    for index, pair in enumerate(value):
        if not isinstance(pair, list) or len(pair) != 2 or any(type(v) is not int for v in pair):
            raise ValueError(f"point {index} must be a list of exactly two integers")
        points.append((pair[0], pair[1]))
    return points


def decode_compact(code: str, n: int | None = None) -> tuple[int, str, list[Point]]:
    """io.py:44-67: (n, symmetry name, points) of a compact code; slot i of
    the payload is a point of row i // 2."""
    if n is not None:
        n = _grid_n(n)
    code = code.strip()
    if not code or code[0] not in SYMMETRY_PREFIXES:
        raise ValueError("compact code has no recognized symmetry prefix")
    payload = code[1:]
    if n is None:
        if len(payload) % 2:
            raise ValueError("compact payload length is not even")
        n = len(payload) // 2
    n = _grid_n(n)
    if len(payload) != 2 * n:
        raise ValueError(f"compact payload has {len(payload)} columns, expected {2 * n}")
    points: list[Point] = []
    for offset, ch in enumerate(payload):
        column = _COLUMN_OF.get(ch)
        if column is None:
            raise ValueError(f"unknown compact character {ch!r}")
        if column >= n:
            raise ValueError(f"column {column} outside n={n}")
        points.append((offset // 2, column))
    return n, SYMMETRY_NAMES[code[0]], points


def encode_compact(points: Iterable[Point], n: int, symmetry: str = "iden") -> str:
    """io.py:70-82: the compact code of a two-points-per-row candidate (each
    row's columns in increasing order)."""
    prefix = {name: p for p, name in SYMMETRY_NAMES.items()}.get(symmetry)
    if prefix is None and symmetry in SYMMETRY_PREFIXES:
        prefix = symmetry
    if prefix is None:
        raise ValueError(f"unknown symmetry {symmetry!r}")
    rows: list[list[int]] = [[] for _ in range(n)]
    for row, column in points:
        if not (0 <= row < n and 0 <= column < n):
            raise ValueError(f"point {(row, column)} outside n={n}")
        rows[row].append(column)
    if any(len(r) != 2 for r in rows):
        raise ValueError("compact format requires exactly two points in every row")
    return prefix + "".join(ALPHABET[c] for r in rows for c in sorted(r))


def load_candidate(path, n: int | None = None) -> tuple[int, list[Point], dict]:
    """io.py:85-126: (n, points, metadata) from a JSON, compact or text
    candidate.  An explicit n overrides the file's (after both are checked);
    otherwise the file's n, else one more than the largest row."""
    if n is not None:
        n = _grid_n(n)
    candidate = Path(path)
    text = candidate.read_text(encoding="utf-8").strip()
    metadata: dict = {"source_path": str(candidate)}
    if candidate.suffix.lower() == ".json" or text.startswith("{") or text.startswith("["):
        payload = json.loads(text)
        if isinstance(payload, dict):
            file_n = _grid_n(payload["n"], "JSON n") if "n" in payload else None
            raw = payload.get("points")
            if raw is None and "P" in payload and "Q" in payload:
                if not isinstance(payload["P"], list) or not isinstance(payload["Q"], list):
                    raise ValueError("JSON P and Q must be lists")
                raw = [[r, c] for r, c in enumerate(payload["P"])] + [[r, c] for r, c in enumerate(payload["Q"])]
            if raw is None:
                raise ValueError("JSON candidate has neither points nor P/Q")
            metadata.update({k: v for k, v in payload.items() if k != "points"})
        elif isinstance(payload, list):
            raw, file_n = payload, None
        else:
            raise ValueError("JSON candidate must be an object or a points list")
        points = _pairs(raw)
        inferred = max((row for row, _ in points), default=-1) + 1
        return _grid_n(n if n is not None else file_n if file_n is not None else inferred), points, metadata
    tokens = text.split()
    if len(tokens) == 1 and tokens[0][0] in SYMMETRY_PREFIXES:
        size, symmetry, points = decode_compact(tokens[0], n=n)
        metadata.update({"compact": tokens[0], "symmetry": symmetry})
        return size, points, metadata
    values = [int(v) for v in re.findall(r"-?\d+", text)]
    if len(values) % 2:
        raise ValueError("text candidate contains an odd number of integers")
    points = list(zip(values[::2], values[1::2]))
    inferred = max((row for row, _ in points), default=-1) + 1
    return _grid_n(n if n is not None else inferred), points, metadata


def write_candidate(path, n: int, points: Iterable[Point], **metadata: object) -> None:
    """io.py:129-132: sorted points, then the metadata keys."""
    ordered = sorted((int(r), int(c)) for r, c in points)
    payload = {"n": int(n), "points": [[r, c] for r, c in ordered], **metadata}
    Path(path).write_text(json.dumps(payload, indent=2, sort_keys=False) + "\n", encoding="utf-8")
