"""Candidate files in the reference's strict JSON shape (io.py:85-132):
{"n": n, "points": [[row, column], ...]} or {"n": n, "P": [...], "Q": [...]},
plus free-form metadata keys on write."""
from __future__ import annotations

import json
from pathlib import Path

from .state import Point


def load_candidate(path, n: int | None = None) -> tuple[int, list[Point], dict]:
    data = json.loads(Path(path).read_text(encoding="utf-8"))
    if not isinstance(data, dict) or "n" not in data:
        raise ValueError("candidate must be a JSON object with an integer n")
    size = data["n"]
    if not isinstance(size, int) or isinstance(size, bool) or size < 1:
        raise ValueError("n must be a positive integer")
    if n is not None and size != n:
        raise ValueError(f"expected n={n}, got n={size}")
    if "points" in data:
        points = []
        for item in data["points"]:
            if not (isinstance(item, list) and len(item) == 2 and all(isinstance(v, int) and not isinstance(v, bool)
                                                                     for v in item)):
                raise ValueError("points must be [row, column] integer pairs")
            points.append((item[0], item[1]))
    elif "P" in data and "Q" in data:
        P, Q = data["P"], data["Q"]
        if len(P) != size or len(Q) != size:
            raise ValueError("P and Q must have length n")
        points = sorted([(r, int(P[r])) for r in range(size)] + [(r, int(Q[r])) for r in range(size)])
    else:
        raise ValueError("candidate needs points or P and Q")
    metadata = {k: v for k, v in data.items() if k not in ("n", "points", "P", "Q")}
    return size, points, metadata


def write_candidate(path, n: int, points: list[Point], **metadata) -> None:
    payload = {"n": n, "points": [[int(r), int(c)] for r, c in points], **metadata}
    Path(path).write_text(json.dumps(payload, indent=2) + "\n", encoding="utf-8")
