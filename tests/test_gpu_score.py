"""Population scoring on the GPU against the CPU oracle (bit-exact)."""
import json
from pathlib import Path

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def _score(pkg, p, q, dev, method):
    tp = torch.from_numpy(np.ascontiguousarray(p, dtype=np.int32)).to(dev)
    tq = torch.from_numpy(np.ascontiguousarray(q, dtype=np.int32)).to(dev)
    bad = torch.empty((tp.shape[0], 2 * tp.shape[1]), dtype=torch.int32, device=dev)
    scores = pkg.score_population_device(tp, tq, badness=bad, method=method)
    return scores.cpu().numpy(), bad.cpu().numpy()


@pytest.mark.parametrize("n,count,seed", [(2, 64, 1), (5, 64, 2), (32, 256, 0), (33, 65, 3), (64, 70, 4),
                                          (75, 130, 5), (100, 40, 6)])
def test_generator_matches_oracle(pkg, oracle, dev, n, count, seed):
    p, q = pkg.generate_states(seed, n, count, device=dev)
    hp, hq = oracle.generate_states(seed, n, count)
    assert np.array_equal(p.cpu().numpy(), hp)
    assert np.array_equal(q.cpu().numpy(), hq)
    assert np.all(np.sort(hp, axis=1) == np.arange(n)) and np.all(hp != hq)


@pytest.mark.parametrize("method", [0, 1, 2])
@pytest.mark.parametrize("n,count,seed", [(2, 8, 1), (3, 16, 9), (7, 64, 2), (32, 256, 0), (33, 64, 3),
                                          (48, 64, 8), (64, 64, 4), (65, 40, 10), (75, 96, 5), (78, 48, 11),
                                          (80, 32, 12), (83, 32, 7), (84, 24, 14), (90, 16, 15), (97, 16, 13),
                                          (100, 24, 6)])
def test_scores_and_badness_match_oracle(pkg, oracle, dev, method, n, count, seed):
    hp, hq = oracle.generate_states(seed, n, count)
    scores, bad = _score(pkg, hp, hq, dev, method)
    ref_scores, ref_bad = oracle.score_population(hp, hq)
    assert np.array_equal(scores, ref_scores)
    assert np.array_equal(bad, ref_bad)
    assert np.array_equal(bad.sum(axis=1), 3 * scores)


def test_config2_subsample(pkg, oracle, dev):
    """n=75, P=65,536 seeded states scored on the GPU; 512-state subsample
    (every 128th state) against the oracle, all states determinant vs slope."""
    n, count, seed = 75, 65536, 0
    p, q = pkg.generate_states(seed, n, count, device=dev)
    bad = torch.empty((count, 2 * n), dtype=torch.int32, device=dev)
    scores = pkg.score_population_device(p, q, badness=bad)
    idx = np.arange(0, count, count // 512)
    hp, hq = oracle.generate_states(seed, n, count)
    assert np.array_equal(p.cpu().numpy(), hp) and np.array_equal(q.cpu().numpy(), hq)
    ref_scores, ref_bad = oracle.score_population(hp[idx], hq[idx])
    s_h, b_h = scores.cpu().numpy(), bad.cpu().numpy()
    assert np.array_equal(s_h[idx], ref_scores)
    assert np.array_equal(b_h[idx], ref_bad)
    assert np.array_equal(b_h.sum(axis=1), 3 * s_h)
    # determinant kernel over the first 4096 states as the second formulation
    det_bad = torch.empty((4096, 2 * n), dtype=torch.int32, device=dev)
    det = pkg.score_population_device(p[:4096].contiguous(), q[:4096].contiguous(), badness=det_bad,
                                      method=pkg.SCORE_DETERMINANT)
    assert torch.equal(det, scores[:4096]) and torch.equal(det_bad, bad[:4096])


def test_near_solution_fixture(pkg, dev):
    data = json.loads((GOLDEN / "best-near-solution-E10.json").read_text())
    n = data["n"]
    state = pkg.points_to_permutations([tuple(x) for x in data["points"]], n)
    for method in (0, 1):
        scores, bad = _score(pkg, np.asarray([state.P]), np.asarray([state.Q]), dev, method)
        assert int(scores[0]) == 10
        assert int(bad.sum()) == 30 and int(bad.max()) == 3 and int((bad > 0).sum()) == 24


def test_coincident_points_and_non_permutations(pkg, oracle, dev):
    """Inputs outside the state model: P[r] == Q[r] (coincident points) and
    rows that are not permutations are counted exactly as the reference's
    determinant loop counts them."""
    rng = np.random.default_rng(11)
    for n in (4, 9, 31, 33, 75, 78, 100):
        p = rng.integers(0, n, size=(20, n)).astype(np.int32)
        q = rng.integers(0, n, size=(20, n)).astype(np.int32)
        q[:, : n // 3] = p[:, : n // 3]
        for method in (0, 1, 2):
            scores, bad = _score(pkg, p, q, dev, method)
            ref_scores, ref_bad = oracle.score_population(p, q)
            assert np.array_equal(scores, ref_scores)
            assert np.array_equal(bad, ref_bad)


@pytest.mark.parametrize("n", [33, 64, 75, 78, 83, 90, 100])
def test_lattice_states_with_long_lines(pkg, oracle, dev, n):
    """States built from lines, P[r] = a r + b and Q[r] = c r + d (mod n),
    without coincident points: up to n points lie on one line and most
    pairs are collinear with others, so the histogram counters reach their
    largest values (n - 1 earlier points on a line)."""
    r = np.arange(n)
    rows = []
    for a, b, c, d in [(1, 0, 1, 1), (1, 0, n - 1, 0), (2, 0, 2, 1), (1, 0, 3, 0), (0, 0, 1, 0), (1, 5, 1, 7)]:
        pr, qr = (a * r + b) % n, (c * r + d) % n
        if np.any(pr == qr):
            qr = (qr + 1 + (n // 2)) % n
            if np.any(pr == qr):
                continue
        rows.append((pr, qr))
    p = np.stack([x for x, _ in rows]).astype(np.int32)
    q = np.stack([y for _, y in rows]).astype(np.int32)
    for method in (0, 2):
        scores, bad = _score(pkg, p, q, dev, method)
        ref_scores, ref_bad = oracle.score_population(p, q)
        assert np.array_equal(scores, ref_scores)
        assert np.array_equal(bad, ref_bad)


def test_out_of_range_columns_raise(pkg, dev):
    p = np.tile(np.arange(10, dtype=np.int32), (3, 1))
    q = np.roll(p, 1, axis=1)
    q[1, 4] = 10
    with pytest.raises(ValueError):
        _score(pkg, p, q, dev, 0)
    with pytest.raises(ValueError):
        pkg.score_population_gpu(p[:, :5], q)


def test_host_entry_point(pkg, oracle, dev):
    hp, hq = oracle.generate_states(91001, 75, 1024)
    ref_scores, ref_bad = oracle.score_population(hp, hq)
    assert np.array_equal(pkg.score_population_gpu(hp, hq), ref_scores)
    s, b = pkg.score_population_badness(hp, hq)
    assert np.array_equal(s, ref_scores) and np.array_equal(b, ref_bad)


def test_host_entry_points_large_population(pkg, oracle, dev):
    """A larger host batch (odd size): every state against the oracle, an
    out-of-range column deep in the batch raises, and the arrays of one call
    (views of page-locked buffers) stay valid after the next calls."""
    chunk = 8192
    count = 2 * chunk + 37
    hp, hq = oracle.generate_states(5, 33, count)
    ref_scores, ref_bad = oracle.score_population(hp, hq)
    s1, b1 = pkg.score_population_badness(hp, hq)
    s2 = pkg.score_population_gpu(hp[::-1].copy(), hq[::-1].copy())
    assert np.array_equal(s1, ref_scores) and np.array_equal(b1, ref_bad)
    assert np.array_equal(s2, ref_scores[::-1])
    empty = pkg.score_population_gpu(hp[:0], hq[:0])
    assert empty.shape == (0,)
    bad_q = hq.copy()
    bad_q[chunk + 5, 3] = 33
    with pytest.raises(ValueError):
        pkg.score_population_badness(hp, bad_q)
    assert np.array_equal(pkg.score_population_gpu(hp, hq), ref_scores)


def test_empty_population(pkg, dev):
    p = torch.empty((0, 75), dtype=torch.int32, device=dev)
    out = pkg.score_population_device(p, p.clone())
    assert out.shape == (0,)


def test_large_n_determinant_path(pkg, oracle, dev):
    """n above the direction-table limit (100) uses the determinant kernel."""
    hp, hq = oracle.generate_states(3, 128, 8)
    scores, bad = _score(pkg, hp, hq, dev, 0)
    ref_scores, ref_bad = oracle.score_population(hp, hq)
    assert np.array_equal(scores, ref_scores) and np.array_equal(bad, ref_bad)


def test_reflections_leave_scores_invariant_full_population(pkg, dev):
    """n=75, P=65,536: mirroring columns (c -> n-1-c) keeps every slot's
    badness; flipping rows (r -> n-1-r) moves each slot's badness with its
    row; both keep every score.  A wrong sign or canonicalisation in the
    direction table breaks these symmetries."""
    n, count = 75, 65536
    p, q = pkg.generate_states(0, n, count, device=dev)
    bad = torch.empty((count, 2 * n), dtype=torch.int32, device=dev)
    scores = pkg.score_population_device(p, q, badness=bad)
    mp, mq = (n - 1 - p).contiguous(), (n - 1 - q).contiguous()
    mbad = torch.empty_like(bad)
    assert torch.equal(pkg.score_population_device(mp, mq, badness=mbad), scores)
    assert torch.equal(mbad, bad)
    fp, fq = p.flip(1).contiguous(), q.flip(1).contiguous()
    fbad = torch.empty_like(bad)
    assert torch.equal(pkg.score_population_device(fp, fq, badness=fbad), scores)
    assert torch.equal(fbad.view(count, n, 2).flip(1).reshape(count, 2 * n), bad)
