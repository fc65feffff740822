"""The learned walk's Metropolis-Hastings law, pinned on the CPU.

The oracle's learned walk (n3o_advance_learned, which the device walk matches
bit for bit) is run for ONE step from one fixed state in many replicas at
once; the Philox counters differ by replica slot and sweep, so the replicas
are independent samples of one transition.  Their tally is compared with the
transition law built here from the sampling mechanism alone:

    proposal of the unordered pair {u, v}:
        q(x -> x') = 0.2 * 2 / m^2 + 0.8 * 2 w(u) w(v) / S1^2
    acceptance:
        min(1, exp(-dE / T) * q(x' -> x) / q(x -> x'))

with w(k) = clamp(table[E][min(badness_k, 63)], 1, 2^15), S1 = sum_k w(k),
and x' the state after the switch, recounted here by brute force.  The
detailed-balance identity pi(x) P(x -> x') = pi(x') P(x' -> x), pi ~
exp(-E/T), is then checked on the measured frequencies of both directions.
Basis: the Metropolis rule of tempering.py:92-95, extended by the Hastings
ratio of the learned proposals.

Mistakes planted in the oracle (S1 and S1' swapped, the uniform term dropped,
the new state's weights read at the old energy) must fail these checks.
"""
import math
from itertools import combinations

import numpy as np
import pytest

N = 6
M = 2 * N
T = 1.5
MAX_E = 24
K64 = 32
REPLICAS = 100_000


def _table():
    # weights whose shape over the badness depends on the energy, so that a
    # wrong energy row is not a uniform rescaling (which would cancel)
    wt = np.zeros((MAX_E + 1, 64), dtype=np.uint32)
    for e in range(MAX_E + 1):
        for b in range(64):
            wt[e, b] = min(2 ** 15, 2 ** min(15, b * (1 + e % 3)) + 3 * (e % 4))
    return wt


def _recount(cols):
    """E and badness by the determinant over all slot triples."""
    pts = [(k >> 1, int(cols[k])) for k in range(M)]
    bad = np.zeros(M, dtype=np.int64)
    e = 0
    for i, j, k in combinations(range(M), 3):
        (a0, a1), (b0, b1), (c0, c1) = pts[i], pts[j], pts[k]
        if (b0 - a0) * (c1 - a1) - (b1 - a1) * (c0 - a0) == 0:
            e += 1
            bad[[i, j, k]] += 1
    return e, bad


def _weights(wt, e, bad):
    return np.clip(wt[min(e, MAX_E), np.minimum(bad, 63)].astype(np.float64), 1, 2 ** 15)


def _switched(cols, u, v):
    out = cols.copy()
    out[u], out[v] = cols[v], cols[u]
    return out


def _law(wt, cols):
    """P(x -> x') for every valid pair {u, v}, from the mechanism."""
    e, bad = _recount(cols)
    w = _weights(wt, e, bad)
    s1 = w.sum()
    law = {}
    for u, v in combinations(range(M), 2):
        cu, cv = cols[u], cols[v]
        if e == 0 or u >> 1 == v >> 1 or cu == cv or cols[u ^ 1] == cv or cols[v ^ 1] == cu:
            continue
        x2 = _switched(cols, u, v)
        e2, bad2 = _recount(x2)
        w2 = _weights(wt, e2, bad2)
        q = 0.4 / M ** 2 + 1.6 * w[u] * w[v] / s1 ** 2
        q_back = 0.4 / M ** 2 + 1.6 * w2[u] * w2[v] / w2.sum() ** 2
        law[(u, v)] = q * min(1.0, math.exp(-(e2 - e) / T) * q_back / q)
    return e, law


def _tally(oracle, wt, cols, sweeps, seed=5):
    """Outcome counts of one learned step from `cols`, REPLICAS x sweeps
    samples: {(u, v): count} plus the count of unchanged replicas."""
    rows = np.tile(cols, (REPLICAS, 1)).astype(np.int32)
    replay = oracle.TemperingReplay(rows[:2], np.full(2, T, np.float32), seed, k_move=64, k_exchange=64)
    e, bad = _recount(cols)
    thr64 = oracle.metropolis_thresholds64(np.full(REPLICAS, T, np.float32), K64)
    temp_index = np.zeros(REPLICAS, dtype=np.int32)
    counts, stay = {}, 0
    for sweep in range(sweeps):
        replay.columns = rows.copy()
        replay.badness = np.tile(bad.astype(np.int32), (REPLICAS, 1))
        replay.scores = np.full(REPLICAS, e, np.int32)
        replay.best_columns = rows.copy()
        replay.best_scores = replay.scores.copy()
        replay.accepted = np.zeros(REPLICAS, np.int32)
        replay.removed_points = np.zeros(REPLICAS, np.int32)
        replay.replicas = REPLICAS
        replay.sweep = 1000 + sweep
        replay.advance_learned(1, wt, temp_index, MAX_E, thr64)
        changed = replay.columns != rows
        moved = changed.sum(axis=1)
        assert set(np.unique(moved)) <= {0, 2}
        stay += int((moved == 0).sum())
        idx = np.nonzero(moved == 2)[0]
        pairs = np.sort(np.nonzero(changed[idx])[1].reshape(-1, 2), axis=1)
        keys, cnt = np.unique(pairs, axis=0, return_counts=True)
        for (u, v), c in zip(keys, cnt):
            counts[(int(u), int(v))] = counts.get((int(u), int(v)), 0) + int(c)
        # the replay's own bookkeeping agrees with a recount of every move
        for (u, v) in {tuple(map(int, k)) for k in keys}:
            r = idx[np.nonzero((pairs[:, 0] == u) & (pairs[:, 1] == v))[0][0]]
            e2, bad2 = _recount(replay.columns[r])
            assert replay.scores[r] == e2 and np.array_equal(replay.badness[r], bad2)
    return counts, stay


def _start_state(oracle):
    """A seeded n=6 state with a few collinear triples and many valid moves."""
    p, q = oracle.generate_states(3, N, 64)
    for i in range(64):
        cols = np.empty(M, dtype=np.int32)
        cols[0::2], cols[1::2] = p[i], q[i]
        e, _ = _recount(cols)
        if 3 <= e <= 8:
            return cols
    raise AssertionError("no suitable start state")


def _worst_z(counts, stay, law, samples):
    z = []
    for pair, prob in law.items():
        var = samples * prob * (1 - prob)
        z.append(abs(counts.get(pair, 0) - samples * prob) / math.sqrt(var))
    p_stay = 1 - sum(law.values())
    z.append(abs(stay - samples * p_stay) / math.sqrt(samples * p_stay * (1 - p_stay)))
    assert set(counts) <= set(law), "a move outside the valid set was made"
    return max(z)


@pytest.fixture
def clean_oracle(oracle):
    oracle.set_learned_variant(0)
    yield oracle
    oracle.set_learned_variant(0)


def test_learned_step_follows_the_hastings_law(clean_oracle):
    oracle = clean_oracle
    wt = _table()
    cols = _start_state(oracle)
    e, law = _law(wt, cols)
    assert e > 0 and len(law) >= 20
    sweeps = 10
    counts, stay = _tally(oracle, wt, cols, sweeps)
    samples = REPLICAS * sweeps
    assert _worst_z(counts, stay, law, samples) < 5.0
    # the law is far from the plain uniform-proposal Metropolis law, so the
    # check above does discriminate
    uniform = {k: 2 / M ** 2 * min(1.0, math.exp(-(_recount(_switched(cols, *k))[0] - e) / T)) for k in law}
    assert max(abs(law[k] - uniform[k]) / uniform[k] for k in law) > 0.5


def test_learned_walk_is_in_detailed_balance(clean_oracle):
    """pi(x) f(x -> x') = pi(x') f(x' -> x) on measured frequencies, for the
    most frequent moves out of x whose target has E > 0."""
    oracle = clean_oracle
    wt = _table()
    cols = _start_state(oracle)
    e, law = _law(wt, cols)
    sweeps = 6
    samples = REPLICAS * sweeps
    counts, _ = _tally(oracle, wt, cols, sweeps)
    checked = 0
    for (u, v), _ in sorted(counts.items(), key=lambda kv: -kv[1]):
        x2 = _switched(cols, u, v)
        e2, _ = _recount(x2)
        if e2 == 0:
            continue
        back, _ = _tally(oracle, wt, x2, sweeps, seed=7)
        f_fwd, f_back = counts[(u, v)] / samples, back.get((u, v), 0) / samples
        assert f_back > 0
        ratio = (math.exp(-e / T) * f_fwd) / (math.exp(-e2 / T) * f_back)
        sigma = math.sqrt(1 / counts[(u, v)] + 1 / back[(u, v)])
        assert abs(math.log(ratio)) < 5 * sigma, ((u, v), e, e2, ratio, sigma)
        checked += 1
        if checked == 4:
            break
    assert checked == 4


@pytest.mark.parametrize("variant", [1, 2, 3])
def test_planted_mistakes_break_the_law(clean_oracle, variant):
    oracle = clean_oracle
    wt = _table()
    cols = _start_state(oracle)
    _, law = _law(wt, cols)
    oracle.set_learned_variant(variant)
    sweeps = 4
    counts, stay = _tally(oracle, wt, cols, sweeps)
    assert _worst_z(counts, stay, law, REPLICAS * sweeps) > 10.0, variant


# ---- point-network proposals: delayed (two-stage) acceptance ----------------

LT_Q = int(round(32 * math.log(T)))


def _point_net(pkg, seed=4):
    """A random point network whose first 32 hidden units set the energy code
    (E/64) against the count through the ReLU, so the weight profile over the
    points changes with the energy (a uniform rescaling would cancel in
    w(u) w(v) / S1^2)."""
    rng = np.random.default_rng(seed)
    w1 = rng.normal(0, 0.6, (128, 8))
    b1 = rng.normal(0, 0.3, 128)
    w2 = rng.normal(0, 0.25, 128)
    w1[:32, :] = 0
    w1[:32, 7] = 1.9
    w1[:32, 0] = -1.9 * rng.uniform(0.5, 1.5, 32)
    b1[:32] = rng.uniform(-0.3, 0.0, 32)
    w2[:32] = 3.0 * rng.choice([-1, 1], 32)
    master = np.concatenate([w1.reshape(-1), b1, w2, [0.1]]).astype(np.float32)
    return pkg.learned.network_bf16(master), pkg.learned.point_exp_table()


def _point_w(oracle, params, expw, cols):
    e, bad = _recount(cols)
    w = oracle.point_weights(cols[None, :], bad[None, :].astype(np.int32), np.array([e], np.int32),
                             np.array([LT_Q], np.int32), params, expw)[0]
    return e, w.astype(np.float64)


def _law_point(oracle, params, expw, cols):
    """P(x -> x') = q min(1, exp(-dE/T)) min(1, q'/q), q from the weights of
    x and q' from those of x' (the oracle's point weights, taken as given)."""
    e, w = _point_w(oracle, params, expw, cols)
    s1 = w.sum()
    law = {}
    for u, v in combinations(range(M), 2):
        cu, cv = cols[u], cols[v]
        if e == 0 or u >> 1 == v >> 1 or cu == cv or cols[u ^ 1] == cv or cols[v ^ 1] == cu:
            continue
        x2 = _switched(cols, u, v)
        e2, w2 = _point_w(oracle, params, expw, x2)
        q = 0.4 / M ** 2 + 1.6 * w[u] * w[v] / s1 ** 2
        q_back = 0.4 / M ** 2 + 1.6 * w2[u] * w2[v] / w2.sum() ** 2
        law[(u, v)] = q * min(1.0, math.exp(-(e2 - e) / T)) * min(1.0, q_back / q)
    return e, law


def _tally_point(oracle, params, expw, cols, sweeps, seed=5):
    rows = np.tile(cols, (REPLICAS, 1)).astype(np.int32)
    replay = oracle.TemperingReplay(rows[:2], np.full(2, T, np.float32), seed, k_move=64, k_exchange=64)
    e, bad = _recount(cols)
    thr64 = oracle.metropolis_thresholds64(np.full(REPLICAS, T, np.float32), K64)
    lt_q = np.full(REPLICAS, LT_Q, np.int32)
    counts, stay = {}, 0
    for sweep in range(sweeps):
        replay.columns = rows.copy()
        replay.badness = np.tile(bad.astype(np.int32), (REPLICAS, 1))
        replay.scores = np.full(REPLICAS, e, np.int32)
        replay.best_columns = rows.copy()
        replay.best_scores = replay.scores.copy()
        replay.accepted = np.zeros(REPLICAS, np.int32)
        replay.removed_points = np.zeros(REPLICAS, np.int32)
        replay.replicas = REPLICAS
        replay.sweep = 2000 + sweep
        replay.advance_point(1, params, expw, lt_q, thr64)
        changed = replay.columns != rows
        moved = changed.sum(axis=1)
        assert set(np.unique(moved)) <= {0, 2}
        stay += int((moved == 0).sum())
        idx = np.nonzero(moved == 2)[0]
        pairs = np.sort(np.nonzero(changed[idx])[1].reshape(-1, 2), axis=1)
        keys, cnt = np.unique(pairs, axis=0, return_counts=True)
        for (u, v), c in zip(keys, cnt):
            counts[(int(u), int(v))] = counts.get((int(u), int(v)), 0) + int(c)
    return counts, stay


def test_point_weights_match_a_float64_forward(pkg, oracle):
    """The oracle's point weights (exact hidden sums, fixed-order fp32 output
    chain) against a float64 forward of the same bf16 network on the same
    feature codes: equal wherever 64 g is not within 1e-3 of an integer, one
    table step apart otherwise."""
    params, expw = _point_net(pkg, seed=8)
    rng = np.random.default_rng(3)
    assert np.all(params[:1152] == pkg.learned.network_bf16(pkg.learned.from_bf16(params))[:1152])
    n, count = 40, 300
    p, q = oracle.generate_states(17, n, count)
    cols = np.stack((p, q), axis=2).reshape(count, 2 * n).astype(np.int32)
    bad = rng.integers(0, 40, (count, 2 * n)).astype(np.int32)
    energy = rng.integers(0, 700, count).astype(np.int32)
    lt_q = rng.integers(-96, 23, count).astype(np.int32)
    got = oracle.point_weights(cols, bad, energy, lt_q, params, expw)
    rows = np.broadcast_to(np.arange(2 * n) >> 1, cols.shape)
    codes = pkg.learned.point_codes(bad, rows, cols, lt_q[:, None], energy[:, None]).reshape(-1, 8)
    w1, b1, w2, b2 = oracle.mlp_unpack(params)
    h = np.maximum(codes / 128.0 @ w1.T + b1, 0.0)
    g = np.clip(h @ w2 + b2, -7, 6)
    q64 = np.clip(np.floor(64 * g), -384, 352).astype(np.int64)
    ref = expw[q64 + 384].reshape(got.shape)
    inside = ((g > -6) & (g < 5.5)).reshape(got.shape)
    near = inside & (np.abs(64 * g - np.rint(64 * g)).reshape(got.shape) < 1e-3)
    assert np.array_equal(got[~near], ref[~near])
    assert near.mean() < 0.01 and inside.mean() > 0.2
    step = np.abs(np.log(got[near].astype(float)) - np.log(ref[near].astype(float)))
    assert np.all(step <= 1 / 64 + 1e-2)
    assert len(np.unique(got)) > 50  # the weights do vary


def test_grid_makes_fp32_hidden_sums_order_free(pkg):
    """With W1, b1 and the features on the 1/128 grid, every fp32 partial
    sum of a hidden pre-activation is exact: summing in any order gives the
    float64 value."""
    params, _ = _point_net(pkg, seed=2)
    w1 = pkg.learned.from_bf16(params[:1024]).reshape(128, 8)
    b1 = pkg.learned.from_bf16(params[1024:1152])
    rng = np.random.default_rng(5)
    codes = pkg.learned.point_codes(rng.integers(0, 63, 500), rng.integers(0, 100, 500), rng.integers(0, 100, 500),
                                    rng.integers(-255, 255, 500), rng.integers(0, 2000, 500))
    x = (codes / 128.0).astype(np.float32)
    exact = x.astype(np.float64) @ w1.astype(np.float64).T + b1
    for order in (range(8), range(7, -1, -1), rng.permutation(8)):
        acc = np.broadcast_to(b1, exact.shape).astype(np.float32).copy()
        for k in order:
            acc = (acc + x[:, k:k + 1] * w1[None, :, k]).astype(np.float32)
        assert np.array_equal(acc.astype(np.float64), exact)


def test_point_step_follows_the_delayed_hastings_law(pkg, clean_oracle):
    oracle = clean_oracle
    params, expw = _point_net(pkg)
    cols = _start_state(oracle)
    e, law = _law_point(oracle, params, expw, cols)
    assert e > 0 and len(law) >= 20
    sweeps = 10
    counts, stay = _tally_point(oracle, params, expw, cols, sweeps)
    assert _worst_z(counts, stay, law, REPLICAS * sweeps) < 5.0
    # the second stage matters: its factor min(1, q'/q) is far from 1 somewhere
    plain = {k: law[k] for k in law}
    _, w = _point_w(oracle, params, expw, cols)
    ratios = []
    for (u, v) in law:
        e2, w2 = _point_w(oracle, params, expw, _switched(cols, u, v))
        q = 0.4 / M ** 2 + 1.6 * w[u] * w[v] / w.sum() ** 2
        ratios.append((0.4 / M ** 2 + 1.6 * w2[u] * w2[v] / w2.sum() ** 2) / q)
    assert min(ratios) < 0.7 and plain


def test_point_walk_is_in_detailed_balance(pkg, clean_oracle):
    oracle = clean_oracle
    params, expw = _point_net(pkg)
    cols = _start_state(oracle)
    e, _ = _law_point(oracle, params, expw, cols)
    sweeps = 6
    samples = REPLICAS * sweeps
    counts, _ = _tally_point(oracle, params, expw, cols, sweeps)
    checked = 0
    for (u, v), _ in sorted(counts.items(), key=lambda kv: -kv[1]):
        x2 = _switched(cols, u, v)
        e2, _ = _recount(x2)
        if e2 == 0:
            continue
        back, _ = _tally_point(oracle, params, expw, x2, sweeps, seed=7)
        f_fwd, f_back = counts[(u, v)] / samples, back.get((u, v), 0) / samples
        assert f_back > 0
        ratio = (math.exp(-e / T) * f_fwd) / (math.exp(-e2 / T) * f_back)
        sigma = math.sqrt(1 / counts[(u, v)] + 1 / back[(u, v)])
        assert abs(math.log(ratio)) < 5 * sigma, ((u, v), e, e2, ratio, sigma)
        checked += 1
        if checked == 4:
            break
    assert checked == 4


@pytest.mark.parametrize("variant", [1, 2, 3, 4])
def test_point_planted_mistakes_break_the_law(pkg, clean_oracle, variant):
    oracle = clean_oracle
    params, expw = _point_net(pkg)
    cols = _start_state(oracle)
    _, law = _law_point(oracle, params, expw, cols)
    oracle.set_learned_variant(variant)
    counts, stay = _tally_point(oracle, params, expw, cols, 4)
    assert _worst_z(counts, stay, law, REPLICAS * 4) > 10.0, variant
