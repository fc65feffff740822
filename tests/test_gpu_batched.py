"""Batched multi-size tempering against per-size runs (bit-exact) and the
launch-shape tuner."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

FIELDS = ("columns", "badness", "scores", "best_columns", "best_scores", "accepted", "removed_points", "counters",
          "proposals")


def test_batched_equals_separate_runs(pkg, dev):
    spec = [(10, 64, 16), (21, 32, 16), (33, 48, 16), (75, 64, 64), (100, 32, 32)]
    batched, separate = [], []
    for i, (n, replicas, ladder) in enumerate(spec):
        a = pkg.seeded_chains(n, replicas, ladder, seed=100 + i)
        b = pkg.seeded_chains(n, replicas, ladder, seed=100 + i)
        b.block_waves = [1, 2, 4, 6, 0][i]
        batched.append(pkg.BatchGroup(a, ladder, 2 * n))
        separate.append((b, ladder, 2 * n))
    pkg.run_batched(batched, 21)
    for b, ladder, steps in separate:
        b.run(21, steps, ladder)
    torch.cuda.synchronize()
    for g, (b, _, _) in zip(batched, separate):
        for name in FIELDS:
            assert torch.equal(getattr(g.chains, name), getattr(b, name)), (g.chains.n, name)
        g.chains.validate()
    # runs continue from the device counters
    pkg.run_batched(batched, 5)
    for b, ladder, steps in separate:
        b.run(5, steps, ladder)
    torch.cuda.synchronize()
    for g, (b, _, _) in zip(batched, separate):
        assert torch.equal(g.chains.columns, b.columns) and torch.equal(g.chains.scores, b.scores)


CONFIG4 = {40: 8192, 52: 4096, 61: 4096, 75: 2048, 83: 2048, 100: 1024}


def _columns(chains_or_replay):
    c = chains_or_replay.columns
    return c.cpu().numpy() if hasattr(c, "cpu") else c


def test_config4_sizes_match_the_replay(pkg, oracle, dev):
    """Config 4's six sizes in ONE batched run (64 replicas each, ladders of
    16 temperatures 0.05..2.0, 2n proposals per sweep, exchanges every
    sweep): every group's replicas bit-exact against the CPU replay of the
    reference walk after every run of sweeps."""
    groups, replays = [], []
    for i, n in enumerate(CONFIG4):
        replicas, ladder = 64, 16
        chains = pkg.seeded_chains(n, replicas, ladder, seed=500 + i)
        cols = chains.columns.cpu().numpy()
        temps = chains.temperatures_host
        replay = oracle.TemperingReplay(cols, temps, chains.seed, k_move=chains.k_move,
                                        k_exchange=chains._exchange_table(ladder)[2])
        groups.append(pkg.BatchGroup(chains, ladder, 2 * n))
        replays.append((replay, ladder, 2 * n))
    for sweeps in (2, 1):
        print(f"config-4 sizes vs replay: {sweeps} sweep(s)", flush=True)
        pkg.run_batched(groups, sweeps)
        torch.cuda.synchronize()
        for g, (replay, ladder, steps) in zip(groups, replays):
            for _ in range(sweeps):
                replay.advance(steps, threads=oracle.cpu_count())
                replay.exchange(ladder)
            for name in ("columns", "badness", "scores", "best_columns", "best_scores", "accepted",
                         "removed_points"):
                got = getattr(g.chains, name).cpu().numpy()
                assert np.array_equal(got, getattr(replay, name)), (g.chains.n, name)


def test_config4_full_counts(pkg, dev):
    """Config 4 at its full replica counts (21,504 replicas over six sizes,
    ladders of 64, 2n proposals per sweep, tuned-by-hand shapes): a few
    batched sweeps, then a determinant recount of a sample of every group."""
    shapes = {40: 16, 52: 16, 61: 16, 75: 8, 83: 8, 100: 4}
    groups = []
    for i, (n, replicas) in enumerate(CONFIG4.items()):
        chains = pkg.seeded_chains(n, replicas, 64, seed=900 + i)
        chains.block_waves = shapes[n]
        groups.append(pkg.BatchGroup(chains, 64, 2 * n))
    start = [g.chains.scores.clone() for g in groups]
    pkg.run_batched(groups, 6)
    torch.cuda.synchronize()
    rng = np.random.default_rng(0)
    for g, s0 in zip(groups, start):
        chains = g.chains
        total = chains.columns.shape[0]
        chains.validate(rng.choice(total, size=min(total, 256), replace=False))
        assert int(chains.counters[0]) == 6 and int(chains.counters[1]) == 6
        assert int(chains.proposals[0]) > 0 and int(chains.accepted.sum()) > 0
        assert float(chains.scores.float().mean()) < float(s0.float().mean())


def test_batched_argument_errors(pkg, dev):
    a = pkg.seeded_chains(12, 32, 16, seed=1)
    with pytest.raises(ValueError):
        pkg.run_batched([pkg.BatchGroup(a, 5, 24)], 1)
    with pytest.raises(ValueError):
        pkg.run_batched([], 1)
    a.block_waves = 64
    with pytest.raises(ValueError):
        a.run(1, 24, 16)


def test_autotune_picks_a_fitting_shape(pkg, dev, tmp_path):
    result = pkg.autotune.tune_size(40, 256, sweeps=1, candidates=(1, 4, 16))
    assert result["block_waves"] in (1, 4, 16)
    table = pkg.autotune.tune({30: 128}, cache=tmp_path / "t.json", sweeps=1, candidates=(2, 8))
    assert table[30] in (2, 8) and (tmp_path / "t.json").exists()


def test_mixed_autotune_tunes_the_batched_run(pkg, dev, tmp_path):
    """The mixed tuner starts from the most waves that fit in LDS for every
    n and returns, per n, a candidate that fits; its choice is cached per size
    mix, and a batched run with it matches the replay-checked form."""
    sizes = {40: 512, 100: 256}
    res = pkg.autotune.tune_mixed_shapes(sizes, sweeps=1, candidates=(4, 8, 16))
    shapes = {int(n): w for n, w in res["block_waves"].items()}
    assert shapes[40] in (4, 8, 16) and shapes[100] in (4,) and res["trials"] >= 2
    table = pkg.autotune.tune_mixed(sizes, cache=tmp_path / "m.json", sweeps=1, candidates=(4, 8, 16))
    assert set(table) == {40, 100} and all(w in (4, 8, 16) for w in table.values())
    assert (tmp_path / "m.json").exists()
    groups = []
    for i, (n, r) in enumerate(sizes.items()):
        chains = pkg.seeded_chains(n, r, 64, seed=70 + i)
        chains.block_waves = table[n]
        groups.append(pkg.BatchGroup(chains, 64, 2 * n))
    pkg.run_batched(groups, 3, persistent=True)
    torch.cuda.synchronize()
    for g in groups:
        g.chains.validate(range(0, len(g.chains.temperatures_host), 37))


def test_persistent_launch_equals_per_sweep_launches(pkg, dev):
    """All sweeps of config 4's six sizes in one cooperative launch give the
    per-sweep launches' trajectories and counters bit for bit; a group
    without exchanges and a badness-biased group ride along, and runs
    continue from the device counters in either form."""
    spec = [(n, 64, 16, 0) for n in CONFIG4] + [(20, 48, 0, 0), (30, 64, 16, 40)]
    persistent, stepped = [], []
    for i, (n, replicas, ladder, bias) in enumerate(spec):
        for out in (persistent, stepped):
            chains = pkg.seeded_chains(n, replicas, max(ladder, 16), seed=300 + i)
            chains.block_waves = [4, 8, 16][i % 3] if n <= 75 else 4
            out.append(pkg.BatchGroup(chains, ladder, 2 * n, bias))
    for sweeps in (7, 3):
        print(f"persistent vs per-sweep launches: {sweeps} sweeps", flush=True)
        pkg.run_batched(persistent, sweeps, persistent=True)
        pkg.run_batched(stepped, sweeps)
        torch.cuda.synchronize()
        for a, b in zip(persistent, stepped):
            for name in FIELDS:
                assert torch.equal(getattr(a.chains, name), getattr(b.chains, name)), (a.chains.n, name)
    for g in persistent:
        g.chains.validate()
    # mixed: per-sweep launches continue a persistent run and vice versa
    pkg.run_batched(persistent, 2)
    pkg.run_batched(stepped, 2, persistent=True)
    torch.cuda.synchronize()
    for a, b in zip(persistent, stepped):
        for name in FIELDS:
            assert torch.equal(getattr(a.chains, name), getattr(b.chains, name)), (a.chains.n, name)
