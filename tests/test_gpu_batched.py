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
