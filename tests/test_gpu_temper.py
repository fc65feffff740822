"""Tempering trajectories on the GPU against the CPU replay (bit-exact)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _states(oracle, n, replicas, seed):
    p, q = oracle.generate_states(seed, n, replicas)
    return np.stack([np.column_stack((p[i], q[i])).reshape(-1) for i in range(replicas)]).astype(np.int32)


def _compare(chains, replay):
    for name in ("columns", "badness", "scores", "best_columns", "best_scores", "accepted", "removed_points"):
        got = getattr(chains, name).cpu().numpy()
        assert np.array_equal(got, getattr(replay, name)), name


@pytest.mark.parametrize("n,replicas,ladder,sweeps,steps,bias", [
    (8, 16, 8, 12, 16, 0),
    (12, 32, 16, 10, 24, 30),
    (33, 64, 32, 6, 66, 0),
    (75, 64, 64, 4, 150, 0),
    (75, 64, 16, 3, 150, 50),
])
def test_trajectory_matches_replay(pkg, oracle, dev, n, replicas, ladder, sweeps, steps, bias):
    cols = _states(oracle, n, replicas, seed=n)
    temps = np.tile(np.geomspace(0.05, 2.0, ladder), replicas // ladder).astype(np.float32)
    seed = 1234 + n
    chains = pkg.TemperingChains(cols, temps, seed)
    replay = oracle.TemperingReplay(cols, temps, seed, k_move=chains.k_move,
                                    k_exchange=chains._exchange_table(ladder)[2])
    assert np.array_equal(chains.move_thresholds.cpu().numpy(), replay.thr)
    for _ in range(sweeps):
        chains.advance(steps, bias)
        replay.advance(steps, bias)
        _compare(chains, replay)
        assert chains.exchange(ladder) == replay.exchange(ladder)
        _compare(chains, replay)
    chains.validate()


def test_graph_run_matches_step_by_step(pkg, oracle, dev):
    n, replicas, ladder = 20, 64, 16
    cols = _states(oracle, n, replicas, seed=5)
    temps = np.tile(np.geomspace(0.05, 2.0, ladder), replicas // ladder).astype(np.float32)
    a = pkg.TemperingChains(cols, temps, 99)
    b = pkg.TemperingChains(cols, temps, 99)
    for _ in range(45):
        a.advance(40)
        a.exchange(ladder)
    b.run(45, 40, ladder)
    torch.cuda.synchronize()
    for name in ("columns", "badness", "scores", "best_columns", "best_scores", "accepted", "removed_points",
                 "counters", "proposals"):
        assert torch.equal(getattr(a, name), getattr(b, name)), name
    b.validate()
    # a second run continues from the device counters (graph replayed)
    a2 = pkg.TemperingChains(cols, temps, 99)
    a2.run(45, 40, ladder)
    a2.run(45, 40, ladder)
    b.run(45, 40, ladder)
    torch.cuda.synchronize()
    assert torch.equal(a2.columns, b.columns) and torch.equal(a2.scores, b.scores)


def test_long_run_validates_and_improves(pkg, oracle, dev):
    n, replicas, ladder = 75, 256, 64
    cols = _states(oracle, n, replicas, seed=2)
    temps = np.tile(np.geomspace(0.05, 2.0, ladder), replicas // ladder).astype(np.float32)
    chains = pkg.TemperingChains(cols, temps, 7)
    start = int(chains.scores.min().item())
    chains.run(200, 150, ladder)
    torch.cuda.synchronize()
    chains.validate()
    assert int(chains.best_scores.min().item()) < start
    valid, accepted = chains.proposals.cpu().numpy()
    assert 0 < accepted < valid <= 200 * 150 * replicas


def test_save_restore(pkg, oracle, dev, tmp_path):
    n, replicas, ladder = 10, 16, 8
    cols = _states(oracle, n, replicas, seed=3)
    temps = np.tile(np.geomspace(0.1, 3.0, ladder), 2).astype(np.float32)
    a = pkg.TemperingChains(cols, temps, 5)
    a.run(7, 20, ladder)
    a.save(tmp_path / "c.npz")
    b = pkg.TemperingChains.restore(tmp_path / "c.npz")
    a.run(5, 20, ladder)
    b.run(5, 20, ladder)
    torch.cuda.synchronize()
    assert torch.equal(a.columns, b.columns) and torch.equal(a.best_scores, b.best_scores)


def test_corrupted_state_is_detected(pkg, oracle, dev):
    cols = _states(oracle, 10, 8, seed=4)
    chains = pkg.TemperingChains(cols, np.full(8, 0.5, np.float32), 1)
    chains.scores[3] += 1
    with pytest.raises(AssertionError):
        chains.validate()
