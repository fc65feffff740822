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
    # three chunks of slots past n=64, four chunks from n=97 on
    (83, 64, 32, 3, 166, 0),
    (97, 32, 16, 2, 194, 0),
    (100, 32, 32, 2, 200, 30),
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


@pytest.mark.parametrize("n,replicas,ladder,sweeps,steps,bias", [
    (12, 32, 8, 8, 24, 0),
    (33, 64, 16, 5, 66, 30),
    (75, 64, 16, 3, 150, 0),
])
def test_umbrella_trajectory_matches_replay(pkg, oracle, dev, n, replicas, ladder, sweeps, steps, bias):
    """tempering.py:92-95 with the driver's umbrella set-up (one center per
    ladder, linspace(0, 20), strength 0.2): every sweep bit-exact against the
    replay, by advance() and by a graph run without exchanges."""
    cols = _states(oracle, n, replicas, seed=n + 1)
    temps = np.tile(np.geomspace(0.05, 2.0, ladder), replicas // ladder).astype(np.float32)
    centers = np.repeat(np.linspace(0, 20, replicas // ladder), ladder).astype(np.float32)
    seed = 777 + n
    chains = pkg.TemperingChains(cols, temps, seed)
    chains.centers = torch.from_numpy(centers).to(dev)
    chains.strengths = torch.full((replicas,), 0.2, dtype=torch.float32, device=dev)
    replay = oracle.TemperingReplay(cols, temps, seed, k_move=chains.k_move, k_exchange=2)
    replay.centers[:] = centers
    replay.strengths[:] = np.float32(0.2)
    for _ in range(sweeps):
        chains.advance(steps, bias)
        replay.advance(steps, bias)
        _compare(chains, replay)
    graph = pkg.TemperingChains(cols, temps, seed)
    graph.centers = torch.from_numpy(centers).to(dev)
    graph.strengths = torch.full((replicas,), 0.2, dtype=torch.float32, device=dev)
    graph.run(sweeps, steps, 0, bias)
    torch.cuda.synchronize()
    for name in ("columns", "badness", "scores", "best_columns", "best_scores", "accepted", "removed_points"):
        assert torch.equal(getattr(graph, name), getattr(chains, name)), name
    chains.validate()
    with pytest.raises(ValueError):
        chains.exchange(ladder)
    with pytest.raises(ValueError):
        chains.run(1, steps, ladder)
    with pytest.raises(NotImplementedError):
        pkg.run_batched([pkg.BatchGroup(chains, ladder, steps)], 1)


def test_config3_full_size_matches_replay(pkg, oracle, dev):
    """Config 3 at full size (4,096 replicas as 64 ladders of 64
    temperatures, 0.05..2.0, n=75, 150 proposals per sweep): three sweeps
    with exchanges, every replica bit-exact against the replay."""
    n, replicas, ladder, steps = 75, 4096, 64, 150
    cols = _states(oracle, n, replicas, seed=31)
    temps = np.tile(np.geomspace(0.05, 2.0, ladder), replicas // ladder).astype(np.float32)
    chains = pkg.TemperingChains(cols, temps, 3)
    replay = oracle.TemperingReplay(cols, temps, 3, k_move=chains.k_move,
                                    k_exchange=chains._exchange_table(ladder)[2])
    for _ in range(3):
        chains.advance(steps)
        replay.advance(steps, threads=oracle.cpu_count())
        assert chains.exchange(ladder) == replay.exchange(ladder)
        _compare(chains, replay)


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


@pytest.mark.parametrize("n,replicas,ladder", [(75, 4096, 64), (40, 2048, 64)])
def test_block_mapping_leaves_trajectories_unchanged(pkg, dev, n, replicas, ladder):
    """A run with exchanges takes each block's replicas from temperatures
    spread over one ladder (4,096 replicas: 16 per block, every fourth
    temperature; 2,048: 8 per block, every eighth); advance() without the
    ladder keeps consecutive slots per block.  Both give the same
    trajectories: the slot, not the block, keys the random stream."""
    a = pkg.seeded_chains(n, replicas, ladder, seed=9)
    b = pkg.seeded_chains(n, replicas, ladder, seed=9)
    a.run(3, 2 * n, ladder)
    for _ in range(3):
        b.advance(2 * n)
        b.exchange(ladder)
    torch.cuda.synchronize()
    for name in ("columns", "badness", "scores", "best_columns", "best_scores", "accepted", "removed_points",
                 "counters", "proposals"):
        assert torch.equal(getattr(a, name), getattr(b, name)), name
    assert b.exchanges_total() > 0


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


REFERENCE_CHECKPOINT_KEYS = {"columns", "badness", "scores", "rngs", "best_columns", "best_scores", "accepted",
                             "reference", "removed_points", "centers", "strengths", "temperatures", "round",
                             "host_rng"}


def test_checkpoint_carries_the_reference_schema(pkg, dev, tmp_path):
    """tempering.py:206-225: a checkpoint of this engine holds every key the
    reference's restore reads, with its dtypes; the reference's restore steps
    (np.load without pickles, host_rng JSON into a PCG64 generator) succeed."""
    import json

    chains = pkg.seeded_chains(20, 64, 16, seed=4)
    chains.run(7, 40, 16)
    torch.cuda.synchronize()
    chains.save(tmp_path / "c.npz")
    with np.load(tmp_path / "c.npz", allow_pickle=False) as saved:
        assert REFERENCE_CHECKPOINT_KEYS <= set(saved.files)
        assert saved["columns"].dtype == np.int32 and saved["columns"].shape == (64, 40)
        assert saved["rngs"].dtype == np.uint64 and saved["rngs"].shape == (64,) and np.all(saved["rngs"] > 0)
        assert saved["centers"].dtype == np.float32 and saved["temperatures"].dtype == np.float32
        assert int(saved["round"]) == chains.round == 7
        generator = np.random.default_rng(0)
        generator.bit_generator.state = json.loads(str(saved["host_rng"]))
        generator.random()


def test_restore_reads_a_reference_checkpoint(pkg, oracle, dev, tmp_path):
    """A checkpoint in the reference's schema only (no seed or counters): the
    restored chains validate against a full recount, start their Philox
    counters at the saved round and run on."""
    import json

    n, replicas = 12, 32
    p, q = oracle.generate_states(3, n, replicas)
    columns = np.stack((p, q), axis=2).reshape(replicas, 2 * n).astype(np.int32)
    scores, badness = oracle.score_population(p, q)
    host = np.random.default_rng(99)
    arrays = {
        "columns": columns, "badness": badness.astype(np.int32), "scores": scores.astype(np.int32),
        "rngs": host.integers(1, 2**64 - 1, size=replicas, dtype=np.uint64),
        "best_columns": columns, "best_scores": scores.astype(np.int32),
        "accepted": np.zeros(replicas, np.int32), "reference": columns[0],
        "centers": np.zeros(replicas, np.float32), "strengths": np.zeros(replicas, np.float32),
        "temperatures": np.tile(np.geomspace(0.2, 4.0, 16), 2).astype(np.float32),
        "round": np.asarray(12), "host_rng": np.asarray(json.dumps(host.bit_generator.state)),
    }
    ref = set((i // 2, int(c)) for i, c in enumerate(columns[0]))
    arrays["removed_points"] = np.array([len(ref - set((i // 2, int(c)) for i, c in enumerate(row)))
                                         for row in columns], np.int32)
    np.savez_compressed(tmp_path / "ref.npz", **arrays)
    chains = pkg.TemperingChains.restore(tmp_path / "ref.npz")
    assert chains.round == 12 and chains.sweep == 12
    assert np.array_equal(chains.columns.cpu().numpy(), columns)
    chains.run(5, 24, 16)
    torch.cuda.synchronize()
    chains.validate()
    assert chains.round == 17


def test_driver_writes_the_reference_outputs(pkg, dev, tmp_path):
    """tempering.main (tempering.py:228-319): parameters.json, best-E*.json
    candidates, history.jsonl, checkpoint.npz and summary.json with the
    reference's fields; --resume continues from the checkpoint."""
    import json
    from random import Random

    state = pkg.random_state(14, Random(2))
    initial = tmp_path / "initial.json"
    pkg.io.write_candidate(initial, 14, state.points())
    out = tmp_path / "run"
    args = ["--initial", str(initial), "--output", str(out), "--seed", "5", "--replicas", "32", "--ladder-size", "8",
            "--rounds", "20", "--steps", "30"]
    assert pkg.tempering.main(args) == 0
    assert json.loads((out / "parameters.json").read_text())["seed"] == 5
    lines = [json.loads(x) for x in (out / "history.jsonl").read_text().splitlines()]
    assert [x["epoch"] for x in lines] == [10, 20]
    for x in lines:
        assert {"epoch", "proposals", "elapsed_seconds", "best", "scores", "distances", "exchanges",
                "accepted"} <= set(x)
        assert len(x["scores"]) == len(x["distances"]) == 32
    summary = json.loads((out / "summary.json").read_text())
    assert {"best", "proposals", "elapsed_seconds", "accepted", "distance_frontier"} <= set(summary)
    best = sorted(out.glob("best-E*.json"), key=lambda f: int(f.stem[6:]))[0]
    n, points, meta = pkg.io.load_candidate(best)
    assert n == 14 and meta["conflicts"] == summary["best"] == int(best.stem[6:])
    again = pkg.points_to_permutations(points, 14)
    assert pkg.score_population_gpu(np.array([again.P]), np.array([again.Q]))[0] == summary["best"]
    assert pkg.tempering.main(args[:-4] + ["--rounds", "10", "--steps", "30", "--resume",
                                           str(out / "checkpoint.npz")]) == 0
    # umbrella sampling (tempering.py:265-267, 296-303): no exchanges, a
    # frontier file per distance bin of 8
    umb = tmp_path / "umbrella"
    uargs = ["--initial", str(initial), "--output", str(umb), "--seed", "5", "--replicas", "32", "--ladder-size",
             "8", "--rounds", "20", "--steps", "30", "--umbrella-max", "12"]
    assert pkg.tempering.main(uargs) == 0
    lines = [json.loads(x) for x in (umb / "history.jsonl").read_text().splitlines()]
    assert all(x["exchanges"] == 0 for x in lines)
    frontier = json.loads((umb / "summary.json").read_text())["distance_frontier"]
    assert frontier and set(frontier) == {f.stem[len("frontier-bin"):].lstrip("0") or "0"
                                          for f in umb.glob("frontier-bin*.json")}
    for f in umb.glob("frontier-bin*.json"):
        n, points, meta = pkg.io.load_candidate(f)
        again = pkg.points_to_permutations(points, 14)
        assert pkg.score_population_gpu(np.array([again.P]), np.array([again.Q]))[0] == meta["conflicts"]
        assert meta["symmetric_difference"] // 8 == int(f.stem[len("frontier-bin"):])
    saved = np.load(umb / "checkpoint.npz")
    assert np.allclose(saved["strengths"], 0.2) and saved["centers"].max() == np.float32(12)


def test_long_line_states_match_replay(pkg, oracle, dev):
    """States with a 38-point line: P is the identity except rows 20 and 21
    (P[20] = 21, P[21] = 20), so the diagonal holds 38 points, and the
    2-switch of slots 40 and 42 puts both moved points back on it
    ((20, 20) and (21, 21)).  The moved-state histograms then count 38
    points in one direction, beyond what a 4-bit counter holds, and the walk
    must take its byte-counter path for those proposals; every replica
    stays bit-exact against the replay (the pair is drawn about once in
    3,200 proposals: about 6 times per sweep over 256 replicas of 80)."""
    n, replicas, ladder, sweeps = 40, 256, 16, 6
    p = np.arange(n)
    p[20], p[21] = 21, 20
    q = (np.arange(n) + 7) % n
    assert np.all(p != q) and q[20] != 20 and q[21] != 21
    row = np.column_stack((p, q)).reshape(-1).astype(np.int32)
    cols = np.tile(row, (replicas, 1))
    temps = np.tile(np.geomspace(0.05, 2.0, ladder), replicas // ladder).astype(np.float32)
    chains = pkg.TemperingChains(cols, temps, 2024)
    replay = oracle.TemperingReplay(cols, temps, 2024, k_move=chains.k_move,
                                    k_exchange=chains._exchange_table(ladder)[2])
    for sweep in range(sweeps):
        print(f"long-line walk vs replay: sweep {sweep}", flush=True)
        chains.advance(2 * n)
        replay.advance(2 * n, threads=oracle.cpu_count())
        _compare(chains, replay)
        assert chains.exchange(ladder) == replay.exchange(ladder)
    chains.validate()
