"""Point-network proposals on the GPU: the in-walk MFMA weights and the
two-stage walk against the oracle (bit-exact), and training on the
proposal log with the point features."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _network(pkg, dev, seed):
    """A random point network with energy-count coupled hidden units (as in
    the CPU law tests), bf16 on the grid."""
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
    net = pkg.learned.ProposalNetwork(seed=seed, device=dev)
    net.master.copy_(torch.from_numpy(master))
    net.params.copy_(torch.from_numpy(pkg.learned.network_bf16(master).view(np.int16)))
    return net


def _cols(oracle, n, count, seed):
    p, q = oracle.generate_states(seed, n, count)
    return np.stack((p, q), axis=2).reshape(count, 2 * n).astype(np.int32)


@pytest.mark.parametrize("n", [6, 12, 33, 40, 64, 75, 100])
def test_point_weights_match_oracle(pkg, oracle, dev, n):
    net = _network(pkg, dev, seed=n)
    rng = np.random.default_rng(n)
    count = 96
    cols = _cols(oracle, n, count, seed=n + 3)
    bad = rng.integers(0, 45, (count, 2 * n)).astype(np.int32)
    scores = rng.integers(0, 900, count).astype(np.int32)
    lt_q = rng.integers(-100, 30, count).astype(np.int32)
    chains = pkg.TemperingChains(cols[:1], np.ones(1, np.float32), 1)
    props = pkg.learned.prepare_point(chains, net)
    print(f"point weights: n={n} states={count}", flush=True)
    got = pkg.learned.point_weights_device(props, torch.from_numpy(cols).to(dev), torch.from_numpy(bad).to(dev),
                                           torch.from_numpy(scores).to(dev), torch.from_numpy(lt_q).to(dev))
    ref = oracle.point_weights(cols, bad, scores, lt_q, net.params.cpu().numpy().view(np.uint16),
                               pkg.learned.point_exp_table())
    assert np.array_equal(got.cpu().numpy().view(np.uint32), ref)
    assert len(np.unique(ref)) > 10


@pytest.mark.parametrize("n,replicas,ladder,sweeps", [(12, 32, 16, 6), (33, 64, 32, 4), (75, 64, 64, 3),
                                                       (100, 32, 32, 2)])
def test_point_walk_matches_replay(pkg, oracle, dev, n, replicas, ladder, sweeps):
    net = _network(pkg, dev, seed=100 + n)
    cols = _cols(oracle, n, replicas, seed=n + 1)
    temps = np.tile(np.geomspace(0.05, 2.0, ladder), replicas // ladder).astype(np.float32)
    chains = pkg.TemperingChains(cols, temps, seed=91)
    props = pkg.learned.prepare_point(chains, net)
    replay = oracle.TemperingReplay(cols, temps, 91, k_move=chains.k_move, k_exchange=chains._exchange_table(ladder)[2])
    params = net.params.cpu().numpy().view(np.uint16)
    expw = pkg.learned.point_exp_table()
    lt_q = props.lt_q.cpu().numpy()
    thr64 = props.thr64.cpu().numpy().view(np.uint64)
    names = ("columns", "badness", "scores", "best_columns", "best_scores", "accepted", "removed_points")
    for sweep in range(sweeps):
        print(f"point walk vs replay: n={n} sweep {sweep}", flush=True)
        pkg.learned.run_learned_point(chains, props, 1, 2 * n, ladder)
        replay.advance_point(2 * n, params, expw, lt_q, thr64, threads=oracle.cpu_count())
        replay.exchange(ladder)
        torch.cuda.synchronize()
        for name in names:
            assert np.array_equal(getattr(chains, name).cpu().numpy(), getattr(replay, name)), (sweep, name)
    assert int(chains.accepted.sum()) > 0
    chains.validate()
    # a graph-captured run of the same sweeps gives the same trajectories
    again = pkg.TemperingChains(cols, temps, seed=91)
    pkg.learned.run_learned_point(again, pkg.learned.prepare_point(again, net), sweeps, 2 * n, ladder)
    torch.cuda.synchronize()
    for name in names:
        assert torch.equal(getattr(again, name), getattr(chains, name)), name


def test_point_training_on_logged_proposals(pkg, dev):
    chains = pkg.seeded_chains(40, 512, 64, seed=5)
    chains.run(5, 80, 64)
    log = pkg.learned.ProposalLog(2 * 80 * 512, dev)
    chains.log = log
    chains.run(2, 80, 64)
    chains.log = None
    torch.cuda.synchronize()
    r = log.rows().cpu().numpy()
    u, v, cu, cv = r[:, 6] & 0xFFFF, r[:, 6] >> 16, r[:, 7] & 0xFFFF, r[:, 7] >> 16
    assert len(r) > 10000
    assert np.all((u < 80) & (v < 80) & ((u >> 1) != (v >> 1)) & (cu < 40) & (cv < 40) & (cu != cv))
    xu, xv, y = pkg.learned.training_tensors(log.rows(), chains.temperatures_host, dev, features="point")
    # the device codes are the host formula's, exactly
    lt_q = pkg.learned.temperature_codes(chains.temperatures_host)[r[:, 3]]
    host = pkg.learned.point_codes(r[:, 0], u >> 1, cu, lt_q, r[:, 2]) / 128.0
    assert np.array_equal(pkg.learned.from_bf16(xu.cpu().numpy().view(np.uint16)).reshape(-1, 8), host)
    net = pkg.learned.ProposalNetwork(seed=2, device=dev)
    losses = net.train(xu, xv, y, steps=200)
    assert losses[-1] < losses[0]
    props = pkg.learned.prepare_point(chains, net)
    pkg.learned.run_learned_point(chains, props, 10, 80, 64)
    torch.cuda.synchronize()
    chains.validate()
    # online retraining rewrites the constants in place
    before = props.consts.clone()
    net.train(xu, xv, y, steps=50, seed=3)
    pkg.learned.refresh_point(props, net)
    assert not torch.equal(before, props.consts)
    pkg.learned.run_learned_point(chains, props, 10, 80, 64)
    torch.cuda.synchronize()
    chains.validate()


def test_trained_point_network_reaches_low_energy_in_fewer_sweeps(pkg, dev):
    """What the trained point network gains over uniform proposals, counted
    in sweeps: from the reference's E=10 near-solution with every replica
    perturbed by 12 random 2-switches (4,096 replicas, 64 ladders of 64
    temperatures 0.05..2.0), the point walk, pretrained on another
    perturbation's plain logs, reaches E <= 12 within fewer sweeps than the
    uniform walk from the same states and seed."""
    from pathlib import Path

    fixture = Path(__file__).parent / "golden" / "best-near-solution-E10.json"
    _, points, _ = pkg.io.load_candidate(fixture, 75)
    state = pkg.points_to_permutations(points, 75)
    base = np.repeat(np.stack((np.array(state.P), np.array(state.Q)), 1).reshape(1, 150).astype(np.int32), 4096, 0)
    net, _ = pkg.learned.pretrain(75, 4096, 64, seed=11, start=pkg.random_two_switches(base, 12, seed=99),
                                  features="point")
    start = pkg.random_two_switches(base, 12, seed=12)
    temps = np.tile(np.geomspace(0.05, 2.0, 64), 64).astype(np.float32)

    def sweeps_to(run, chains, limit=400, chunk=10):
        done = 0
        while done < limit:
            run(chains, chunk)
            done += chunk
            if int(chains.best_scores.min().item()) <= 12:
                return done
        return None

    uniform = pkg.TemperingChains(start, temps, 5)
    learned = pkg.TemperingChains(start, temps, 5)
    props = pkg.learned.prepare_point(learned, net)
    s_uniform = sweeps_to(lambda c, k: c.run(k, 150, 64), uniform)
    s_learned = sweeps_to(lambda c, k: pkg.learned.run_learned_point(c, props, k, 150, 64), learned)
    print(f"sweeps to E<=12: point {s_learned}, uniform {s_uniform}", flush=True)
    assert s_learned is not None
    assert s_uniform is None or s_learned < s_uniform, (s_learned, s_uniform)
    learned.validate(range(0, 4096, 257))


def test_block_mapping_leaves_trajectories_unchanged(pkg, dev):
    """As for the table mode: 2,048 replicas run 8 per block, and the walk
    with the ladder (each block's replicas spread over a ladder) equals the
    walk without it plus separate exchanges."""
    net = _network(pkg, dev, seed=4)
    a = pkg.seeded_chains(40, 2048, 64, seed=9)
    b = pkg.seeded_chains(40, 2048, 64, seed=9)
    pa, pb = pkg.learned.prepare_point(a, net), pkg.learned.prepare_point(b, net)
    pkg.learned.run_learned_point(a, pa, 3, 80, 64)
    for _ in range(3):
        pkg.learned.run_learned_point(b, pb, 1, 80)
        b.exchange(64)
    torch.cuda.synchronize()
    for name in ("columns", "badness", "scores", "best_columns", "best_scores", "accepted", "removed_points",
                 "counters", "proposals"):
        assert torch.equal(getattr(a, name), getattr(b, name)), name
    assert b.exchanges_total() > 0
