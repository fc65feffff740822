"""Proposal network kernels against the float64 oracle, and the learned walk
against its CPU replay (bit-exact)."""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _params(rng, scale=1.0):
    w1 = rng.normal(0, scale * math.sqrt(2 / 8), (128, 8))
    b1 = rng.normal(0, 0.1, 128)
    w2 = rng.normal(0, scale * math.sqrt(1 / 128), 128)
    return np.concatenate([w1.reshape(-1), b1, w2, [0.3]]).astype(np.float32)


def _net(pkg, dev, rng):
    net = pkg.learned.ProposalNetwork(seed=0, device=dev)
    master = _params(rng)
    net.master.copy_(torch.from_numpy(master))
    net.params.copy_(torch.from_numpy(pkg.learned.to_bf16(master).view(np.int16)))
    return net


def _features(pkg, rng, rows):
    f = pkg.learned.point_features(rng.integers(0, 40, rows), np.log(rng.uniform(0.05, 2.0, rows)),
                                   rng.integers(0, 600, rows))
    return pkg.learned.to_bf16(f)


def test_forward_matches_float64(pkg, oracle, dev):
    rng = np.random.default_rng(1)
    net = _net(pkg, dev, rng)
    for rows in (1, 31, 32, 33, 1000, 4099):
        x = _features(pkg, rng, rows)
        got = net.forward(torch.from_numpy(x.view(np.int16)).to(dev)).cpu().numpy()
        ref = oracle.mlp_forward(net.params.cpu().numpy().view(np.uint16), x)
        # bf16 inputs and weights on both sides; the kernel accumulates in fp32
        np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-4)


def test_oracle_gradient_is_the_loss_derivative(oracle):
    """Finite differences of the float64 loss against mlp_pair_grad."""
    rng = np.random.default_rng(2)
    p = _params(rng)
    xu = rng.normal(0, 1, (50, 8)).astype(np.float32)
    xv = rng.normal(0, 1, (50, 8)).astype(np.float32)
    y = (rng.random(50) < 0.4).astype(np.float64)
    from no_three_in_line_amd.learned import to_bf16

    pb, xub, xvb = to_bf16(p), to_bf16(xu), to_bf16(xv)
    g, _ = oracle.mlp_pair_grad(pb, xub, xvb, y)
    base = oracle.bf16_to_f64(pb)

    def loss_at(vec):
        w1 = vec[:1024].reshape(128, 8)
        b1, w2, b2 = vec[1024:1152], vec[1152:1280], vec[1280]
        a = oracle.bf16_to_f64(xub).reshape(-1, 8)
        b = oracle.bf16_to_f64(xvb).reshape(-1, 8)
        z = np.maximum(a @ w1.T + b1, 0) @ w2 + np.maximum(b @ w1.T + b1, 0) @ w2 + b2
        return np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))

    for i in list(rng.choice(1024, 12, replace=False)) + [1030, 1100, 1160, 1279, 1280]:
        e = np.zeros_like(base)
        e[i] = 1e-6
        fd = (loss_at(base + e) - loss_at(base - e)) / 2e-6
        assert abs(fd - g[i]) < 1e-6 + 1e-4 * abs(g[i])


def test_gradient_matches_float64(pkg, oracle, dev):
    rng = np.random.default_rng(3)
    net = _net(pkg, dev, rng)
    for pairs in (7, 64, 1000, 999):
        xu, xv = _features(pkg, rng, pairs), _features(pkg, rng, pairs)
        # 0/1 outcomes, and soft Metropolis-probability labels (last case)
        y = (rng.random(pairs) < 0.3).astype(np.float32) if pairs != 999 else rng.random(pairs).astype(np.float32)
        grad, loss = net.gradient(torch.from_numpy(xu.view(np.int16)).to(dev),
                                  torch.from_numpy(xv.view(np.int16)).to(dev), torch.from_numpy(y).to(dev))
        ref, ref_loss = oracle.mlp_pair_grad(net.params.cpu().numpy().view(np.uint16), xu, xv, y)
        got = grad.cpu().numpy()
        # dH is rounded to bf16 before the dW1 MFMA (8 significant bits):
        # tolerance 2% of the gradient's scale per block of parameters
        for lo, hi in ((0, 1024), (1024, 1152), (1152, 1280), (1280, 1281)):
            scale = np.abs(ref[lo:hi]).max() + 1e-12
            assert np.abs(got[lo:hi] - ref[lo:hi]).max() <= 2e-2 * scale, (pairs, lo)
        assert abs(loss - ref_loss / pairs) < 1e-4 * max(1.0, abs(ref_loss / pairs))


def test_gradient_is_deterministic_and_additive(pkg, oracle, dev):
    rng = np.random.default_rng(9)
    net = _net(pkg, dev, rng)
    pairs = 70000  # more tiles than the partial-sum rows: grid-stride tiles
    xu, xv = _features(pkg, rng, pairs), _features(pkg, rng, pairs)
    y = rng.random(pairs).astype(np.float32)
    args = [torch.from_numpy(xu.view(np.int16)).to(dev), torch.from_numpy(xv.view(np.int16)).to(dev),
            torch.from_numpy(y).to(dev)]
    g1, l1 = net.gradient(*args)
    g2, l2 = net.gradient(*args)
    assert torch.equal(g1, g2) and l1 == l2
    acc = torch.zeros_like(g1)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    net.gradient(*args, loss=loss)
    _lib = pkg._lib
    _lib.check(_lib.lib().n3_mlp_grad(net.params.data_ptr(), args[0].data_ptr(), args[1].data_ptr(),
                                      args[2].data_ptr(), pairs, acc.data_ptr(), loss.data_ptr(),
                                      _lib.stream_handle(dev)))
    _lib.check(_lib.lib().n3_mlp_grad(net.params.data_ptr(), args[0].data_ptr(), args[1].data_ptr(),
                                      args[2].data_ptr(), pairs, acc.data_ptr(), loss.data_ptr(),
                                      _lib.stream_handle(dev)))
    torch.cuda.synchronize()
    assert torch.allclose(acc, 2 * g1, rtol=1e-6, atol=1e-9)
    ref, _ = oracle.mlp_pair_grad(net.params.cpu().numpy().view(np.uint16), xu, xv, y)
    got = g1.cpu().numpy()
    for lo, hi in ((0, 1024), (1024, 1152), (1152, 1280), (1280, 1281)):
        scale = np.abs(ref[lo:hi]).max() + 1e-12
        assert np.abs(got[lo:hi] - ref[lo:hi]).max() <= 2e-2 * scale, lo


def test_weight_table(pkg, oracle, dev):
    rng = np.random.default_rng(4)
    net = _net(pkg, dev, rng)
    log_temps = np.log(np.array([0.05, 0.3, 2.0], np.float32))
    table = net.weight_table(log_temps, max_energy=40).cpu().numpy().view(np.uint32)
    assert table.shape == (3, 41, 64) and table.min() >= 1 and table.max() <= 1 << 15
    L = pkg.learned
    for t, lt in enumerate(log_temps):
        for e in (0, 7, 40):
            x = L.to_bf16(L.point_features(np.arange(64), lt, e))
            g = np.clip(oracle.mlp_forward(net.params.cpu().numpy().view(np.uint16), x), -6, 5.5)
            w = np.clip(np.rint(128 * np.exp(g)), 1, 1 << 15)
            # the device computes the features with its own sqrt/log before
            # bf16 rounding: allow a few bf16 ulps of the logit (and one unit
            # of the integer rounding)
            np.testing.assert_allclose(table[t, e], w, rtol=2e-2, atol=1)


def _cols(oracle, n, replicas, seed):
    p, q = oracle.generate_states(seed, n, replicas)
    return np.stack([np.column_stack((p[i], q[i])).reshape(-1) for i in range(replicas)]).astype(np.int32)


@pytest.mark.parametrize("n,replicas,ladder,sweeps", [(12, 32, 16, 8), (33, 64, 32, 4), (75, 64, 64, 3)])
def test_learned_walk_matches_replay(pkg, oracle, dev, n, replicas, ladder, sweeps):
    rng = np.random.default_rng(n)
    net = _net(pkg, dev, rng)
    cols = _cols(oracle, n, replicas, seed=n + 1)
    temps = np.tile(np.geomspace(0.05, 2.0, ladder), replicas // ladder).astype(np.float32)
    chains = pkg.TemperingChains(cols, temps, seed=77)
    props = pkg.learned.prepare(chains, net)
    replay = oracle.TemperingReplay(cols, temps, 77, k_move=chains.k_move, k_exchange=chains._exchange_table(ladder)[2])
    thr64 = props.thr64.cpu().numpy().view(np.uint64)
    assert np.array_equal(thr64, oracle.metropolis_thresholds64(temps, chains.k_move))
    weights = props.weights.cpu().numpy().view(np.uint32)
    index = props.temp_index.cpu().numpy()
    for _ in range(sweeps):
        pkg.learned.run_learned(chains, props, 1, 2 * n, ladder)
        replay.advance_learned(2 * n, weights, index, props.max_energy, thr64)
        replay.exchange(ladder)
        torch.cuda.synchronize()
        for name in ("columns", "badness", "scores", "best_columns", "best_scores", "accepted", "removed_points"):
            assert np.array_equal(getattr(chains, name).cpu().numpy(), getattr(replay, name)), name
    chains.validate()


def test_training_on_logged_proposals(pkg, dev):
    chains = pkg.seeded_chains(40, 512, 64, seed=5)
    chains.run(5, 80, 64)
    torch.cuda.synchronize()
    valid0, acc0 = (int(x) for x in chains.proposals.cpu())
    log = pkg.learned.ProposalLog(4 * 80 * 512, dev)
    chains.log = log
    chains.run(4, 80, 64)
    chains.log = None
    torch.cuda.synchronize()
    valid1, acc1 = (int(x) for x in chains.proposals.cpu())
    rows = log.rows()
    r = rows.cpu().numpy()
    # every valid proposal of the logged sweeps is logged exactly once
    assert int(log.count.item()) == 4 * 80 * 512
    assert len(r) == valid1 - valid0 > 10000
    assert int(r[:, 5].sum()) == acc1 - acc0
    assert set(np.unique(r[:, 5])) <= {0, 1} and np.all(r[r[:, 4] <= 0, 5] == 1)
    assert np.bincount(r[:, 3], minlength=512).max() <= 4 * 80 and r[:, 2].min() >= 0
    # a log too small for every replica's reservation keeps whole sweeps only
    small = pkg.learned.ProposalLog(80 * 100 + 7, dev)
    chains.log = small
    chains.run(1, 80, 64)
    chains.log = None
    torch.cuda.synchronize()
    part = small.rows().cpu().numpy()
    assert int(small.count.item()) == 80 * 512 and 0 < len(part) <= 80 * 100
    assert np.all(small.entries[80 * 100:].cpu().numpy() == -1)
    net = pkg.learned.ProposalNetwork(seed=1, device=dev)
    xu, xv, y = pkg.learned.training_tensors(rows, chains.temperatures_host, dev)
    # device features equal the host formula up to one bf16 rounding step
    host = pkg.learned.point_features(r[:, 0], np.log(chains.temperatures_host.astype(np.float32))[r[:, 3]], r[:, 2])
    dev_f = pkg.learned.from_bf16(xu.cpu().numpy().view(np.uint16))
    assert np.allclose(dev_f, pkg.learned.from_bf16(pkg.learned.to_bf16(host)), rtol=2 ** -7, atol=1e-6)
    losses = net.train(xu, xv, y, epochs=4)
    assert losses[-1] < losses[0]
    props = pkg.learned.prepare(chains, net)
    pkg.learned.run_learned(chains, props, 20, 80, 64)
    torch.cuda.synchronize()
    chains.validate()
    # online retraining rewrites the table in place (the captured graph is reused)
    before = props.weights.clone()
    net.train(xu, xv, y, steps=50, seed=3)
    again = pkg.learned.refresh(props, chains, net)
    assert again.weights.data_ptr() == props.weights.data_ptr() and not torch.equal(before, props.weights)
    fresh = pkg.learned.prepare(chains, net)
    assert torch.equal(fresh.weights, props.weights)
    pkg.learned.run_learned(chains, props, 20, 80, 64)
    torch.cuda.synchronize()
    chains.validate()


def test_trained_network_reaches_low_energy_in_fewer_sweeps(pkg, dev):
    """What the trained network gains over uniform proposals, counted in
    sweeps (not wall clock): from the reference's E=10 near-solution with
    every replica perturbed by 12 random 2-switches (4,096 replicas, 64
    ladders of 64 temperatures 0.05..2.0), the learned walk, pretrained on
    another perturbation's plain logs, reaches E <= 12 within fewer sweeps
    than the uniform walk from the same states and seed."""
    from pathlib import Path

    fixture = Path(__file__).parent / "golden" / "best-near-solution-E10.json"
    _, points, _ = pkg.io.load_candidate(fixture, 75)
    state = pkg.points_to_permutations(points, 75)
    base = np.repeat(np.stack((np.array(state.P), np.array(state.Q)), 1).reshape(1, 150).astype(np.int32), 4096, 0)
    net, _ = pkg.learned.pretrain(75, 4096, 64, seed=11, start=pkg.random_two_switches(base, 12, seed=99))
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
    props = pkg.learned.prepare(learned, net)
    s_uniform = sweeps_to(lambda c, k: c.run(k, 150, 64), uniform)
    s_learned = sweeps_to(lambda c, k: pkg.learned.run_learned(c, props, k, 150, 64), learned)
    assert s_learned is not None
    assert s_uniform is None or s_learned < s_uniform, (s_learned, s_uniform)
    learned.validate(range(0, 4096, 257))


def test_pretraining_is_reproducible(pkg, dev):
    """The same plain run logged and trained twice gives the same network:
    the log is read in slot order and the gradient has no float atomics."""
    a, _ = pkg.learned.pretrain(40, 1024, 64, seed=3, sweeps=20, log_sweeps=2, steps=60)
    b, _ = pkg.learned.pretrain(40, 1024, 64, seed=3, sweeps=20, log_sweeps=2, steps=60)
    assert torch.equal(a.master, b.master) and torch.equal(a.params, b.params)


def test_block_mapping_leaves_trajectories_unchanged(pkg, dev):
    """With 2,048 replicas the learned walk runs 8 replicas per block, and a
    run with exchanges takes each block's replicas from every eighth
    temperature of a 64-temperature ladder.  Launching the walk without the
    ladder (consecutive slots per block) and the exchanges separately gives
    the same trajectories: the slot, not the block, keys the random stream."""
    net = _net(pkg, dev, np.random.default_rng(3))
    a = pkg.seeded_chains(40, 2048, 64, seed=9)
    b = pkg.seeded_chains(40, 2048, 64, seed=9)
    pa, pb = pkg.learned.prepare(a, net), pkg.learned.prepare(b, net)
    pkg.learned.run_learned(a, pa, 3, 80, 64)
    for _ in range(3):
        pkg.learned.run_learned(b, pb, 1, 80)
        b.exchange(64)
    torch.cuda.synchronize()
    for name in ("columns", "badness", "scores", "best_columns", "best_scores", "accepted", "removed_points",
                 "counters", "proposals"):
        assert torch.equal(getattr(a, name), getattr(b, name)), name
    assert b.exchanges_total() > 0
