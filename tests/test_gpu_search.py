"""Device GA / beam engines: the mutation and enumeration kernels against the
oracle, the elite keys and ranking, and the two search drivers."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _parents(oracle, n, count, seed):
    p, q = oracle.generate_states(seed, n, count)
    _, bad = oracle.score_population(p, q)
    return p, q, bad.astype(np.int32)


@pytest.mark.parametrize("n,elites,population,active,maximum,bias", [
    (4, 3, 64, 1, 1, 85),
    (9, 6, 500, 2, 6, 0),
    (33, 16, 1000, 3, 12, 85),
    (75, 64, 4096, 4, 20, 85),
    (75, 8, 777, 2, 3, 100),
    (100, 5, 300, 6, 9, 50),
])
def test_mutation_matches_oracle(pkg, oracle, dev, n, elites, population, active, maximum, bias):
    p, q, bad = _parents(oracle, n, elites, seed=n)
    quality = max(1, elites // 4)
    tp, tq, tb = (torch.from_numpy(a).to(dev) for a in (p, q, bad))
    for generation in (1, 7):
        cp, cq = pkg.gpu_device_search.mutate_population(
            tp, tq, tb, quality_count=quality, population=population, active_strength=active,
            max_strength=maximum, bad_bias=bias, seed=12345 + n, generation=generation)
        rp, rq = oracle.mutate_population(p, q, bad, quality_count=quality, population=population,
                                          active_strength=active, max_strength=maximum, bad_bias=bias,
                                          seed=12345 + n, generation=generation)
        assert np.array_equal(cp.cpu().numpy(), rp) and np.array_equal(cq.cpu().numpy(), rq), (n, generation)


def test_mutation_host_wrapper_and_errors(pkg, oracle, dev):
    p, q, bad = _parents(oracle, 20, 8, seed=3)
    cp, cq = pkg.mutate_population_device(p, q, population=100, active_strength=2, max_strength=4, seed=9,
                                          generation=2)
    rp, rq = oracle.mutate_population(p, q, bad, quality_count=2, population=100, active_strength=2,
                                      max_strength=4, bad_bias=85, seed=9, generation=2)
    assert np.array_equal(cp, rp) and np.array_equal(cq, rq)
    tp, tq, tb = (torch.from_numpy(a).to(dev) for a in (p, q, bad))
    mutate = pkg.gpu_device_search.mutate_population
    with pytest.raises(ValueError):
        mutate(tp, tq, tb, quality_count=9, population=10, active_strength=1, max_strength=1, bad_bias=0, seed=0,
               generation=0)
    with pytest.raises(ValueError):
        mutate(tp, tq, tb, quality_count=1, population=10, active_strength=0, max_strength=1, bad_bias=0, seed=0,
               generation=0)
    with pytest.raises(ValueError):
        mutate(tp, tq, tb, quality_count=1, population=10, active_strength=1, max_strength=1, bad_bias=101, seed=0,
               generation=0)
    empty = mutate(tp, tq, tb, quality_count=1, population=0, active_strength=1, max_strength=1, bad_bias=0,
                   seed=0, generation=0)
    assert empty[0].shape == (0, 20)


@pytest.mark.parametrize("n", [4, 32, 33, 75, 100])
def test_enumeration_matches_oracle(pkg, oracle, dev, n):
    p, q, bad = _parents(oracle, n, 5, seed=n + 1)
    rng = np.random.default_rng(n)
    desc = np.stack((rng.integers(0, 5, 3000), rng.integers(0, 2 * n, 3000), rng.integers(0, 2 * n, 3000)),
                    axis=1).astype(np.int32)
    desc = np.concatenate((desc, pkg.candidate_descriptors(bad)[:2000]), axis=0)
    tp, tq = torch.from_numpy(p).to(dev), torch.from_numpy(q).to(dev)
    cp, cq, valid = pkg.gpu_exhaustive_beam.enumerate_switches(tp, tq, torch.from_numpy(desc).to(dev))
    rp, rq, rvalid = oracle.enumerate_switches(p, q, desc)
    assert np.array_equal(cp.cpu().numpy(), rp) and np.array_equal(cq.cpu().numpy(), rq)
    assert np.array_equal(valid.cpu().numpy().astype(bool), rvalid)
    with pytest.raises(ValueError):
        pkg.gpu_exhaustive_beam.enumerate_switches(tp, tq, torch.tensor([[5, 0, 1]], dtype=torch.int32, device=dev))
    with pytest.raises(ValueError):
        pkg.gpu_exhaustive_beam.enumerate_switches(tp, tq, torch.tensor([[0, 2 * n, 1]], dtype=torch.int32,
                                                                        device=dev))


def test_candidate_scores_equal_enumerated_recounts(pkg, oracle, dev):
    """The fused beam path: scores and legality read from the 2-switch table
    equal the reference's path of materialising every child and recounting
    it (enumerate_switches + score_population)."""
    n = 75
    p, q, bad = _parents(oracle, n, 4, seed=21)
    scores, _ = oracle.score_population(p, q)
    desc = pkg.candidate_descriptors(bad)
    tp, tq = torch.from_numpy(p).to(dev), torch.from_numpy(q).to(dev)
    td = torch.from_numpy(desc).to(dev)
    got, valid = pkg.gpu_exhaustive_beam.candidate_scores(tp, tq, torch.from_numpy(scores).to(dev),
                                                           torch.from_numpy(bad).to(dev), td)
    cp, cq, cvalid = pkg.gpu_exhaustive_beam.enumerate_switches(tp, tq, td)
    recount = pkg.score_population_device(cp, cq).cpu().numpy()
    cvalid = cvalid.cpu().numpy().astype(bool)
    assert np.array_equal(valid, cvalid)
    assert np.array_equal(got[valid], recount[valid])
    assert np.all(got[~valid] == 1 << 29)


def test_elite_keys_and_device_ranking(pkg, oracle, dev):
    rng = np.random.default_rng(2)
    scores = rng.integers(0, 5, 3000).astype(np.int32)
    bad = rng.integers(0, 4, (3000, 150)).astype(np.int32)
    tb = torch.from_numpy(bad).to(dev)
    mx, sq = pkg.gpu_device_search.elite_keys_device(tb)
    assert np.array_equal(mx.cpu().numpy(), bad.max(axis=1))
    assert np.array_equal(sq.cpu().numpy(), (bad.astype(np.int64) ** 2).sum(axis=1))
    order = pkg.rank_elites_device(torch.from_numpy(scores).to(dev), tb).cpu().numpy()
    assert np.array_equal(order, pkg.rank_elites(scores, bad))


def test_device_elite_selection_matches_host(pkg, dev):
    """The device sort and window search pick the same children, in the same
    order and with the same rng calls, as the host argsort form, with many
    ties and windows larger and smaller than the beam."""
    from random import Random

    rng = np.random.default_rng(4)
    for trial in range(12):
        population = int(rng.integers(50, 5000))
        scores = rng.integers(0, int(rng.integers(3, 60)), size=population).astype(np.int32)
        best = int(scores.min()) + int(rng.integers(0, 3))
        window = int(rng.integers(0, 4))
        elite = int(rng.integers(1, min(population, 300)))
        host = pkg.gpu_device_search.select_elites(scores, best, window, elite, Random(trial))
        device = pkg.gpu_device_search.select_elites_device(torch.from_numpy(scores).to(dev), best, window, elite,
                                                            Random(trial))
        assert device == host
        assert len(host) >= min(elite, population) and len(set(host)) == len(host)


def test_device_beam_search_runs_and_improves(pkg, oracle, dev):
    from random import Random

    state = pkg.random_state(40, Random(5))
    snapshots = []
    result = pkg.gpu_device_beam_search(state, seed=3, generations=30, population=2048, elite_size=64,
                                        max_strength=8, energy_window=2, callback=snapshots.append)
    result.state.validate()
    start = oracle.score_population(np.array([state.P]), np.array([state.Q]))[0][0]
    final = oracle.score_population(np.array([result.state.P]), np.array([result.state.Q]))[0][0]
    assert result.score == final < start
    assert result.generations == 30 and result.evaluations == 1 + 30 * 2048
    assert snapshots and all(s.best_score >= result.score for s in snapshots)
    assert [s.best_score for s in snapshots] == sorted((s.best_score for s in snapshots), reverse=True)
    # the same seed gives the same run
    again = pkg.gpu_device_beam_search(state, seed=3, generations=30, population=2048, elite_size=64,
                                       max_strength=8, energy_window=2)
    assert again.score == result.score and again.state.points() == result.state.points()


def test_exhaustive_beam_search_runs_and_improves(pkg, oracle, dev):
    from random import Random

    state = pkg.random_state(30, Random(8))
    result = pkg.gpu_exhaustive_beam_search(state, seed=4, generations=6, elite_size=8, energy_window=1,
                                            global_every=3)
    result.state.validate()
    start = oracle.score_population(np.array([state.P]), np.array([state.Q]))[0][0]
    final = oracle.score_population(np.array([result.state.P]), np.array([result.state.Q]))[0][0]
    assert result.score == final < start
    assert result.generations >= 1 and result.evaluations > 1
