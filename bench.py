"""Benchmark of the no-three-in-line hot path on one MI355X.

Headline (value): n=75 population scoring -- P = 65,536 seeded states,
exact collinear-triple count plus the per-slot badness vector int32[150] of
every state -- in states scored per second, inputs resident in HBM.  One
step is one scoring pass over the whole population.

Also reported: the same population's kernel roofline (INT32 VALU), a CPU
baseline (the oracle's OpenMP restatement of the reference loop, timed on the
host's cores on a bounded sample), the replica-exchange tempering rate (64
ladders x 64 temperatures at n=75) and the batched uneven-size rate, each in
proposals per second with its walk kernel's roofline and a CPU replay
baseline, the learned-proposal A/B, the search engines and the move scores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

N = 75
POPULATION = 65536
SEED = 0
PEAK_VALU_TOPS = 256 * 64 * 2.4e9 / 1e12  # 256 CUs x 64 lanes x 2.4 GHz, int32 lane-ops
# LDS: 128 bytes per clock per CU, i.e. 32 conflict-free 4-byte lane accesses
PEAK_LDS_TACCESS = 256 * 32 * 2.4e9 / 1e12


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def slope_ops_per_state(n: int) -> int:
    """Algorithmic work of the one-sided direction-histogram scorer (the
    kernel for 33 <= n <= 100): for every slot pair b < a, key difference,
    table lookup, histogram increment, count read-back, counter clear and
    accumulate -- 6 integer operations per unordered pair."""
    m = 2 * n
    return 6 * (m * (m - 1) // 2)


def slope_lds_accesses_per_state(n: int) -> int:
    """LDS lane accesses of the same kernel: lookup, increment, read-back and
    clear per unordered slot pair."""
    m = 2 * n
    return 4 * (m * (m - 1) // 2)


def determinant_ops_per_state(n: int) -> int:
    """6 C(m,3) + 2 C(m,2): the reference loop's integer work (accelerated.py:27-36)."""
    m = 2 * n
    return 6 * (m * (m - 1) * (m - 2) // 6) + 2 * (m * (m - 1) // 2)


def walk_ops_per_valid_proposal(n: int) -> int:
    """Algorithmic work of the walk's exact delta for one valid proposal
    (the rejection path, which 97-98% of valid proposals take): for each of
    the m - 2 unmoved slots, the histograms seen from u' and from v' in the
    moved state cost a key difference, a table lookup, an increment that
    returns the rank, a clear and an accumulate each (5 + 5), and joint(u, v)
    in the old state a key difference, a lookup and a compare (3):
    13 (m - 2) integer operations (1,924 at n=75)."""
    return 13 * (2 * n - 2)


def _walk_roofline(valid: float, ops_weighted: float, kernel_ms: float) -> dict:
    achieved = ops_weighted / (kernel_ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": PEAK_VALU_TOPS, "unit": "TOP/s",
            "frac": achieved / PEAK_VALU_TOPS, "traffic": None, "kernel_ms": kernel_ms,
            "valid_proposals_per_launch": valid}


def _profile_values() -> dict:
    """`name: value` lines of the committed rocprofv3 PMC summary of the
    scoring kernel (HBM bytes per launch and counter-derived fractions)."""
    path = ROOT / "profiles" / "score_onesided_n75_P65536_pmc.txt"
    out = {}
    if path.exists():
        for line in path.read_text().splitlines():
            key, sep, value = line.partition(":")
            if sep and key and " " not in key:
                try:
                    out[key] = float(value)
                except ValueError:
                    pass
    return out


def traffic_from_profile() -> float | None:
    """HBM bytes per scoring launch from the committed rocprofv3 PMC summary."""
    return _profile_values().get("traffic_bytes_per_launch")


def cpu_baseline(seconds: float = 12.0) -> dict:
    import numpy as np

    import __graft_entry__ as entry

    oracle = entry.load_oracle()
    cores = oracle.cpu_count()
    p, q = oracle.generate_states(SEED, N, 4096)
    oracle.score_population(p[:64], q[:64], threads=cores)  # warm up the pool
    t0 = time.perf_counter()
    oracle.score_population(p[:1024], q[:1024], threads=cores)
    probe = time.perf_counter() - t0
    reps = max(1, int(seconds / max(probe * 4, 1e-6)))
    t0 = time.perf_counter()
    for _ in range(reps):
        oracle.score_population(p, q, threads=cores)
    elapsed = time.perf_counter() - t0
    states = reps * len(p)
    t0 = time.perf_counter()
    oracle.score_population(p[:128], q[:128], threads=1)
    single = 128 / (time.perf_counter() - t0)
    return {"value": states / elapsed, "unit": "states/s", "cores": cores, "kind": "port",
            "sample": f"{states} n=75 seeded states (the first 4096 of the population, {reps} passes), "
                      f"oracle C restatement of accelerated.score_and_badness, OpenMP over states",
            "single_thread_states_per_s": single}


def tempering_rate(pkg, torch, sweeps: int = 10000) -> dict:
    """Config 3 in full: 64 ladders x 64 temperatures (geometric 0.05..2.0)
    at n=75, 10,000 sweeps of 150 proposals, exchange every sweep,
    graph-captured; a sample of replicas is recounted afterwards."""
    import numpy as np

    replicas, ladder, steps = 4096, 64, 150
    dev = pkg._lib.device(0)
    p, q = pkg.generate_states(SEED + 1, N, replicas, device=dev)
    cols = torch.stack((p, q), dim=2).reshape(replicas, 2 * N).cpu().numpy()
    temps = np.tile(np.geomspace(0.05, 2.0, ladder), replicas // ladder).astype(np.float32)
    chains = pkg.TemperingChains(cols, temps, seed=3)
    chains.run(32, steps, ladder)  # captures the 32-sweep graph outside the timed window
    torch.cuda.synchronize()
    before = chains.proposals.cpu().numpy().copy()
    # roofline of the walk: HIP events on the launch stream around the timed
    # sweeps themselves (per sweep one walk launch, whose blocks take their
    # replicas spread over a ladder, and one exchange launch of about 9 us)
    stream = torch.cuda.current_stream(dev)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    chains.run(sweeps, steps, ladder)
    ev1.record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    after = chains.proposals.cpu().numpy()
    chains.validate(range(0, replicas, 97))
    proposals = sweeps * steps * replicas
    out = {"replicas": replicas, "ladders": replicas // ladder, "sweeps": sweeps, "steps_per_sweep": steps,
           "seconds": elapsed, "proposals_per_s": proposals / elapsed,
           "valid_proposals_per_s": float(after[0] - before[0]) / elapsed,
           "accepted_per_s": float(after[1] - before[1]) / elapsed,
           "best_E": int(chains.best_scores.min().item())}
    valid = float(after[0] - before[0]) / sweeps
    out["roofline"] = _walk_roofline(valid, valid * walk_ops_per_valid_proposal(N), ev0.elapsed_time(ev1) / sweeps)
    out["roofline"]["timed"] = "walk + exchange launch per sweep, events around the timed sweeps"
    out["roofline"]["ops_per_valid_proposal"] = walk_ops_per_valid_proposal(N)
    return out


def batched_rate(pkg, torch, sweeps: int = 1000, persistent: bool = True) -> dict:
    """Config 4: mixed sizes n in {40, 52, 61, 75, 83, 100} with {8192, 4096,
    4096, 2048, 2048, 1024} replicas (ladders of 64, geometric 0.05..2.0), 2n
    proposals per sweep, replicas per workgroup from the autotuner (tuned on
    the mixed run itself, in the form timed); all
    1,000 sweeps in ONE persistent cooperative launch (the leg's rate and
    roofline, HIP events around that launch), and the same sweeps as one
    batched walk launch per sweep replayed from captured graphs.  Without
    `persistent` (for profilers, which fail on programs that made a
    cooperative launch), the shapes are tuned on and the leg reports the
    per-sweep graphs only."""
    sizes = {40: 8192, 52: 4096, 61: 4096, 75: 2048, 83: 2048, 100: 1024}
    t0 = time.perf_counter()
    shapes = pkg.autotune.tune_mixed(sizes, sweeps=8, candidates=(4, 6, 8, 12, 16), refresh=True, persistent=persistent)
    tune_seconds = time.perf_counter() - t0
    groups = []
    for i, (n, replicas) in enumerate(sizes.items()):
        chains = pkg.seeded_chains(n, replicas, 64, seed=40 + i)
        chains.block_waves = shapes[n]
        groups.append(pkg.BatchGroup(chains, 64, 2 * n))
    proposals = sum(g.chains.temperatures_host.size * g.steps for g in groups) * sweeps
    # per-sweep launches from graphs
    pkg.run_batched(groups, 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pkg.run_batched(groups, sweeps)
    torch.cuda.synchronize()
    graph_seconds = time.perf_counter() - t0
    if not persistent:
        return {"sizes": {str(n): r for n, r in sizes.items()}, "sweeps": sweeps, "launch": "per-sweep graphs",
                "per_sweep_graphs": {"seconds": graph_seconds, "proposals_per_s": proposals / graph_seconds},
                "proposals_per_s": proposals / graph_seconds,
                "block_waves": {str(n): w for n, w in shapes.items()}, "autotune_seconds": tune_seconds}
    # the persistent launch (its plan is built by a short first run)
    pkg.run_batched(groups, 2, persistent=True)
    torch.cuda.synchronize()
    before = [g.chains.proposals.cpu().numpy()[0] for g in groups]
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    pkg.run_batched(groups, sweeps, persistent=True)
    ev1.record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    valid = [float(g.chains.proposals.cpu().numpy()[0] - b) for g, b in zip(groups, before)]
    for g in groups:
        g.chains.validate(range(0, len(g.chains.temperatures_host), 211))
    kernel_ms = ev0.elapsed_time(ev1)
    ops = sum(v * walk_ops_per_valid_proposal(g.chains.n) for v, g in zip(valid, groups))
    roofline = _walk_roofline(sum(valid), ops, kernel_ms)
    roofline["ops_per_valid_proposal"] = {str(g.chains.n): walk_ops_per_valid_proposal(g.chains.n) for g in groups}
    roofline["launch"] = f"one cooperative launch of {sweeps} sweeps (walk and exchange phases)"
    return {"sizes": {str(n): r for n, r in sizes.items()}, "sweeps": sweeps, "seconds": elapsed,
            "proposals_per_s": proposals / elapsed, "launch": "persistent (one launch for all sweeps)",
            "per_sweep_graphs": {"seconds": graph_seconds, "proposals_per_s": proposals / graph_seconds},
            "block_waves": {str(n): w for n, w in shapes.items()}, "autotune_seconds": tune_seconds,
            "best_E": {str(g.chains.n): int(g.chains.best_scores.min().item()) for g in groups},
            "roofline": roofline}


def batched_cpu_baseline(pkg, oracle, replicas_per_size: int = 256) -> dict:
    """The oracle's replay of the reference walk for every size of the
    batched workload: one sweep of 2n proposals for `replicas_per_size`
    seeded replicas of each n (ladders of 64 temperatures 0.05..2.0),
    OpenMP over replicas on all host cores; the config's rate is its
    proposals over the time the measured per-size rates give it."""
    import numpy as np

    sizes = {40: 8192, 52: 4096, 61: 4096, 75: 2048, 83: 2048, 100: 1024}
    cores = oracle.cpu_count()
    per_size, seconds_per_sweep, proposals_per_sweep = {}, 0.0, 0
    for n, full in sizes.items():
        p, q = oracle.generate_states(40 + n, n, replicas_per_size)
        cols = np.stack((p, q), axis=2).reshape(replicas_per_size, 2 * n)
        temps = np.tile(np.geomspace(0.05, 2.0, 64), replicas_per_size // 64).astype(np.float32)
        replay = oracle.TemperingReplay(cols, temps, 7, k_move=pkg.tempering.move_table_size(temps), k_exchange=2)
        t0 = time.perf_counter()
        replay.advance(2 * n, threads=cores)
        rate = replicas_per_size * 2 * n / (time.perf_counter() - t0)
        per_size[str(n)] = rate
        seconds_per_sweep += full * 2 * n / rate
        proposals_per_sweep += full * 2 * n
    return {"value": proposals_per_sweep / seconds_per_sweep, "unit": "proposals/s", "cores": cores, "kind": "port",
            "sample": f"one sweep of 2n proposals for {replicas_per_size} replicas of each n, oracle C replay of "
                      f"tempering.walk with its pair-loop delta, OpenMP over replicas; the rate of the full "
                      f"mix weighted by its replica counts",
            "per_size_proposals_per_s": per_size}


AB_SIDES = (("uniform", "uniform"), ("learned", "learned_table"), ("learned_point", "learned_point"))


def _ab_summary(res: dict) -> dict:
    out = {"budget_seconds_each": res["config"]["seconds"]}
    for key, name in AB_SIDES:
        if key not in res:
            continue
        r = res[key]
        side = {"sweeps": r["sweeps"], "best_E": min(e for _, e in r["trace"]), "seconds_to_E": r["threshold_seconds"]}
        if key != "uniform":
            side.update({"training_tuples": r["training_tuples"], "retrains": r["retrains"],
                         "seconds_before_learned_proposals": r["seconds_until_learned_proposals"]})
        out[name] = side
    out["time_to_E12"] = {name: res[key]["threshold_seconds"].get("12") for key, name in AB_SIDES if key in res}
    return out


AB_THRESHOLDS = (80, 70, 60, 55, 50, 47, 45, 44, 43, 42, 41, 40, 38, 35, 30, 12)


def learned_ab(pkg, seconds: float) -> dict:
    """n=75, 4096 replicas as 64 ladders x 64 temperatures from seeded random
    states: uniform proposals against both learned modes (the tabulated
    count weights and the point network with row/column features), same
    seeds, replicas and GPU, equal wall-clock budget; each learned side pays
    for its plain warm-up, logging and online training inside its budget.
    The config-5 result is the time to each energy level."""
    res = pkg.learned.ab_run(75, 4096, 64, seconds=seconds, seed=11, state_seed=1, train_sweeps=300, chunk=100,
                             thresholds=AB_THRESHOLDS, modes=("table", "point"))
    out = _ab_summary(res)
    out["start"] = "seeded random states (state seed 1)"
    return out


NEAR_SOLUTION = ROOT / "tests" / "golden" / "best-near-solution-E10.json"


def learned_ab_near_solution(pkg, seconds: float, moves: int = 12, seeds=(12, 13, 14)) -> dict | None:
    """The A/B from the reference's saved E=10 near-solution, each replica
    perturbed by `moves` random legal 2-switches (seeded), where E <= 12 is
    reachable within the budget so time-to-E<=12 is measured.  Each learned
    side trains its network inside its own budget: 60 plain sweeps from the
    race's start, the last 2 logged, 300 Adam steps, then learned proposals
    with online retraining.  One race per seed."""
    import numpy as np

    if not NEAR_SOLUTION.exists():
        return None
    _, points, _ = pkg.io.load_candidate(NEAR_SOLUTION, 75)
    state = pkg.points_to_permutations(points, 75)
    base = np.repeat(np.stack((np.array(state.P), np.array(state.Q)), 1).reshape(1, 150).astype(np.int32), 4096, 0)
    races = []
    for s in seeds:
        res = pkg.learned.ab_run(75, 4096, 64, seconds=seconds, seed=11, state_seed=1, chunk=20, train_sweeps=60,
                                 train_steps=300, thresholds=(30, 20, 15, 12, 11, 10),
                                 start=pkg.random_two_switches(base, moves, seed=s), modes=("table", "point"))
        summary = _ab_summary(res)
        summary["perturbation_seed"] = s
        races.append(summary)

    def median(values):
        values = sorted(v if v is not None else float("inf") for v in values)
        mid = values[len(values) // 2]
        return None if mid == float("inf") else mid

    names = [name for _, name in AB_SIDES]
    out = {"start": f"reference near-solution (E=10) + {moves} random 2-switches per replica",
           "budget_seconds_each": seconds, "training": "inside each learned side's budget",
           "races": races,
           "time_to_E12_median": {name: median([r["time_to_E12"][name] for r in races]) for name in names}}
    for name in names[1:]:
        out[f"{name}_first"] = sum(1 for r in races if r["time_to_E12"][name] is not None and
                                   (r["time_to_E12"]["uniform"] is None or
                                    r["time_to_E12"][name] < r["time_to_E12"]["uniform"]))
    return out


def host_array_rates(pkg, torch, p, q, repeats: int = 5) -> dict:
    """The drop-ins that take host NumPy arrays, timed with their PCIe copies
    (pageable H2D of p, q = 39.3 MB; D2H of the scores, plus 39.3 MB of
    badness for score_population_badness) on the same 65,536 states: the
    PCIe-inclusive rate of the reference's boundary, not the headline."""
    hp, hq = p.cpu().numpy(), q.cpu().numpy()
    out = {}
    for name in ("score_population_gpu", "score_population_badness"):
        fn = getattr(pkg, name)
        fn(hp, hq)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            fn(hp, hq)
        elapsed = (time.perf_counter() - t0) / repeats
        out[name] = {"seconds_per_call": elapsed, "states_per_s": len(hp) / elapsed}
    return out


def search_rates(pkg) -> dict:
    """The device GA (population 65,536 children per generation, 256 elites)
    and the exhaustive 2-switch beam (32 elites) from one seeded n=75 state:
    children scored per second and the best energy reached."""
    from random import Random

    state = pkg.random_state(75, Random(7))
    # warm-up at the timed sizes: code objects (the library sorts pick their
    # kernels by size), allocator
    pkg.gpu_device_beam_search(state, seed=1, generations=2, population=65536, elite_size=256, max_strength=8,
                               energy_window=2)
    pkg.gpu_exhaustive_beam_search(state, seed=1, generations=1, elite_size=32, energy_window=1)
    ga = pkg.gpu_device_beam_search(state, seed=1, generations=20, population=65536, elite_size=256,
                                    max_strength=8, energy_window=2)
    beam = pkg.gpu_exhaustive_beam_search(state, seed=1, generations=5, elite_size=32, energy_window=1)
    return {
        "ga": {"n": 75, "population": 65536, "elite_size": 256, "generations": ga.generations,
               "seconds": ga.elapsed_seconds, "children_per_s": ga.evaluations / ga.elapsed_seconds,
               "best_E": ga.score},
        "exhaustive_beam": {"n": 75, "elite_size": 32, "generations": beam.generations,
                            "seconds": beam.elapsed_seconds,
                            "children_scored_per_s": beam.evaluations / beam.elapsed_seconds,
                            "best_E": beam.score},
    }


MOVE_ANCHORS = (0, 37, 74, 111, 149)
MOVE_REPEATS = 100
MOVE_TABLE_STATES = 4096


def _move_state(pkg):
    import numpy as np

    from random import Random

    state = pkg.random_state(N, Random(91001))
    p = np.asarray(state.P, dtype=np.int32)
    q = np.asarray(state.Q, dtype=np.int32)
    score, bad = pkg.score_and_badness(p, q)
    return p, q, int(score), np.asarray(bad)


def move_rates(pkg, torch) -> dict:
    """Move scores at n=75.
    - The reference's delta harness (gpu_delta.py:131-157): one seeded state,
      5 anchors u, 100 repeats of the host entry point score_two_switches_gpu
      (copies included), and of the same call on device tensors.
    - The full 2-switch table of 4,096 resident states (150 x 150 scores each).
    - 3-cycle scores of every legal candidate of one anchor."""
    import numpy as np

    n, m = N, 2 * N
    dev = pkg._lib.device(0)
    lib = pkg._lib.lib()
    stream = pkg._lib.stream_handle(dev)
    p, q, score, bad = _move_state(pkg)
    anchors, repeats = MOVE_ANCHORS, MOVE_REPEATS
    out = {"n": n, "anchors": len(anchors), "repeats": repeats}
    for u in anchors:
        pkg.score_two_switches_gpu(p, q, u, score, bad)  # warm-up
    t0 = time.perf_counter()
    for u in anchors:
        for _ in range(repeats):
            pkg.score_two_switches_gpu(p, q, u, score, bad)
    out["host_call_us"] = (time.perf_counter() - t0) / (len(anchors) * repeats) * 1e6
    tp, tq = torch.from_numpy(p).to(dev), torch.from_numpy(q).to(dev)
    tb = torch.from_numpy(np.asarray(bad, dtype=np.int32)).to(dev)
    res = torch.empty(m, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for u in anchors:
        for _ in range(repeats):
            pkg._lib.check(lib.n3_two_switch_scores(tp.data_ptr(), tq.data_ptr(), tb.data_ptr(), n, u, int(score),
                                                    res.data_ptr(), stream))
    torch.cuda.synchronize()
    out["device_call_us"] = (time.perf_counter() - t0) / (len(anchors) * repeats) * 1e6
    # full tables of a resident population
    count = MOVE_TABLE_STATES
    pp, pq = pkg.generate_states(5, n, count, device=dev)
    pbad = torch.empty((count, m), dtype=torch.int32, device=dev)
    pscores = pkg.score_population_device(pp, pq, badness=pbad)
    table = torch.empty((count, m, m), dtype=torch.int32, device=dev)
    pkg.delta_table_device(pp, pq, pscores, pbad, out=table)
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        pkg.delta_table_device(pp, pq, pscores, pbad, out=table)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / reps
    out["table"] = {"states": count, "scores_per_state": m * m, "seconds": el,
                    "move_scores_per_s": count * m * m / el}
    # 3-cycles of one anchor
    cand = pkg.legal_three_cycle_candidates(p, q, 0)
    pkg.evaluate_three_cycles(p, q, cand, score, bad)
    t0 = time.perf_counter()
    for _ in range(10):
        pkg.evaluate_three_cycles(p, q, cand, score, bad)
    el = (time.perf_counter() - t0) / 10
    out["three_cycles"] = {"candidates": int(len(cand)), "seconds_per_call": el,
                           "candidates_per_s": len(cand) / el}
    return out


def tempering_cpu_baseline(pkg, oracle, seconds: float = 10.0) -> dict:
    """The oracle's replay of the reference walk (tempering.py:48-110, its
    O(m^2) pair-loop delta) on config 3's replicas and temperatures, OpenMP
    over replicas on all host cores: whole sweeps of 150 proposals, as many as
    fit in about `seconds`."""
    import numpy as np

    replicas, ladder, steps = 4096, 64, 150
    cores = oracle.cpu_count()
    p, q = oracle.generate_states(SEED + 1, N, replicas)
    cols = np.stack((p, q), axis=2).reshape(replicas, 2 * N)
    temps = np.tile(np.geomspace(0.05, 2.0, ladder), replicas // ladder).astype(np.float32)
    k_move = pkg.tempering.move_table_size(temps)
    import torch

    chains = pkg.TemperingChains(cols, temps, seed=3)
    replay = oracle.TemperingReplay(cols, temps, 3, k_move=k_move, k_exchange=chains._exchange_table(ladder)[2])
    t0 = time.perf_counter()
    replay.advance(steps, threads=cores)
    sweeps = 1
    while time.perf_counter() - t0 < seconds * 0.5 and sweeps < 8:
        replay.advance(steps, threads=cores)
        sweeps += 1
    elapsed = time.perf_counter() - t0
    # one thread, as the reference's serial harnesses run: one sweep of the
    # first 256 replicas
    single = oracle.TemperingReplay(cols[:256], temps[:256], 3, k_move=k_move, k_exchange=2)
    t1 = time.perf_counter()
    single.advance(steps, threads=1)
    single_rate = 256 * steps / (time.perf_counter() - t1)
    # the device walk over the same sweeps, states, temperatures and seed,
    # then both continue with alternating sweeps and exchanges: every
    # replica's trajectory must match the replay's
    for _ in range(sweeps):
        chains.advance(steps)
    names = ("columns", "badness", "scores", "best_columns", "best_scores", "accepted", "removed_points")
    exact = all(np.array_equal(getattr(chains, name).cpu().numpy(), getattr(replay, name)) for name in names)
    for _ in range(2):
        exact &= chains.exchange(ladder) == replay.exchange(ladder)
        chains.advance(steps)
        replay.advance(steps, threads=cores)
    exact &= chains.exchange(ladder) == replay.exchange(ladder)
    torch.cuda.synchronize()
    exact &= all(np.array_equal(getattr(chains, name).cpu().numpy(), getattr(replay, name)) for name in names)
    return {"value": sweeps * steps * replicas / elapsed, "unit": "proposals/s", "cores": cores, "kind": "port",
            "sample": f"{sweeps} sweep(s) x {steps} proposals x {replicas} replicas (config 3's ladder), oracle C "
                      f"replay of tempering.walk with its pair-loop delta, OpenMP over replicas",
            "single_thread_proposals_per_s": single_rate, "trajectory_exact": exact,
            "trajectory_checked": f"{sweeps} sweeps, then 3 exchanges and 2 more sweeps alternating, all "
                                  f"{replicas} replicas"}


def move_cpu_baseline(pkg, oracle) -> dict:
    """CPU side of the delta harness: the oracle's C loop of
    accelerated.evaluate_two_switches (one thread) on the same state and
    anchors, timed per call, and exactness checks of the GPU outputs (the
    single-state scores and a sample of the resident tables)."""
    import numpy as np

    p, q, score, bad = _move_state(pkg)
    bad64 = np.asarray(bad, dtype=np.int64)
    exact = True
    for u in MOVE_ANCHORS:
        ref = oracle.evaluate_two_switches(p, q, u, score, bad64)
        got = pkg.score_two_switches_gpu(p, q, u, score, bad)
        legal = ref < (1 << 59)
        exact &= bool(np.array_equal(got >= (1 << 29), ~legal) and np.array_equal(got[legal], ref[legal]))
    t0 = time.perf_counter()
    for u in MOVE_ANCHORS:
        for _ in range(MOVE_REPEATS):
            oracle.evaluate_two_switches(p, q, u, score, bad64)
    call_us = (time.perf_counter() - t0) / (len(MOVE_ANCHORS) * MOVE_REPEATS) * 1e6
    idx = [0, MOVE_TABLE_STATES // 2, MOVE_TABLE_STATES - 1]
    hp, hq = oracle.generate_states(5, N, MOVE_TABLE_STATES)
    table = pkg.delta_table(hp[idx], hq[idx])
    table_exact = bool(np.array_equal(table, oracle.delta_table(hp[idx], hq[idx])))
    return {"call_us": call_us, "cores": 1, "kind": "port",
            "sample": "oracle C loop of accelerated.evaluate_two_switches, n=75, 5 anchors x 100 repeats",
            "two_switch_exact": exact, "table_sample_exact": table_exact}


def calibrate_copy() -> int:
    """A plain device-to-device copy of 256 MiB, five times: profiled with
    FETCH_SIZE / WRITE_SIZE it gives the counters' reading for known bytes."""
    import torch

    x = torch.ones(64 << 20, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": y.numel() * 4, "launches": 5}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-temper", action="store_true", help="skip the tempering, batched and learned legs")
    ap.add_argument("--temper-sweeps", type=int, default=10000,
                    help="sweeps of the tempering leg (config 3 is 10,000; fewer for counter profiles)")
    ap.add_argument("--legs", default="temper,batched,learned,near,search,moves",
                    help="comma list of the extra legs to run: temper, batched, learned, near, search, moves")
    ap.add_argument("--ab-seconds", type=float, default=20.0, help="wall-clock budget of each A/B side")
    ap.add_argument("--near-seconds", type=float, default=4.0,
                    help="wall-clock budget of each side of the near-solution A/B")
    ap.add_argument("--no-persistent", action="store_true",
                    help="config 4 without the persistent cooperative launch (per-sweep graphs only; for profilers)")
    ap.add_argument("--method", type=int, default=0, help="0 auto (direction histograms), 1 determinant")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device copy (the reference kernel for PMC byte counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the engine shards by whole "
              f"populations/ladders and runs no collectives", file=sys.stderr)
        return 2

    import torch

    import __graft_entry__ as entry

    pkg = entry.load_package()
    dev = pkg._lib.device(0)
    log(f"bench: device {torch.cuda.get_device_name(dev)}; generating n={N} P={POPULATION}")
    p, q = pkg.generate_states(SEED, N, POPULATION, device=dev)
    scores = torch.empty(POPULATION, dtype=torch.int32, device=dev)
    bad = torch.empty((POPULATION, 2 * N), dtype=torch.int32, device=dev)
    for _ in range(args.warmup):
        pkg.score_population_device(p, q, scores, bad, method=args.method, check=False)
    pkg.check_status(dev)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        pkg.score_population_device(p, q, scores, bad, method=args.method, check=False)
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    pkg.check_status(dev)
    kernel_ms = start.elapsed_time(end) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    value = POPULATION * args.steps / wall
    log(f"bench: {args.steps} steps, {ms_per_step:.3f} ms/step wall, {kernel_ms:.3f} ms/launch (events)")

    ops = slope_ops_per_state(N) if args.method != 1 else determinant_ops_per_state(N)
    achieved = POPULATION * ops / (kernel_ms * 1e-3) / 1e12
    roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_VALU_TOPS, "unit": "TOP/s",
                "frac": achieved / PEAK_VALU_TOPS, "traffic": traffic_from_profile(),
                "ops_per_state": ops, "kernel_ms": kernel_ms,
                "triple_tests_per_s": POPULATION * (2 * N) * (2 * N - 1) * (2 * N - 2) / 6 / (kernel_ms * 1e-3)}
    if args.method != 1:
        # the unit that binds the kernel: LDS lane accesses (bank conflicts
        # make most of the gap to the conflict-free peak)
        lds = POPULATION * slope_lds_accesses_per_state(N) / (kernel_ms * 1e-3) / 1e12
        roofline["lds"] = {"achieved": lds, "peak": PEAK_LDS_TACCESS, "unit": "T accesses/s",
                           "frac": lds / PEAK_LDS_TACCESS, "accesses_per_state": slope_lds_accesses_per_state(N)}
        # hardware counters of the same kernel (SQ_INSTS_VALU, SQ_LDS_*), as
        # fractions of the launch's cycles per CU
        prof = _profile_values()
        roofline["counters"] = {k: prof.get(k) for k in ("valu_issue_fraction", "lds_active_fraction",
                                                          "lds_bank_conflict_fraction")}

    extra = {"host_arrays": host_array_rates(pkg, torch, p, q)}
    legs = set() if args.no_temper else {x.strip() for x in args.legs.split(",") if x.strip()}
    unknown = legs - {"temper", "batched", "learned", "near", "search", "moves"}
    if unknown:
        log(f"bench: unknown legs {sorted(unknown)}")
        return 2
    if "temper" in legs:
        log("bench: tempering rate")
        extra["tempering_n75_64x64"] = tempering_rate(pkg, torch, sweeps=args.temper_sweeps)
    if "batched" in legs:
        log("bench: batched uneven sizes")
        extra["batched_uneven"] = batched_rate(pkg, torch, persistent=not args.no_persistent)
    if "learned" in legs:
        log("bench: learned vs uniform proposals")
        extra["learned_vs_uniform"] = learned_ab(pkg, args.ab_seconds)
    if "near" in legs:
        log("bench: learned vs uniform proposals from a perturbed near-solution")
        near = learned_ab_near_solution(pkg, args.near_seconds)
        if near is not None:
            extra["learned_vs_uniform_near_solution"] = near
    if "search" in legs:
        log("bench: device GA and exhaustive beam")
        extra["search_engines"] = search_rates(pkg)
    if "moves" in legs:
        log("bench: move scores")
        extra["move_scores"] = move_rates(pkg, torch)
    cpu = None
    exact = None
    if not args.no_cpu:
        # the CPU leg: the oracle as the baseline and as the parity checker
        # of this run's outputs
        import numpy as np

        oracle = entry.load_oracle()
        idx = np.arange(0, POPULATION, POPULATION // 512)  # the 512-state subsample
        hp, hq = p[idx].cpu().numpy(), q[idx].cpu().numpy()
        ref_s, ref_b = oracle.score_population(hp, hq)
        exact = bool(np.array_equal(scores[idx].cpu().numpy(), ref_s) and
                     np.array_equal(bad[idx].cpu().numpy(), ref_b))
        if not exact:
            log("bench: PARITY FAILURE on the subsample")
        log("bench: CPU baseline")
        cpu = cpu_baseline()
        cpu["gpu_over_cpu"] = value / cpu["value"]
        if "tempering_n75_64x64" in extra:
            log("bench: tempering CPU baseline")
            t_cpu = tempering_cpu_baseline(pkg, oracle)
            extra["tempering_n75_64x64"]["cpu_baseline"] = t_cpu
            extra["tempering_n75_64x64"]["gpu_over_cpu"] = (extra["tempering_n75_64x64"]["proposals_per_s"] /
                                                           t_cpu["value"])
            if not t_cpu["trajectory_exact"]:
                log("bench: PARITY FAILURE in the tempering trajectories")
                exact = False
        if "batched_uneven" in extra:
            log("bench: batched CPU baseline")
            b_cpu = batched_cpu_baseline(pkg, oracle)
            extra["batched_uneven"]["cpu_baseline"] = b_cpu
            extra["batched_uneven"]["gpu_over_cpu"] = extra["batched_uneven"]["proposals_per_s"] / b_cpu["value"]
        if "move_scores" in extra:
            moves_cpu = move_cpu_baseline(pkg, oracle)
            extra["move_scores"]["cpu_baseline"] = moves_cpu
            extra["move_scores"]["host_call_speedup"] = moves_cpu["call_us"] / extra["move_scores"]["host_call_us"]
            if not (moves_cpu["two_switch_exact"] and moves_cpu["table_sample_exact"]):
                log("bench: PARITY FAILURE in the move scores")
                exact = False

    line = {
        "metric": "n=75 states scored/s + time-to-E≤12 (learned vs uniform proposals), 1 GPU",
        "value": value, "unit": "states/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic",
        "config": {"workload": "n=75 population scoring: exact collinear-triple count + badness int32[150] per state",
                   "n": N, "population": POPULATION, "seed": SEED,
                   "method": "direction-histogram" if args.method != 1 else "determinant"},
        "parity_subsample_exact": exact,
        "device": {"name": torch.cuda.get_device_name(dev),
                   "compute_units": torch.cuda.get_device_properties(dev).multi_processor_count},
        "roofline": roofline,
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(line), flush=True)
    return 0 if exact is not False else 1


if __name__ == "__main__":
    raise SystemExit(main())
