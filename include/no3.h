/*
 * no3 — C ABI of the gfx950 (MI355X) engine for the no-three-in-line hot
 * path: exact collinear-triple scoring of P/Q permutation states, 2-switch
 * move deltas, and replica-exchange tempering.
 *
 * Conventions (all entry points):
 *   - returns 0 on success, a negative status otherwise (N3_E*); nothing
 *     throws across the ABI; n3_last_error() describes the last failure;
 *   - array arguments are DEVICE pointers owned by the caller unless the name
 *     says host_; `stream` is a hipStream_t (NULL = the null stream); calls
 *     enqueue work and return without synchronising unless stated;
 *   - a state of size n is two int32 permutations P, Q of 0..n-1 stored as
 *     rows of planar arrays p[count][n], q[count][n]; slot s of the state is
 *     the point (s >> 1, s & 1 ? Q[s >> 1] : P[s >> 1]) and m = 2n;
 *   - columns outside [0, n) are rejected with N3_EINVAL; nothing else about
 *     the states is assumed (coincident points P[r] == Q[r] are counted
 *     exactly as the reference's determinant loops count them).
 */
#ifndef NO3_H
#define NO3_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define N3_OK 0
#define N3_EINVAL (-22)   /* bad argument (shape, range, n) */
#define N3_ENODEV (-19)   /* no usable gfx950 device */
#define N3_ENOMEM (-12)
#define N3_EHIP (-1000)   /* a HIP call failed; -1000 - hipError_t */

/* Library version (major * 10000 + minor * 100 + patch). */
int n3_version(void);

/* Text of the last error recorded by this thread (never NULL). */
const char* n3_last_error(void);

/* Layout checks for bindings (no counterpart in the reference): the size in
 * bytes of n3_chains (which = 0) or n3_group (which = 1), -1 otherwise; and
 * the byte offsets of n3_chains' members in declaration order (writes at most
 * `capacity`, returns the member count).  Host only, no device needed. */
int64_t n3_struct_size(int which);
int n3_chains_offsets(int64_t* offsets, int capacity);

/* Checks that `device` is a gfx950 GPU and makes it current.  No reference
 * counterpart beyond gpu.py:50-51 ("no CUDA device is available"). */
int n3_device_init(int device);

/* Direction table of size n (see csrc/no3_common.h): fills `host_table`
 * (uint16[(2n-1)^2]) when not NULL and reports the number of primitive
 * directions.  Host-only.  Basis: verifier_b.py:8-19. */
int n3_dirtable(int n, uint16_t* host_table, int* directions, int* hist_words);

/* Seeded population of `count` states, state i being the generator's state
 * (first_index + i) for `seed` (Philox4x32-10 Fisher-Yates shuffles of P and
 * Q, each collision P[r] == Q[r] repaired by swapping Q[r] with Q of a random
 * other row).  Replaces the host loop of cli.py:203-205 / state.py:88-95. */
int n3_generate_states(uint64_t seed, int n, int64_t count, int64_t first_index, int32_t* p, int32_t* q,
                       void* stream);

#define N3_SCORE_AUTO 0          /* direction histograms, determinant for coincident points */
#define N3_SCORE_DETERMINANT 1   /* i<j<k determinant loop only */
#define N3_SCORE_SLOPE 2         /* direction histograms (as AUTO) */

/* Exact collinear-triple count of every state, and optionally the number of
 * collinear triples through every slot (int32 badness[count][2n]).
 * Replaces the kernels score_population (gpu.py:13-42, called by
 * score_population_gpu gpu.py:69-77) and score_population_badness
 * (gpu_device_search.py:118-161); CPU equivalent accelerated.py:14-37. */
int n3_score_population(const int32_t* p, const int32_t* q, int n, int64_t count, int32_t* scores,
                        int32_t* badness_or_null, int method, void* stream);

/* Synchronises `stream` and reports whether the scoring launches since the
 * last call met a column outside [0, n) (N3_EINVAL; the affected scores are
 * -1).  The flag is one word per device, shared by every stream and thread:
 * a launch on another stream that has not finished is not covered, and an
 * error it sets may be reported here.  The reference has no range check
 * (gpu.py:61-66 checks shapes only). */
int n3_score_status(void* stream);

/* Score after every 2-switch of slot u with slot v of ONE state, int32 with
 * the sentinel 1<<30 for illegal v.  Replaces two_switch_scores
 * (gpu_delta.py:18-75; host wrapper score_two_switches_gpu gpu_delta.py:86-116). */
int n3_two_switch_scores(const int32_t* p, const int32_t* q, const int32_t* badness, int n, int u,
                         int current_score, int32_t* scores, void* stream);

/* n3_two_switch_scores without any synchronisation: the call only enqueues
 * the kernel, which ORs 1 into *status (a device int32 the caller zeroes
 * beforehand) when a column lies outside [0, n).  A caller that stages its
 * inputs, the zeroed status word and the scores in one buffer then needs
 * one copy in, one copy out and one synchronisation per call.  Same kernel
 * and scores as n3_two_switch_scores (gpu_delta.py:18-116). */
int n3_two_switch_scores_status(const int32_t* p, const int32_t* q, const int32_t* badness, int n, int u,
                                int current_score, int32_t* scores, int32_t* status, void* stream);

/* accelerated.evaluate_two_switches (accelerated.py:122-195) on the device:
 * int64 scores[2n] with the sentinel 1<<60; badness is int64[2n]. */
int n3_evaluate_two_switches(const int32_t* p, const int32_t* q, const int64_t* badness, int n, int u,
                             int64_t current_score, int64_t* scores, void* stream);

/* accelerated.evaluate_transpositions (accelerated.py:40-119) on the device:
 * int64 scores[n] with the sentinel 1<<60; badness is int64[2n]. */
int n3_evaluate_transpositions(const int32_t* p, const int32_t* q, const int64_t* badness, int n, int layer,
                               int row, int64_t current_score, int64_t* scores, void* stream);

/* accelerated.evaluate_three_cycles (accelerated.py:198-276) on the device:
 * int64 scores[count] for candidates int32[count][4] = (u, v, w,
 * orientation), orientation 0 moving u <- col v, v <- col w, w <- col u and
 * 1 moving u <- col w, w <- col v, v <- col u (search.py:38-65 lists the
 * legal ones); badness is int64[2n]. */
int n3_evaluate_three_cycles(const int32_t* p, const int32_t* q, const int64_t* badness, int n,
                             const int32_t* candidates, int64_t count, int64_t current_score, int64_t* scores,
                             void* stream);

/* Full 2-switch table of a population: table[s][u][v] is
 * evaluate_two_switches(state s, u, scores[s], badness[s])[v] as int32 with
 * the sentinel 1<<30 (the all-anchor sweep of analysis.py:57-74).  `scores`
 * and `badness` (int32) are the states' current values. */
int n3_delta_table(const int32_t* p, const int32_t* q, const int32_t* scores, const int32_t* badness, int n,
                   int64_t count, int32_t* table, void* stream);

/* ---- replica-exchange tempering (tempering.py:19-225) ------------------ */

/* Device state of R replicas of size n, slot-ordered like the reference's
 * TemperingChains arrays (tempering.py:138-152).  All pointers are device
 * pointers owned by the caller. */
typedef struct n3_chains {
    int n;
    int replicas;
    int32_t* columns;        /* [R][2n] interleaved P0,Q0,P1,Q1,... */
    int32_t* badness;        /* [R][2n] */
    int32_t* scores;         /* [R] */
    int32_t* best_columns;   /* [R][2n] */
    int32_t* best_scores;    /* [R] */
    int32_t* accepted;       /* [R] */
    const int32_t* reference;/* [2n] umbrella reference state */
    int32_t* removed_points; /* [R] */
    const uint32_t* move_thresholds;  /* [R][k_move]: floor(2^32 exp(-d/T_r)), non-increasing in d (the walk's
                                         early rejection relies on it) */
    int k_move;
    const uint32_t* exchange_thresholds; /* [R][k_exchange] for pair (r, r+1) */
    const int8_t* exchange_sign;         /* [R] sign of 1/T_r - 1/T_{r+1} */
    int k_exchange;
    uint32_t* counters;      /* [4]: sweep, round, exchanges (low), exchanges (high) */
    int64_t* proposals;      /* [2]: valid proposals (total), accepted moves (total) */
    int block_waves;         /* replicas (wavefronts) per workgroup; 0 = automatic */
    int32_t* log;            /* optional proposal log (NULL = off): per valid proposal
                                8 int32 (badness u, badness v, energy, slot, delta,
                                accepted, u | v << 16, column u | column v << 16);
                                each replica reserves `steps` consecutive entries per
                                sweep (entry = reservation + step); invalid proposals
                                and reservations past the capacity write nothing */
    int64_t log_capacity;    /* entries */
    uint64_t* log_count;     /* entries reserved (may exceed the capacity) */
    /* umbrella potentials (tempering.py:92-95; NULL strengths = none): a
       move's energy is E = delta + strengths[r] (2 (removed_r - centers[r]) dd
       + dd^2), dd the change of removed_points, and it is accepted when
       E <= 0 or its 32-bit word is below floor(2^32 exp(-E inverse_temperatures[r])).
       Accepted by n3_temper_advance and by n3_temper_run without exchanges
       only (the reference disables exchanges with umbrellas, tempering.py:173-174). */
    const float* centers;                /* [R] */
    const float* strengths;              /* [R] or NULL */
    const double* inverse_temperatures;  /* [R]: 1 / (double)T_r, required with strengths */
} n3_chains;

/* `steps` proposals per replica (one sweep), TemperingChains.advance
 * (tempering.py:158-167, kernel walk tempering.py:29-115), with the umbrella
 * term when chains->strengths is set.  Proposal u, v and the acceptance word
 * come from Philox keyed by `seed` with counter (step, sweep, slot, "MOVE");
 * `sweep` is read from counters[0], which the call then increments. */
int n3_temper_advance(const n3_chains* chains, int steps, int bad_bias, uint64_t seed, void* stream);

/* Adjacent-temperature exchange inside ladders of `ladder` slots
 * (TemperingChains.exchange, tempering.py:169-191) on the device; pairs start
 * at round % 2 where round = counters[1], which the call then increments.
 * Accepted pairs swap columns, badness, scores and removed_points. */
int n3_temper_exchange(const n3_chains* chains, int ladder, uint64_t seed, void* stream);

/* `sweeps` x (advance(steps); exchange(ladder) if ladder > 1) replayed from a
 * captured hipGraph (the epoch loop of tempering.py:275-278).  The graph is
 * cached per (chains, steps, ladder, bad_bias, seed, stream) and rebuilt when
 * they change.  With umbrella strengths, ladder must be 0 or 1. */
int n3_temper_run(const n3_chains* chains, int sweeps, int steps, int ladder, int bad_bias, uint64_t seed,
                  void* stream);

/* One population of a batched run: its chains, exchange ladder size (0 or 1
 * = no exchange), proposals per replica per sweep, bad_bias and seed. */
typedef struct n3_group {
    n3_chains chains;
    int ladder;
    int steps;
    int bad_bias;
    uint64_t seed;
} n3_group;

#define N3_MAX_GROUPS 8

/* Populations of different sizes (2 <= n <= 100 each) advanced together:
 * every sweep is ONE walk launch over all groups' replicas (each group with
 * its own LDS direction table and chains.block_waves replicas per
 * workgroup), one exchange launch and one counter update, replayed `sweeps`
 * times from a captured hipGraph.  Group g's trajectory is bit-identical to
 * n3_temper_run on that group alone.  No reference counterpart (the
 * reference caps n at 75 and runs one size per process, tempering.py:124). */
int n3_temper_batched_run(const n3_group* groups, int group_count, int sweeps, void* stream);

/* The same run as n3_temper_batched_run in ONE persistent cooperative
 * launch for all `sweeps` sweeps: one block per compute unit, each sweep a
 * queue of work items (W_g replicas of group g, largest n first, W_g =
 * chains.block_waves, or as many as fit in LDS when 0) drained from a
 * device counter, a grid barrier, the exchanges of every group inside the
 * kernel, and a second barrier.
 * Trajectories and counters are bit-identical to n3_temper_batched_run's.
 * No reference counterpart (config 4 of the benchmark). */
int n3_temper_batched_persistent(const n3_group* groups, int group_count, int sweeps, void* stream);

/* Drops cached graphs and persistent-run plans (all of them); synchronises
 * the current device first. */
void n3_temper_release(void);

/* ---- learned move proposals (no reference counterpart) ----------------- */

/* The proposal network: 8 features -> 128 hidden (ReLU) -> 1 logit.  Its
 * parameters are one bf16 vector of n3_mlp_params() entries laid out as
 * W1[128][8] | b1[128] | w2[128] | b2; every product accumulates in fp32 on
 * MFMA (v_mfma_f32_32x32x8_bf16). */
#define N3_MLP_INPUTS 8
#define N3_MLP_HIDDEN 128
int n3_mlp_params(void);

/* Features (bf16[8]) of a candidate point with collinear-pair count `count`
 * at temperature exp(log_temperature) in a state of energy `energy`.  Host. */
int n3_mlp_features(int count, float log_temperature, int energy, uint16_t* host_out);

/* Logits of `rows` feature rows x (bf16[rows][8]) into out (fp32[rows]). */
int n3_mlp_forward(const uint16_t* params, const uint16_t* x, int64_t rows, float* out, void* stream);

/* Gradient of the mean pair loss BCE(sigmoid(g(xu) + g(xv) - b2), y) over
 * `pairs` tuples, ADDED to grad (fp32[n3_mlp_params()]); the summed loss is
 * added to loss[0].  Per-wavefront partial sums are added in a fixed order,
 * so the result is deterministic.  Calls that overlap on one device race on
 * the library's partial-sum buffer. */
int n3_mlp_grad(const uint16_t* params, const uint16_t* xu, const uint16_t* xv, const float* y, int64_t pairs,
                float* grad, float* loss, void* stream);

/* n3_mlp_grad without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight training loops).  The per-block partial sums live in a
 * file-scope device array of the library (641 KiB per GPU, held until the
 * process ends; its address is looked up by n3_device_init or by the first
 * call on a device) instead of a buffer the library allocates.  The scratch
 * holds the partials of at most 128 blocks and the gradient grid never
 * exceeds 128 blocks (they stride over the tuples), so every pairs >= 1
 * fits.  Same result bits as n3_mlp_grad.  All calls on one GPU share the
 * array: calls that overlap on one GPU race. */
int n3_mlp_grad_noalloc(const uint16_t* params, const uint16_t* xu, const uint16_t* xv, const float* y,
                        int64_t pairs, float* grad, float* loss, void* stream);

/* One Adam step on the fp32 master copy; writes the bf16 parameters. */
int n3_mlp_adam(float* master, uint16_t* params, const float* grad, float* m, float* v, int step, float lr,
                float beta1, float beta2, float eps, void* stream);

/* Integer proposal weights table[t][e][b] = clamp(round(128 exp(clamp(g,
 * -6, 5.5))), 1, 2^15) of the network's logit g for temperature index t
 * (log_temps[t]), energy e in [0, max_energy] and count b in [0, 63]. */
int n3_mlp_weight_table(const uint16_t* params, const float* log_temps, int temps, int max_energy, uint32_t* table,
                        void* stream);

/* host_out[r][d + k] = floor(2^32 exp(-d / T_r)), saturating at 2^63, for
 * d in [-k, k): the acceptance words of the learned walk. */
int n3_metropolis_thresholds64(const float* host_temperatures, int replicas, int k, uint64_t* host_out);

/* `sweeps` x (learned-proposal walk of `steps` proposals; exchange(ladder)
 * if ladder > 1), graph-captured.  A proposal of slot r is, with probability
 * 0.2, a uniform pair of slots, otherwise both slots drawn in proportion to
 * w = clamp(weights[temp_index[r]][E][badness], 1, 2^15) (the softmax of the
 * network's additive pair logit over all pairs); acceptance is the
 * Metropolis-Hastings test on the exact integer delta, evaluated exactly in
 * integers (see csrc/learned.hip).  No counterpart in the reference (the
 * nearest is bad_bias, tempering.py:51-55). */
int n3_temper_run_learned(const n3_chains* chains, const uint32_t* weights, const int32_t* temp_index,
                          int max_energy, const uint64_t* thr64, int k64, int sweeps, int steps, int ladder,
                          uint64_t seed, void* stream);

/* ---- point-network proposals (no reference counterpart) ----------------
 * The network scores a point from 8 features: its collinear-pair count
 * (b/32 up to 63, min(b, 7)/4, b > 0), its row/128 and column/128, ln T/4
 * (temperature code lt_q = round(32 ln T) per replica) and the state's
 * energy (E/256 up to 510, E/64 up to 63), each a multiple of 1/128.  With
 * W1 and b1 on the same grid (n3_mlp_adam keeps them there), the hidden
 * layer's fp32 MFMA sums are exact and the output layer is a fixed-order fma
 * chain, so a CPU replays every weight bit for bit; a point's weight is
 * expw[clamp(floor(64 g), -384, 352) + 384].  n <= 127. */

/* expw (uint32[737]): clamp(round(128 exp(q / 64)), 1, 2^15), q = -384..352.
 * Host. */
int n3_point_exp_table(uint32_t* host_out);

/* Walk constants (float[260]) from the bf16 parameters: b1 and w2 in the
 * order the MFMA output lanes read them, then b2.  N3_EINVAL when W1 or b1
 * is off the 1/128 grid.  Host. */
int n3_point_constants(const uint16_t* host_params, float* host_consts);

/* Integer weight of every slot of `count` states (columns, badness int32
 * [count][2n], scores [count], lt_q [count]) into weights int32[count][2n],
 * from the same device function the walk uses. */
int n3_point_weights(const int32_t* columns, const int32_t* badness, const int32_t* scores, const int32_t* lt_q,
                     int n, int64_t count, const uint16_t* params, const float* consts, const uint32_t* expw,
                     int32_t* weights, void* stream);

/* `sweeps` x (point-proposal walk of `steps` proposals; exchange(ladder) if
 * ladder > 1), graph-captured.  A proposal is, with probability 0.2, a
 * uniform pair of slots, otherwise both slots drawn in proportion to the
 * point weights of the current state (64-bit inverse-CDF targets).
 * Acceptance is delayed: the Metropolis test on the exact integer delta
 * (Philox word < thr64(delta)), then, for moves that pass it, the moved
 * state's weights and the proposal-ratio test h N S1'^2 < 2^32 N' S1^2
 * (N = S1^2 + 4 m^2 w(u) w(v)), i.e. acceptance min(1, e^(-delta/T))
 * min(1, q'/q).  Philox counters (step, sweep, slot, "LRND") and
 * (step, sweep, slot, "HAST").  No counterpart in the reference (the
 * nearest is bad_bias, tempering.py:51-55). */
int n3_temper_run_learned_point(const n3_chains* chains, const uint16_t* params, const float* consts,
                                const uint32_t* expw, const int32_t* lt_q, const uint64_t* thr64, int k64, int sweeps,
                                int steps, int ladder, uint64_t seed, void* stream);

/* Drops the learned walk's cached graphs. */
void n3_learned_release(void);

/* Host tables of the integer acceptance tests (no device work).
 * Metropolis: host_out[r][d] = floor(2^32 exp(-d / T_r)) for 1 <= d < k
 * (clamped to 2^32-1); a proposal with integer energy change d > 0 is
 * accepted iff its Philox word < host_out[r][min(d, k-1)] (the law of
 * log(U) < -d/T in tempering.py:92-95).
 * Exchange: for the pair (a, a+1) of a ladder, c = 1/T_a - 1/T_{a+1},
 * host_sign[a] = sign(c), host_out[a][d] = floor(2^32 exp(-|c| d)); the pair
 * swaps iff c (E_a - E_{a+1}) >= 0 or the word < host_out[a][min(|E_a -
 * E_{a+1}|, k-1)] (tempering.py:181-182). */
int n3_metropolis_thresholds(const float* host_temperatures, int replicas, int k, uint32_t* host_out);
int n3_exchange_thresholds(const float* host_temperatures, int replicas, int ladder, int k, uint32_t* host_out,
                           int8_t* host_sign);

/* ---- device search engines (GA / beam) over the scorer ------------------ */

/* Kernel mutate_population (gpu_device_search.py:27-113): child c < elite_count
 * copies parent c; every other child copies a parent drawn from the first
 * quality_count parents (probability 160/256) or from all elite_count, then
 * applies `strength` random legal 2-switches, strength = 1 + [0,
 * active_strength) or, one time in 32, 1 + [0, max_strength).  Each switch
 * tries up to 32 draws of (u, v); u is drawn in proportion to the parent's
 * badness int32[elite_count][2n] with probability bad_bias/100, uniformly
 * otherwise, and v uniformly.  Draws are Philox keyed by `seed` with counter
 * (child, generation, draw, "MUTA") in place of the reference's xorshift64*.
 * Parents and children are P/Q arrays int32[count][n]. */
int n3_mutate_population(const int32_t* parent_p, const int32_t* parent_q, const int32_t* parent_badness,
                         int elite_count, int quality_count, int n, int64_t population, int active_strength,
                         int max_strength, int bad_bias, uint64_t seed, int generation, int32_t* child_p,
                         int32_t* child_q, void* stream);

/* Kernel enumerate_switches (gpu_exhaustive_beam.py:17-57): child i copies
 * parent descriptors[i][0] and applies the 2-switch of slots
 * (descriptors[i][1], descriptors[i][2]) when it is legal; valid[i] says
 * whether it was.  A descriptor out of range is an error (N3_EINVAL);
 * synchronises `stream`. */
int n3_enumerate_switches(const int32_t* parent_p, const int32_t* parent_q, int parent_count,
                          const int32_t* descriptors, int64_t count, int n, int32_t* child_p, int32_t* child_q,
                          uint8_t* valid, void* stream);

/* Secondary keys of rank_elites (gpu_device_search.py:185-197) for badness
 * int32[count][2n]: the maximum and the sum of squares of each row. */
int n3_elite_keys(const int32_t* badness, int64_t count, int n, int32_t* max_badness, int64_t* sum_squares,
                  void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NO3_H */
