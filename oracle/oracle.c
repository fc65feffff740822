/*
 * CPU oracle for the no-three-in-line hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker (or as the timed CPU
 * baseline), never as the computation being measured or shipped.
 *
 * Every routine is a plain restatement of the reference algorithm it names
 * (file:line in the reference repository MilliganCo/no-three-in-line), kept in
 * the reference's own loop order so that integer results are bit-identical:
 *
 *   n3o_score_and_badness        accelerated.py:14-37   (i<j<k determinant loop)
 *   n3o_score_slope              verifier_b.py:8-33     (normalized-direction counting form)
 *   n3o_evaluate_transpositions  accelerated.py:40-119
 *   n3o_evaluate_two_switches    accelerated.py:122-195
 *   n3o_evaluate_three_cycles    accelerated.py:198-276
 *   n3o_walk (tempering replay)  tempering.py:48-110    (walk kernel step semantics)
 *   n3o_exchange                 tempering.py:169-191   (adjacent-temperature swaps)
 *   n3o_mutate_population        gpu_device_search.py:27-113 (GA children)
 *   n3o_enumerate_switches       gpu_exhaustive_beam.py:17-57 (enumerated 2-switch children)
 *   n3o_advance_learned          the walk above with learned pair proposals
 *                                and the Metropolis-Hastings test (no reference
 *                                counterpart; the rule is stated at the function)
 *   n3o_point_weights,           the point network's exact integer weights and
 *   n3o_advance_point            the walk with point-network proposals and the
 *                                delayed (two-stage) Metropolis-Hastings test
 *
 * Deliberate, documented differences from the reference (the same in the
 * product and in this oracle, so parity is exact):
 *   - random numbers come from Philox4x32-10 (Salmon et al., SC'11) keyed by
 *     the seed, with counters (move, sweep, slot, tag), instead of xorshift64*
 *     (gpu_device_search.py:18-25) and the host PCG64 (tempering.py:145,182);
 *   - Metropolis / exchange acceptance compares a 32-bit Philox word against a
 *     precomputed integer threshold floor(2^32 * exp(-x)) instead of
 *     log(U) < -x in double (tempering.py:92-95,181-182): same acceptance law
 *     to within 2^-32, and free of libm differences between host and device;
 *   - states for the benchmark populations come from the two-permutation
 *     generator below (seeded Fisher-Yates + collision repair by a swap with a
 *     random other row), not state.random_state's rejection loop.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define N3O_SENTINEL64 ((int64_t)1 << 60)
#define N3O_MAXM 512

/* ------------------------------------------------------------------------ */
/* Philox4x32-10                                                              */
/* ------------------------------------------------------------------------ */

void n3o_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

static inline uint32_t mulhi32(uint32_t a, uint32_t b)
{
    return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

#define TAG_GEN 0x47454E53u  /* "GENS" */
#define TAG_MOVE 0x4D4F5645u /* "MOVE" */
#define TAG_BIAS 0x42494153u /* "BIAS" */
#define TAG_EXCH 0x45584348u /* "EXCH" */
#define TAG_LEARN 0x4C524E44u /* "LRND" */

/* ------------------------------------------------------------------------ */
/* Seeded two-permutation state generator                                   */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint32_t key[2];
    uint32_t ctr[4];
    uint32_t buf[4];
    int pos;
} gen_stream;

static uint32_t gen_draw(gen_stream* s)
{
    if (s->pos == 4) {
        n3o_philox4x32_10(s->ctr, s->key, s->buf);
        s->ctr[2] += 1u;
        s->pos = 0;
    }
    return s->buf[s->pos++];
}

/* State `index` of the population for `seed`: s1, s2 are Fisher-Yates
 * shuffles (j = floor(u32 * (i+1) / 2^32)); every row with s1[i] == s2[i] is
 * repaired, in increasing i, by swapping s2[i] with s2[r] for a random row
 * r != i.  Such a swap can never create a new collision (s2[r] != s2[i] ==
 * s1[i], and row r receives s1[i] != s1[r]), so one pass suffices. */
void n3o_generate_state(uint64_t seed, int64_t index, int n, int32_t* p, int32_t* q)
{
    gen_stream s;
    s.key[0] = (uint32_t)seed;
    s.key[1] = (uint32_t)(seed >> 32);
    s.ctr[0] = (uint32_t)index;
    s.ctr[1] = (uint32_t)((uint64_t)index >> 32);
    s.ctr[2] = 0;
    s.ctr[3] = TAG_GEN;
    s.pos = 4;
    for (int i = 0; i < n; ++i) {
        p[i] = i;
        q[i] = i;
    }
    for (int i = n - 1; i >= 1; --i) {
        int j = (int)mulhi32(gen_draw(&s), (uint32_t)(i + 1));
        int32_t t = p[i];
        p[i] = p[j];
        p[j] = t;
    }
    for (int i = n - 1; i >= 1; --i) {
        int j = (int)mulhi32(gen_draw(&s), (uint32_t)(i + 1));
        int32_t t = q[i];
        q[i] = q[j];
        q[j] = t;
    }
    for (int i = 0; i < n; ++i) {
        if (q[i] == p[i]) {
            int r = (int)mulhi32(gen_draw(&s), (uint32_t)(n - 1));
            if (r >= i) r += 1;
            int32_t t = q[i];
            q[i] = q[r];
            q[r] = t;
        }
    }
}

void n3o_generate_states(uint64_t seed, int n, int64_t count, int64_t first_index, int32_t* p, int32_t* q)
{
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < count; ++s) n3o_generate_state(seed, first_index + s, n, p + s * n, q + s * n);
}

/* ------------------------------------------------------------------------ */
/* Full scores                                                                */
/* ------------------------------------------------------------------------ */

static void interleave(const int32_t* p, const int32_t* q, int n, int32_t* rows, int32_t* cols)
{
    for (int r = 0; r < n; ++r) {
        rows[2 * r] = r;
        cols[2 * r] = p[r];
        rows[2 * r + 1] = r;
        cols[2 * r + 1] = q[r];
    }
}

/* accelerated.py:14-37.  The collinear test and the (i, j, k) order are the
 * reference's; the k loop keeps per-k hit flags so that the compiler can
 * vectorise it, and the i and j counters are credited once per (i, j). */
int64_t n3o_score_and_badness(const int32_t* p, const int32_t* q, int n, int64_t* badness)
{
    const int m = 2 * n;
    int32_t rows[N3O_MAXM], cols[N3O_MAXM], bad[N3O_MAXM];
    interleave(p, q, n, rows, cols);
    for (int i = 0; i < m; ++i) bad[i] = 0;
    int64_t score = 0;
    for (int i = 0; i < m - 2; ++i) {
        const int32_t ri = rows[i], ci = cols[i];
        for (int j = i + 1; j < m - 1; ++j) {
            const int32_t dx1 = rows[j] - ri;
            const int32_t dy1 = cols[j] - ci;
            int32_t hits = 0;
            for (int k = j + 1; k < m; ++k) {
                const int32_t h = (dx1 * (cols[k] - ci) == dy1 * (rows[k] - ri));
                bad[k] += h;
                hits += h;
            }
            bad[i] += hits;
            bad[j] += hits;
            score += hits;
        }
    }
    if (badness)
        for (int i = 0; i < m; ++i) badness[i] = bad[i];
    return score;
}

/* Population form of the above (the reference's benchmark_gpu CPU leg,
 * gpu.py:106-109, is the same loop over states), parallel over states. */
void n3o_score_population(const int32_t* p, const int32_t* q, int n, int64_t count, int32_t* scores,
                          int32_t* badness, int threads)
{
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t s = 0; s < count; ++s) {
        int64_t bad[N3O_MAXM];
        const int64_t e = n3o_score_and_badness(p + s * n, q + s * n, n, bad);
        scores[s] = (int32_t)e;
        if (badness)
            for (int i = 0; i < 2 * n; ++i) badness[s * 2 * n + i] = (int32_t)bad[i];
    }
}

static int64_t gcd64(int64_t a, int64_t b)
{
    if (a < 0) a = -a;
    if (b < 0) b = -b;
    while (b) {
        int64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

/* Counting form of verifier_b.find_collinear_triple (verifier_b.py:8-33):
 * per anchor, every other point's primitive undirected direction is counted;
 * the anchor lies in sum_d C(cnt_d, 2) collinear triples and every triple is
 * seen from its three members.  Returns -1 if two points coincide (the
 * reference raises "duplicate point has no direction"). */
int64_t n3o_score_slope(const int32_t* p, const int32_t* q, int n, int64_t* badness)
{
    const int m = 2 * n;
    int32_t rows[N3O_MAXM], cols[N3O_MAXM];
    interleave(p, q, n, rows, cols);
    int32_t span = 1;
    for (int i = 0; i < m; ++i) {
        int32_t a = cols[i] < 0 ? -cols[i] : cols[i];
        if (a + 1 > span) span = a + 1;
    }
    if (n > span) span = n;
    const int32_t w = 2 * span + 1;
    int32_t* counts = (int32_t*)calloc((size_t)w * (size_t)w, sizeof(int32_t));
    int64_t total = 0;
    for (int a = 0; a < m; ++a) {
        int64_t pairs = 0;
        for (int b = 0; b < m; ++b) {
            if (b == a) continue;
            int64_t dx = rows[b] - rows[a], dy = cols[b] - cols[a];
            int64_t g = gcd64(dx, dy);
            if (g == 0) {
                free(counts);
                return -1;
            }
            dx /= g;
            dy /= g;
            if (dx < 0 || (dx == 0 && dy < 0)) {
                dx = -dx;
                dy = -dy;
            }
            int32_t* c = &counts[(dx + span) * w + (dy + span)];
            pairs += *c;
            *c += 1;
        }
        for (int b = 0; b < m; ++b) {
            if (b == a) continue;
            int64_t dx = rows[b] - rows[a], dy = cols[b] - cols[a];
            int64_t g = gcd64(dx, dy);
            dx /= g;
            dy /= g;
            if (dx < 0 || (dx == 0 && dy < 0)) {
                dx = -dx;
                dy = -dy;
            }
            counts[(dx + span) * w + (dy + span)] = 0;
        }
        if (badness) badness[a] = pairs;
        total += pairs;
    }
    free(counts);
    return total / 3;
}

/* ------------------------------------------------------------------------ */
/* Incremental move scores                                                    */
/* ------------------------------------------------------------------------ */

static inline int det0(int32_t ar, int32_t ac, int32_t br, int32_t bc, int32_t cr, int32_t cc)
{
    return (br - ar) * (cc - ac) == (bc - ac) * (cr - ar);
}

/* new_affected of accelerated.py:161-193 for slots u, v moving to
 * (rows[u], cols[v]) and (rows[v], cols[u]). */
static int64_t new_affected_count(const int32_t* rows, const int32_t* cols, int m, int u, int v)
{
    const int32_t nur = rows[u], nuc = cols[v], nvr = rows[v], nvc = cols[u];
    int64_t count = 0;
    for (int i = 0; i < m - 1; ++i) {
        if (i == u) continue;
        int32_t ir, ic;
        if (i == v) {
            ir = nvr;
            ic = nvc;
        } else {
            ir = rows[i];
            ic = cols[i];
        }
        for (int j = i + 1; j < m; ++j) {
            if (j == u) continue;
            int32_t jr, jc;
            if (j == v) {
                jr = nvr;
                jc = nvc;
            } else {
                jr = rows[j];
                jc = cols[j];
            }
            if ((ir - nur) * (jc - nuc) == (ic - nuc) * (jr - nur)) count += 1;
        }
    }
    for (int i = 0; i < m - 1; ++i) {
        if (i == u || i == v) continue;
        for (int j = i + 1; j < m; ++j) {
            if (j == u || j == v) continue;
            if ((rows[i] - nvr) * (cols[j] - nvc) == (cols[i] - nvc) * (rows[j] - nvr)) count += 1;
        }
    }
    return count;
}

static int64_t joint_old_count(const int32_t* rows, const int32_t* cols, int m, int u, int v)
{
    int64_t joint = 0;
    for (int o = 0; o < m; ++o) {
        if (o == u || o == v) continue;
        if ((rows[v] - rows[u]) * (cols[o] - cols[u]) == (cols[v] - cols[u]) * (rows[o] - rows[u])) joint += 1;
    }
    return joint;
}

/* accelerated.py:40-119 */
void n3o_evaluate_transpositions(const int32_t* p, const int32_t* q, int n, int layer, int selected_row,
                                 int64_t current_score, const int64_t* badness, int64_t* out)
{
    const int m = 2 * n;
    int32_t rows[N3O_MAXM], cols[N3O_MAXM];
    interleave(p, q, n, rows, cols);
    const int u = 2 * selected_row + layer;
    for (int partner = 0; partner < n; ++partner) {
        out[partner] = N3O_SENTINEL64;
        if (partner == selected_row) continue;
        if (layer == 0) {
            if (p[partner] == q[selected_row] || p[selected_row] == q[partner]) continue;
        } else {
            if (q[partner] == p[selected_row] || q[selected_row] == p[partner]) continue;
        }
        const int v = 2 * partner + layer;
        const int64_t old_affected = badness[u] + badness[v] - joint_old_count(rows, cols, m, u, v);
        out[partner] = current_score - old_affected + new_affected_count(rows, cols, m, u, v);
    }
}

/* accelerated.py:122-195 */
void n3o_evaluate_two_switches(const int32_t* p, const int32_t* q, int n, int selected_index, int64_t current_score,
                               const int64_t* badness, int64_t* out)
{
    const int m = 2 * n;
    int32_t rows[N3O_MAXM], cols[N3O_MAXM];
    interleave(p, q, n, rows, cols);
    const int u = selected_index;
    const int selected_row = rows[u];
    const int selected_other = 2 * selected_row + (1 - (u & 1));
    for (int v = 0; v < m; ++v) {
        out[v] = N3O_SENTINEL64;
        if (v == u || rows[v] == selected_row) continue;
        const int partner_other = 2 * rows[v] + (1 - (v & 1));
        if (cols[v] == cols[selected_other] || cols[u] == cols[partner_other]) continue;
        const int64_t old_affected = badness[u] + badness[v] - joint_old_count(rows, cols, m, u, v);
        out[v] = current_score - old_affected + new_affected_count(rows, cols, m, u, v);
    }
}

/* accelerated.py:198-276: candidates (u, v, w, orientation); orientation 0
 * moves u <- col v, v <- col w, w <- col u; orientation 1 moves u <- col w,
 * w <- col v, v <- col u.  Sentinel 1<<60 is never produced (every candidate
 * is scored, as in the reference). */
void n3o_evaluate_three_cycles(const int32_t* p, const int32_t* q, int n, const int32_t* candidates, int64_t count,
                               int64_t current_score, const int64_t* badness, int64_t* out)
{
    const int m = 2 * n;
    int32_t rows[N3O_MAXM], cols[N3O_MAXM], changed[N3O_MAXM];
    interleave(p, q, n, rows, cols);
    for (int64_t ci = 0; ci < count; ++ci) {
        const int u = candidates[4 * ci], v = candidates[4 * ci + 1], w = candidates[4 * ci + 2];
        const int orientation = candidates[4 * ci + 3];
        int64_t joint_uv = 0, joint_uw = 0, joint_vw = 0;
        for (int o = 0; o < m; ++o) {
            if (o != u && o != v &&
                (rows[v] - rows[u]) * (cols[o] - cols[u]) == (cols[v] - cols[u]) * (rows[o] - rows[u]))
                joint_uv += 1;
            if (o != u && o != w &&
                (rows[w] - rows[u]) * (cols[o] - cols[u]) == (cols[w] - cols[u]) * (rows[o] - rows[u]))
                joint_uw += 1;
            if (o != v && o != w &&
                (rows[w] - rows[v]) * (cols[o] - cols[v]) == (cols[w] - cols[v]) * (rows[o] - rows[v]))
                joint_vw += 1;
        }
        const int64_t triple_uvw =
            (rows[v] - rows[u]) * (cols[w] - cols[u]) == (cols[v] - cols[u]) * (rows[w] - rows[u]) ? 1 : 0;
        const int64_t old_affected =
            badness[u] + badness[v] + badness[w] - joint_uv - joint_uw - joint_vw + triple_uvw;
        for (int i = 0; i < m; ++i) changed[i] = cols[i];
        if (orientation == 0) {
            changed[u] = cols[v];
            changed[v] = cols[w];
            changed[w] = cols[u];
        } else {
            changed[u] = cols[w];
            changed[w] = cols[v];
            changed[v] = cols[u];
        }
        int64_t new_affected = 0;
        for (int i = 0; i < m - 1; ++i) {
            if (i == u) continue;
            for (int j = i + 1; j < m; ++j) {
                if (j == u) continue;
                if ((rows[i] - rows[u]) * (changed[j] - changed[u]) == (changed[i] - changed[u]) * (rows[j] - rows[u]))
                    new_affected += 1;
            }
        }
        for (int i = 0; i < m - 1; ++i) {
            if (i == u || i == v) continue;
            for (int j = i + 1; j < m; ++j) {
                if (j == u || j == v) continue;
                if ((rows[i] - rows[v]) * (changed[j] - changed[v]) == (changed[i] - changed[v]) * (rows[j] - rows[v]))
                    new_affected += 1;
            }
        }
        for (int i = 0; i < m - 1; ++i) {
            if (i == u || i == v || i == w) continue;
            for (int j = i + 1; j < m; ++j) {
                if (j == u || j == v || j == w) continue;
                if ((rows[i] - rows[w]) * (changed[j] - changed[w]) == (changed[i] - changed[w]) * (rows[j] - rows[w]))
                    new_affected += 1;
            }
        }
        out[ci] = current_score - old_affected + new_affected;
    }
}

/* The full m x m 2-switch table of a population (analysis.py:57-74 evaluates
 * the same function for every anchor): table[s][u][v] =
 * evaluate_two_switches(state s, u)[v] with the int32 sentinel 1<<30 of
 * gpu_delta.py:30.  Scores and badness are recomputed per state. */
void n3o_delta_table(const int32_t* p, const int32_t* q, int n, int64_t count, int32_t* table)
{
    const int m = 2 * n;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t s = 0; s < count; ++s) {
        int64_t bad[N3O_MAXM], out[N3O_MAXM];
        const int64_t e = n3o_score_and_badness(p + s * n, q + s * n, n, bad);
        for (int u = 0; u < m; ++u) {
            n3o_evaluate_two_switches(p + s * n, q + s * n, n, u, e, bad, out);
            for (int v = 0; v < m; ++v)
                table[(s * m + u) * m + v] = out[v] >= N3O_SENTINEL64 ? (1 << 30) : (int32_t)out[v];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Replica-exchange tempering replay                                          */
/* ------------------------------------------------------------------------ */

/* Acceptance thresholds: thr[r][d] = floor(2^32 * exp(-d / T_r)) for d >= 1
 * (clamped to 2^32-1), thr[r][0] unused.  A move with integer energy change
 * d > 0 is accepted iff its Philox word is < thr[r][min(d, k-1)]. */
void n3o_metropolis_thresholds(const float* temperatures, int replicas, int k, uint32_t* thr)
{
    for (int r = 0; r < replicas; ++r) {
        const double t = (double)temperatures[r];
        thr[(int64_t)r * k] = 0xFFFFFFFFu;
        for (int d = 1; d < k; ++d) {
            double x = ldexp(exp(-(double)d / t), 32);
            thr[(int64_t)r * k + d] = x >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)floor(x);
        }
    }
}

/* Exchange thresholds for the pair (a, a+1) inside each ladder:
 * c = 1/T_a - 1/T_b (double from float32 temperatures), x = c * (E_a - E_b);
 * accept iff x >= 0 or word < thr[a][min(|E_a - E_b|, k-1)] with
 * thr[a][d] = floor(2^32 * exp(-|c| * d)).  The sign of c is stored in
 * sign[a] (+1 / -1 / 0). */
void n3o_exchange_thresholds(const float* temperatures, int replicas, int ladder, int k, uint32_t* thr, int8_t* sign)
{
    for (int a = 0; a < replicas; ++a) {
        int b = a + 1;
        if ((a % ladder) == ladder - 1) b = a;
        const double c = 1.0 / (double)temperatures[a] - 1.0 / (double)temperatures[b];
        sign[a] = c > 0 ? 1 : (c < 0 ? -1 : 0);
        const double ac = fabs(c);
        thr[(int64_t)a * k] = 0xFFFFFFFFu;
        for (int d = 1; d < k; ++d) {
            double x = ldexp(exp(-ac * (double)d), 32);
            thr[(int64_t)a * k + d] = x >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)floor(x);
        }
    }
}

/* floor(2^32 exp(-x)) for x > 0: the engine's umbrella acceptance word.
 * Cody-Waite reduction -x = k ln2 + f with fdlibm's two-part ln2, then the
 * degree-13 Taylor series of exp(f) by Horner's rule, every step an fma, a
 * product, rint or ldexp, each rounded once under IEEE-754, so the device
 * computes the same word.  Pinned against exp() in the tests (within one unit
 * of the word). */
__attribute__((optimize("fp-contract=off"))) uint32_t n3o_exp_word(double x)
{
    if (!(x < 23.0)) return 0u; /* 2^32 exp(-23) < 1 */
    const double y = -x;
    const double k = rint(y * 1.4426950408889634);
    double f = fma(k, -6.93147180369123816490e-01, y);
    f = fma(k, -1.90821492927058770002e-10, f);
    static const double inv_fact[14] = {1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0,
                                        1.0 / 362880.0,     1.0 / 40320.0,     1.0 / 5040.0,      1.0 / 720.0,
                                        1.0 / 120.0,        1.0 / 24.0,        1.0 / 6.0,         0.5,
                                        1.0,                1.0};
    double p = inv_fact[0];
    for (int i = 1; i < 14; ++i) p = fma(p, f, inv_fact[i]);
    const double v = ldexp(p, (int)k + 32);
    return v >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)v;
}

/* The umbrella energy of a move (tempering.py:93-94): delta + strength
 * (2 (removed - center) dd + dd^2), with the products and sums in the
 * order the device uses. */
__attribute__((optimize("fp-contract=off"))) static double umbrella_energy(int delta, int dd, int removed,
                                                                          float strength, float center)
{
    const double offset = (double)removed - (double)center;
    const double a = fma(2.0 * offset, (double)dd, (double)(dd * dd));
    return fma((double)strength, a, (double)delta);
}

/* One replica (slot) running `steps` moves of the walk kernel
 * (tempering.py:48-110).  Arrays are the slot's own rows of the caller's
 * arrays; with `umbrella` the move energy carries the slot's umbrella term and
 * is accepted against n3o_exp_word(E / T) (inverse_t = 1 / T).  Returns the
 * number of valid proposals. */
static int64_t walk_slot(int n, int steps, int bad_bias, uint64_t seed, uint32_t sweep, uint32_t slot,
                         int32_t* col, int32_t* bad, int32_t* score_io, int32_t* best_col, int32_t* best_score,
                         int32_t* accepted, const int32_t* reference, int32_t* removed_io, const uint32_t* thr, int k,
                         int umbrella, float strength, float center, double inverse_t)
{
    const int m = 2 * n;
    int32_t score = *score_io, removed = *removed_io;
    int32_t db[N3O_MAXM];
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    int64_t valid_count = 0;
    for (int step = 0; step < steps; ++step) {
        uint32_t ctr[4] = {(uint32_t)step, sweep, slot, TAG_MOVE}, r[4];
        n3o_philox4x32_10(ctr, key, r);
        int u = (int)mulhi32(r[0], (uint32_t)m);
        if (bad_bias > 0 && score > 0 && (int)mulhi32(r[3], 100u) < bad_bias) {
            uint32_t cb[4] = {(uint32_t)step, sweep, slot, TAG_BIAS}, rb[4];
            n3o_philox4x32_10(cb, key, rb);
            int64_t target = (int64_t)mulhi32(rb[0], (uint32_t)(3 * score));
            u = 0;
            while (u < m - 1 && target >= bad[u]) {
                target -= bad[u];
                ++u;
            }
        }
        const int v = (int)mulhi32(r[1], (uint32_t)m);
        const int32_t cu = col[u], cv = col[v];
        const int valid = score > 0 && (u >> 1) != (v >> 1) && cu != cv && col[u ^ 1] != cv && col[v ^ 1] != cu;
        const int ur = (u >> 1) * 2, vr = (v >> 1) * 2;
        const int dd = (cu == reference[ur] || cu == reference[ur + 1]) + (cv == reference[vr] || cv == reference[vr + 1]) -
                       (cv == reference[ur] || cv == reference[ur + 1]) - (cu == reference[vr] || cu == reference[vr + 1]);
        if (!valid) continue;
        valid_count += 1;
        for (int i = 0; i < m; ++i) db[i] = 0;
        int32_t delta = 0;
        const int ru = u >> 1, rv = v >> 1;
        /* tempering.py:71-82: triples with exactly one moved point */
        for (int i = 0; i < m - 1; ++i) {
            if (i == u || i == v) continue;
            const int ri = i >> 1, ci = col[i];
            for (int j = i + 1; j < m; ++j) {
                if (j == u || j == v) continue;
                const int rj = j >> 1, cj = col[j];
                int d = det0(ru, cv, ri, ci, rj, cj) - det0(ru, cu, ri, ci, rj, cj);
                if (d) {
                    delta += d;
                    db[u] += d;
                    db[i] += d;
                    db[j] += d;
                }
                d = det0(rv, cu, ri, ci, rj, cj) - det0(rv, cv, ri, ci, rj, cj);
                if (d) {
                    delta += d;
                    db[v] += d;
                    db[i] += d;
                    db[j] += d;
                }
            }
        }
        /* tempering.py:84-88: triples with both moved points */
        for (int kk = 0; kk < m; ++kk) {
            if (kk == u || kk == v) continue;
            const int d = det0(ru, cv, rv, cu, kk >> 1, col[kk]) - det0(ru, cu, rv, cv, kk >> 1, col[kk]);
            if (d) {
                delta += d;
                db[u] += d;
                db[v] += d;
                db[kk] += d;
            }
        }
        int accept;
        if (umbrella) {
            const double e = umbrella_energy(delta, dd, removed, strength, center);
            accept = e <= 0.0 || r[2] < n3o_exp_word(e * inverse_t);
        } else {
            accept = delta <= 0;
            if (!accept) {
                const int idx = delta < k ? delta : k - 1;
                accept = r[2] < thr[idx];
            }
        }
        if (accept) {
            col[u] = cv;
            col[v] = cu;
            score += delta;
            removed += dd;
            *accepted += 1;
            for (int i = 0; i < m; ++i) bad[i] += db[i];
            if (score < *best_score) {
                *best_score = score;
                for (int i = 0; i < m; ++i) best_col[i] = col[i];
            }
        }
    }
    *score_io = score;
    *removed_io = removed;
    return valid_count;
}

/* TemperingChains.advance (tempering.py:158-167) for every slot; `strengths`
 * NULL means no umbrella potentials, else `centers` and `inverse_t` (1 / T per
 * slot) are read too. */
int64_t n3o_advance(int n, int replicas, int steps, int bad_bias, uint64_t seed, uint32_t sweep, int32_t* columns,
                    int32_t* badness, int32_t* scores, int32_t* best_columns, int32_t* best_scores, int32_t* accepted,
                    const int32_t* reference, int32_t* removed, const uint32_t* thr, int k, int threads,
                    const float* centers, const float* strengths, const double* inverse_t)
{
    const int m = 2 * n;
    int64_t valid = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : valid)
    for (int r = 0; r < replicas; ++r) {
        valid += walk_slot(n, steps, bad_bias, seed, sweep, (uint32_t)r, columns + (int64_t)r * m,
                           badness + (int64_t)r * m, scores + r, best_columns + (int64_t)r * m, best_scores + r,
                           accepted + r, reference, removed + r, thr + (int64_t)r * k, k, strengths != NULL,
                           strengths ? strengths[r] : 0.0f, strengths ? centers[r] : 0.0f,
                           strengths ? inverse_t[r] : 0.0);
    }
    return valid;
}

/* TemperingChains.exchange (tempering.py:169-191): per ladder, pairs
 * (a, a+1) from start + round%2 in steps of 2; an accepted pair swaps the
 * slots' columns, badness, scores and removed counters.  Returns the number
 * of accepted swaps. */
int n3o_exchange(int n, int replicas, int ladder, uint64_t seed, uint32_t round, int32_t* columns, int32_t* badness,
                 int32_t* scores, int32_t* removed, const uint32_t* thr, const int8_t* sign, int k)
{
    const int m = 2 * n;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    int exchanged = 0;
    for (int start = 0; start < replicas; start += ladder) {
        for (int a = start + (int)(round % 2u); a < start + ladder - 1; a += 2) {
            const int b = a + 1;
            const int64_t diff = (int64_t)scores[a] - (int64_t)scores[b];
            int accept;
            if (sign[a] == 0 || (sign[a] > 0 && diff >= 0) || (sign[a] < 0 && diff <= 0)) {
                accept = 1;
            } else {
                uint32_t ctr[4] = {(uint32_t)a, round, (uint32_t)(start / ladder), TAG_EXCH}, r[4];
                n3o_philox4x32_10(ctr, key, r);
                int64_t d = diff < 0 ? -diff : diff;
                if (d > k - 1) d = k - 1;
                accept = r[0] < thr[(int64_t)a * k + d];
            }
            if (accept) {
                for (int i = 0; i < m; ++i) {
                    int32_t t = columns[(int64_t)a * m + i];
                    columns[(int64_t)a * m + i] = columns[(int64_t)b * m + i];
                    columns[(int64_t)b * m + i] = t;
                    t = badness[(int64_t)a * m + i];
                    badness[(int64_t)a * m + i] = badness[(int64_t)b * m + i];
                    badness[(int64_t)b * m + i] = t;
                }
                int32_t t = scores[a];
                scores[a] = scores[b];
                scores[b] = t;
                t = removed[a];
                removed[a] = removed[b];
                removed[b] = t;
                exchanged += 1;
            }
        }
    }
    return exchanged;
}

/* ------------------------------------------------------------------------ */
/* Learned-proposal walk replay                                               */
/* ------------------------------------------------------------------------ */

/* thr64[r][d + k] = floor(2^32 exp(-d / T_r)) saturating at 2^63, d in [-k, k). */
void n3o_metropolis_thresholds64(const float* temperatures, int replicas, int k, uint64_t* thr)
{
    const double cap = 9223372036854775808.0;
    for (int r = 0; r < replicas; ++r) {
        const double t = (double)temperatures[r];
        for (int d = -k; d < k; ++d) {
            const double x = ldexp(exp(-(double)d / t), 32);
            thr[(int64_t)r * 2 * k + (d + k)] = x >= cap ? ((uint64_t)1 << 63) : (uint64_t)floor(x);
        }
    }
}

/* The walk of walk_slot with learned pair proposals: weights
 * w(k) = clamp(wt[E][min(bad[k], 63)], 1, 2^15); a 32-bit coin r0 below
 * floor(0.2 * 2^32) draws both slots uniformly (u = mulhi(r1, m),
 * v = mulhi(r2, m)), otherwise each slot is the first k whose inclusive
 * prefix of w exceeds mulhi(r1, S1) (resp. mulhi(r2, S1)); both slots in one
 * row make the proposal invalid.  The Metropolis-Hastings test is
 *   r3 * N * S1'^2 < thr64(delta) * N' * S1^2,  N = S1^2 + 4 m^2 w(u) w(v),
 * evaluated exactly (primes for the moved state).  Philox counter
 * (step, sweep, slot, "LRND"). */
/* Planted mistakes in the learned walk's Metropolis-Hastings rule, for the
 * tests that show a detailed-balance check catches them (0 = the rule as
 * specified; 1 = S1 and S1' swapped in the acceptance test; 2 = the uniform
 * term dropped from N and N'; 3 = the moved state's weights read at the old
 * energy; 4 = the point walk's second, proposal-ratio stage skipped).
 * Test infrastructure only. */
static int learned_variant = 0;
void n3o_set_learned_variant(int variant) { learned_variant = variant; }

/* Learned walk weights are clamped to [1, 2^15]. */
static inline uint32_t clamp_weight(uint32_t w) { return w < 1u ? 1u : (w > 32768u ? 32768u : w); }

/* x * y for x < 2^128, y < 2^64 as three 64-bit limbs (schoolbook), then
 * a < x * y for a < 2^128. */
static int below_product(unsigned __int128 a, unsigned __int128 x, uint64_t y)
{
    const uint64_t x0 = (uint64_t)x, x1 = (uint64_t)(x >> 64);
    const unsigned __int128 t0 = (unsigned __int128)x0 * y;
    const unsigned __int128 t1 = (unsigned __int128)x1 * y + (uint64_t)(t0 >> 64);
    const uint64_t limb0 = (uint64_t)t0, limb1 = (uint64_t)t1, limb2 = (uint64_t)(t1 >> 64);
    if (limb2 != 0) return 1;
    const uint64_t a1 = (uint64_t)(a >> 64), a0 = (uint64_t)a;
    return a1 < limb1 || (a1 == limb1 && a0 < limb0);
}

static int64_t walk_slot_learned(int n, int steps, uint64_t seed, uint32_t sweep, uint32_t slot, int32_t* col,
                                 int32_t* bad, int32_t* score_io, int32_t* best_col, int32_t* best_score,
                                 int32_t* accepted, const int32_t* reference, int32_t* removed_io, const uint32_t* wt,
                                 int max_energy, const uint64_t* thr64, int k64)
{
    const int m = 2 * n;
    int32_t score = *score_io, removed = *removed_io;
    int32_t db[N3O_MAXM], nb[N3O_MAXM];
    uint32_t w[N3O_MAXM], w2[N3O_MAXM];
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    int64_t valid_count = 0;
#define ROW(e) ((int64_t)((e) < 0 ? 0 : ((e) > max_energy ? max_energy : (e))) * 64)
    uint32_t s1 = 0;
    for (int i = 0; i < m; ++i) {
        w[i] = clamp_weight(wt[ROW(score) + (bad[i] < 63 ? bad[i] : 63)]);
        s1 += w[i];
    }
    for (int step = 0; step < steps; ++step) {
        uint32_t ctr[4] = {(uint32_t)step, sweep, slot, TAG_LEARN}, r[4];
        n3o_philox4x32_10(ctr, key, r);
        /* 0.2: a uniform pair; 0.8: both slots drawn in proportion to w (the
         * first slot whose running weight sum exceeds the target) */
        int u, v;
        if (r[0] < 858993459u) {
            u = (int)mulhi32(r[1], (uint32_t)m);
            v = (int)mulhi32(r[2], (uint32_t)m);
        } else {
            const uint32_t target_u = mulhi32(r[1], s1), target_v = mulhi32(r[2], s1);
            uint32_t acc = 0;
            u = m - 1;
            v = m - 1;
            int found_u = 0, found_v = 0;
            for (int i = 0; i < m; ++i) {
                acc += w[i];
                if (!found_u && acc > target_u) {
                    u = i;
                    found_u = 1;
                }
                if (!found_v && acc > target_v) {
                    v = i;
                    found_v = 1;
                }
            }
        }
        const int32_t cu = col[u], cv = col[v];
        const int valid = score > 0 && (u >> 1) != (v >> 1) && cu != cv && col[u ^ 1] != cv && col[v ^ 1] != cu;
        const int ur = (u >> 1) * 2, vr = (v >> 1) * 2;
        const int dd = (cu == reference[ur] || cu == reference[ur + 1]) + (cv == reference[vr] || cv == reference[vr + 1]) -
                       (cv == reference[ur] || cv == reference[ur + 1]) - (cu == reference[vr] || cu == reference[vr + 1]);
        if (!valid) continue;
        valid_count += 1;
        for (int i = 0; i < m; ++i) db[i] = 0;
        int32_t delta = 0;
        const int ru = u >> 1, rv = v >> 1;
        for (int i = 0; i < m - 1; ++i) {
            if (i == u || i == v) continue;
            const int ri = i >> 1, ci = col[i];
            for (int j = i + 1; j < m; ++j) {
                if (j == u || j == v) continue;
                const int rj = j >> 1, cj = col[j];
                int d = det0(ru, cv, ri, ci, rj, cj) - det0(ru, cu, ri, ci, rj, cj);
                if (d) {
                    delta += d;
                    db[u] += d;
                    db[i] += d;
                    db[j] += d;
                }
                d = det0(rv, cu, ri, ci, rj, cj) - det0(rv, cv, ri, ci, rj, cj);
                if (d) {
                    delta += d;
                    db[v] += d;
                    db[i] += d;
                    db[j] += d;
                }
            }
        }
        for (int kk = 0; kk < m; ++kk) {
            if (kk == u || kk == v) continue;
            const int d = det0(ru, cv, rv, cu, kk >> 1, col[kk]) - det0(ru, cu, rv, cv, kk >> 1, col[kk]);
            if (d) {
                delta += d;
                db[u] += d;
                db[v] += d;
                db[kk] += d;
            }
        }
        uint32_t s1n = 0;
        for (int i = 0; i < m; ++i) {
            nb[i] = bad[i] + db[i];
            const int32_t e2 = learned_variant == 3 ? score : score + delta;
            w2[i] = clamp_weight(wt[ROW(e2) + (nb[i] < 63 ? nb[i] : 63)]);
            s1n += w2[i];
        }
        /* pair probability 2 (0.2/m^2 + 0.8 w(u) w(v)/S1^2) = 0.4 N / (m^2 S1^2);
         * accept iff r N S1'^2 < thr64(delta) N' S1^2 */
        const uint64_t mm4 = 4ull * (uint64_t)m * (uint64_t)m;
        const uint64_t uniform_sq = learned_variant == 2 ? 0 : 1;
        const uint64_t big_n = uniform_sq * s1 * s1 + mm4 * w[u] * w[v];
        const uint64_t big_n2 = uniform_sq * s1n * s1n + mm4 * w2[u] * w2[v];
        const int d = delta < -k64 ? -k64 : (delta > k64 - 1 ? k64 - 1 : delta);
        uint64_t old_sq = (uint64_t)s1 * s1, new_sq = (uint64_t)s1n * s1n;
        if (learned_variant == 1) {
            const uint64_t t = old_sq;
            old_sq = new_sq;
            new_sq = t;
        }
        const unsigned __int128 lhs = (unsigned __int128)r[3] * big_n * new_sq;
        if (!below_product(lhs, (unsigned __int128)thr64[d + k64] * big_n2, old_sq)) continue;
        col[u] = cv;
        col[v] = cu;
        score += delta;
        removed += dd;
        *accepted += 1;
        for (int i = 0; i < m; ++i) {
            bad[i] = nb[i];
            w[i] = w2[i];
        }
        s1 = s1n;
        if (score < *best_score) {
            *best_score = score;
            for (int i = 0; i < m; ++i) best_col[i] = col[i];
        }
    }
#undef ROW
    *score_io = score;
    *removed_io = removed;
    return valid_count;
}

int64_t n3o_advance_learned(int n, int replicas, int steps, uint64_t seed, uint32_t sweep, int32_t* columns,
                            int32_t* badness, int32_t* scores, int32_t* best_columns, int32_t* best_scores,
                            int32_t* accepted, const int32_t* reference, int32_t* removed, const uint32_t* weights,
                            const int32_t* temp_index, int max_energy, const uint64_t* thr64, int k64, int threads)
{
    const int m = 2 * n;
    int64_t valid = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : valid)
    for (int r = 0; r < replicas; ++r) {
        valid += walk_slot_learned(n, steps, seed, sweep, (uint32_t)r, columns + (int64_t)r * m,
                                   badness + (int64_t)r * m, scores + r, best_columns + (int64_t)r * m,
                                   best_scores + r, accepted + r, reference, removed + r,
                                   weights + (int64_t)temp_index[r] * (max_energy + 1) * 64, max_energy,
                                   thr64 + (int64_t)r * 2 * k64, k64);
    }
    return valid;
}

/* ------------------------------------------------------------------------- */
/* Learned proposals from the point network (row/column features), with the  */
/* delayed (two-stage) Metropolis-Hastings test.  No reference counterpart;   */
/* the nearest heuristic is bad_bias (tempering.py:51-55).                    */
/* ------------------------------------------------------------------------- */

#define TAG_HAST 0x48415354u /* "HAST" */
#define POINT_QMIN (-384)
#define POINT_QMAX 352
#define MLP_HIDDEN 128
#define MLP_B1 1024
#define MLP_W2 1152
#define MLP_B2 1280

static float bf16_value(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, sizeof f);
    return f;
}

/* Integer proposal weight of a point with collinear-pair count b at (row,
 * col), temperature code lt_q = round(32 ln T), state energy e.  Features
 * are the codes k / 128 of
 *   (4 min(b,63), 32 min(b,7), 128 [b>0], row, col, lt_q, min(e,510)/2, 2 min(e,63)).
 * W1 and b1 lie on the grid of multiples of 1/128, so every hidden
 * pre-activation b1 + W1 x is a multiple of 2^-14 below 2^5 in magnitude and
 * its double sum here is exact (as is the device's fp32 MFMA sum).  The
 * logit is two fp32 fused multiply-add chains over the ReLU outputs, hidden
 * units 32 t + 8 (r / 4) + 4 h + r % 4 for h = 0 (from b2) and h = 1 (from
 * 0), t = 0..3, r = 0..15 in that order, added together; the weight is
 * expw[clamp(floor(64 g), -384, 352) + 384] (g first clamped to [-7, 6],
 * which changes no weight). */
static uint32_t point_weight(const uint16_t* params, const uint32_t* expw, int b, int row, int col, int lt_q, int e)
{
    const int bb = b < 63 ? b : 63, b7 = b < 7 ? b : 7;
    const int code[8] = {4 * bb, 32 * b7, b > 0 ? 128 : 0, row, col, lt_q, (e < 510 ? e : 510) >> 1,
                         2 * (e < 63 ? e : 63)};
    float hidden[MLP_HIDDEN];
    for (int i = 0; i < MLP_HIDDEN; ++i) {
        double acc = bf16_value(params[MLP_B1 + i]);
        for (int k = 0; k < 8; ++k) acc += (double)bf16_value(params[i * 8 + k]) * ((double)code[k] / 128.0);
        hidden[i] = (float)acc;
    }
    float g0 = bf16_value(params[MLP_B2]), g1 = 0.0f;
    for (int t = 0; t < 4; ++t)
        for (int r = 0; r < 16; ++r) {
            const int i = 32 * t + 8 * (r >> 2) + (r & 3);
            g0 = fmaf(bf16_value(params[MLP_W2 + i]), fmaxf(hidden[i], 0.0f), g0);
            g1 = fmaf(bf16_value(params[MLP_W2 + i + 4]), fmaxf(hidden[i + 4], 0.0f), g1);
        }
    const float g = fminf(fmaxf(g0 + g1, -7.0f), 6.0f);
    int q = (int)floorf(g * 64.0f);
    q = q < POINT_QMIN ? POINT_QMIN : (q > POINT_QMAX ? POINT_QMAX : q);
    return expw[q - POINT_QMIN];
}

static void point_weights_state(int m, const int32_t* col, const int32_t* bad, int energy, int lt_q,
                                const uint16_t* params, const uint32_t* expw, uint32_t* w)
{
    for (int k = 0; k < m; ++k) w[k] = point_weight(params, expw, bad[k], k >> 1, col[k], lt_q, energy);
}

void n3o_point_weights(int n, int64_t count, const int32_t* columns, const int32_t* badness, const int32_t* scores,
                       const int32_t* lt_q, const uint16_t* params, const uint32_t* expw, uint32_t* out)
{
    const int m = 2 * n;
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < count; ++s)
        point_weights_state(m, columns + s * m, badness + s * m, scores[s], lt_q[s], params, expw, out + s * m);
}

/* The exact energy change of swapping the columns of slots u and v, with the
 * per-slot badness changes db[] (the pair loops of walk_slot_learned). */
static int32_t switch_delta(int m, const int32_t* col, int u, int v, int32_t* db)
{
    for (int i = 0; i < m; ++i) db[i] = 0;
    int32_t delta = 0;
    const int ru = u >> 1, rv = v >> 1, cu = col[u], cv = col[v];
    for (int i = 0; i < m - 1; ++i) {
        if (i == u || i == v) continue;
        const int ri = i >> 1, ci = col[i];
        for (int j = i + 1; j < m; ++j) {
            if (j == u || j == v) continue;
            const int rj = j >> 1, cj = col[j];
            int d = det0(ru, cv, ri, ci, rj, cj) - det0(ru, cu, ri, ci, rj, cj);
            if (d) {
                delta += d;
                db[u] += d;
                db[i] += d;
                db[j] += d;
            }
            d = det0(rv, cu, ri, ci, rj, cj) - det0(rv, cv, ri, ci, rj, cj);
            if (d) {
                delta += d;
                db[v] += d;
                db[i] += d;
                db[j] += d;
            }
        }
    }
    for (int k = 0; k < m; ++k) {
        if (k == u || k == v) continue;
        const int d = det0(ru, cv, rv, cu, k >> 1, col[k]) - det0(ru, cu, rv, cv, k >> 1, col[k]);
        if (d) {
            delta += d;
            db[u] += d;
            db[v] += d;
            db[k] += d;
        }
    }
    return delta;
}

/* The first slot whose inclusive prefix of w exceeds target. */
static int inverse_cdf(int m, const uint32_t* w, uint64_t target)
{
    uint64_t acc = 0;
    for (int i = 0; i < m; ++i) {
        acc += w[i];
        if (acc > target) return i;
    }
    return m - 1;
}

/* One replica's `steps` proposals with point-network weights w (recomputed
 * for every state the walk visits).  Philox words r = (step, sweep, slot,
 * "LRND") and h = (step, sweep, slot, "HAST"):
 *   - with r0 < floor(0.2 * 2^32) both slots are uniform, u = mulhi(r1, m),
 *     v = mulhi(r2, m); otherwise each slot is drawn in proportion to w by
 *     the inverse CDF at the 64-bit targets floor(((h0 << 32) | r1) S1 / 2^64)
 *     and floor(((h1 << 32) | r2) S1 / 2^64);
 *   - stage 1 (Metropolis on the exact delta): continue only if
 *     r3 < thr64(delta) = floor(2^32 exp(-delta/T));
 *   - stage 2 (proposal ratio): accept iff h2 N S1'^2 < 2^32 N' S1^2 with
 *     N = S1^2 + 4 m^2 w(u) w(v) and primes for the moved state.
 * The unordered pair's probability is q = 0.4 N / (m^2 S1^2); the move is
 * accepted with min(1, exp(-delta/T)) min(1, q'/q), a valid
 * Metropolis-Hastings kernel (delayed acceptance: pi q min(1, pi'/pi)
 * min(1, q'/q) = min(pi, pi') min(q, q') is symmetric). */
static int64_t walk_slot_point(int n, int steps, uint64_t seed, uint32_t sweep, uint32_t slot, int32_t* col,
                               int32_t* bad, int32_t* score_io, int32_t* best_col, int32_t* best_score,
                               int32_t* accepted, const int32_t* reference, int32_t* removed_io,
                               const uint16_t* params, const uint32_t* expw, int lt_q, const uint64_t* thr64, int k64)
{
    const int m = 2 * n;
    int32_t score = *score_io, removed = *removed_io;
    int32_t db[N3O_MAXM], nb[N3O_MAXM], col2[N3O_MAXM];
    uint32_t w[N3O_MAXM], w2[N3O_MAXM];
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    int64_t valid_count = 0;
    point_weights_state(m, col, bad, score, lt_q, params, expw, w);
    uint64_t s1 = 0;
    for (int i = 0; i < m; ++i) s1 += w[i];
    for (int step = 0; step < steps; ++step) {
        const uint32_t ctr[4] = {(uint32_t)step, sweep, slot, TAG_LEARN}, ctr_h[4] = {(uint32_t)step, sweep, slot, TAG_HAST};
        uint32_t r[4], h[4];
        n3o_philox4x32_10(ctr, key, r);
        n3o_philox4x32_10(ctr_h, key, h);
        int u, v;
        if (r[0] < 858993459u) {
            u = (int)mulhi32(r[1], (uint32_t)m);
            v = (int)mulhi32(r[2], (uint32_t)m);
        } else {
            const uint64_t xu = ((uint64_t)h[0] << 32) | r[1], xv = ((uint64_t)h[1] << 32) | r[2];
            u = inverse_cdf(m, w, (uint64_t)(((unsigned __int128)xu * s1) >> 64));
            v = inverse_cdf(m, w, (uint64_t)(((unsigned __int128)xv * s1) >> 64));
        }
        const int32_t cu = col[u], cv = col[v];
        const int valid = score > 0 && (u >> 1) != (v >> 1) && cu != cv && col[u ^ 1] != cv && col[v ^ 1] != cu;
        if (!valid) continue;
        valid_count += 1;
        const int ur = (u >> 1) * 2, vr = (v >> 1) * 2;
        const int dd = (cu == reference[ur] || cu == reference[ur + 1]) + (cv == reference[vr] || cv == reference[vr + 1]) -
                       (cv == reference[ur] || cv == reference[ur + 1]) - (cu == reference[vr] || cu == reference[vr + 1]);
        const int32_t delta = switch_delta(m, col, u, v, db);
        const int d = delta < -k64 ? -k64 : (delta > k64 - 1 ? k64 - 1 : delta);
        if (!((uint64_t)r[3] < thr64[d + k64])) continue;
        for (int i = 0; i < m; ++i) {
            nb[i] = bad[i] + db[i];
            col2[i] = col[i];
        }
        col2[u] = cv;
        col2[v] = cu;
        point_weights_state(m, col2, nb, learned_variant == 3 ? score : score + delta, lt_q, params, expw, w2);
        uint64_t s1n = 0;
        for (int i = 0; i < m; ++i) s1n += w2[i];
        const uint64_t mm4 = 4ull * (uint64_t)m * (uint64_t)m;
        const uint64_t uniform_sq = learned_variant == 2 ? 0 : 1;
        const uint64_t big_n = uniform_sq * s1 * s1 + mm4 * w[u] * w[v];
        const uint64_t big_n2 = uniform_sq * s1n * s1n + mm4 * w2[u] * w2[v];
        uint64_t old_sq = s1 * s1, new_sq = s1n * s1n;
        if (learned_variant == 1) {
            const uint64_t t = old_sq;
            old_sq = new_sq;
            new_sq = t;
        }
        const unsigned __int128 lhs = (unsigned __int128)h[2] * ((unsigned __int128)big_n * new_sq);
        const unsigned __int128 rhs = ((unsigned __int128)big_n2 * old_sq) << 32;
        if (learned_variant != 4 && !(lhs < rhs)) continue;
        col[u] = cv;
        col[v] = cu;
        score += delta;
        removed += dd;
        *accepted += 1;
        for (int i = 0; i < m; ++i) {
            bad[i] = nb[i];
            w[i] = w2[i];
        }
        s1 = s1n;
        if (score < *best_score) {
            *best_score = score;
            for (int i = 0; i < m; ++i) best_col[i] = col[i];
        }
    }
    *score_io = score;
    *removed_io = removed;
    return valid_count;
}

int64_t n3o_advance_point(int n, int replicas, int steps, uint64_t seed, uint32_t sweep, int32_t* columns,
                          int32_t* badness, int32_t* scores, int32_t* best_columns, int32_t* best_scores,
                          int32_t* accepted, const int32_t* reference, int32_t* removed, const uint16_t* params,
                          const uint32_t* expw, const int32_t* lt_q, const uint64_t* thr64, int k64, int threads)
{
    const int m = 2 * n;
    int64_t valid = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : valid)
    for (int r = 0; r < replicas; ++r) {
        valid += walk_slot_point(n, steps, seed, sweep, (uint32_t)r, columns + (int64_t)r * m, badness + (int64_t)r * m,
                                 scores + r, best_columns + (int64_t)r * m, best_scores + r, accepted + r, reference,
                                 removed + r, params, expw, lt_q[r], thr64 + (int64_t)r * 2 * k64, k64);
    }
    return valid;
}

/* ------------------------------------------------------------------------- */
/* Device search engines: mutate_population (gpu_device_search.py:27-113)    */
/* and enumerate_switches (gpu_exhaustive_beam.py:17-57), restated slot by   */
/* slot with the Philox draws (child, generation, draw, "MUTA").             */
/* ------------------------------------------------------------------------- */

#define TAG_MUTATE 0x4D555441u

static void mutate_draw(uint64_t seed, uint32_t child, uint32_t generation, uint32_t draw, uint32_t out[4])
{
    const uint32_t ctr[4] = {child, generation, draw, TAG_MUTATE};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    n3o_philox4x32_10(ctr, key, out);
}

/* gpu_exhaustive_beam.py:39-54: swap the columns of slots u and v (slot s is
 * P[s>>1] for even s, Q[s>>1] for odd s) unless the rows coincide, the
 * columns coincide, or a row would hold one column twice. */
static int apply_switch(int32_t* p, int32_t* q, int u, int v)
{
    const int row_u = u >> 1, row_v = v >> 1;
    if (row_u == row_v) return 0;
    int32_t* slot_u = (u & 1) ? q + row_u : p + row_u;
    int32_t* slot_v = (v & 1) ? q + row_v : p + row_v;
    const int32_t column_u = *slot_u, column_v = *slot_v;
    if (column_u == column_v) return 0;
    const int32_t other_u = (u & 1) ? p[row_u] : q[row_u];
    const int32_t other_v = (v & 1) ? p[row_v] : q[row_v];
    if (column_v == other_u || column_u == other_v) return 0;
    *slot_u = column_v;
    *slot_v = column_u;
    return 1;
}

void n3o_mutate_population(const int32_t* parent_p, const int32_t* parent_q, const int32_t* parent_badness,
                           int elite_count, int quality_count, int n, int64_t population, int active_strength,
                           int max_strength, int bad_bias, uint64_t seed, int generation, int32_t* child_p,
                           int32_t* child_q)
{
    const int m = 2 * n;
    for (int64_t child = 0; child < population; ++child) {
        uint32_t a[4];
        mutate_draw(seed, (uint32_t)child, (uint32_t)generation, 0u, a);
        int parent = (int)child;
        if (child >= elite_count) {
            const int pool = (a[0] >> 24) < 160u ? quality_count : elite_count;
            parent = (int)mulhi32(a[1], (uint32_t)pool);
        }
        int32_t* cp = child_p + child * n;
        int32_t* cq = child_q + child * n;
        memcpy(cp, parent_p + (int64_t)parent * n, sizeof(int32_t) * n);
        memcpy(cq, parent_q + (int64_t)parent * n, sizeof(int32_t) * n);
        if (child < elite_count) continue;
        const int span = mulhi32(a[2], 32u) == 0u ? max_strength : active_strength;
        const int strength = 1 + (int)mulhi32(a[3], (uint32_t)span);
        const int32_t* bad = parent_badness + (int64_t)parent * m;
        uint32_t total = 0;
        for (int e = 0; e < m; ++e) total += (uint32_t)bad[e];
        for (int step = 0; step < strength; ++step) {
            for (int attempt = 0; attempt < 32; ++attempt) {
                uint32_t b[4];
                mutate_draw(seed, (uint32_t)child, (uint32_t)generation, 1u + (uint32_t)(step * 32 + attempt), b);
                int u = (int)mulhi32(b[2], (uint32_t)m);
                if ((int)mulhi32(b[0], 100u) < bad_bias && total > 0u) {
                    /* gpu_device_search.py:84-88: first slot whose running
                     * badness sum exceeds the threshold */
                    int64_t threshold = (int64_t)mulhi32(b[1], total);
                    u = 0;
                    for (; u < m - 1; ++u) {
                        threshold -= bad[u];
                        if (threshold < 0) break;
                    }
                }
                const int v = (int)mulhi32(b[3], (uint32_t)m);
                if (apply_switch(cp, cq, u, v)) break;
            }
        }
    }
}

void n3o_enumerate_switches(const int32_t* parent_p, const int32_t* parent_q, const int32_t* descriptors,
                            int64_t count, int n, int32_t* child_p, int32_t* child_q, uint8_t* valid)
{
    for (int64_t i = 0; i < count; ++i) {
        const int parent = descriptors[3 * i], u = descriptors[3 * i + 1], v = descriptors[3 * i + 2];
        memcpy(child_p + i * n, parent_p + (int64_t)parent * n, sizeof(int32_t) * n);
        memcpy(child_q + i * n, parent_q + (int64_t)parent * n, sizeof(int32_t) * n);
        valid[i] = (uint8_t)apply_switch(child_p + i * n, child_q + i * n, u, v);
    }
}
