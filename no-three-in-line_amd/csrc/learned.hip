// Tempering walk with learned move proposals.
//
// The proposal network scores a move {u, v} by the additive pair logit
// g(u) + g(v) - b2, so the softmax over all pairs factorises: a pair is drawn
// with probability proportional to w(u) w(v), where w_x(k) = W[t][E(x)][b]
// is the network's tabulated integer weight of slot k (temperature index t
// of the replica slot, energy E(x), the slot's collinear-pair count b),
// clamped to [1, 2^15].  The proposal mixes 20% uniform pairs with 80% of
// this softmax: with probability 0.2 both slots are uniform, otherwise both
// are drawn from w / S1 (S1 = sum_k w(k)) by the inverse CDF; a draw with
// both slots in one row is an invalid proposal.  The unordered pair {u, v}
// has probability
//   q(x -> x') = 2 (0.2 / m^2 + 0.8 w(u) w(v) / S1^2) = 0.4 N / (m^2 S1^2),
//   N = S1^2 + 4 m^2 w(u) w(v),
// and the reverse move from x' is the same pair, so the Hastings ratio is
// (N' S1^2) / (N S1'^2).  Acceptance stays on the exact integer delta:
//   accept  iff  r N S1'^2 < thr64(delta) N' S1^2
// with r the move's 32-bit Philox word and thr64(d) = floor(2^32 exp(-d/T))
// (saturating at 2^63), evaluated exactly (a 128-bit left side against a
// 192-bit right side) -- so a CPU replay with the same weight table
// reproduces every trajectory bit for bit.  Evaluating the reverse
// probability needs the badness of every slot in x' (two more direction
// histograms, from u and v in x); only a proposal whose acceptance word is 0
// is rejected before them.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "temper_internal.h"
#include "wave_ops.h"

namespace no3 {

constexpr uint32_t kTagLearn = 0x4C524E44u;  // "LRND"
constexpr uint32_t kUniformShare = 858993459u;  // floor(0.2 * 2^32)

struct LearnedArgs {
  const uint32_t* weights;             // [temps][max_energy + 1][64]
  const int32_t* temp_index;           // [R]
  int max_energy;
  const unsigned long long* thr64;     // [R][2 k64], entry d + k64
  int k64;
};

// Weight of a slot with collinear-pair count `count` from a weight row held
// one entry per lane: per-lane counts use a lane permute, uniform counts a
// scalar read.
constexpr uint32_t kMaxWeight = 1u << 15;

__device__ inline uint32_t clamp_weight(uint32_t w) { return w < 1u ? 1u : (w > kMaxWeight ? kMaxWeight : w); }
__device__ inline uint32_t weight_of(uint32_t row, int count) {
  return (uint32_t)__shfl((int)row, count < 63 ? count : 63, kWave);
}
__device__ inline uint32_t weight_at(uint32_t row, int count) {
  return (uint32_t)__builtin_amdgcn_readlane((int)row, count < 63 ? count : 63);
}

// a < p s for a, p < 2^128 and s < 2^64: p s = p_hi s 2^64 + p_lo s may
// need 192 bits, so it is compared limb by limb.
__device__ inline bool less_than_product(unsigned __int128 a, unsigned __int128 p, uint64_t s) {
  const unsigned __int128 lo = (unsigned __int128)(uint64_t)p * s;
  const unsigned __int128 hi = (unsigned __int128)(uint64_t)(p >> 64) * s;
  if ((uint64_t)(hi >> 64) != 0) return true;
  const unsigned __int128 shifted = hi << 64;
  const unsigned __int128 sum = shifted + lo;
  if (sum < shifted) return true;
  return a < sum;
}

template <int CH>
__device__ void walk_replica_learned(const WalkArgs& a, const LearnedArgs& L, const uint16_t* table, uint32_t* hist,
                                     const DirTableInfo& info, int slot_in) {
  const int lane = lane_id();
  const int slot = __builtin_amdgcn_readfirstlane(slot_in);  // uniform: scalar registers
  const int n = a.n, m = 2 * n, width = info.width, offset = info.offset;
  const uint32_t sweep = a.counters[0];
  const long long base = (long long)slot * m;

  int key[CH], bad[CH], ref[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int k = lane + kWave * c;
    key[c] = k < m ? (k >> 1) * width + a.columns[base + k] : 0;
    bad[c] = k < m ? a.badness[base + k] : 0;
    ref[c] = k < m ? a.reference[k] : -1;
  }
  // both slots of a row, packed per lane: current columns (bytes 0, 1) and
  // reference columns (bytes 2, 3) of slots 2r and 2r+1
  int row_word[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int k = lane + kWave * c;
    const int col = key[c] - (k >> 1) * width, partner_col = __shfl_xor(col, 1, kWave);
    const int partner_ref = __shfl_xor(ref[c], 1, kWave);
    const int col0 = (k & 1) ? partner_col : col, col1 = (k & 1) ? col : partner_col;
    const int ref0 = (k & 1) ? partner_ref : ref[c], ref1 = (k & 1) ? ref[c] : partner_ref;
    row_word[c] = (col0 & 0xFF) | (col1 & 0xFF) << 8 | (ref0 & 0xFF) << 16 | (ref1 & 0xFF) << 24;
  }
  {
    const int key0 = __builtin_amdgcn_readfirstlane(key[0]);
#pragma unroll
    for (int c = 0; c < CH; ++c)
      if (lane + kWave * c >= m) key[c] = key0;
  }
  const int trash = trash_word(info, lane);
  const bool tail_in = lane + kWave * (CH - 1) < m;
  int score = a.scores[slot];
  int removed = a.removed[slot];
  int best = a.best_scores[slot];
  int accepted = 0;
  unsigned long long valid_moves = 0;
  const uint32_t* wt = L.weights + (long long)L.temp_index[slot] * (L.max_energy + 1) * 64;
  const unsigned long long* thr = L.thr64 + (long long)slot * 2 * L.k64;
  // acceptance words thr[j], j = d + k64 < 128, in registers
  const unsigned long long thr_lo = lane < 2 * L.k64 ? thr[lane] : 0ull;
  const unsigned long long thr_hi = lane + kWave < 2 * L.k64 ? thr[lane + kWave] : 0ull;
  auto weight_row = [&](int energy) -> uint32_t {
    const int e = energy < 0 ? 0 : (energy > L.max_energy ? L.max_energy : energy);
    return clamp_weight(wt[(long long)e * 64 + lane]);
  };
  int w[CH];
  uint32_t s1;
  {
    const uint32_t row = weight_row(score);
    uint32_t part = 0;
    // the row exchange runs on every lane (a permute from a lane outside the
    // exec mask reads nothing); idle lanes are zeroed afterwards
#pragma unroll
    for (int c = 0; c < CH; ++c) w[c] = (int)weight_of(row, bad[c]);
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      if (lane + kWave * c >= m) w[c] = 0;
      part += (uint32_t)w[c];
    }
    s1 = wave_sum_dpp(part);
  }
  // inclusive prefix sums of w in slot order (the inverse-CDF table of the
  // proposal draw); they change only when a move is accepted
  uint32_t cdf[CH];
  auto build_cdf = [&]() {
    uint32_t carry = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const uint32_t x = wave_scan_dpp((uint32_t)w[c]);
      cdf[c] = carry + x;
      carry += (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
    }
  };
  build_cdf();

  // Philox outputs of 64 consecutive steps, one step per lane, computed on
  // the VALU in one pass and read back with readlane (the per-step counters,
  // so the same stream as one draw per step)
  uint32_t learn_r[4];
  int batch = -kWave;  // first step of the current batch
  for (int step = 0; step < a.steps; ++step) {
    if (step - batch >= kWave) {
      batch = step;
      philox4x32_10((uint32_t)(step + lane), sweep, (uint32_t)slot, kTagLearn, a.seed_lo, a.seed_hi, learn_r);
    }
    uint32_t r[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = (uint32_t)__builtin_amdgcn_readlane((int)learn_r[i], step - batch);
    int u, v;
    if (r[0] < kUniformShare) {
      u = (int)mulhi32(r[1], (uint32_t)m);
      v = (int)mulhi32(r[2], (uint32_t)m);
    } else {
      // both slots by the inverse CDF of w over the slots in slot order
      const uint32_t target_u = mulhi32(r[1], s1), target_v = mulhi32(r[2], s1);
      int below_u = 0, below_v = 0;
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const bool live = lane + kWave * c < m;
        below_u += __popcll(lanes(live) & lanes(cdf[c] <= target_u));
        below_v += __popcll(lanes(live) & lanes(cdf[c] <= target_v));
      }
      u = below_u < m - 1 ? below_u : m - 1;
      v = below_v < m - 1 ? below_v : m - 1;
    }
    const int ru = u >> 1, rv = v >> 1;
    const uint32_t word_ru = (uint32_t)slot_value<CH>(row_word, u & ~1);
    const uint32_t word_rv = (uint32_t)slot_value<CH>(row_word, v & ~1);
    const int cu = (int)((word_ru >> (8 * (u & 1))) & 0xFF), cu_other = (int)((word_ru >> (8 * (~u & 1))) & 0xFF);
    const int cv = (int)((word_rv >> (8 * (v & 1))) & 0xFF), cv_other = (int)((word_rv >> (8 * (~v & 1))) & 0xFF);
    const int key_u = ru * width + cu, key_v = rv * width + cv;
    const bool valid = score > 0 && ru != rv && cu != cv && cu_other != cv && cv_other != cu;
    if (!valid) continue;
    valid_moves += 1;
    const int ref_u0 = (int)((word_ru >> 16) & 0xFF), ref_u1 = (int)(word_ru >> 24);
    const int ref_v0 = (int)((word_rv >> 16) & 0xFF), ref_v1 = (int)(word_rv >> 24);
    const int dd = (cu == ref_u0 || cu == ref_u1) + (cv == ref_v0 || cv == ref_v1) - (cv == ref_u0 || cv == ref_u1) -
                   (cu == ref_v0 || cu == ref_v1);

    // bad'(u') and bad'(v') from the ranks the increments return (both
    // histograms back to back, one wait); per-slot counts are read back only
    // for the full evaluation.  The moved slot itself looks up the zero
    // vector (rank 0), lanes past the last slot use their trash word.
    const int id_uv = table[key_v - key_u + offset];
    const int key_u2 = ru * width + cv, key_v2 = rv * width + cu;
    const int id_u2v2 = table[key_v2 - key_u2 + offset];
    int idu[CH], idu2[CH], idv2[CH];
    bool others[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int k = lane + kWave * c;
      others[c] = (k < m) && k != u && k != v;
      idu[c] = table[key[c] - key_u + offset];
      const int moved = (k == v) ? key_v2 : ((k == u) ? key_u2 : key[c]);
      idu2[c] = table[moved - key_u2 + offset];
      idv2[c] = table[moved - key_v2 + offset];
    }
    int joint_old = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) joint_old += __popcll(lanes(others[c]) & lanes(idu[c] == id_uv));
    uint32_t word_u2[CH], word_v2[CH], old_u2[CH], old_v2[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      word_u2[c] = (c < CH - 1 || tail_in) ? ((uint32_t)idu2[c] >> 2) : (uint32_t)trash;
      old_u2[c] = atomicAdd(&hist[word_u2[c]], byte_one((uint32_t)idu2[c]));
    }
    wave_sync();
#pragma unroll
    for (int c = 0; c < CH; ++c) hist[word_u2[c]] = 0u;
    wave_sync();
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      word_v2[c] = (c < CH - 1 || tail_in) ? ((uint32_t)idv2[c] >> 2) : (uint32_t)trash;
      old_v2[c] = atomicAdd(&hist[word_v2[c]], byte_one((uint32_t)idv2[c]));
    }
    wave_sync();
#pragma unroll
    for (int c = 0; c < CH; ++c) hist[word_v2[c]] = 0u;
    wave_sync();
    // keep the rank extraction below every LDS step above
    __builtin_amdgcn_sched_barrier(0);
    uint32_t part = 0;
    int joint_new = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      part += __builtin_amdgcn_ubfe(old_u2[c], byte_shift((uint32_t)idu2[c]), 8) +
              (__builtin_amdgcn_ubfe(old_v2[c], byte_shift((uint32_t)idv2[c]), 8) << 16);
      joint_new += __popcll(lanes(others[c]) & lanes(idu2[c] == id_u2v2));
    }
    const uint32_t packed = wave_sum_dpp(part);
    const int bad_u2 = (int)(packed & 0xFFFFu), bad_v2 = (int)(packed >> 16);
    const int bad_u = slot_value<CH>(bad, u), bad_v = slot_value<CH>(bad, v);
    const int delta = (bad_u2 + bad_v2 - joint_new) - (bad_u + bad_v - joint_old);
    const int d = delta < -L.k64 ? -L.k64 : (delta > L.k64 - 1 ? L.k64 - 1 : delta);
    const unsigned long long t64 = register_table64(thr_lo, thr_hi, thr, d + L.k64);
    // thr64 = 0: r N S1' < 0 is false whatever the weights, so the move is
    // rejected without reading the moved state's weight row
    if (t64 == 0ull) continue;
    const uint32_t row2 = weight_row(score + delta);
    const uint64_t wu = (uint32_t)slot_value<CH>(w, u), wv = (uint32_t)slot_value<CH>(w, v);
    const uint64_t mm4 = 4ull * (uint64_t)m * (uint64_t)m;
    const uint64_t s1sq = (uint64_t)s1 * s1;
    const uint64_t big_n = s1sq + mm4 * wu * wv;  // < 2^48 (w <= 2^15, m <= 200)
    // badness of every slot in the moved state (needed for the reverse move)
    int cnt_u2[CH], cnt_v2[CH], cnt_u[CH], cnt_v[CH], idv[CH];
    // the old state's directions from v, looked up here so that the reads
    // overlap the first histograms instead of waiting behind the last one
#pragma unroll
    for (int c = 0; c < CH; ++c) idv[c] = table[key[c] - key_v + offset];
    bool include[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) include[c] = (lane + kWave * c < m) && (lane + kWave * c) != u;
    direction_counts_known<CH>(hist, trash, idu2, include, cnt_u2);
    direction_counts_known<CH>(hist, trash, idu, include, cnt_u);
#pragma unroll
    for (int c = 0; c < CH; ++c) include[c] = (lane + kWave * c < m) && (lane + kWave * c) != v;
    direction_counts_known<CH>(hist, trash, idv2, include, cnt_v2);
    direction_counts_known<CH>(hist, trash, idv, include, cnt_v);
    int nb[CH], w2[CH];
    uint32_t part2 = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int k = lane + kWave * c;
      const int bit_old = idu[c] == id_uv ? 1 : 0;
      const int bit_new = idu2[c] == id_u2v2 ? 1 : 0;
      nb[c] = bad[c];
      if (others[c]) nb[c] += (cnt_u2[c] + cnt_v2[c] - bit_new) - (cnt_u[c] + cnt_v[c] - bit_old);
      if (k == u) nb[c] = bad_u2;
      if (k == v) nb[c] = bad_v2;
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) w2[c] = (int)weight_of(row2, nb[c]);
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      if (lane + kWave * c >= m) w2[c] = 0;
      part2 += (uint32_t)w2[c];
    }
    const uint32_t s1_new = wave_sum_dpp(part2);
    const uint64_t wu2 = (uint32_t)slot_value<CH>(w2, u), wv2 = (uint32_t)slot_value<CH>(w2, v);
    const uint64_t s1sq_new = (uint64_t)s1_new * s1_new;
    const uint64_t big_n2 = s1sq_new + mm4 * wu2 * wv2;
    // r N S1'^2 < 2^126; thr N' < 2^111 times S1^2 < 2^46 may need 192 bits
    const unsigned __int128 lhs = (unsigned __int128)r[3] * big_n * s1sq_new;
    if (!less_than_product(lhs, (unsigned __int128)t64 * big_n2, s1sq)) continue;

#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int k = lane + kWave * c;
      bad[c] = nb[c];
      w[c] = w2[c];
      if (k == u) key[c] = key_u2;
      if (k == v) key[c] = key_v2;
      // u now holds column cv, v column cu: both lanes of each row
      if ((k >> 1) == ru) row_word[c] = (int)(((uint32_t)row_word[c] & ~(0xFFu << (8 * (u & 1)))) | ((uint32_t)cv << (8 * (u & 1))));
      if ((k >> 1) == rv) row_word[c] = (int)(((uint32_t)row_word[c] & ~(0xFFu << (8 * (v & 1)))) | ((uint32_t)cu << (8 * (v & 1))));
    }
    s1 = s1_new;
    build_cdf();
    score += delta;
    removed += dd;
    accepted += 1;
    if (score < best) {
      best = score;
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int k = lane + kWave * c;
        if (k < m) a.best_columns[base + k] = key[c] - (k >> 1) * width;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int k = lane + kWave * c;
    if (k < m) {
      a.columns[base + k] = key[c] - (k >> 1) * width;
      a.badness[base + k] = bad[c];
    }
  }
  if (lane == 0) {
    a.scores[slot] = score;
    a.removed[slot] = removed;
    a.best_scores[slot] = best;
    a.accepted[slot] += accepted;
    atomicAdd(&a.proposals[0], valid_moves);
    atomicAdd(&a.proposals[1], (unsigned long long)accepted);
  }
}

template <int CH>
__global__ void __launch_bounds__(1024) walk_learned_kernel(WalkArgs a, LearnedArgs L,
                                                           const uint16_t* __restrict__ table_g, DirTableInfo info) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int waves = blockDim.x / kWave;
  const int wave = threadIdx.x / kWave, lane = lane_id();
  uint16_t* table = reinterpret_cast<uint16_t*>(smem);
  uint32_t* hist = reinterpret_cast<uint32_t*>(smem + table_lds_bytes(info)) + wave * info.hist_words;
  stage_table(table_g, table, info.entries);
  for (int i = lane; i < info.hist_words; i += kWave) hist[i] = 0;
  __syncthreads();
  const int slot = blockIdx.x * waves + wave;
  if (slot < a.replicas) walk_replica_learned<CH>(a, L, table, hist, info, slot);
}

__global__ void bump_sweep_kernel(uint32_t* counters) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counters[0] += 1u;
}

struct LearnedPlan {
  const uint16_t* table;
  DirTableInfo info;
  int waves;
  unsigned blocks;
  size_t lds;
  void (*kernel)(WalkArgs, LearnedArgs, const uint16_t*, DirTableInfo);
};

static int learned_plan(const n3_chains* c, LearnedPlan* plan) {
  int st = device_dirtable(c->n, &plan->table, &plan->info);
  if (st) return st;
  int lds_limit = 0, compute_units = 0;
  st = device_limits(&lds_limit, &compute_units);
  if (st) return st;
  const int table_bytes = table_lds_bytes(plan->info), hist_bytes = plan->info.hist_words * 4;
  const int fit = (lds_limit - table_bytes) / hist_bytes;
  if (fit < 1) return set_error(N3_EINVAL, "direction table for n=%d does not fit in LDS", c->n);
  int waves = fit < 16 ? fit : 16;
  if (c->block_waves > 0) {
    if (c->block_waves > waves)
      return set_error(N3_EINVAL, "block_waves=%d exceeds the %d that fit in LDS for n=%d", c->block_waves, waves,
                       c->n);
    waves = c->block_waves;
  } else {
    const int per_cu = (c->replicas + compute_units - 1) / compute_units;
    if (waves > per_cu) waves = per_cu < 1 ? 1 : per_cu;
  }
  plan->waves = waves;
  plan->blocks = (unsigned)((c->replicas + waves - 1) / waves);
  plan->lds = (size_t)table_bytes + (size_t)waves * hist_bytes;
  switch ((2 * c->n + kWave - 1) / kWave) {
    case 1: plan->kernel = walk_learned_kernel<1>; break;
    case 2: plan->kernel = walk_learned_kernel<2>; break;
    case 3: plan->kernel = walk_learned_kernel<3>; break;
    default: plan->kernel = walk_learned_kernel<4>; break;
  }
  return N3_OK;
}

static int enqueue_learned(const n3_chains* c, const LearnedPlan& plan, const LearnedArgs& L, int steps, uint64_t seed,
                           hipStream_t stream) {
  hipLaunchKernelGGL(plan.kernel, dim3(plan.blocks), dim3(plan.waves * kWave), plan.lds, stream,
                     walk_args(c, steps, 0, seed), L, plan.table, plan.info);
  int st = hip_status(hipGetLastError(), "walk_learned_kernel launch");
  if (st) return st;
  hipLaunchKernelGGL(bump_sweep_kernel, dim3(1), dim3(64), 0, stream, c->counters);
  return hip_status(hipGetLastError(), "bump_sweep_kernel launch");
}

struct LearnedEntry {
  std::vector<unsigned char> key;
  hipGraph_t graph;
  hipGraphExec_t exec;
  int per_graph;
};
static std::mutex g_learned_mutex;
static std::vector<LearnedEntry> g_learned;

}  // namespace no3

using namespace no3;

extern "C" int n3_metropolis_thresholds64(const float* host_temperatures, int replicas, int k, uint64_t* host_out) {
  if (!host_temperatures || !host_out || replicas < 1 || k < 2) return set_error(N3_EINVAL, "bad threshold arguments");
  const double cap = 9223372036854775808.0;  // 2^63
  for (int r = 0; r < replicas; ++r) {
    const double t = (double)host_temperatures[r];
    if (!(t > 0.0) || !std::isfinite(t)) return set_error(N3_EINVAL, "temperatures must be positive and finite");
    for (int d = -k; d < k; ++d) {
      const double x = std::ldexp(std::exp(-(double)d / t), 32);
      host_out[(long long)r * 2 * k + (d + k)] = x >= cap ? (uint64_t)1 << 63 : (uint64_t)std::floor(x);
    }
  }
  return N3_OK;
}

extern "C" int n3_temper_run_learned(const n3_chains* chains, const uint32_t* weights, const int32_t* temp_index,
                                     int max_energy, const uint64_t* thr64, int k64, int sweeps, int steps, int ladder,
                                     uint64_t seed, void* stream) {
  int st = check_chains(chains);
  if (st) return st;
  if (!weights || !temp_index || !thr64 || max_energy < 0 || k64 < 2 || steps < 1 || sweeps < 0)
    return set_error(N3_EINVAL, "bad learned-walk arguments");
  if (chains->strengths)
    return set_error(N3_EINVAL, "umbrella potentials are supported by n3_temper_advance and n3_temper_run only");
  if (ladder > 1) {
    st = check_exchange(chains, ladder);
    if (st) return st;
  }
  if (sweeps == 0) return N3_OK;
  LearnedPlan plan;
  st = learned_plan(chains, &plan);
  if (st) return st;
  const LearnedArgs L{weights, temp_index, max_energy, reinterpret_cast<const unsigned long long*>(thr64), k64};
  hipStream_t s = as_stream(stream);
  const int per_graph = sweeps < 16 ? sweeps : 16;
  int device = 0;
  (void)hipGetDevice(&device);
  std::vector<unsigned char> key(sizeof(n3_chains) + sizeof(LearnedArgs) + 6 * sizeof(intptr_t));
  std::memcpy(key.data(), chains, sizeof(n3_chains));
  std::memcpy(key.data() + sizeof(n3_chains), &L, sizeof(LearnedArgs));
  const intptr_t extra[6] = {(intptr_t)s, (intptr_t)device, per_graph, steps, ladder, (intptr_t)seed};
  std::memcpy(key.data() + sizeof(n3_chains) + sizeof(LearnedArgs), extra, sizeof(extra));
  std::lock_guard<std::mutex> lock(g_learned_mutex);
  LearnedEntry* entry = nullptr;
  for (auto& e : g_learned)
    if (e.key == key) entry = &e;
  if (!entry) {
    LearnedEntry e;
    e.key = key;
    e.per_graph = per_graph;
    hipStream_t cap = nullptr;
    st = hip_status(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreate");
    if (st) return st;
    st = hip_status(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (!st) {
      for (int i = 0; i < per_graph && !st; ++i) {
        st = enqueue_learned(chains, plan, L, steps, seed, cap);
        if (!st && ladder > 1) st = enqueue_exchange(chains, ladder, seed, cap);
      }
      hipGraph_t g = nullptr;
      const int end = hip_status(hipStreamEndCapture(cap, &g), "hipStreamEndCapture");
      if (!st) st = end;
      if (!st) {
        st = hip_status(hipGraphInstantiate(&e.exec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
        e.graph = g;
        if (st) (void)hipGraphDestroy(g);
      } else if (g) {
        (void)hipGraphDestroy(g);
      }
    }
    (void)hipStreamDestroy(cap);
    if (st) return st;
    g_learned.push_back(e);
    entry = &g_learned.back();
  }
  const int full = sweeps / entry->per_graph, rest = sweeps % entry->per_graph;
  for (int i = 0; i < full; ++i) {
    st = hip_status(hipGraphLaunch(entry->exec, s), "hipGraphLaunch");
    if (st) return st;
  }
  for (int i = 0; i < rest && !st; ++i) {
    st = enqueue_learned(chains, plan, L, steps, seed, s);
    if (!st && ladder > 1) st = enqueue_exchange(chains, ladder, seed, s);
  }
  return st;
}

extern "C" void n3_learned_release(void) {
  std::lock_guard<std::mutex> lock(g_learned_mutex);
  for (auto& e : g_learned) {
    (void)hipGraphExecDestroy(e.exec);
    (void)hipGraphDestroy(e.graph);
  }
  g_learned.clear();
}
