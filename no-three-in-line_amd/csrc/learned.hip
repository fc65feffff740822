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

#include "mlp_common.h"
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
  const uint32_t sweep = a.counters[0] + a.sweep_offset;
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
      // opaque copies of the key: its round schedule is derived here, once
      // per batch, instead of being kept live in scalar registers all loop
      uint32_t k0 = a.seed_lo, k1 = a.seed_hi;
      asm volatile("" : "+s"(k0), "+s"(k1));
      philox4x32_10((uint32_t)(step + lane), sweep, (uint32_t)slot, kTagLearn, k0, k1, learn_r);
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
      part += byte_counter(old_u2[c], (uint32_t)idu2[c]) +
              (byte_counter(old_v2[c], (uint32_t)idv2[c]) << 16);
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
  const int slot = walk_slot((int)blockIdx.x, wave, waves, a.ladder);
  if (slot < a.replicas) walk_replica_learned<CH>(a, L, table, hist, info, slot);
}

// ---- point-network proposals ---------------------------------------------
//
// The weight of slot k is the point network's integer weight of its own
// features (csrc/mlp_common.h: collinear-pair count, row, column, the
// replica's temperature code and the state's energy), evaluated inside the
// walk for every state it visits: 32 slots per tile as H^T = W1 X^T on
// v_mfma_f32_32x32x8_bf16 (exact: W1, b1 and the features lie on the 1/128
// grid), then a fixed-order fp32 fma chain per lane half for the output
// layer.  Proposals draw pairs as the table walk does, with 64-bit
// inverse-CDF targets; acceptance is delayed (two stages): the Metropolis
// test on the exact delta first, and only for moves that pass it the moved
// state's weights and the proposal-ratio test h2 N S1'^2 < 2^32 N' S1^2.

constexpr uint32_t kTagHast = 0x48415354u;  // "HAST"

struct PointArgs {
  const float* consts;                 // kPointConsts floats (b1, w2 by lane half, b2)
  const uint16_t* params;              // bf16 network; W1 is staged for the A operand
  const uint32_t* expw;                // kPointExpEntries weights by floor(64 g)
  const int32_t* lt_q;                 // [R] round(32 ln T)
  const unsigned long long* thr64;     // [R][2 k64], entry d + k64
  int k64;
};

// LDS bytes of the point network's staged constants and W1.
constexpr int kPointLdsBytes = kPointConsts * 4 + kHidden * kIn * 2;

// w[c]: integer weight of slot lane + 64 c of the state (key, bad, energy),
// 0 past the last slot.  `cst` and `w1` are the LDS copies.
template <int CH>
__device__ inline void point_weights(const float* cst, const uint16_t* w1, const uint32_t* __restrict__ expw, int m,
                                     int width, int lt_q, int energy, const int (&key)[CH], const int (&bad)[CH],
                                     int (&w)[CH]) {
  const int lane = lane_id(), half = lane >> 5, j = lane & 31;
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int bad_other = __shfl_xor(bad[c], 32, kWave), key_other = __shfl_xor(key[c], 32, kWave);
    int wt[2] = {0, 0};
#pragma unroll
    for (int sub = 0; sub < 2; ++sub) {
      const int tile = 2 * c + sub;
      if (32 * tile >= m) continue;  // uniform
      // slot s of this tile row lives in lane 32 sub + j of chunk c
      const int s = 32 * tile + j;
      const int b = half == sub ? bad[c] : bad_other;
      const int k = half == sub ? key[c] : key_other;
      const int row = s >> 1;
      int code[kIn];
      point_codes(b, row, k - row * width, lt_q, energy, code);
      bf16x4_t x;
#pragma unroll
      for (int i = 0; i < 4; ++i) x[i] = (short)code_bf16(half ? code[4 + i] : code[i]);
      float g = half ? 0.0f : cst[kPointB2];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const bf16x4_t a = *reinterpret_cast<const bf16x4_t*>(w1 + (32 * t + j) * kIn + 4 * half);
        floatx16_t acc;
        const float* b1 = cst + kPointB1 + (2 * t + half) * 16;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = b1[r];
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, x, acc, 0, 0, 0);
        const float* w2 = cst + kPointW2 + (2 * t + half) * 16;
#pragma unroll
        for (int r = 0; r < 16; ++r) g = fmaf(w2[r], fmaxf(acc[r], 0.0f), g);
      }
      // both halves hold the same sum (fp32 addition commutes); the clamp
      // to [-7, 6] keeps the conversion defined and changes no weight
      g += __shfl_xor(g, 32, kWave);
      g = fminf(fmaxf(g, -7.0f), 6.0f);
      int q = (int)floorf(g * 64.0f);
      q = q < kPointQMin ? kPointQMin : (q > kPointQMax ? kPointQMax : q);
      wt[sub] = s < m ? (int)expw[q - kPointQMin] : 0;
    }
    // lane 32 h + j holds slot 64 c + 32 h + j: tile 2 c + h, row j
    w[c] = half ? wt[1] : wt[0];
  }
}

template <int CH>
__device__ void walk_replica_point(const WalkArgs& a, const PointArgs& P, const uint16_t* table, uint32_t* hist,
                                   const float* cst, const uint16_t* w1, const DirTableInfo& info, int slot_in) {
  const int lane = lane_id();
  const int slot = __builtin_amdgcn_readfirstlane(slot_in);
  const int n = a.n, m = 2 * n, width = info.width, offset = info.offset;
  const uint32_t sweep = a.counters[0] + a.sweep_offset;
  const long long base = (long long)slot * m;

  int key[CH], bad[CH], ref[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int k = lane + kWave * c;
    key[c] = k < m ? (k >> 1) * width + a.columns[base + k] : 0;
    bad[c] = k < m ? a.badness[base + k] : 0;
    ref[c] = k < m ? a.reference[k] : -1;
  }
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
  const int lt_q = P.lt_q[slot];
  const unsigned long long* thr = P.thr64 + (long long)slot * 2 * P.k64;
  const unsigned long long thr_lo = lane < 2 * P.k64 ? thr[lane] : 0ull;
  const unsigned long long thr_hi = lane + kWave < 2 * P.k64 ? thr[lane + kWave] : 0ull;

  int w[CH];
  point_weights<CH>(cst, w1, P.expw, m, width, lt_q, score, key, bad, w);
  uint32_t s1;
  uint32_t cdf[CH];
  auto build_cdf = [&]() {
    uint32_t carry = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const uint32_t x = wave_scan_dpp((uint32_t)w[c]);
      cdf[c] = carry + x;
      carry += (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
    }
    s1 = carry;
  };
  build_cdf();

  uint32_t learn_r[4], hast_r[4];
  int batch = -kWave;
  for (int step = 0; step < a.steps; ++step) {
    if (step - batch >= kWave) {
      batch = step;
      // opaque copies of the key (see walk_replica_learned)
      uint32_t k0 = a.seed_lo, k1 = a.seed_hi;
      asm volatile("" : "+s"(k0), "+s"(k1));
      philox4x32_10((uint32_t)(step + lane), sweep, (uint32_t)slot, kTagLearn, k0, k1, learn_r);
      philox4x32_10((uint32_t)(step + lane), sweep, (uint32_t)slot, kTagHast, k0, k1, hast_r);
    }
    uint32_t r[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = (uint32_t)__builtin_amdgcn_readlane((int)learn_r[i], step - batch);
    int u, v;
    if (r[0] < kUniformShare) {
      u = (int)mulhi32(r[1], (uint32_t)m);
      v = (int)mulhi32(r[2], (uint32_t)m);
    } else {
      const uint32_t h0 = (uint32_t)__builtin_amdgcn_readlane((int)hast_r[0], step - batch);
      const uint32_t h1 = (uint32_t)__builtin_amdgcn_readlane((int)hast_r[1], step - batch);
      const uint32_t target_u = (uint32_t)__umul64hi(((unsigned long long)h0 << 32) | r[1], s1);
      const uint32_t target_v = (uint32_t)__umul64hi(((unsigned long long)h1 << 32) | r[2], s1);
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
    __builtin_amdgcn_sched_barrier(0);
    uint32_t part = 0;
    int joint_new = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      part += byte_counter(old_u2[c], (uint32_t)idu2[c]) +
              (byte_counter(old_v2[c], (uint32_t)idv2[c]) << 16);
      joint_new += __popcll(lanes(others[c]) & lanes(idu2[c] == id_u2v2));
    }
    const uint32_t packed = wave_sum_dpp(part);
    const int bad_u2 = (int)(packed & 0xFFFFu), bad_v2 = (int)(packed >> 16);
    const int bad_u = slot_value<CH>(bad, u), bad_v = slot_value<CH>(bad, v);
    const int delta = (bad_u2 + bad_v2 - joint_new) - (bad_u + bad_v - joint_old);
    // stage 1: Metropolis on the exact delta
    const int d = delta < -P.k64 ? -P.k64 : (delta > P.k64 - 1 ? P.k64 - 1 : delta);
    if (!((unsigned long long)r[3] < register_table64(thr_lo, thr_hi, thr, d + P.k64))) continue;

    // stage 2: every slot's badness and weight in the moved state
    int cnt_u2[CH], cnt_v2[CH], cnt_u[CH], cnt_v[CH], idv[CH];
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
    int nb[CH], key2[CH], w2[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int k = lane + kWave * c;
      const int bit_old = idu[c] == id_uv ? 1 : 0;
      const int bit_new = idu2[c] == id_u2v2 ? 1 : 0;
      nb[c] = bad[c];
      if (others[c]) nb[c] += (cnt_u2[c] + cnt_v2[c] - bit_new) - (cnt_u[c] + cnt_v[c] - bit_old);
      if (k == u) nb[c] = bad_u2;
      if (k == v) nb[c] = bad_v2;
      key2[c] = k == u ? key_u2 : (k == v ? key_v2 : key[c]);
    }
    point_weights<CH>(cst, w1, P.expw, m, width, lt_q, score + delta, key2, nb, w2);
    uint32_t part2 = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) part2 += (uint32_t)w2[c];
    const uint64_t s1_new = wave_sum_dpp(part2);
    const uint64_t wu = (uint32_t)slot_value<CH>(w, u), wv = (uint32_t)slot_value<CH>(w, v);
    const uint64_t wu2 = (uint32_t)slot_value<CH>(w2, u), wv2 = (uint32_t)slot_value<CH>(w2, v);
    const uint64_t mm4 = 4ull * (uint64_t)m * (uint64_t)m;
    const uint64_t big_n = (uint64_t)s1 * s1 + mm4 * wu * wv;        // < 2^48
    const uint64_t big_n2 = s1_new * s1_new + mm4 * wu2 * wv2;
    const uint32_t h2 = (uint32_t)__builtin_amdgcn_readlane((int)hast_r[2], step - batch);
    // h2 N S1'^2 < 2^126 and 2^32 N' S1^2 < 2^126
    const unsigned __int128 lhs = (unsigned __int128)h2 * ((unsigned __int128)big_n * (s1_new * s1_new));
    const unsigned __int128 rhs = ((unsigned __int128)big_n2 * ((uint64_t)s1 * s1)) << 32;
    if (!(lhs < rhs)) continue;

#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int k = lane + kWave * c;
      bad[c] = nb[c];
      w[c] = w2[c];
      key[c] = key2[c];
      if ((k >> 1) == ru) row_word[c] = (int)(((uint32_t)row_word[c] & ~(0xFFu << (8 * (u & 1)))) | ((uint32_t)cv << (8 * (u & 1))));
      if ((k >> 1) == rv) row_word[c] = (int)(((uint32_t)row_word[c] & ~(0xFFu << (8 * (v & 1)))) | ((uint32_t)cu << (8 * (v & 1))));
    }
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

// Stages the point network's constants and W1 after the direction table.
__device__ inline void stage_point(const PointArgs& P, float* cst, uint16_t* w1) {
  for (int i = threadIdx.x; i < kPointConsts; i += blockDim.x) cst[i] = P.consts[i];
  const uint2* src = reinterpret_cast<const uint2*>(P.params + kW1);
  uint2* dst = reinterpret_cast<uint2*>(w1);
  for (int i = threadIdx.x; i < kHidden * kIn / 4; i += blockDim.x) dst[i] = src[i];
}

template <int CH>
__global__ void __launch_bounds__(1024) walk_point_kernel(WalkArgs a, PointArgs P, const uint16_t* __restrict__ table_g,
                                                         DirTableInfo info) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int waves = blockDim.x / kWave;
  const int wave = threadIdx.x / kWave, lane = lane_id();
  uint16_t* table = reinterpret_cast<uint16_t*>(smem);
  float* cst = reinterpret_cast<float*>(smem + table_lds_bytes(info));
  uint16_t* w1 = reinterpret_cast<uint16_t*>(cst + kPointConsts);
  uint32_t* hist = reinterpret_cast<uint32_t*>(smem + table_lds_bytes(info) + kPointLdsBytes) + wave * info.hist_words;
  stage_table(table_g, table, info.entries);
  stage_point(P, cst, w1);
  for (int i = lane; i < info.hist_words; i += kWave) hist[i] = 0;
  __syncthreads();
  const int slot = walk_slot((int)blockIdx.x, wave, waves, a.ladder);
  if (slot < a.replicas) walk_replica_point<CH>(a, P, table, hist, cst, w1, info, slot);
}

// Weights of every slot of `count` states (one wavefront per state), from
// the same device function as the walk.
template <int CH>
__global__ void __launch_bounds__(256) point_weights_kernel(PointArgs P, const int32_t* __restrict__ columns,
                                                             const int32_t* __restrict__ badness,
                                                             const int32_t* __restrict__ scores, int n, long long count,
                                                             int32_t* __restrict__ out) {
  __shared__ float cst[kPointConsts];
  __shared__ __align__(16) uint16_t w1[kHidden * kIn];
  stage_point(P, cst, w1);
  __syncthreads();
  const int m = 2 * n, width = 2 * n - 1, lane = lane_id();
  const long long s = (long long)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (s >= count) return;  // whole wavefronts leave together
  int key[CH], bad[CH], w[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int k = lane + kWave * c;
    key[c] = k < m ? (k >> 1) * width + columns[s * m + k] : 0;
    bad[c] = k < m ? badness[s * m + k] : 0;
  }
  point_weights<CH>(cst, w1, P.expw, m, width, P.lt_q[s], scores[s], key, bad, w);
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int k = lane + kWave * c;
    if (k < m) out[s * m + k] = w[c];
  }
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

// The walk of sweep counters[0] + sweep_offset; with `advance` a counter
// kernel advances counters[0] after it.
static int enqueue_learned(const n3_chains* c, const LearnedPlan& plan, const LearnedArgs& L, int steps, uint64_t seed,
                           hipStream_t stream, int ladder, uint32_t sweep_offset = 0, bool advance = true) {
  WalkArgs a = walk_args(c, steps, 0, seed);
  a.ladder = ladder;
  a.sweep_offset = sweep_offset;
  hipLaunchKernelGGL(plan.kernel, dim3(plan.blocks), dim3(plan.waves * kWave), plan.lds, stream, a, L, plan.table,
                     plan.info);
  int st = hip_status(hipGetLastError(), "walk_learned_kernel launch");
  if (st || !advance) return st;
  hipLaunchKernelGGL(bump_sweep_kernel, dim3(1), dim3(64), 0, stream, c->counters);
  return hip_status(hipGetLastError(), "bump_sweep_kernel launch");
}

struct PointPlan {
  const uint16_t* table;
  DirTableInfo info;
  int waves;
  unsigned blocks;
  size_t lds;
  void (*kernel)(WalkArgs, PointArgs, const uint16_t*, DirTableInfo);
};

static int point_plan(const n3_chains* c, PointPlan* plan) {
  int st = device_dirtable(c->n, &plan->table, &plan->info);
  if (st) return st;
  int lds_limit = 0, compute_units = 0;
  st = device_limits(&lds_limit, &compute_units);
  if (st) return st;
  const int fixed = table_lds_bytes(plan->info) + kPointLdsBytes, hist_bytes = plan->info.hist_words * 4;
  const int fit = (lds_limit - fixed) / hist_bytes;
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
  plan->lds = (size_t)fixed + (size_t)waves * hist_bytes;
  switch ((2 * c->n + kWave - 1) / kWave) {
    case 1: plan->kernel = walk_point_kernel<1>; break;
    case 2: plan->kernel = walk_point_kernel<2>; break;
    case 3: plan->kernel = walk_point_kernel<3>; break;
    default: plan->kernel = walk_point_kernel<4>; break;
  }
  return N3_OK;
}

// The walk of sweep counters[0] + sweep_offset; with `advance` a counter
// kernel advances counters[0] after it.
static int enqueue_point(const n3_chains* c, const PointPlan& plan, const PointArgs& P, int steps, uint64_t seed,
                         hipStream_t stream, int ladder, uint32_t sweep_offset = 0, bool advance = true) {
  WalkArgs a = walk_args(c, steps, 0, seed);
  a.ladder = ladder;
  a.sweep_offset = sweep_offset;
  hipLaunchKernelGGL(plan.kernel, dim3(plan.blocks), dim3(plan.waves * kWave), plan.lds, stream, a, P, plan.table,
                     plan.info);
  int st = hip_status(hipGetLastError(), "walk_point_kernel launch");
  if (st || !advance) return st;
  hipLaunchKernelGGL(bump_sweep_kernel, dim3(1), dim3(64), 0, stream, c->counters);
  return hip_status(hipGetLastError(), "bump_sweep_kernel launch");
}

struct LearnedEntry {
  std::vector<unsigned char> key;
  hipGraph_t graph;
  hipGraphExec_t exec;
  int per_graph;
  hipStream_t stream;
};

// At most this many cached learned-walk graphs; the oldest is dropped (once
// its stream has drained) to make room.
constexpr size_t kMaxLearnedGraphs = 16;
static std::mutex g_learned_mutex;
static std::vector<LearnedEntry> g_learned;

// `sweeps` sweeps of a learned walk: whole graphs of up to 16 sweeps
// captured once per (chains, walk arguments, steps, ladder, seed, stream,
// device) and replayed, the remainder enqueued directly.  `one_sweep`
// enqueues one sweep (walk, counter bump, exchange) on the given stream.
template <class OneSweep>
static int run_graphs(const n3_chains* chains, const void* args, size_t args_size, int sweeps, int steps, int ladder,
                      uint64_t seed, hipStream_t s, OneSweep one_sweep) {
  const int per_graph = sweeps < 16 ? sweeps : 16;
  int device = 0;
  (void)hipGetDevice(&device);
  std::vector<unsigned char> key(sizeof(n3_chains) + args_size + 6 * sizeof(intptr_t));
  std::memcpy(key.data(), chains, sizeof(n3_chains));
  std::memcpy(key.data() + sizeof(n3_chains), args, args_size);
  const intptr_t extra[6] = {(intptr_t)s, (intptr_t)device, per_graph, steps, ladder, (intptr_t)seed};
  std::memcpy(key.data() + sizeof(n3_chains) + args_size, extra, sizeof(extra));
  std::lock_guard<std::mutex> lock(g_learned_mutex);
  LearnedEntry* entry = nullptr;
  for (auto& e : g_learned)
    if (e.key == key) entry = &e;
  int st = N3_OK;
  if (!entry) {
    LearnedEntry e;
    e.key = key;
    e.per_graph = per_graph;
    e.stream = s;
    hipStream_t cap = nullptr;
    st = hip_status(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreate");
    if (st) return st;
    st = hip_status(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (!st) {
      // the k-th sweep of the graph reads its sweep and round numbers as the
      // counters plus k; one counter kernel advances them at the end
      for (int i = 0; i < per_graph && !st; ++i) st = one_sweep(cap, (uint32_t)i, false);
      if (!st) st = enqueue_counters(chains, (uint32_t)per_graph, ladder > 1 ? (uint32_t)per_graph : 0u, cap);
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
    if (g_learned.size() >= kMaxLearnedGraphs) {
      LearnedEntry& old = g_learned.front();
      (void)hipStreamSynchronize(old.stream);
      (void)hipGraphExecDestroy(old.exec);
      (void)hipGraphDestroy(old.graph);
      g_learned.erase(g_learned.begin());
    }
    g_learned.push_back(e);
    entry = &g_learned.back();
  }
  const int full = sweeps / entry->per_graph, rest = sweeps % entry->per_graph;
  for (int i = 0; i < full; ++i) {
    st = hip_status(hipGraphLaunch(entry->exec, s), "hipGraphLaunch");
    if (st) return st;
  }
  for (int i = 0; i < rest && !st; ++i) st = one_sweep(s, 0u, true);
  return st;
}

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
  return run_graphs(chains, &L, sizeof L, sweeps, steps, ladder, seed, as_stream(stream),
                    [&](hipStream_t s, uint32_t k, bool advance) {
                      int e = enqueue_learned(chains, plan, L, steps, seed, s, ladder, k, advance);
                      if (!e && ladder > 1) e = enqueue_exchange(chains, ladder, seed, s, k, advance);
                      return e;
                    });
}

extern "C" int n3_temper_run_learned_point(const n3_chains* chains, const uint16_t* params, const float* consts,
                                           const uint32_t* expw, const int32_t* lt_q, const uint64_t* thr64, int k64,
                                           int sweeps, int steps, int ladder, uint64_t seed, void* stream) {
  int st = check_chains(chains);
  if (st) return st;
  if (!params || !consts || !expw || !lt_q || !thr64 || k64 < 2 || steps < 1 || sweeps < 0)
    return set_error(N3_EINVAL, "bad point-walk arguments");
  if (chains->strengths)
    return set_error(N3_EINVAL, "umbrella potentials are supported by n3_temper_advance and n3_temper_run only");
  if (chains->n > 127) return set_error(N3_EINVAL, "the point network's features hold rows and columns below 128");
  if (ladder > 1) {
    st = check_exchange(chains, ladder);
    if (st) return st;
  }
  if (sweeps == 0) return N3_OK;
  PointPlan plan;
  st = point_plan(chains, &plan);
  if (st) return st;
  const PointArgs P{consts, params, expw, lt_q, reinterpret_cast<const unsigned long long*>(thr64), k64};
  return run_graphs(chains, &P, sizeof P, sweeps, steps, ladder, seed, as_stream(stream),
                    [&](hipStream_t s, uint32_t k, bool advance) {
                      int e = enqueue_point(chains, plan, P, steps, seed, s, ladder, k, advance);
                      if (!e && ladder > 1) e = enqueue_exchange(chains, ladder, seed, s, k, advance);
                      return e;
                    });
}

extern "C" int n3_point_weights(const int32_t* columns, const int32_t* badness, const int32_t* scores,
                                const int32_t* lt_q, int n, int64_t count, const uint16_t* params, const float* consts,
                                const uint32_t* expw, int32_t* weights, void* stream) {
  if (!columns || !badness || !scores || !lt_q || !params || !consts || !expw || !weights || count < 0 || n < 2 ||
      n > 127)
    return set_error(N3_EINVAL, "bad point-weight arguments");
  if (count == 0) return N3_OK;
  const PointArgs P{consts, params, expw, lt_q, nullptr, 0};
  const long long blocks = (count + 3) / 4;
  const dim3 grid((unsigned)blocks), block(256);
  hipStream_t s = as_stream(stream);
  switch ((2 * n + kWave - 1) / kWave) {
    case 1: hipLaunchKernelGGL(point_weights_kernel<1>, grid, block, 0, s, P, columns, badness, scores, n, (long long)count, weights); break;
    case 2: hipLaunchKernelGGL(point_weights_kernel<2>, grid, block, 0, s, P, columns, badness, scores, n, (long long)count, weights); break;
    case 3: hipLaunchKernelGGL(point_weights_kernel<3>, grid, block, 0, s, P, columns, badness, scores, n, (long long)count, weights); break;
    default: hipLaunchKernelGGL(point_weights_kernel<4>, grid, block, 0, s, P, columns, badness, scores, n, (long long)count, weights); break;
  }
  return hip_status(hipGetLastError(), "point_weights_kernel launch");
}

extern "C" int n3_point_exp_table(uint32_t* host_out) {
  if (!host_out) return set_error(N3_EINVAL, "null output");
  for (int q = kPointQMin; q <= kPointQMax; ++q) {
    const double w = std::nearbyint(128.0 * std::exp((double)q / 64.0));
    host_out[q - kPointQMin] = (uint32_t)(w < 1.0 ? 1.0 : (w > 32768.0 ? 32768.0 : w));
  }
  return N3_OK;
}

extern "C" int n3_point_constants(const uint16_t* host_params, float* host_consts) {
  if (!host_params || !host_consts) return set_error(N3_EINVAL, "null argument");
  // W1 and b1 must lie on the grid of multiples of 1/128 within (-2, 2), or
  // the device's fp32 sums would round
  for (int i = kW1; i < kW2; ++i) {
    const float x = bf16_to_float(host_params[i]) * (float)kGridScale;
    if (!(x == std::nearbyint(x)) || std::fabs(x) > (float)kGridMax)
      return set_error(N3_EINVAL, "parameter %d (W1 or b1) is not a multiple of 1/128 in (-2, 2)", i);
  }
  for (int t = 0; t < 4; ++t)
    for (int half = 0; half < 2; ++half)
      for (int r = 0; r < 16; ++r) {
        const int h = hidden_of(t, r, half);
        host_consts[kPointB1 + (2 * t + half) * 16 + r] = bf16_to_float(host_params[kB1 + h]);
        host_consts[kPointW2 + (2 * t + half) * 16 + r] = bf16_to_float(host_params[kW2 + h]);
      }
  host_consts[kPointB2] = bf16_to_float(host_params[kB2]);
  for (int i = kPointB2 + 1; i < kPointConsts; ++i) host_consts[i] = 0.0f;
  return N3_OK;
}

extern "C" void n3_learned_release(void) {
  std::lock_guard<std::mutex> lock(g_learned_mutex);
  for (auto& e : g_learned) {
    (void)hipGraphExecDestroy(e.exec);
    (void)hipGraphDestroy(e.graph);
  }
  g_learned.clear();
}
