// Replica-exchange tempering (tempering.py:19-225) on the device.
//
// walk: one wavefront per replica slot keeps the slot's state in registers
// (lane l holds slots l, l+64, ...: key row*W + column and its badness) and
// runs `steps` proposals of the reference's walk kernel (tempering.py:48-110)
// without leaving the kernel.  The exact energy change of a 2-switch (u, v)
// comes from four direction histograms in LDS instead of the reference's
// O(m^2) determinant pair loops:
//
//   delta = [bad'(u') + bad'(v') - joint'(u',v')] - [bad(u) + bad(v) - joint(u,v)]
//
// bad'(x) = sum over the histogram seen from x of C(cnt, 2); joint counts the
// points on the line through both moved points.  On acceptance the badness of
// every other slot k changes by
//
//   (cnt_u'(k) + cnt_v'(k) - [k on u'v']) - (cnt_u(k) + cnt_v(k) - [k on uv])
//
// where cnt_x(k) is the number of points on the line (x, k) seen from x (the
// slot-level form of the db[] accumulation of tempering.py:71-88).
//
// exchange: adjacent-temperature swaps inside each ladder
// (tempering.py:169-191), one wavefront per candidate pair; an accepted pair
// swaps the slots' columns, badness, scores and removed counters.
#include <cmath>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <vector>

#include <hip/hip_cooperative_groups.h>

#include "runtime.h"
#include "temper_internal.h"
#include "walk.h"
#include "wave_ops.h"

namespace no3 {


// floor(2^32 exp(-x)) for x > 0, from operations that every IEEE-754 double
// unit rounds alike (fma, rint, ldexp, one product), so that a host replay
// computes the same word: Cody-Waite reduction y = -x = k ln2 + f with
// fdlibm's two-part ln2 (|f| <= ln2 / 2), then the Taylor series of exp(f)
// to degree 13 (truncation error below 1e-17) by Horner's rule.
__device__ inline uint32_t exp_word(double x) {
#pragma clang fp contract(off)
  if (!(x < 23.0)) return 0u;  // 2^32 exp(-23) < 1
  const double y = -x;
  const double k = rint(y * 1.4426950408889634);
  double f = fma(k, -6.93147180369123816490e-01, y);
  f = fma(k, -1.90821492927058770002e-10, f);
  double p = 1.0 / 6227020800.0;
  p = fma(p, f, 1.0 / 479001600.0);
  p = fma(p, f, 1.0 / 39916800.0);
  p = fma(p, f, 1.0 / 3628800.0);
  p = fma(p, f, 1.0 / 362880.0);
  p = fma(p, f, 1.0 / 40320.0);
  p = fma(p, f, 1.0 / 5040.0);
  p = fma(p, f, 1.0 / 720.0);
  p = fma(p, f, 1.0 / 120.0);
  p = fma(p, f, 1.0 / 24.0);
  p = fma(p, f, 1.0 / 6.0);
  p = fma(p, f, 0.5);
  p = fma(p, f, 1.0);
  p = fma(p, f, 1.0);
  const double v = ldexp(p, (int)k + 32);
  return v >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)v;
}

// Umbrella energy of a move (tempering.py:93-94): delta + strength (2 offset
// dd + dd^2) with offset = removed - center, rounded the same way on the host.
__device__ inline double umbrella_energy(int delta, int dd, int removed, float strength, float center) {
#pragma clang fp contract(off)
  const double offset = (double)removed - (double)center;
  const double a = fma(2.0 * offset, (double)dd, (double)(dd * dd));
  return fma((double)strength, a, (double)delta);
}

// One replica slot's `a.steps` proposals; `table` and `hist` are this
// wavefront's LDS direction table and histogram.  kPlain: the caller
// guarantees no proposal log and no badness bias, and the instantiation
// carries neither (fewer live scalar values in the loop).  kCounted (plain
// only): the histogram holds only the counted directions (DirTableInfo) and
// a trash word per lane and chunk, a quarter of the LDS, so that more waves
// fit beside the larger tables.
template <int CH, bool kUmbrella = false, bool kPlain = false, bool kCounted = false>
__device__ void walk_replica(const WalkArgs& a, const uint16_t* table, uint32_t* hist, const DirTableInfo& info,
                             int slot_in, uint32_t sweep) {
  const int lane = lane_id();
  // the slot is the same on every lane: keep it (and the Philox counters,
  // proposals and table indexes derived from it) in scalar registers
  const int slot = __builtin_amdgcn_readfirstlane(slot_in);
  const int n = a.n, m = 2 * n, width = info.width, offset = info.offset;
  const long long base = (long long)slot * m;

  // per lane: keys, badness and reference columns of slots lane + 64 c
  int key[CH], bad[CH], ref[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int k = lane + kWave * c;
    key[c] = k < m ? (k >> 1) * width + a.columns[base + k] : 0;
    bad[c] = k < m ? a.badness[base + k] : 0;
    ref[c] = k < m ? a.reference[k] : -1;
  }
  // both slots of a row, packed per lane: current columns (bytes 0, 1) and
  // reference columns (bytes 2, 3) of slots 2r and 2r+1; one readlane per
  // chunk gives all four for a row
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
  // lanes past the last slot hold slot 0's key (always a valid point); in
  // the histograms they update their own trash word, in the last chunk only
  {
    const int key0 = __builtin_amdgcn_readfirstlane(key[0]);
#pragma unroll
    for (int c = 0; c < CH; ++c)
      if (lane + kWave * c >= m) key[c] = key0;
  }
  const int trash = trash_word(info, lane);
  // kCounted: trash word of chunk c at counted_trash + 64 c
  const int counted_trash = info.counted_words - kCountedTrashWords + lane;
  const bool tail_in = lane + kWave * (CH - 1) < m;
  int score = a.scores[slot];
  int removed = a.removed[slot];
  // umbrella potential of this slot (tempering.py:92-95), uniform values;
  // a separate instantiation, so the plain walk keeps its registers
  float strength = 0.0f, center = 0.0f;
  double inverse_t = 0.0;
  if constexpr (kUmbrella) {
    strength = a.strengths[slot];
    center = a.centers[slot];
    inverse_t = a.inverse_temperatures[slot];
  }
  int best = a.best_scores[slot];
  int accepted = 0;
  unsigned long long valid_moves = 0;
  const uint32_t* thr = a.thr + (long long)slot * a.k_move;
  // Metropolis words thr[d] for d < 128 in two registers per lane
  const uint32_t thr_lo = lane < a.k_move ? thr[lane] : 0u;
  const uint32_t thr_hi = lane + kWave < a.k_move ? thr[lane + kWave] : 0u;
  // proposal log: the wavefront reserves `steps` entries for this sweep with
  // one atomic (entry log_base + step); a reservation that does not fit in
  // the capacity logs nothing, and invalid proposals leave their entry as is
  int32_t* log_base = nullptr;
  if (!kPlain && a.log) {
    unsigned long long first = 0;
    if (lane == 0) first = atomicAdd(a.log_count, (unsigned long long)a.steps);
    first = ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(first >> 32), 0) << 32) |
            (uint32_t)__builtin_amdgcn_readlane((int)first, 0);
    if (first + (unsigned long long)a.steps <= (unsigned long long)a.log_capacity) log_base = a.log + kLogFields * first;
  }

  // Philox outputs of 64 consecutive steps, one step per lane, computed on
  // the VALU in one pass and read back with readlane: the counters are the
  // per-step ones, so the stream is the same as one draw per step, and the
  // scalar unit is left to the proposal bookkeeping
  uint32_t move_r[4], bias_r[4] = {0u, 0u, 0u, 0u};
  int batch = -kWave;  // first step of the current batch
  // kPlain: lane j holds the proposal of step batch + j on the current state
  // (u | v << 8 | cu << 16 | cv << 24), the badness of its two slots
  // (bad_u | bad_v << 16, each below 2^16) and, as a lane mask, whether it
  // is valid.  The batch is planned on the VALU in one pass (the row words
  // and badness gathered with ds_bpermute), again after every acceptance,
  // and the walk goes from one valid step to the next (s_ff1), with one
  // readlane per value instead of the scalar unit's per-step draw
  // decoding, row-word selects and validity tests.
  uint32_t plan_uv = 0u, plan_bad = 0u;
  uint64_t plan_valid = 0ull;
  auto plan = [&]() {
    const int pu = (int)mulhi32(move_r[0], (uint32_t)m), pv = (int)mulhi32(move_r[1], (uint32_t)m);
    auto gather = [&](const int(&arr)[CH], int s) -> int {
      const int index = (s & (kWave - 1)) * 4;
      int r = __builtin_amdgcn_ds_bpermute(index, arr[0]);
#pragma unroll
      for (int c = 1; c < CH; ++c) {
        const int t = __builtin_amdgcn_ds_bpermute(index, arr[c]);
        r = ((s >> 6) == c) ? t : r;
      }
      return r;
    };
    const uint32_t wu = (uint32_t)gather(row_word, pu & ~1), wv = (uint32_t)gather(row_word, pv & ~1);
    const int pcu = (int)((wu >> (8 * (pu & 1))) & 0xFF), pcu_other = (int)((wu >> (8 * (~pu & 1))) & 0xFF);
    const int pcv = (int)((wv >> (8 * (pv & 1))) & 0xFF), pcv_other = (int)((wv >> (8 * (~pv & 1))) & 0xFF);
    plan_valid = (score > 0 ? ~0ull : 0ull) & lanes(batch + lane < a.steps) & lanes((pu >> 1) != (pv >> 1)) &
                 lanes(pcu != pcv) & lanes(pcu_other != pcv) & lanes(pcv_other != pcu);
    plan_uv = (uint32_t)pu | (uint32_t)pv << 8 | (uint32_t)pcu << 16 | (uint32_t)pcv << 24;
    plan_bad = (uint32_t)gather(bad, pu) | (uint32_t)gather(bad, pv) << 16;
  };
  for (int step = 0; step < a.steps; ++step) {
    if (step - batch >= kWave) {
      batch = step;
      // opaque copies of the key: its round schedule is derived here, once
      // per batch, instead of being kept live in scalar registers all loop
      uint32_t k0 = a.seed_lo, k1 = a.seed_hi;
      asm volatile("" : "+s"(k0), "+s"(k1));
      philox4x32_10((uint32_t)(step + lane), sweep, (uint32_t)slot, kTagMove, k0, k1, move_r);
      if (!kPlain && a.bad_bias > 0)
        philox4x32_10((uint32_t)(step + lane), sweep, (uint32_t)slot, kTagBias, k0, k1, bias_r);
      if constexpr (kPlain) plan();
    }
    if constexpr (kPlain) {
      // the next valid step of the batch (invalid steps change nothing)
      const uint64_t rest = plan_valid >> (step - batch);
      if (rest == 0ull) {
        step = batch + kWave - 1;
        continue;
      }
      step += (int)__builtin_ctzll(rest);
    }
    const int draw = step - batch;
    uint32_t r[4];
    int u, v, cu, cv, cu_other = 0, cv_other = 0, planned_bad = 0;
    uint32_t word_ru = 0u, word_rv = 0u;
    if constexpr (kPlain) {
      const uint32_t uv = (uint32_t)__builtin_amdgcn_readlane((int)plan_uv, draw);
      planned_bad = __builtin_amdgcn_readlane((int)plan_bad, draw);
      r[2] = (uint32_t)__builtin_amdgcn_readlane((int)move_r[2], draw);
      u = (int)(uv & 0xFFu);
      v = (int)((uv >> 8) & 0xFFu);
      cu = (int)((uv >> 16) & 0xFFu);
      cv = (int)(uv >> 24);
    } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = (uint32_t)__builtin_amdgcn_readlane((int)move_r[i], draw);
    u = (int)mulhi32(r[0], (uint32_t)m);
    if (a.bad_bias > 0 && score > 0 && (int)mulhi32(r[3], 100u) < a.bad_bias) {
      // tempering.py:51-55: u drawn in proportion to badness
      const uint32_t rb0 = (uint32_t)__builtin_amdgcn_readlane((int)bias_r[0], draw);
      const int target = (int)mulhi32(rb0, (uint32_t)(3 * score));
      int below = 0, carry = 0;
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        // inclusive prefix of badness over the slots of this chunk
        const int x = (int)wave_scan_dpp((uint32_t)bad[c]);
        const int k = lane + kWave * c;
        below += __popcll(lanes(k < m) & lanes(carry + x <= target));
        carry += __builtin_amdgcn_readlane(x, 63);
      }
      u = below < m - 1 ? below : m - 1;
    }
    v = (int)mulhi32(r[1], (uint32_t)m);
    word_ru = (uint32_t)slot_value<CH>(row_word, u & ~1);
    word_rv = (uint32_t)slot_value<CH>(row_word, v & ~1);
    cu = (int)((word_ru >> (8 * (u & 1))) & 0xFF);
    cu_other = (int)((word_ru >> (8 * (~u & 1))) & 0xFF);
    cv = (int)((word_rv >> (8 * (v & 1))) & 0xFF);
    cv_other = (int)((word_rv >> (8 * (~v & 1))) & 0xFF);
    }
    const int ru = u >> 1, rv = v >> 1;
    const int key_u = ru * width + cu, key_v = rv * width + cv;
    if constexpr (!kPlain) {
      const bool valid = score > 0 && ru != rv && cu != cv && cu_other != cv && cv_other != cu;
      if (!valid) continue;
    }
    valid_moves += 1;
    // removed-point change (tempering.py:86-88): the reference columns of
    // both rows; the plain walk reads them on acceptance only
    auto removed_change = [&](uint32_t wu, uint32_t wv) {
      const int ref_u0 = (int)((wu >> 16) & 0xFF), ref_u1 = (int)(wu >> 24);
      const int ref_v0 = (int)((wv >> 16) & 0xFF), ref_v1 = (int)(wv >> 24);
      return (cu == ref_u0 || cu == ref_u1) + (cv == ref_v0 || cv == ref_v1) - (cv == ref_u0 || cv == ref_u1) -
             (cu == ref_v0 || cu == ref_v1);
    };
    const int dd = kPlain ? 0 : removed_change(word_ru, word_rv);

    bool others[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int k = lane + kWave * c;
      others[c] = (k < m) && k != u && k != v;
    }
    // joint(u, v) in the current state: the points other than u, v on the
    // line (u, v).  Only the exact delta needs it; when the bound below
    // already rejects, the lookups are skipped.
    int id_uv = 0, idu[CH], joint_old = 0;
    auto old_joint = [&]() {
      id_uv = table[key_v - key_u + offset];
#pragma unroll
      for (int c = 0; c < CH; ++c) idu[c] = table[key[c] - key_u + offset];
#pragma unroll
      for (int c = 0; c < CH; ++c) joint_old += __popcll(lanes(others[c]) & lanes(idu[c] == id_uv));
    };
    // the log records the exact delta of every valid proposal, and the
    // umbrella test has no integer bound
    const bool exact = (!kPlain && log_base != nullptr) || kUmbrella;
    if (exact) old_joint();
    // histograms seen from u' and v' in the moved state.  The rejection path
    // needs only bad'(u') and bad'(v') = sum_d C(cnt_d, 2), the sums of the
    // ranks the increments return, so both histograms are built and cleared
    // back to back with one wait; the per-slot counts are read only when the
    // move is accepted.
    const int key_u2 = ru * width + cv, key_v2 = rv * width + cu;
    const int id_u2v2 = table[key_v2 - key_u2 + offset];
    // the moved slot itself looks up the zero vector (its rank is 0): the
    // lanes need no exclusion mask except past the last slot
    int idu2[CH], idv2[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int k = lane + kWave * c;
      const int moved = (k == v) ? key_v2 : ((k == u) ? key_u2 : key[c]);
      idu2[c] = table[moved - key_u2 + offset];
      idv2[c] = table[moved - key_v2 + offset];
    }
    uint32_t word_u2[CH], word_v2[CH], old_u2[CH], old_v2[CH];
    if constexpr (kCounted) {
      // a lane seen in a direction left out of the histogram (the moved
      // slot's zero vector among them) and the lanes past the last slot are
      // alone on their line: they count in their trash word (rank 0)
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int live = (c < CH - 1 || tail_in) ? 1 : 0;
        word_u2[c] = (live & (int)((uint32_t)idu2[c] < (uint32_t)info.counted)) ? ((uint32_t)idu2[c] >> 2)
                                                                                 : (uint32_t)(counted_trash + kWave * c);
        word_v2[c] = (live & (int)((uint32_t)idv2[c] < (uint32_t)info.counted)) ? ((uint32_t)idv2[c] >> 2)
                                                                                 : (uint32_t)(counted_trash + kWave * c);
      }
    } else {
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        word_u2[c] = (c < CH - 1 || tail_in) ? ((uint32_t)idu2[c] >> 2) : (uint32_t)trash;
        word_v2[c] = (c < CH - 1 || tail_in) ? ((uint32_t)idv2[c] >> 2) : (uint32_t)trash;
      }
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) old_u2[c] = atomicAdd(&hist[word_u2[c]], byte_one((uint32_t)idu2[c]));
    wave_sync();
#pragma unroll
    for (int c = 0; c < CH; ++c) hist[word_u2[c]] = 0u;
    wave_sync();
#pragma unroll
    for (int c = 0; c < CH; ++c) old_v2[c] = atomicAdd(&hist[word_v2[c]], byte_one((uint32_t)idv2[c]));
    wave_sync();
#pragma unroll
    for (int c = 0; c < CH; ++c) hist[word_v2[c]] = 0u;
    wave_sync();
    // keep the rank extraction below every LDS step above, so the
    // increments are not held up by waits for earlier ranks
    __builtin_amdgcn_sched_barrier(0);
    uint32_t part = 0;
    int joint_new = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      // trash words (one per lane, last chunk only) and the zero vector of
      // the moved slot return rank 0
      part += byte_counter(old_u2[c], (uint32_t)idu2[c]) + (byte_counter(old_v2[c], (uint32_t)idv2[c]) << 16);
      joint_new += __popcll(lanes(others[c]) & lanes(idu2[c] == id_u2v2));
    }
    // bad'(u') in the low half, bad'(v') in the high half (each < 2^16)
    const uint32_t packed = wave_sum_dpp(part);
    const int bad_u2 = (int)(packed & 0xFFFFu), bad_v2 = (int)(packed >> 16);
    const int bad_u = kPlain ? (planned_bad & 0xFFFF) : slot_value<CH>(bad, u);
    const int bad_v = kPlain ? (int)((uint32_t)planned_bad >> 16) : slot_value<CH>(bad, v);
    if (!exact) {
      // delta >= delta - joint(u, v), and the Metropolis words floor(2^32
      // exp(-d / T)) do not grow with d: a proposal that the bound rejects
      // is rejected by the exact delta too
      const int low = (bad_u2 + bad_v2 - joint_new) - (bad_u + bad_v);
      if (low > 0 && r[2] >= register_table(thr_lo, thr_hi, thr, low < a.k_move ? low : a.k_move - 1)) continue;
      old_joint();
    }
    const int delta = (bad_u2 + bad_v2 - joint_new) - (bad_u + bad_v - joint_old);

    // Metropolis on the integer delta (tempering.py:91-95)
    bool accept;
    if constexpr (kUmbrella) {
      const double e = umbrella_energy(delta, dd, removed, strength, center);
      accept = e <= 0.0 || r[2] < exp_word(e * inverse_t);
    } else {
      accept = delta <= 0;
      if (!accept) accept = r[2] < register_table(thr_lo, thr_hi, thr, delta < a.k_move ? delta : a.k_move - 1);
    }
    if (!kPlain && log_base && lane == 0) {
      // (badness u, badness v, energy, slot, delta, accepted, u | v << 16,
      // column u | column v << 16): training data for the proposal network
      int32_t* e = log_base + kLogFields * step;
      e[0] = bad_u;
      e[1] = bad_v;
      e[2] = score;
      e[3] = slot;
      e[4] = delta;
      e[5] = accept ? 1 : 0;
      e[6] = u | (v << 16);
      e[7] = cu | (cv << 16);
    }
    if (!accept) continue;
    const int moved_dd =
        kPlain ? removed_change((uint32_t)slot_value<CH>(row_word, u & ~1), (uint32_t)slot_value<CH>(row_word, v & ~1))
               : dd;

    // badness of the unmoved slots: per-slot counts seen from u', v' in the
    // moved state and from u, v now
    int cnt_u2[CH], cnt_v2[CH], cnt_u[CH], cnt_v[CH], id[CH];
    bool include[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) include[c] = (lane + kWave * c < m) && (lane + kWave * c) != u;
    if constexpr (kCounted) {
      direction_counts_counted<CH>(hist, info.counted, counted_trash, idu2, include, cnt_u2);
#pragma unroll
      for (int c = 0; c < CH; ++c) include[c] = (lane + kWave * c < m) && (lane + kWave * c) != v;
      direction_counts_counted<CH>(hist, info.counted, counted_trash, idv2, include, cnt_v2);
#pragma unroll
      for (int c = 0; c < CH; ++c) include[c] = (lane + kWave * c < m) && (lane + kWave * c) != u;
      direction_counts_counted<CH>(hist, info.counted, counted_trash, idu, include, cnt_u);
#pragma unroll
      for (int c = 0; c < CH; ++c) include[c] = (lane + kWave * c < m) && (lane + kWave * c) != v;
#pragma unroll
      for (int c = 0; c < CH; ++c) id[c] = table[key[c] - key_v + offset];
      direction_counts_counted<CH>(hist, info.counted, counted_trash, id, include, cnt_v);
    } else {
      direction_counts_known<CH>(hist, trash, idu2, include, cnt_u2);
#pragma unroll
      for (int c = 0; c < CH; ++c) include[c] = (lane + kWave * c < m) && (lane + kWave * c) != v;
      direction_counts_known<CH>(hist, trash, idv2, include, cnt_v2);
#pragma unroll
      for (int c = 0; c < CH; ++c) include[c] = (lane + kWave * c < m) && (lane + kWave * c) != u;
      direction_counts_known<CH>(hist, trash, idu, include, cnt_u);
#pragma unroll
      for (int c = 0; c < CH; ++c) include[c] = (lane + kWave * c < m) && (lane + kWave * c) != v;
      direction_counts<CH>(table, hist, trash, offset, key_v, key, include, cnt_v, id);
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int k = lane + kWave * c;
      const int bit_old = idu[c] == id_uv ? 1 : 0;
      const int bit_new = idu2[c] == id_u2v2 ? 1 : 0;
      if (others[c]) bad[c] += (cnt_u2[c] + cnt_v2[c] - bit_new) - (cnt_u[c] + cnt_v[c] - bit_old);
      if (k == u) {
        bad[c] = bad_u2;
        key[c] = key_u2;
      }
      if (k == v) {
        bad[c] = bad_v2;
        key[c] = key_v2;
      }
      // u now holds column cv, v column cu: both lanes of each row
      if ((k >> 1) == ru) row_word[c] = (int)(((uint32_t)row_word[c] & ~(0xFFu << (8 * (u & 1)))) | ((uint32_t)cv << (8 * (u & 1))));
      if ((k >> 1) == rv) row_word[c] = (int)(((uint32_t)row_word[c] & ~(0xFFu << (8 * (v & 1)))) | ((uint32_t)cu << (8 * (v & 1))));
    }
    score += delta;
    removed += moved_dd;
    accepted += 1;
    if (score < best) {
      best = score;
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int k = lane + kWave * c;
        if (k < m) a.best_columns[base + k] = key[c] - (k >> 1) * width;
      }
    }
    // the rest of the batch on the new state
    if constexpr (kPlain) plan();
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

template <int CH, bool kUmbrella = false, bool kPlain = false>
__global__ void __launch_bounds__(1024) walk_kernel(WalkArgs a, const uint16_t* __restrict__ table_g,
                                                   DirTableInfo info) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int waves = blockDim.x / kWave;
  const int wave = threadIdx.x / kWave, lane = lane_id();
  uint16_t* table = reinterpret_cast<uint16_t*>(smem);
  uint32_t* hist = reinterpret_cast<uint32_t*>(smem + table_lds_bytes(info)) + wave * info.hist_words;
  stage_table(table_g, table, info.entries);
  for (int i = lane; i < info.hist_words; i += kWave) hist[i] = 0;
  __syncthreads();
  const int slot = walk_slot((int)blockIdx.x, wave, waves, a.ladder);
  if (slot < a.replicas) walk_replica<CH, kUmbrella, kPlain>(a, table, hist, info, slot, a.counters[0] + a.sweep_offset);
}

// Several populations of different sizes in one launch: block b belongs to
// the group whose block range holds b; each group has its own direction
// table, histogram size and number of working waves per block.
struct BatchGroup {
  WalkArgs walk;
  const uint16_t* table;
  DirTableInfo info;
  int ch;
  int waves;
  int block_begin;
  int pad;
};

// Histogram words per wavefront of the batched walks: plain runs count only
// the counted directions (walk_replica's kCounted), so that the larger sizes
// fit more waves per block.
__host__ __device__ inline int batch_hist_words(const DirTableInfo& info, bool plain) {
  return plain ? info.counted_words : info.hist_words;
}

// kPlain: no group has a proposal log or a badness bias (see walk_replica).
// Every group walks sweep counters[0] + sweep_offset.
template <bool kPlain>
__global__ void __launch_bounds__(1024) walk_batched_kernel(const BatchGroup* __restrict__ groups, int group_count,
                                                            uint32_t sweep_offset) {
  extern __shared__ __align__(16) unsigned char smem[];
  int g = 0;
  while (g + 1 < group_count && (int)blockIdx.x >= groups[g + 1].block_begin) ++g;
  const BatchGroup& G = groups[g];
  const DirTableInfo info = G.info;
  const int wave = threadIdx.x / kWave, lane = lane_id();
  uint16_t* table = reinterpret_cast<uint16_t*>(smem);
  const int hist_words = batch_hist_words(info, kPlain);
  uint32_t* hist = reinterpret_cast<uint32_t*>(smem + table_lds_bytes(info)) + wave * hist_words;
  stage_table(G.table, table, info.entries);
  if (wave < G.waves)
    for (int i = lane; i < hist_words; i += kWave) hist[i] = 0;
  __syncthreads();
  if (wave >= G.waves) return;
  const int slot = walk_slot((int)blockIdx.x - G.block_begin, wave, G.waves, G.walk.ladder);
  if (slot >= G.walk.replicas) return;
  const uint32_t sweep = G.walk.counters[0] + sweep_offset;
  switch (G.ch) {
    case 1: walk_replica<1, false, kPlain, kPlain>(G.walk, table, hist, info, slot, sweep); break;
    case 2: walk_replica<2, false, kPlain, kPlain>(G.walk, table, hist, info, slot, sweep); break;
    case 3: walk_replica<3, false, kPlain, kPlain>(G.walk, table, hist, info, slot, sweep); break;
    default: walk_replica<4, false, kPlain, kPlain>(G.walk, table, hist, info, slot, sweep); break;
  }
}

struct ExchangeArgs {
  int n;
  int replicas;
  int ladder;
  uint32_t seed_lo, seed_hi;
  int32_t* columns;
  int32_t* badness;
  int32_t* scores;
  int32_t* removed;
  const uint32_t* thr;
  const int8_t* sign;
  int k_exchange;
  uint32_t* counters;
  uint32_t round_offset;  // the round number is counters[1] + round_offset
};

// One wavefront per candidate pair (a, a+1); pairs start at round % 2 inside
// every ladder and do not overlap, so the swaps need no synchronisation.
// Returns whether the pair swapped.
__device__ bool exchange_pair(const ExchangeArgs& x, int wave, uint32_t round) {
  const int lane = lane_id();
  const int per_ladder = x.ladder / 2;  // candidate pairs per ladder (upper bound)
  const int ladders = x.replicas / x.ladder;
  if (wave >= ladders * per_ladder) return false;
  const int ladder = wave / per_ladder;
  const int start = ladder * x.ladder;
  const int a = start + (int)(round & 1u) + 2 * (wave % per_ladder);
  if (a >= start + x.ladder - 1) return false;
  const int b = a + 1;
  const long long diff = (long long)x.scores[a] - (long long)x.scores[b];
  const int sg = x.sign[a];
  bool accept;
  if (sg == 0 || (sg > 0 && diff >= 0) || (sg < 0 && diff <= 0)) {
    accept = true;
  } else {
    uint32_t r[4];
    philox4x32_10((uint32_t)a, round, (uint32_t)ladder, kTagExch, x.seed_lo, x.seed_hi, r);
    long long d = diff < 0 ? -diff : diff;
    if (d > x.k_exchange - 1) d = x.k_exchange - 1;
    accept = r[0] < x.thr[(long long)a * x.k_exchange + d];
  }
  if (!accept) return false;
  const int m = 2 * x.n;
  for (int i = lane; i < m; i += kWave) {
    const long long ia = (long long)a * m + i, ib = (long long)b * m + i;
    const int32_t ca = x.columns[ia], cb = x.columns[ib];
    x.columns[ia] = cb;
    x.columns[ib] = ca;
    const int32_t ba = x.badness[ia], bb = x.badness[ib];
    x.badness[ia] = bb;
    x.badness[ib] = ba;
  }
  if (lane == 0) {
    const int32_t sa = x.scores[a];
    x.scores[a] = x.scores[b];
    x.scores[b] = sa;
    const int32_t ra = x.removed[a];
    x.removed[a] = x.removed[b];
    x.removed[b] = ra;
  }
  return true;
}

// The block's swaps are counted with one atomic; the round counter is
// advanced by a counter kernel (once per captured graph of many sweeps).
__global__ void __launch_bounds__(1024) exchange_kernel(ExchangeArgs x) {
  __shared__ uint32_t swaps;
  if (threadIdx.x == 0) swaps = 0;
  __syncthreads();
  const bool swapped =
      exchange_pair(x, (int)((blockIdx.x * blockDim.x + threadIdx.x) / kWave), x.counters[1] + x.round_offset);
  if (swapped && lane_id() == 0) atomicAdd(&swaps, 1u);
  __syncthreads();
  if (threadIdx.x == 0 && swaps)
    atomicAdd(reinterpret_cast<unsigned long long*>(x.counters + 2), (unsigned long long)swaps);
}

// Device-side description of a batched run (one allocation per plan).
struct BatchTables {
  BatchGroup walk[N3_MAX_GROUPS];
  ExchangeArgs exchange[N3_MAX_GROUPS];
  int exchange_begin[N3_MAX_GROUPS + 1];  // first exchange wave of each group
  uint32_t* counters[N3_MAX_GROUPS];
  uint32_t round_step[N3_MAX_GROUPS];     // 1 if the group exchanges
  int count;
  uint32_t done;  // blocks of the running exchange launch that finished
  int total_items;  // walk blocks of one sweep over all groups (the persistent run's work items)
};

// As exchange_kernel, for every group of a batched run: swaps are counted per
// block and group, and the last block to finish advances every group's sweep
// counter and, for exchanging groups, its round counter.
__global__ void __launch_bounds__(1024) exchange_batched_kernel(BatchTables* __restrict__ t, uint32_t round_offset,
                                                                int advance) {
  __shared__ uint32_t swaps[N3_MAX_GROUPS];
  if (threadIdx.x < N3_MAX_GROUPS) swaps[threadIdx.x] = 0;
  __syncthreads();
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
  int g = 0;
  while (g + 1 < t->count && wave >= t->exchange_begin[g + 1]) ++g;
  if (wave < t->exchange_begin[t->count] && t->round_step[g] != 0) {
    if (exchange_pair(t->exchange[g], wave - t->exchange_begin[g], t->exchange[g].counters[1] + round_offset) &&
        lane_id() == 0)
      atomicAdd(&swaps[g], 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < t->count; ++i)
      if (swaps[i]) atomicAdd(reinterpret_cast<unsigned long long*>(t->counters[i] + 2), (unsigned long long)swaps[i]);
    if (!advance) return;
    __threadfence();
    if (atomicAdd(&t->done, 1u) == gridDim.x - 1) {
      for (int i = 0; i < t->count; ++i) {
        atomicAdd(&t->counters[i][0], 1u);
        atomicAdd(&t->counters[i][1], t->round_step[i]);
      }
      atomicExch(&t->done, 0u);
    }
  }
}

// All `sweeps` sweeps of a batched run in ONE cooperative launch with one
// block per CU (the launch guarantees that every block is resident).  Each
// sweep drains a queue of work items -- item i is walk block i of the
// per-sweep launch: W_g replicas of group g, groups largest n first -- that
// blocks take one at a time from a device counter, restaging the direction
// table in LDS when the group changes; then a grid barrier, the exchanges of
// every group (one wavefront per candidate pair, strided over the grid), and
// a second barrier.  Sweep and round numbers are the groups' counters at
// launch plus the sweep index; block 0 advances the counters at the end.
// queue[2] holds the item counters of even and odd sweeps (zero at launch;
// each is cleared for its next use during the sweep before it, and both are
// zero again when the launch ends).  Every trajectory is bit-identical to
// the per-sweep launches'.
template <bool kPlain>
__global__ void __launch_bounds__(1024) walk_persistent_kernel(BatchTables* __restrict__ t, int sweeps,
                                                                unsigned int* __restrict__ queue) {
  namespace cg = cooperative_groups;
  cg::grid_group grid = cg::this_grid();
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int item_sh;
  __shared__ uint32_t sweep0[N3_MAX_GROUPS], round0[N3_MAX_GROUPS], swaps[N3_MAX_GROUPS];
  const int wave = threadIdx.x / kWave, lane = lane_id(), block_waves = blockDim.x / kWave;
  const int count = t->count, total_items = t->total_items;
  if (threadIdx.x < count) {
    sweep0[threadIdx.x] = t->counters[threadIdx.x][0];
    round0[threadIdx.x] = t->counters[threadIdx.x][1];
  }
  __syncthreads();
  int staged = -1;
  for (int s = 0; s < sweeps; ++s) {
    unsigned int* next = queue + (s & 1);
    for (;;) {
      if (threadIdx.x == 0) item_sh = (int)atomicAdd(next, 1u);
      __syncthreads();
      const int item = item_sh;
      __syncthreads();  // every thread has read the item before the next grab
      if (item >= total_items) break;
      int g = 0;
      while (g + 1 < count && item >= t->walk[g + 1].block_begin) ++g;
      const BatchGroup& G = t->walk[g];
      const DirTableInfo info = G.info;
      uint16_t* table = reinterpret_cast<uint16_t*>(smem);
      const int hist_words = batch_hist_words(info, kPlain);
      uint32_t* hist = reinterpret_cast<uint32_t*>(smem + table_lds_bytes(info)) + wave * hist_words;
      if (g != staged) {
        stage_table(G.table, table, info.entries);
        if (wave < G.waves)
          for (int i = lane; i < hist_words; i += kWave) hist[i] = 0;
        staged = g;
        __syncthreads();
      }
      // consecutive slots: the block waits for the slowest wave of each
      // item, so an item of neighbouring temperatures (alike in cost) keeps
      // the waits short; spreading an item over a ladder was 2% slower here
      const int slot = (item - G.block_begin) * G.waves + wave;
      if (wave < G.waves && slot < G.walk.replicas) {
        const uint32_t sweep = sweep0[g] + (uint32_t)s;
        switch (G.ch) {
          case 1: walk_replica<1, false, kPlain, kPlain>(G.walk, table, hist, info, slot, sweep); break;
          case 2: walk_replica<2, false, kPlain, kPlain>(G.walk, table, hist, info, slot, sweep); break;
          case 3: walk_replica<3, false, kPlain, kPlain>(G.walk, table, hist, info, slot, sweep); break;
          default: walk_replica<4, false, kPlain, kPlain>(G.walk, table, hist, info, slot, sweep); break;
        }
      }
      __syncthreads();  // the item is done before the next one may restage the table
    }
    grid.sync();
    if (threadIdx.x < N3_MAX_GROUPS) swaps[threadIdx.x] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) queue[(s + 1) & 1] = 0u;  // last used by sweep s - 1
    __syncthreads();
    const int pairs = t->exchange_begin[count], stride = (int)gridDim.x * block_waves;
    for (int w = (int)blockIdx.x * block_waves + wave; w < pairs; w += stride) {
      int g = 0;
      while (g + 1 < count && w >= t->exchange_begin[g + 1]) ++g;
      if (t->round_step[g] != 0 &&
          exchange_pair(t->exchange[g], w - t->exchange_begin[g], round0[g] + (uint32_t)s * t->round_step[g]) &&
          lane == 0)
        atomicAdd(&swaps[g], 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int i = 0; i < count; ++i)
        if (swaps[i]) atomicAdd(reinterpret_cast<unsigned long long*>(t->counters[i] + 2), (unsigned long long)swaps[i]);
    grid.sync();
  }
  if (blockIdx.x == 0) {
    if (threadIdx.x < count) {
      t->counters[threadIdx.x][0] = sweep0[threadIdx.x] + (uint32_t)sweeps;
      t->counters[threadIdx.x][1] = round0[threadIdx.x] + (uint32_t)sweeps * t->round_step[threadIdx.x];
    }
    if (threadIdx.x == 0 && sweeps > 0) queue[(sweeps - 1) & 1] = 0u;
  }
}

__global__ void bump_batched_kernel(const BatchTables* __restrict__ t, uint32_t sweeps) {
  const int g = threadIdx.x;
  if (blockIdx.x == 0 && g < t->count) {
    t->counters[g][0] += sweeps;
    t->counters[g][1] += sweeps * t->round_step[g];
  }
}

__global__ void bump_kernel(uint32_t* counters, uint32_t d_sweep, uint32_t d_round) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    counters[0] += d_sweep;
    counters[1] += d_round;
  }
}

int enqueue_counters(const n3_chains* c, uint32_t d_sweep, uint32_t d_round, hipStream_t stream) {
  hipLaunchKernelGGL(bump_kernel, dim3(1), dim3(64), 0, stream, c->counters, d_sweep, d_round);
  return hip_status(hipGetLastError(), "bump_kernel launch");
}

int check_chains(const n3_chains* c) {
  if (!c) return set_error(N3_EINVAL, "null chains");
  if (c->n < 2 || c->n > kMaxNTable) return set_error(N3_EINVAL, "n must be in [2, %d] (got %d)", kMaxNTable, c->n);
  if (c->replicas < 1) return set_error(N3_EINVAL, "replicas must be positive");
  if (!c->columns || !c->badness || !c->scores || !c->best_columns || !c->best_scores || !c->accepted ||
      !c->reference || !c->removed_points || !c->move_thresholds || !c->counters || !c->proposals)
    return set_error(N3_EINVAL, "null chains array");
  if (c->k_move < 2) return set_error(N3_EINVAL, "k_move must be at least 2");
  if (c->strengths && (!c->centers || !c->inverse_temperatures))
    return set_error(N3_EINVAL, "umbrella strengths need centers and inverse temperatures");
  return N3_OK;
}

struct WalkPlan {
  const uint16_t* table;
  DirTableInfo info;
  int waves;
  unsigned blocks;
  size_t lds;
  void (*kernel)(WalkArgs, const uint16_t*, DirTableInfo);
};

static int walk_plan(int n, int replicas, int block_waves, WalkPlan* plan, bool umbrella = false,
                     bool plain = false) {
  int st = device_dirtable(n, &plan->table, &plan->info);
  if (st) return st;
  int lds_limit = 0, compute_units = 0;
  st = device_limits(&lds_limit, &compute_units);
  if (st) return st;
  const int table_bytes = table_lds_bytes(plan->info), hist_bytes = plan->info.hist_words * 4;
  const int fit = (lds_limit - table_bytes) / hist_bytes;
  if (fit < 1) return set_error(N3_EINVAL, "direction table for n=%d does not fit in LDS", n);
  int waves = fit < 16 ? fit : 16;
  if (block_waves > 0) {
    if (block_waves > waves)
      return set_error(N3_EINVAL, "block_waves=%d exceeds the %d that fit in LDS for n=%d", block_waves, waves, n);
    waves = block_waves;
  } else {
    // spread small populations over more CUs
    const int per_cu = (replicas + compute_units - 1) / compute_units;
    if (waves > per_cu) waves = per_cu < 1 ? 1 : per_cu;
  }
  plan->waves = waves;
  plan->blocks = (unsigned)((replicas + waves - 1) / waves);
  plan->lds = (size_t)table_bytes + (size_t)waves * hist_bytes;
  const int ch = (2 * n + kWave - 1) / kWave;
  // the plain instantiation when there is no umbrella, log or badness bias
  switch (ch) {
    case 1: plan->kernel = umbrella ? walk_kernel<1, true> : plain ? walk_kernel<1, false, true> : walk_kernel<1>; break;
    case 2: plan->kernel = umbrella ? walk_kernel<2, true> : plain ? walk_kernel<2, false, true> : walk_kernel<2>; break;
    case 3: plan->kernel = umbrella ? walk_kernel<3, true> : plain ? walk_kernel<3, false, true> : walk_kernel<3>; break;
    default: plan->kernel = umbrella ? walk_kernel<4, true> : plain ? walk_kernel<4, false, true> : walk_kernel<4>; break;
  }
  return N3_OK;
}

WalkArgs walk_args(const n3_chains* c, int steps, int bad_bias, uint64_t seed) {
  return WalkArgs{c->n, c->replicas, steps, bad_bias, (uint32_t)seed, (uint32_t)(seed >> 32), c->columns, c->badness,
                  c->scores, c->best_columns, c->best_scores, c->accepted, c->reference, c->removed_points,
                  c->move_thresholds, c->k_move, c->counters, reinterpret_cast<unsigned long long*>(c->proposals),
                  c->log, (long long)c->log_capacity, reinterpret_cast<unsigned long long*>(c->log_count),
                  c->centers, c->strengths, c->inverse_temperatures};
}

// The walk of sweep counters[0] + sweep_offset; with bump_sweep the sweep
// counter is advanced by a counter kernel after it.
static int enqueue_advance(const n3_chains* c, const WalkPlan& plan, int steps, int bad_bias, uint64_t seed,
                           hipStream_t stream, bool bump_sweep = true, int ladder = 0, uint32_t sweep_offset = 0) {
  WalkArgs a = walk_args(c, steps, bad_bias, seed);
  a.ladder = ladder;
  a.sweep_offset = sweep_offset;
  hipLaunchKernelGGL(plan.kernel, dim3(plan.blocks), dim3(plan.waves * kWave), plan.lds, stream, a, plan.table,
                     plan.info);
  int st = hip_status(hipGetLastError(), "walk_kernel launch");
  if (st || !bump_sweep) return st;
  hipLaunchKernelGGL(bump_kernel, dim3(1), dim3(64), 0, stream, c->counters, 1u, 0u);
  return hip_status(hipGetLastError(), "bump_kernel launch");
}

int enqueue_exchange(const n3_chains* c, int ladder, uint64_t seed, hipStream_t stream, uint32_t round_offset,
                     bool advance) {
  const int pairs = (c->replicas / ladder) * (ladder / 2);
  ExchangeArgs x{c->n, c->replicas, ladder, (uint32_t)seed, (uint32_t)(seed >> 32), c->columns, c->badness,
                 c->scores, c->removed_points, c->exchange_thresholds, c->exchange_sign, c->k_exchange, c->counters,
                 round_offset};
  if (pairs > 0) {
    const int threads = 1024, per_block = threads / kWave;  // few blocks: few counter atomics
    hipLaunchKernelGGL(exchange_kernel, dim3((unsigned)((pairs + per_block - 1) / per_block)), dim3(threads), 0, stream,
                       x);
    int st = hip_status(hipGetLastError(), "exchange_kernel launch");
    if (st) return st;
  }
  if (!advance) return N3_OK;
  hipLaunchKernelGGL(bump_kernel, dim3(1), dim3(64), 0, stream, c->counters, 0u, 1u);
  return hip_status(hipGetLastError(), "bump_kernel launch");
}

int check_exchange(const n3_chains* c, int ladder) {
  if (ladder < 2 || c->replicas % ladder) return set_error(N3_EINVAL, "ladder_size must divide replica count");
  if (!c->exchange_thresholds || !c->exchange_sign || c->k_exchange < 2)
    return set_error(N3_EINVAL, "exchange thresholds missing");
  // tempering.py:173-174
  if (c->strengths) return set_error(N3_EINVAL, "temperature-only exchange is disabled with umbrella potentials");
  return N3_OK;
}

// ---------------------------------------------------------------------------
// graph cache for n3_temper_run
// ---------------------------------------------------------------------------

struct GraphKey {
  n3_chains chains;
  int steps, ladder, bad_bias, sweeps_per_graph;
  uint64_t seed;
  hipStream_t stream;
  int device;
};

static bool same_key(const GraphKey& x, const GraphKey& y) {
  return std::memcmp(&x.chains, &y.chains, sizeof(n3_chains)) == 0 && x.steps == y.steps && x.ladder == y.ladder &&
         x.bad_bias == y.bad_bias && x.sweeps_per_graph == y.sweeps_per_graph && x.seed == y.seed &&
         x.stream == y.stream && x.device == y.device;
}

struct GraphEntry {
  GraphKey key;
  hipGraphExec_t exec;
  hipGraph_t graph;
};

static std::mutex g_graph_mutex;
static std::vector<GraphEntry> g_graphs;

// Cached graphs per kind: the oldest is dropped (after its stream has
// drained, since a replay may still be using it) when a new one would
// exceed the cap, so long drivers do not accumulate graphs.
constexpr size_t kMaxCachedGraphs = 16;

static void drop_oldest_graph() {
  GraphEntry& e = g_graphs.front();
  (void)hipStreamSynchronize(e.key.stream);
  (void)hipGraphExecDestroy(e.exec);
  (void)hipGraphDestroy(e.graph);
  g_graphs.erase(g_graphs.begin());
}

// A graph of `sweeps` x (walk, exchange) and one counter kernel: the k-th
// walk and exchange take their sweep and round numbers as the counters plus
// k, and the counters advance once, by `sweeps`, at the end.
static int capture_graph(const n3_chains* c, const WalkPlan& plan, int sweeps, int steps, int ladder, int bad_bias,
                         uint64_t seed, hipGraph_t* graph, hipGraphExec_t* exec) {
  hipStream_t cap = nullptr;
  int st = hip_status(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreate");
  if (st) return st;
  st = hip_status(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  if (!st) {
    for (int i = 0; i < sweeps && !st; ++i) {
      st = enqueue_advance(c, plan, steps, bad_bias, seed, cap, false, ladder, (uint32_t)i);
      if (!st && ladder > 1) st = enqueue_exchange(c, ladder, seed, cap, (uint32_t)i, false);
    }
    if (!st) st = enqueue_counters(c, (uint32_t)sweeps, ladder > 1 ? (uint32_t)sweeps : 0u, cap);
    hipGraph_t g = nullptr;
    const int end = hip_status(hipStreamEndCapture(cap, &g), "hipStreamEndCapture");
    if (!st) st = end;
    if (!st) {
      st = hip_status(hipGraphInstantiate(exec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
      if (st)
        (void)hipGraphDestroy(g);
      else
        *graph = g;
    } else if (g) {
      (void)hipGraphDestroy(g);
    }
  }
  (void)hipStreamDestroy(cap);
  return st;
}

// ---------------------------------------------------------------------------
// batched runs of several sizes
// ---------------------------------------------------------------------------

struct BatchPlan {
  BatchTables host;
  BatchTables* device = nullptr;
  unsigned walk_blocks = 0;
  int walk_threads = 0;
  size_t walk_lds = 0;
  unsigned exchange_blocks = 0;
  bool plain = true;  // no group has a proposal log or a badness bias
};

// A group's blocks hold block_waves replicas (wavefronts), or as many as fit
// in LDS beside its direction table when block_waves is 0.
static int batch_plan(const n3_group* groups, int count, BatchPlan* plan) {
  plan->plain = true;
  int lds_limit = 0, compute_units = 0;
  int st = device_limits(&lds_limit, &compute_units);
  if (st) return st;
  std::memset(&plan->host, 0, sizeof(BatchTables));
  // largest sizes first: their blocks run longest, so they are dispatched first
  int order[N3_MAX_GROUPS];
  for (int i = 0; i < count; ++i) order[i] = i;
  for (int i = 1; i < count; ++i)
    for (int j = i; j > 0 && groups[order[j]].chains.n > groups[order[j - 1]].chains.n; --j) {
      const int t = order[j];
      order[j] = order[j - 1];
      order[j - 1] = t;
    }
  for (int i = 0; i < count; ++i)
    if (groups[i].chains.log || groups[i].bad_bias > 0) plan->plain = false;
  int blocks = 0, max_waves = 1, exchange_waves = 0;
  size_t lds = 0;
  for (int i = 0; i < count; ++i) {
    const n3_group& gr = groups[order[i]];
    const n3_chains* c = &gr.chains;
    BatchGroup& bg = plan->host.walk[i];
    st = device_dirtable(c->n, &bg.table, &bg.info);
    if (st) return st;
    const int table_bytes = table_lds_bytes(bg.info), hist_bytes = batch_hist_words(bg.info, plan->plain) * 4;
    const int fit = (lds_limit - table_bytes) / hist_bytes;
    int waves = fit < 16 ? fit : 16;
    if (c->block_waves > 0) {
      if (c->block_waves > waves)
        return set_error(N3_EINVAL, "block_waves=%d exceeds the %d that fit in LDS for n=%d", c->block_waves, waves,
                         c->n);
      waves = c->block_waves;
    }
    if (waves < 1) return set_error(N3_EINVAL, "direction table for n=%d does not fit in LDS", c->n);
    bg.walk = walk_args(c, gr.steps, gr.bad_bias, gr.seed);
    bg.walk.ladder = gr.ladder > 1 ? gr.ladder : 0;
    bg.ch = (2 * c->n + kWave - 1) / kWave;
    bg.waves = waves;
    bg.block_begin = blocks;
    blocks += (c->replicas + waves - 1) / waves;
    if (waves > max_waves) max_waves = waves;
    const size_t need = (size_t)table_bytes + (size_t)waves * hist_bytes;
    if (need > lds) lds = need;
    plan->host.exchange[i] = ExchangeArgs{c->n, c->replicas, gr.ladder > 1 ? gr.ladder : 2, (uint32_t)gr.seed,
                                          (uint32_t)(gr.seed >> 32), c->columns, c->badness, c->scores,
                                          c->removed_points, c->exchange_thresholds, c->exchange_sign,
                                          c->k_exchange, c->counters};
    plan->host.exchange_begin[i] = exchange_waves;
    if (gr.ladder > 1) exchange_waves += (c->replicas / gr.ladder) * (gr.ladder / 2);
    plan->host.counters[i] = c->counters;
    plan->host.round_step[i] = gr.ladder > 1 ? 1u : 0u;
  }
  plan->host.exchange_begin[count] = exchange_waves;
  plan->host.count = count;
  plan->host.total_items = blocks;
  plan->walk_blocks = (unsigned)blocks;
  plan->walk_threads = max_waves * kWave;
  plan->walk_lds = lds;
  plan->exchange_blocks = (unsigned)((exchange_waves + 15) / 16);  // 1024-thread blocks
  st = hip_status(hipMalloc(&plan->device, sizeof(BatchTables)), "hipMalloc(batch tables)");
  if (st) return st;
  st = hip_status(hipMemcpy(plan->device, &plan->host, sizeof(BatchTables), hipMemcpyHostToDevice),
                  "hipMemcpy(batch tables)");
  if (st) {
    (void)hipFree(plan->device);
    plan->device = nullptr;
  }
  return st;
}

// One sweep of every group, number counters + offset.  With `advance` the
// counters move on by one sweep after it (the exchange launch's last block,
// or a counter kernel); a captured graph passes offsets 0, 1, ... and
// advances them once at its end instead.
static int enqueue_batched_sweep(const BatchPlan& plan, hipStream_t stream, uint32_t offset = 0, bool advance = true) {
  // the group array is the first member of the device tables
  static_assert(offsetof(BatchTables, walk) == 0, "walk groups lead the batch tables");
  hipLaunchKernelGGL(plan.plain ? walk_batched_kernel<true> : walk_batched_kernel<false>, dim3(plan.walk_blocks),
                     dim3(plan.walk_threads), plan.walk_lds, stream, reinterpret_cast<const BatchGroup*>(plan.device),
                     plan.host.count, offset);
  int st = hip_status(hipGetLastError(), "walk_batched_kernel launch");
  if (st) return st;
  if (plan.exchange_blocks > 0) {
    hipLaunchKernelGGL(exchange_batched_kernel, dim3(plan.exchange_blocks), dim3(1024), 0, stream, plan.device, offset,
                       advance ? 1 : 0);
    return hip_status(hipGetLastError(), "exchange_batched_kernel launch");
  }
  if (!advance) return N3_OK;
  hipLaunchKernelGGL(bump_batched_kernel, dim3(1), dim3(64), 0, stream, plan.device, 1u);
  return hip_status(hipGetLastError(), "bump_batched_kernel launch");
}

struct BatchEntry {
  std::vector<unsigned char> key;
  BatchPlan plan;
  hipGraph_t graph;
  hipGraphExec_t exec;
  int sweeps_per_graph;
  hipStream_t stream;
};

static std::vector<BatchEntry> g_batches;

static void drop_oldest_batch() {
  BatchEntry& e = g_batches.front();
  (void)hipStreamSynchronize(e.stream);
  (void)hipGraphExecDestroy(e.exec);
  (void)hipGraphDestroy(e.graph);
  (void)hipFree(e.plan.device);
  g_batches.erase(g_batches.begin());
}

}  // namespace no3

using namespace no3;

extern "C" int n3_temper_batched_run(const n3_group* groups, int group_count, int sweeps, void* stream) {
  if (!groups || group_count < 1 || group_count > N3_MAX_GROUPS)
    return set_error(N3_EINVAL, "group_count must be in [1, %d]", N3_MAX_GROUPS);
  if (sweeps < 0) return set_error(N3_EINVAL, "negative sweeps");
  for (int i = 0; i < group_count; ++i) {
    int st = check_chains(&groups[i].chains);
    if (st) return st;
    if (groups[i].chains.strengths)
      return set_error(N3_EINVAL, "umbrella potentials are supported by n3_temper_advance and n3_temper_run only");
    if (groups[i].steps < 1 || groups[i].bad_bias < 0 || groups[i].bad_bias > 100)
      return set_error(N3_EINVAL, "steps must be positive and bad_bias in [0,100]");
    if (groups[i].ladder > 1) {
      st = check_exchange(&groups[i].chains, groups[i].ladder);
      if (st) return st;
    }
  }
  if (sweeps == 0) return N3_OK;
  hipStream_t s = as_stream(stream);
  const int per_graph = sweeps < 16 ? sweeps : 16;
  int device = 0;
  (void)hipGetDevice(&device);
  std::vector<unsigned char> key(sizeof(n3_group) * group_count + 3 * sizeof(intptr_t));
  std::memcpy(key.data(), groups, sizeof(n3_group) * group_count);
  const intptr_t extra[3] = {(intptr_t)s, (intptr_t)device, (intptr_t)per_graph};
  std::memcpy(key.data() + sizeof(n3_group) * group_count, extra, sizeof(extra));
  std::lock_guard<std::mutex> lock(g_graph_mutex);
  BatchEntry* entry = nullptr;
  for (auto& e : g_batches)
    if (e.key == key) entry = &e;
  if (!entry) {
    BatchEntry e;
    e.key = key;
    e.sweeps_per_graph = per_graph;
    e.stream = s;
    int st = batch_plan(groups, group_count, &e.plan);
    if (st) return st;
    hipStream_t cap = nullptr;
    st = hip_status(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreate");
    if (st) return st;
    st = hip_status(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (!st) {
      for (int i = 0; i < per_graph && !st; ++i) st = enqueue_batched_sweep(e.plan, cap, (uint32_t)i, false);
      if (!st) {
        hipLaunchKernelGGL(bump_batched_kernel, dim3(1), dim3(64), 0, cap, e.plan.device, (uint32_t)per_graph);
        st = hip_status(hipGetLastError(), "bump_batched_kernel launch");
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
    if (st) {
      (void)hipFree(e.plan.device);
      return st;
    }
    if (g_batches.size() >= kMaxCachedGraphs) drop_oldest_batch();
    g_batches.push_back(std::move(e));
    entry = &g_batches.back();
  }
  const int full = sweeps / entry->sweeps_per_graph, rest = sweeps % entry->sweeps_per_graph;
  for (int i = 0; i < full; ++i) {
    const int st = hip_status(hipGraphLaunch(entry->exec, s), "hipGraphLaunch");
    if (st) return st;
  }
  for (int i = 0; i < rest; ++i) {
    const int st = enqueue_batched_sweep(entry->plan, s);
    if (st) return st;
  }
  return N3_OK;
}

// Plans (device tables and work-queue counters) of persistent batched runs,
// cached like the graphs, at most kMaxCachedGraphs of them.
struct PersistentEntry {
  std::vector<unsigned char> key;
  BatchPlan plan;
  unsigned int* queue;
  unsigned grid;
  hipStream_t stream;
};
static std::vector<PersistentEntry> g_persistent;

extern "C" int n3_temper_batched_persistent(const n3_group* groups, int group_count, int sweeps, void* stream) {
  if (!groups || group_count < 1 || group_count > N3_MAX_GROUPS)
    return set_error(N3_EINVAL, "group_count must be in [1, %d]", N3_MAX_GROUPS);
  if (sweeps < 0) return set_error(N3_EINVAL, "negative sweeps");
  for (int i = 0; i < group_count; ++i) {
    int st = check_chains(&groups[i].chains);
    if (st) return st;
    if (groups[i].chains.strengths)
      return set_error(N3_EINVAL, "umbrella potentials are supported by n3_temper_advance and n3_temper_run only");
    if (groups[i].steps < 1 || groups[i].bad_bias < 0 || groups[i].bad_bias > 100)
      return set_error(N3_EINVAL, "steps must be positive and bad_bias in [0,100]");
    if (groups[i].ladder > 1) {
      st = check_exchange(&groups[i].chains, groups[i].ladder);
      if (st) return st;
    }
  }
  if (sweeps == 0) return N3_OK;
  hipStream_t s = as_stream(stream);
  int device = 0;
  (void)hipGetDevice(&device);
  std::vector<unsigned char> key(sizeof(n3_group) * group_count + 2 * sizeof(intptr_t));
  std::memcpy(key.data(), groups, sizeof(n3_group) * group_count);
  const intptr_t extra[2] = {(intptr_t)s, (intptr_t)device};
  std::memcpy(key.data() + sizeof(n3_group) * group_count, extra, sizeof(extra));
  std::lock_guard<std::mutex> lock(g_graph_mutex);
  PersistentEntry* entry = nullptr;
  for (auto& e : g_persistent)
    if (e.key == key) entry = &e;
  if (!entry) {
    PersistentEntry e;
    e.key = key;
    e.stream = s;
    e.queue = nullptr;
    int st = batch_plan(groups, group_count, &e.plan);
    if (st) return st;
    int coop = 0, compute_units = 0, per_cu = 0;
    st = hip_status(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device), "hipDeviceGetAttribute");
    if (!st && !coop) st = set_error(N3_ENODEV, "device %d does not support cooperative launches", device);
    if (!st)
      st = hip_status(hipDeviceGetAttribute(&compute_units, hipDeviceAttributeMultiprocessorCount, device),
                      "hipDeviceGetAttribute");
    if (!st)
      st = hip_status(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, e.plan.plain ? walk_persistent_kernel<true>
                                                                                   : walk_persistent_kernel<false>,
                                                                   e.plan.walk_threads, e.plan.walk_lds),
                      "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (!st && per_cu < 1) st = set_error(N3_EINVAL, "the persistent walk does not fit on a compute unit");
    if (!st) st = hip_status(hipMalloc(&e.queue, 2 * sizeof(unsigned int)), "hipMalloc(work queue)");
    if (!st) st = hip_status(hipMemset(e.queue, 0, 2 * sizeof(unsigned int)), "hipMemset(work queue)");
    if (st) {
      (void)hipFree(e.plan.device);
      if (e.queue) (void)hipFree(e.queue);
      return st;
    }
    e.grid = (unsigned)(per_cu * compute_units);
    if (g_persistent.size() >= kMaxCachedGraphs) {
      PersistentEntry& old = g_persistent.front();
      (void)hipStreamSynchronize(old.stream);
      (void)hipFree(old.plan.device);
      (void)hipFree(old.queue);
      g_persistent.erase(g_persistent.begin());
    }
    g_persistent.push_back(std::move(e));
    entry = &g_persistent.back();
  }
  BatchTables* tables = entry->plan.device;
  unsigned int* queue = entry->queue;
  void* args[] = {&tables, &sweeps, &queue};
  const void* kernel = entry->plan.plain ? reinterpret_cast<const void*>(walk_persistent_kernel<true>)
                                          : reinterpret_cast<const void*>(walk_persistent_kernel<false>);
  return hip_status(hipLaunchCooperativeKernel(kernel,
                                               dim3(entry->grid), dim3(entry->plan.walk_threads), args,
                                               (unsigned)entry->plan.walk_lds, s),
                    "hipLaunchCooperativeKernel(walk_persistent_kernel)");
}

extern "C" int n3_temper_advance(const n3_chains* chains, int steps, int bad_bias, uint64_t seed, void* stream) {
  int st = check_chains(chains);
  if (st) return st;
  if (steps < 1 || bad_bias < 0 || bad_bias > 100)
    return set_error(N3_EINVAL, "steps must be positive and bad_bias in [0,100]");
  WalkPlan plan;
  st = walk_plan(chains->n, chains->replicas, chains->block_waves, &plan, chains->strengths != nullptr,
                 !chains->strengths && !chains->log && bad_bias == 0);
  if (st) return st;
  return enqueue_advance(chains, plan, steps, bad_bias, seed, as_stream(stream));
}

extern "C" int n3_temper_exchange(const n3_chains* chains, int ladder, uint64_t seed, void* stream) {
  int st = check_chains(chains);
  if (st) return st;
  st = check_exchange(chains, ladder);
  if (st) return st;
  return enqueue_exchange(chains, ladder, seed, as_stream(stream));
}

extern "C" int n3_temper_run(const n3_chains* chains, int sweeps, int steps, int ladder, int bad_bias, uint64_t seed,
                             void* stream) {
  int st = check_chains(chains);
  if (st) return st;
  if (sweeps < 0) return set_error(N3_EINVAL, "negative sweeps");
  if (steps < 1 || bad_bias < 0 || bad_bias > 100)
    return set_error(N3_EINVAL, "steps must be positive and bad_bias in [0,100]");
  if (ladder > 1) {
    st = check_exchange(chains, ladder);
    if (st) return st;
  }
  if (sweeps == 0) return N3_OK;
  WalkPlan plan;
  st = walk_plan(chains->n, chains->replicas, chains->block_waves, &plan, chains->strengths != nullptr,
                 !chains->strengths && !chains->log && bad_bias == 0);
  if (st) return st;
  hipStream_t s = as_stream(stream);
  const int per_graph = sweeps < 32 ? sweeps : 32;
  const int full = sweeps / per_graph, rest = sweeps % per_graph;
  GraphKey key{*chains, steps, ladder, bad_bias, per_graph, seed, s, 0};
  (void)hipGetDevice(&key.device);
  hipGraphExec_t exec = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_graph_mutex);
    for (const auto& e : g_graphs)
      if (same_key(e.key, key)) exec = e.exec;
    if (!exec) {
      hipGraph_t graph = nullptr;
      st = capture_graph(chains, plan, per_graph, steps, ladder, bad_bias, seed, &graph, &exec);
      if (st) return st;
      if (g_graphs.size() >= kMaxCachedGraphs) drop_oldest_graph();
      g_graphs.push_back(GraphEntry{key, exec, graph});
    }
  }
  for (int i = 0; i < full; ++i) {
    st = hip_status(hipGraphLaunch(exec, s), "hipGraphLaunch");
    if (st) return st;
  }
  for (int i = 0; i < rest && !st; ++i) {
    st = enqueue_advance(chains, plan, steps, bad_bias, seed, s, true, ladder);
    if (!st && ladder > 1) st = enqueue_exchange(chains, ladder, seed, s);
  }
  return st;
}

extern "C" void n3_temper_release(void) {
  std::lock_guard<std::mutex> lock(g_graph_mutex);
  // nothing may still be running on the buffers and graphs freed below
  (void)hipDeviceSynchronize();
  for (auto& e : g_graphs) {
    (void)hipGraphExecDestroy(e.exec);
    (void)hipGraphDestroy(e.graph);
  }
  g_graphs.clear();
  for (auto& e : g_persistent) {
    (void)hipFree(e.plan.device);
    (void)hipFree(e.queue);
  }
  g_persistent.clear();
  for (auto& e : g_batches) {
    (void)hipGraphExecDestroy(e.exec);
    (void)hipGraphDestroy(e.graph);
    (void)hipFree(e.plan.device);
  }
  g_batches.clear();
}

// Host-side threshold tables shared by the device kernels and any CPU replay.
extern "C" int n3_metropolis_thresholds(const float* host_temperatures, int replicas, int k, uint32_t* host_out) {
  if (!host_temperatures || !host_out || replicas < 1 || k < 2) return set_error(N3_EINVAL, "bad threshold arguments");
  for (int r = 0; r < replicas; ++r) {
    const double t = (double)host_temperatures[r];
    if (!(t > 0.0) || !std::isfinite(t)) return set_error(N3_EINVAL, "temperatures must be positive and finite");
    host_out[(long long)r * k] = 0xFFFFFFFFu;
    for (int d = 1; d < k; ++d) {
      const double x = std::ldexp(std::exp(-(double)d / t), 32);
      host_out[(long long)r * k + d] = x >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)std::floor(x);
    }
  }
  return N3_OK;
}

extern "C" int n3_exchange_thresholds(const float* host_temperatures, int replicas, int ladder, int k,
                                      uint32_t* host_out, int8_t* host_sign) {
  if (!host_temperatures || !host_out || !host_sign || replicas < 1 || k < 2 || ladder < 1)
    return set_error(N3_EINVAL, "bad threshold arguments");
  for (int a = 0; a < replicas; ++a) {
    int b = a + 1;
    if ((a % ladder) == ladder - 1) b = a;
    const double c = 1.0 / (double)host_temperatures[a] - 1.0 / (double)host_temperatures[b];
    host_sign[a] = c > 0 ? 1 : (c < 0 ? -1 : 0);
    const double ac = std::fabs(c);
    host_out[(long long)a * k] = 0xFFFFFFFFu;
    for (int d = 1; d < k; ++d) {
      const double x = std::ldexp(std::exp(-ac * (double)d), 32);
      host_out[(long long)a * k + d] = x >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)std::floor(x);
    }
  }
  return N3_OK;
}
