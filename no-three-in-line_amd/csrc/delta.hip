// 2-switch move scores: the reference's incremental formula
//
//   score(u, v) = current - (badness[u] + badness[v] - joint_old) + new_affected
//
// (accelerated.py:153-194, gpu_delta.py:42-74) for the switch of slots u, v
// to (row_u, col_v), (row_v, col_u), with the same legality rules and
// sentinels.  joint_old counts the points o != u, v on the line (u, v) (or
// coincident with u); new_affected counts the collinear triples of the moved
// state through u' or v', which is bad'(u') + bad'(v') - joint'(u', v'), each
// bad' the sum of C(cnt, 2) over the direction histogram seen from u' (v').
// One wavefront evaluates one (state, u, v) at a time in O(m) table lookups
// instead of the reference's O(m^2) determinant pair loops.
#include <mutex>

#include "runtime.h"
#include "wave_ops.h"

namespace no3 {

struct SwitchArgs {
  const int32_t* p;
  const int32_t* q;
  int n;
  long long states;
  int u_fixed;               // >= 0: this anchor for every state; -1: every anchor (table)
  int layer;                 // -1: every partner slot; 0 / 1: partners of that layer only
  int splits;                // wavefronts per (state, anchor)
  const void* badness;       // [states][m], int32 or int64
  const void* scores;        // [states], int32 or int64; NULL: use `current`
  long long current;
  void* out;                 // [states][anchors][width]
  int width;                 // m (partner slot v) or n (partner row)
  int* status;
};

template <typename T>
__device__ inline long long load_value(const void* base, long long index) {
  return (long long)reinterpret_cast<const T*>(base)[index];
}

template <int CH, typename BT, typename OT>
__global__ void __launch_bounds__(1024) switch_scores_kernel(SwitchArgs a, const uint16_t* __restrict__ table_g,
                                                            DirTableInfo info) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int waves = blockDim.x / kWave;
  const int wave = threadIdx.x / kWave, lane = lane_id();
  uint16_t* table = reinterpret_cast<uint16_t*>(smem);
  uint32_t* hist = reinterpret_cast<uint32_t*>(smem + table_lds_bytes(info)) + wave * info.hist_words;
  stage_table(table_g, table, info.entries);
  for (int i = lane; i < info.hist_words; i += kWave) hist[i] = 0;
  __syncthreads();

  const int n = a.n, m = 2 * n, width = info.width, offset = info.offset;
  const int zero_id = info.directions;
  const long long anchors = a.u_fixed >= 0 ? 1 : m;
  const long long items = a.states * anchors * a.splits;
  for (long long w = (long long)blockIdx.x * waves + wave; w < items; w += (long long)gridDim.x * waves) {
    const int split = (int)(w % a.splits);
    const long long item = w / a.splits;
    const long long s = item / anchors;
    const int u = a.u_fixed >= 0 ? a.u_fixed : (int)(item % anchors);
    OT* out = reinterpret_cast<OT*>(a.out) + item * a.width;

    int key[CH];
    int outside = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int k = lane + kWave * c;
      key[c] = 0;
      if (k < m) {
        const int col = (k & 1) ? a.q[s * n + (k >> 1)] : a.p[s * n + (k >> 1)];
        outside |= (col < 0 || col >= n);
        key[c] = (k >> 1) * width + col;
      }
    }
    if (__any(outside)) {
      if (lane == 0) atomicOr(a.status, 1);
      continue;
    }
    // lanes past the last slot hold slot 0's key (always a valid point)
    const int key0 = __builtin_amdgcn_readfirstlane(key[0]);
#pragma unroll
    for (int c = 0; c < CH; ++c)
      if (lane + kWave * c >= m) key[c] = key0;
    const int trash = trash_word(info, lane);
    const int ru = u >> 1;
    const int key_u = slot_value<CH>(key, u);
    const int cu = key_u - ru * width;
    const int cu_other = slot_value<CH>(key, u ^ 1) - ru * width;
    // direction from u to every slot (joint_old)
    int idu[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) idu[c] = table[key[c] - key_u + offset];
    const long long bad_u = load_value<BT>(a.badness, s * m + u);
    const long long current = a.scores ? load_value<BT>(a.scores, s) : a.current;

    const int partners = a.layer >= 0 ? n : m;
    for (int t = split; t < partners; t += a.splits) {
      const int v = a.layer >= 0 ? 2 * t + a.layer : t;
      const int rv = v >> 1;
      const int key_v = slot_value<CH>(key, v);
      const int cv = key_v - rv * width;
      const int cv_other = slot_value<CH>(key, v ^ 1) - rv * width;
      // accelerated.py:146-151, gpu_delta.py:35-40
      if (v == u || rv == ru || cv == cu_other || cu == cv_other) {
        if (lane == 0) out[t] = (OT)(sizeof(OT) == 8 ? kSentinel64 : (long long)kSentinel32);
        continue;
      }
      const int id_uv = table[key_v - key_u + offset];
      int joint_old = 0;
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int k = lane + kWave * c;
        const bool on = (k < m) && k != u && k != v && (idu[c] == id_uv || idu[c] == zero_id);
        joint_old += __popcll(__ballot(on));
      }
      const int key_u2 = ru * width + cv;  // u' = (row_u, col_v)
      const int key_v2 = rv * width + cu;  // v' = (row_v, col_u)
      const int id_u2v2 = table[key_v2 - key_u2 + offset];
      // Histograms seen from u' and v' over the moved state: bad'(x) is the
      // sum of the ranks the increments return (sum_d C(cnt_d, 2)), so both
      // are built and cleared back to back with one wait.  The moved slot
      // itself looks up the zero vector (rank 0; the legality rules leave
      // no point on u' or v'); lanes past the last slot use their trash
      // word, in the last chunk only.
      int idu2[CH], idv2[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int k = lane + kWave * c;
        const int moved = (k == v) ? key_v2 : ((k == u) ? key_u2 : key[c]);
        idu2[c] = table[moved - key_u2 + offset];
        idv2[c] = table[moved - key_v2 + offset];
      }
      const bool tail_in = lane + kWave * (CH - 1) < m;
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
        const int k = lane + kWave * c;
        part += byte_counter(old_u2[c], (uint32_t)idu2[c]) +
                byte_counter(old_v2[c], (uint32_t)idv2[c]);
        joint_new += __popcll(lanes((k < m) && k != u && k != v) & lanes(idu2[c] == id_u2v2));
      }
      // bad'(u') + bad'(v') (each below 2^16, so the sum fits one word)
      const long long new_affected = (long long)wave_sum_dpp(part) - joint_new;
      const long long bad_v = load_value<BT>(a.badness, s * m + v);
      const long long value = current - (bad_u + bad_v - joint_old) + new_affected;
      if (lane == 0) out[t] = (OT)value;
    }
  }
}

// 3-cycles (accelerated.py:198-276): candidate (u, v, w, orientation) moves
// u <- col v, v <- col w, w <- col u (orientation 0) or u <- col w,
// w <- col v, v <- col u (1).  The reference's three pair loops count the
// moved state's triples through u', v' or w' by inclusion-exclusion; here
//   new = bad'(u') + bad'(v') + bad'(w') - j'(u'v') - j'(u'w') - j'(v'w') + t'
// from three direction histograms (j' the points on the line through two
// moved points, t' whether the three are collinear), and the old side
//   bad[u] + bad[v] + bad[w] - j(uv) - j(uw) - j(vw) + t
// uses the caller's badness like the reference.
template <int CH>
__global__ void __launch_bounds__(1024) three_cycle_kernel(const int32_t* __restrict__ p,
                                                          const int32_t* __restrict__ q, int n,
                                                          const int32_t* __restrict__ candidates, long long count,
                                                          long long current, const long long* __restrict__ badness,
                                                          long long* __restrict__ out, int* __restrict__ status,
                                                          const uint16_t* __restrict__ table_g, DirTableInfo info) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int waves = blockDim.x / kWave;
  const int wave = threadIdx.x / kWave, lane = lane_id();
  uint16_t* table = reinterpret_cast<uint16_t*>(smem);
  uint32_t* hist = reinterpret_cast<uint32_t*>(smem + table_lds_bytes(info)) + wave * info.hist_words;
  stage_table(table_g, table, info.entries);
  for (int i = lane; i < info.hist_words; i += kWave) hist[i] = 0;
  __syncthreads();
  const int m = 2 * n, width = info.width, offset = info.offset, zero_id = info.directions;
  int key[CH];
  int outside = 0;
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int k = lane + kWave * c;
    key[c] = 0;
    if (k < m) {
      const int col = (k & 1) ? q[k >> 1] : p[k >> 1];
      outside |= (col < 0 || col >= n);
      key[c] = (k >> 1) * width + col;
    }
  }
  if (__any(outside)) {
    if (lane == 0) atomicOr(status, 1);
    return;
  }
  const int key0 = __builtin_amdgcn_readfirstlane(key[0]);
#pragma unroll
  for (int c = 0; c < CH; ++c)
    if (lane + kWave * c >= m) key[c] = key0;
  const int trash = trash_word(info, lane);
  for (long long ci = (long long)blockIdx.x * waves + wave; ci < count; ci += (long long)gridDim.x * waves) {
    const int u = candidates[4 * ci], v = candidates[4 * ci + 1], w = candidates[4 * ci + 2];
    const int orientation = candidates[4 * ci + 3];
    if (u < 0 || u >= m || v < 0 || v >= m || w < 0 || w >= m) {
      if (lane == 0) atomicOr(status, 2);
      continue;
    }
    const int key_u = slot_value<CH>(key, u), key_v = slot_value<CH>(key, v), key_w = slot_value<CH>(key, w);
    const int ru = u >> 1, rv = v >> 1, rw = w >> 1;
    const int cu = key_u - ru * width, cv = key_v - rv * width, cw = key_w - rw * width;
    // old side
    const int id_uv = table[key_v - key_u + offset], id_uw = table[key_w - key_u + offset];
    const int id_vw = table[key_w - key_v + offset];
    int joint = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int k = lane + kWave * c;
      const bool live = k < m;
      const int a = table[key[c] - key_u + offset];
      const int b = table[key[c] - key_v + offset];
      joint += __popcll(lanes(live && k != u && k != v) & lanes(a == id_uv || a == zero_id));
      joint += __popcll(lanes(live && k != u && k != w) & lanes(a == id_uw || a == zero_id));
      joint += __popcll(lanes(live && k != v && k != w) & lanes(b == id_vw || b == zero_id));
    }
    const int triple = (key_v != key_u && key_w != key_u && id_uv == id_uw) ? 1 : 0;
    const long long old_affected = badness[u] + badness[v] + badness[w] - joint + triple;
    // moved state
    int nu, nv, nw;
    if (orientation == 0) {
      nu = cv;
      nv = cw;
      nw = cu;
    } else {
      nu = cw;
      nw = cv;
      nv = cu;
    }
    const int ku = ru * width + nu, kv = rv * width + nv, kw = rw * width + nw;
    const int id2_uv = table[kv - ku + offset], id2_uw = table[kw - ku + offset], id2_vw = table[kw - kv + offset];
    // Histograms seen from u', v' and w' over the moved state (each without
    // its own anchor), built and cleared back to back: bad'(x) is the sum of
    // the ranks the increments return.  Excluded lanes update their trash
    // word, which a lane shares across its chunks, and add no rank.
    int idu2[CH], idv2[CH], idw2[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int k = lane + kWave * c;
      const int moved = (k == u) ? ku : ((k == v) ? kv : ((k == w) ? kw : key[c]));
      idu2[c] = table[moved - ku + offset];
      idv2[c] = table[moved - kv + offset];
      idw2[c] = table[moved - kw + offset];
    }
    uint32_t rank_u[CH], rank_v[CH], rank_w[CH];
    auto build = [&](const int(&id)[CH], int anchor_slot, uint32_t(&old)[CH]) {
      uint32_t word[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int k = lane + kWave * c;
        word[c] = (k < m && k != anchor_slot) ? ((uint32_t)id[c] >> 2) : (uint32_t)trash;
        old[c] = atomicAdd(&hist[word[c]], byte_one((uint32_t)id[c]));
      }
      wave_sync();
#pragma unroll
      for (int c = 0; c < CH; ++c) hist[word[c]] = 0u;
      wave_sync();
    };
    build(idu2, u, rank_u);
    build(idv2, v, rank_v);
    build(idw2, w, rank_w);
    // keep the rank extraction below every LDS step above
    __builtin_amdgcn_sched_barrier(0);
    uint32_t part_uv = 0, part_w = 0;
    int joint_new = 0;
    auto rank_of = [&](uint32_t old, int id, int k, int anchor_slot) -> uint32_t {
      return (k < m && k != anchor_slot) ? byte_counter(old, (uint32_t)id) : 0u;
    };
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int k = lane + kWave * c;
      const bool live = k < m;
      part_uv += rank_of(rank_u[c], idu2[c], k, u) + (rank_of(rank_v[c], idv2[c], k, v) << 16);
      part_w += rank_of(rank_w[c], idw2[c], k, w);
      joint_new += __popcll(lanes(live && k != u && k != v) & lanes(idu2[c] == id2_uv));
      joint_new += __popcll(lanes(live && k != u && k != w) & lanes(idu2[c] == id2_uw));
      joint_new += __popcll(lanes(live && k != v && k != w) & lanes(idv2[c] == id2_vw));
    }
    const uint32_t packed = wave_sum_dpp(part_uv), total_w = wave_sum_dpp(part_w);
    const int triple_new = (id2_uv == id2_uw) ? 1 : 0;
    const long long new_affected = (long long)(packed & 0xFFFFu) + (long long)(packed >> 16) + (long long)total_w -
                                   joint_new + triple_new;
    if (lane == 0) out[ci] = current - old_affected + new_affected;
  }
}

template <int CH, typename BT, typename OT>
static int launch_switch(const SwitchArgs& args, hipStream_t stream) {
  const uint16_t* table = nullptr;
  DirTableInfo info{};
  int st = device_dirtable(args.n, &table, &info);
  if (st) return st;
  int lds_limit = 0, compute_units = 0;
  st = device_limits(&lds_limit, &compute_units);
  if (st) return st;
  const int table_bytes = table_lds_bytes(info), hist_bytes = info.hist_words * 4;
  // up to 16 waves per block (n=75: the 44 KB table and 16 histograms of
  // 7 KB fill one CU's LDS); a single-state call (150 items at n=75) then
  // stages the table in 10 blocks of 1,024 threads
  int waves = (lds_limit - table_bytes) / hist_bytes;
  if (waves > 16) waves = 16;
  if (waves < 1) return set_error(N3_EINVAL, "direction table for n=%d does not fit in LDS", args.n);
  const long long anchors = args.u_fixed >= 0 ? 1 : 2 * args.n;
  const long long items = args.states * anchors * args.splits;
  long long blocks = (items + waves - 1) / waves;
  const long long max_blocks = (long long)compute_units * 2;
  if (blocks > max_blocks) blocks = max_blocks;
  if (blocks < 1) blocks = 1;
  const size_t lds = (size_t)table_bytes + (size_t)waves * hist_bytes;
  hipLaunchKernelGGL((switch_scores_kernel<CH, BT, OT>), dim3((unsigned)blocks), dim3(waves * kWave), lds, stream,
                     args, table, info);
  return hip_status(hipGetLastError(), "switch_scores_kernel launch");
}

template <typename BT, typename OT>
static int dispatch_switch(const SwitchArgs& args, hipStream_t stream) {
  const int ch = (2 * args.n + kWave - 1) / kWave;
  switch (ch) {
    case 1: return launch_switch<1, BT, OT>(args, stream);
    case 2: return launch_switch<2, BT, OT>(args, stream);
    case 3: return launch_switch<3, BT, OT>(args, stream);
    default: return launch_switch<4, BT, OT>(args, stream);
  }
}

int switch_status_word(int** word);

static int check_switch(int n, const void* p, const void* q, const void* badness, const void* out) {
  if (n < 2 || n > kMaxNTable) return set_error(N3_EINVAL, "n must be in [2, %d] (got %d)", kMaxNTable, n);
  if (!p || !q || !badness || !out) return set_error(N3_EINVAL, "null pointer");
  return N3_OK;
}

static int run_switch(SwitchArgs args, bool wide, hipStream_t stream) {
  int* status = nullptr;
  int st = switch_status_word(&status);
  if (st) return st;
  args.status = status;
  st = hip_status(hipMemsetAsync(status, 0, sizeof(int), stream), "hipMemsetAsync(status)");
  if (st) return st;
  st = wide ? dispatch_switch<long long, long long>(args, stream) : dispatch_switch<int32_t, int32_t>(args, stream);
  if (st) return st;
  int host = 0;
  st = hip_status(hipMemcpyAsync(&host, status, sizeof(int), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync(status)");
  if (st) return st;
  st = hip_status(hipStreamSynchronize(stream), "hipStreamSynchronize");
  if (st) return st;
  if (host & 1) return set_error(N3_EINVAL, "a column lies outside [0, n)");
  return N3_OK;
}

int switch_status_word(int** word) {
  static int* words[64] = {nullptr};
  static std::mutex mutex;
  int device = 0;
  (void)hipGetDevice(&device);
  if (device < 0 || device >= 64) return set_error(N3_ENODEV, "device index %d out of range", device);
  std::lock_guard<std::mutex> lock(mutex);
  if (!words[device]) {
    const int st = hip_status(hipMalloc(&words[device], sizeof(int)), "hipMalloc(status)");
    if (st) return st;
  }
  *word = words[device];
  return N3_OK;
}

}  // namespace no3

using namespace no3;

extern "C" int n3_two_switch_scores(const int32_t* p, const int32_t* q, const int32_t* badness, int n, int u,
                                    int current_score, int32_t* scores, void* stream) {
  int st = check_switch(n, p, q, badness, scores);
  if (st) return st;
  if (u < 0 || u >= 2 * n) return set_error(N3_EINVAL, "selected index %d outside [0, %d)", u, 2 * n);
  // one wavefront per partner: a single state's call is latency-bound
  SwitchArgs a{p, q, n, 1, u, -1, 2 * n, badness, nullptr, current_score, scores, 2 * n, nullptr};
  return run_switch(a, false, as_stream(stream));
}

extern "C" int n3_two_switch_scores_status(const int32_t* p, const int32_t* q, const int32_t* badness, int n, int u,
                                           int current_score, int32_t* scores, int32_t* status, void* stream) {
  int st = check_switch(n, p, q, badness, scores);
  if (st) return st;
  if (!status) return set_error(N3_EINVAL, "null status");
  if (u < 0 || u >= 2 * n) return set_error(N3_EINVAL, "selected index %d outside [0, %d)", u, 2 * n);
  SwitchArgs a{p, q, n, 1, u, -1, 2 * n, badness, nullptr, current_score, scores, 2 * n, status};
  return dispatch_switch<int32_t, int32_t>(a, as_stream(stream));
}

extern "C" int n3_evaluate_two_switches(const int32_t* p, const int32_t* q, const int64_t* badness, int n, int u,
                                        int64_t current_score, int64_t* scores, void* stream) {
  int st = check_switch(n, p, q, badness, scores);
  if (st) return st;
  if (u < 0 || u >= 2 * n) return set_error(N3_EINVAL, "selected index %d outside [0, %d)", u, 2 * n);
  SwitchArgs a{p, q, n, 1, u, -1, 2 * n, badness, nullptr, (long long)current_score, scores, 2 * n, nullptr};
  return run_switch(a, true, as_stream(stream));
}

extern "C" int n3_evaluate_transpositions(const int32_t* p, const int32_t* q, const int64_t* badness, int n, int layer,
                                          int row, int64_t current_score, int64_t* scores, void* stream) {
  int st = check_switch(n, p, q, badness, scores);
  if (st) return st;
  if (layer != 0 && layer != 1) return set_error(N3_EINVAL, "layer must be 0 (P) or 1 (Q)");
  if (row < 0 || row >= n) return set_error(N3_EINVAL, "row %d outside [0, %d)", row, n);
  SwitchArgs a{p, q, n, 1, 2 * row + layer, layer, n, badness, nullptr, (long long)current_score, scores, n, nullptr};
  return run_switch(a, true, as_stream(stream));
}

extern "C" int n3_delta_table(const int32_t* p, const int32_t* q, const int32_t* scores, const int32_t* badness, int n,
                              int64_t count, int32_t* table, void* stream) {
  int st = check_switch(n, p, q, badness, table);
  if (st) return st;
  if (!scores) return set_error(N3_EINVAL, "null pointer");
  if (count < 0) return set_error(N3_EINVAL, "negative count");
  if (count == 0) return N3_OK;
  SwitchArgs a{p, q, n, (long long)count, -1, -1, 1, badness, scores, 0, table, 2 * n, nullptr};
  return run_switch(a, false, as_stream(stream));
}

extern "C" int n3_evaluate_three_cycles(const int32_t* p, const int32_t* q, const int64_t* badness, int n,
                                        const int32_t* candidates, int64_t count, int64_t current_score,
                                        int64_t* scores, void* stream) {
  int st = check_switch(n, p, q, badness, scores);
  if (count == 0) return N3_OK;
  if (st) return st;
  if (!candidates || count < 0) return set_error(N3_EINVAL, "bad candidates");
  const uint16_t* table = nullptr;
  DirTableInfo info{};
  st = device_dirtable(n, &table, &info);
  if (st) return st;
  int* status = nullptr;
  st = switch_status_word(&status);
  if (st) return st;
  hipStream_t s = as_stream(stream);
  st = hip_status(hipMemsetAsync(status, 0, sizeof(int), s), "hipMemsetAsync(status)");
  if (st) return st;
  // up to 16 waves per block beside the table (a CU's LDS holds one block)
  int lds_limit = 0, compute_units = 0;
  st = device_limits(&lds_limit, &compute_units);
  if (st) return st;
  int waves = (lds_limit - table_lds_bytes(info)) / (info.hist_words * 4);
  if (waves > 16) waves = 16;
  if (waves < 1) return set_error(N3_EINVAL, "direction table for n=%d does not fit in LDS", n);
  const size_t lds = (size_t)table_lds_bytes(info) + (size_t)waves * info.hist_words * 4;
  long long blocks = ((long long)count + waves - 1) / waves;
  if (blocks > 2 * compute_units) blocks = 2 * compute_units;
  const int ch = (2 * n + kWave - 1) / kWave;
  void (*kernel)(const int32_t*, const int32_t*, int, const int32_t*, long long, long long, const long long*,
                 long long*, int*, const uint16_t*, DirTableInfo) =
      ch == 1 ? three_cycle_kernel<1> : (ch == 2 ? three_cycle_kernel<2> : (ch == 3 ? three_cycle_kernel<3> : three_cycle_kernel<4>));
  hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(waves * kWave), lds, s, p, q, n, candidates,
                     (long long)count, (long long)current_score, reinterpret_cast<const long long*>(badness),
                     reinterpret_cast<long long*>(scores), status, table, info);
  st = hip_status(hipGetLastError(), "three_cycle_kernel launch");
  if (st) return st;
  int host = 0;
  st = hip_status(hipMemcpyAsync(&host, status, sizeof(int), hipMemcpyDeviceToHost, s), "hipMemcpyAsync(status)");
  if (st) return st;
  st = hip_status(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (st) return st;
  if (host & 1) return set_error(N3_EINVAL, "a column lies outside [0, n)");
  if (host & 2) return set_error(N3_EINVAL, "a candidate slot lies outside [0, 2n)");
  return N3_OK;
}
