// Population kernels: the seeded state generator and exact collinear-triple
// scoring with per-slot badness.
//
// Scoring (direction histograms).  One wavefront owns one state; lane l holds
// slots l, l+64, l+128, ... in registers as keys row*W + column, and a per-n
// LDS table maps the key difference of two slots to the id of their
// primitive undirected direction.
//  - score_onesided_kernel (n <= 100): anchor a builds a histogram of the
//    directions to the slots b < a, each b reads back its count, and every
//    collinear triple is counted once, at its last slot (see the kernel).
//    C(m, 2) lookups per state; slots that are alone on their line through
//    the anchor skip the histogram.
// The result is the reference's per-point badness (accelerated.py:34-36)
// and E = sum_b badness_b / 3 (accelerated.py:33).
//
// States with coincident points (P[r] == Q[r]) have no direction for the
// pair; those states take the determinant path, which counts them exactly
// like the reference (gpu.py:19-33).
#include <algorithm>
#include <mutex>
#include <numeric>
#include <vector>

#include "runtime.h"
#include "wave_ops.h"

namespace no3 {

// ---------------------------------------------------------------------------
// generator
// ---------------------------------------------------------------------------

constexpr int kGenThreads = 64;
constexpr int kGenStride = 2 * kMaxN + 4;  // bytes per thread in LDS

struct GenStream {
  uint32_t k0, k1, c0, c1, c2;
  uint32_t buf[4];
  int pos;
  __device__ uint32_t draw() {
    if (pos == 4) {
      philox4x32_10(c0, c1, c2, kTagGen, k0, k1, buf);
      c2 += 1u;
      pos = 0;
    }
    return buf[pos++];
  }
};

__global__ void __launch_bounds__(kGenThreads) generate_kernel(uint64_t seed, int n, long long count,
                                                               long long first_index, int32_t* __restrict__ p,
                                                               int32_t* __restrict__ q) {
  __shared__ uint8_t space[kGenThreads * kGenStride];
  const long long s = (long long)blockIdx.x * kGenThreads + threadIdx.x;
  if (s >= count) return;
  uint8_t* P = space + threadIdx.x * kGenStride;
  uint8_t* Q = P + kMaxN;
  const long long index = first_index + s;
  GenStream g{(uint32_t)seed, (uint32_t)(seed >> 32), (uint32_t)index, (uint32_t)((unsigned long long)index >> 32), 0u,
              {0, 0, 0, 0}, 4};
  for (int i = 0; i < n; ++i) {
    P[i] = (uint8_t)i;
    Q[i] = (uint8_t)i;
  }
  for (int i = n - 1; i >= 1; --i) {
    const int j = (int)mulhi32(g.draw(), (uint32_t)(i + 1));
    const uint8_t t = P[i];
    P[i] = P[j];
    P[j] = t;
  }
  for (int i = n - 1; i >= 1; --i) {
    const int j = (int)mulhi32(g.draw(), (uint32_t)(i + 1));
    const uint8_t t = Q[i];
    Q[i] = Q[j];
    Q[j] = t;
  }
  for (int i = 0; i < n; ++i) {
    if (Q[i] == P[i]) {
      int r = (int)mulhi32(g.draw(), (uint32_t)(n - 1));
      if (r >= i) r += 1;
      const uint8_t t = Q[i];
      Q[i] = Q[r];
      Q[r] = t;
    }
  }
  for (int i = 0; i < n; ++i) {
    p[s * n + i] = P[i];
    q[s * n + i] = Q[i];
  }
}

// ---------------------------------------------------------------------------
// determinant path (one wavefront per state)
// ---------------------------------------------------------------------------

// Exact reference count for the state whose columns are in cols[0..m) (LDS):
// for every pair i < j (uniform), lanes test the slots k > j with
// dr1 * (c_k - c_i) == dc1 * (r_k - r_i) (accelerated.py:32, gpu.py:30).
// Hits for k stay in the lane's register; hits for i and j are credited in
// the LDS array credit[0..m).  Returns, per lane and chunk, the badness.
template <int CH>
__device__ void determinant_state(const int32_t* cols, int32_t* credit, int m, int bad[CH]) {
  const int lane = lane_id();
  int colk[CH], rowk[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int k = lane + kWave * c;
    colk[c] = k < m ? cols[k] : 0;
    rowk[c] = k >> 1;
    bad[c] = 0;
  }
  for (int i = lane; i < m; i += kWave) credit[i] = 0;
  wave_sync();
  for (int i = 0; i < m - 2; ++i) {
    const int ri = i >> 1;
    const int ci = __builtin_amdgcn_readfirstlane(cols[i]);
    for (int j = i + 1; j < m - 1; ++j) {
      const int dr1 = (j >> 1) - ri;
      const int dc1 = __builtin_amdgcn_readfirstlane(cols[j]) - ci;
      const int key = dr1 * ci - dc1 * ri;
      int hits = 0;
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        if (kWave * c + kWave - 1 <= j) continue;  // whole chunk at k <= j
        const int k = lane + kWave * c;
        const bool h = (k > j) && (k < m) && (__mul24(dr1, colk[c]) - __mul24(dc1, rowk[c]) == key);
        bad[c] += h ? 1 : 0;
        hits += __popcll(__ballot(h));
      }
      if (hits) {
        if (lane == 0) {
          credit[i] += hits;
          credit[j] += hits;
        }
        wave_sync();
      }
    }
  }
  wave_sync();
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int k = lane + kWave * c;
    if (k < m) bad[c] += credit[k];
  }
  wave_sync();
}

template <int CH>
__global__ void __launch_bounds__(256) score_determinant_kernel(const int32_t* __restrict__ p,
                                                                const int32_t* __restrict__ q, int n, long long count,
                                                                int32_t* __restrict__ scores,
                                                                int32_t* __restrict__ badness,
                                                                int* __restrict__ status) {
  __shared__ int32_t space[4][2 * 2 * kMaxN];
  const int waves = blockDim.x / kWave;
  const int wave = threadIdx.x / kWave, lane = lane_id();
  int32_t* cols = space[wave];
  int32_t* credit = cols + 2 * kMaxN;
  const int m = 2 * n;
  for (long long s = (long long)blockIdx.x * waves + wave; s < count; s += (long long)gridDim.x * waves) {
    int outside = 0;
    for (int k = lane; k < m; k += kWave) {
      const int col = (k & 1) ? q[s * n + (k >> 1)] : p[s * n + (k >> 1)];
      outside |= (col < 0 || col >= n);
      cols[k] = col;
    }
    wave_sync();
    if (__any(outside)) {
      if (lane == 0) {
        scores[s] = -1;
        atomicOr(status, 1);
      }
      continue;
    }
    int bad[CH];
    determinant_state<CH>(cols, credit, m, bad);
    int sum = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int k = lane + kWave * c;
      if (k < m) {
        sum += bad[c];
        if (badness) badness[s * m + k] = bad[c];
      }
    }
    sum = wave_sum(sum);
    if (lane == 0) scores[s] = sum / 3;
  }
}

// ---------------------------------------------------------------------------
// one-sided direction-histogram path (one wavefront per state)
// ---------------------------------------------------------------------------

// Anchor a sees only the slots b < a, all in rows <= row(a), so the path
// stages only the direction-table rows dr = -(n-1) .. 0.  An entry is 4 bytes:
//   counter word's byte offset << 16 | lone-row bound << 8 | counter shift,
// laid out with dr fastest and a column stride S = 16 (mod 32) words, so the
// points of the 16 consecutive rows that a half-wave looks up fall in
// distinct bank pairs (row-major entries fall in random banks).  Byte
// counters, four per word, exist only for the directions whose lines can
// meet three cells of the grid (see lone_row_bound).
struct OneSidedTable {
  const uint32_t* data;  // entries of rows dr = -(n-1) .. 0
  int bytes;             // staged bytes (multiple of 16)
  int stride;            // entries between columns
  int hist_words;        // 32-bit words of one wave's histogram
};

// After anchor a's histogram of the directions to the slots b < a is built,
// each b reads back cnt_b, the number of slots < a on the line (a, b): b
// lies in cnt_b - 1 collinear triples whose last slot is a, and a lies in
// sum_d C(cnt_d, 2) = (sum_b (cnt_b - 1)) / 2 of them.  Every triple is
// counted once, at its last slot, so the badness of a slot is its credits
// from later anchors plus its own last-slot count (accelerated.py:34-36),
// with half the pairs of the two-sided histogram.  Anchors in the last,
// partly filled chunk are the only ones that touch it.
template <int CH>
__global__ void __launch_bounds__(1024) score_onesided_kernel(const int32_t* __restrict__ p,
                                                             const int32_t* __restrict__ q, int n, long long count,
                                                             int32_t* __restrict__ scores,
                                                             int32_t* __restrict__ badness, OneSidedTable tab,
                                                             DirTableInfo info, int* __restrict__ status) {
  const int waves = blockDim.x / kWave;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = lane_id();
  extern __shared__ __align__(16) unsigned char lds[];
  const uint32_t hist_base = (uint32_t)(tab.bytes + wave * tab.hist_words * 4);  // byte address
  uint32_t* hist = reinterpret_cast<uint32_t*>(lds + hist_base);
  using LdsWord = __attribute__((address_space(3))) uint32_t;
  // an entry's counter word; the low five bits of the entry are the
  // counter's shift, and v_lshlrev and v_bfe read only those
  auto counter = [&](uint32_t e) { return (LdsWord*)reinterpret_cast<uint32_t*>(lds + hist_base + (e >> 16)); };
  {
    const uint2* src = reinterpret_cast<const uint2*>(tab.data);
    uint2* dst = reinterpret_cast<uint2*>(lds);
    for (int i = threadIdx.x; i < tab.bytes / 8; i += blockDim.x) dst[i] = src[i];
  }
  for (int i = lane; i < tab.hist_words; i += kWave) hist[i] = 0;
  __syncthreads();

  const int m = 2 * n;
  const int width = info.width;
  const int span = n - 1;
  // byte offset of the entry of (dr, dc) = key_b - key_a + zero
  const int zero = 4 * (span + span * tab.stride);
  for (long long s = (long long)blockIdx.x * waves + wave; s < count; s += (long long)gridDim.x * waves) {
    int key[CH];
    int outside = 0, coincide = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int k = lane + kWave * c;
      key[c] = 0;
      if (k < m) {
        const int col = (k & 1) ? q[s * n + (k >> 1)] : p[s * n + (k >> 1)];
        outside |= (col < 0 || col >= n);
        key[c] = (k >> 1) * width + col;
      }
      // slots 2r and 2r+1 sit in neighbouring lanes of the same chunk
      const int partner = __shfl_xor(key[c], 1, kWave);
      coincide |= (k < m) && (partner == key[c]);
    }
    if (__any(outside)) {
      if (lane == 0) {
        scores[s] = -1;
        atomicOr(status, 1);
      }
      continue;
    }
    if (__any(coincide)) {
      // coincident points: exact determinant count, with the histogram area
      // as scratch for the state's columns (it is cleared again below)
      int32_t* cols = reinterpret_cast<int32_t*>(hist);
      int32_t* credit = cols + m;
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int k = lane + kWave * c;
        if (k < m) cols[k] = key[c] - (k >> 1) * width;
      }
      wave_sync();
      int bad[CH];
      determinant_state<CH>(cols, credit, m, bad);
      int sum = 0;
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int k = lane + kWave * c;
        if (k < m) {
          sum += bad[c];
          if (badness) badness[s * m + k] = bad[c];
        }
      }
      sum = wave_sum(sum);
      if (lane == 0) scores[s] = sum / 3;
      for (int i = lane; i < 2 * m; i += kWave) hist[i] = 0;
      wave_sync();
      continue;
    }

    // at[c]: byte offset of the lane's point in the staged table's layout;
    // row[c]: its row
    int at[CH], row[CH], counts[CH], own[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int k = lane + kWave * c;
      row[c] = k >> 1;
      at[c] = 4 * (row[c] + (key[c] - row[c] * width) * tab.stride);
      counts[c] = own[c] = 0;
    }
    // Lanes without an earlier slot look up the zero vector's entry (one
    // broadcast address), read back and clear its word and skip only the
    // increment: no lane increments that word, so their read-back adds 0
    // and no LDS step waits behind a branch.
    const unsigned char* idle_entry = lds + zero;
    // One anchor (slot 64 ca + l) is: lookup of the earlier slots'
    // directions, histogram increments, read-back of each slot's count and
    // clear.  Chunks before ca are all earlier slots, chunks after it none;
    // in chunk ca the lanes before l are.  The next anchor's lookups (table
    // only) issue together with this anchor's read-backs and clears, so each
    // anchor waits on LDS once.
    // lane < l as a scalar lane mask (l is uniform), so the tests below cost
    // no vector compare
    auto before = [](int l) { return __builtin_amdgcn_inverse_ballot_w64((1ull << l) - 1ull); };
    auto lookup = [&](int ca, int l, uint32_t(&e)[CH]) {
      const unsigned char* from = lds + (zero - __builtin_amdgcn_readlane(at[ca], l));
#pragma unroll
      for (int c = 0; c <= ca; ++c) {
        const unsigned char* entry = (c < ca || before(l)) ? from + at[c] : idle_entry;
        e[c] = *reinterpret_cast<const uint32_t*>(entry);
      }
    };
    // Byte 1 of an entry is the least row of an earlier slot b whose line
    // through the anchor can hold another earlier point (255: no row can,
    // 0: every row can).  A lane below it is alone on its line and skips the
    // increment.  Its read-back finds 0 (no other lane shares its direction;
    // the entries of 255 point at the word that no lane increments), so the
    // lanes accumulate count - 1 clamped at 0: 0 for lone lanes and for
    // lanes without an earlier slot alike.
    auto alone = [&](int c, uint32_t e) -> bool { return ((e >> 8) & 0xFFu) > (uint32_t)row[c]; };
    // ctr[c]: the lane's counter word, computed once for the increment and
    // the read-back (the empty asm keeps the compiler from recomputing it on
    // both sides of the increment's mask)
    auto increment = [&](int ca, int l, const uint32_t(&e)[CH], LdsWord* (&ctr)[CH]) {
#pragma unroll
      for (int c = 0; c <= ca; ++c) {
        ctr[c] = counter(e[c]);
        asm volatile("" : "+v"(ctr[c]));
        const uint32_t one = 1u << (e[c] & 31u);
        // one combined mask (a short-circuit && nests two exec masks)
        if ((int)(c < ca || before(l)) & (int)!alone(c, e[c]))
          __hip_atomic_fetch_add(ctr[c], one, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      wave_sync();
    };
    auto collect = [&](int ca, LdsWord* const (&ctr)[CH], uint32_t(&word)[CH]) {
#pragma unroll
      for (int c = 0; c <= ca; ++c) word[c] = *ctr[c];
#pragma unroll
      for (int c = 0; c <= ca; ++c) *ctr[c] = 0u;
    };
    // sum of the lanes' counts - 1; each slot keeps its own running sum
    auto accumulate = [&](int ca, const uint32_t(&e)[CH], const uint32_t(&word)[CH]) {
      uint32_t part = 0;
#pragma unroll
      for (int c = 0; c <= ca; ++c) {
        const uint32_t cnt = __builtin_elementwise_sub_sat(__builtin_amdgcn_ubfe(word[c], e[c], 8), 1u);
        counts[c] += (int)cnt;
        part += cnt;
      }
      return part;
    };
    // Anchors in pairs, from lane 0 of each chunk: the chunks hold 64 slots
    // or, the last one, m - 64 (CH - 1), so every count is even.  Anchor 0
    // has no earlier slot and adds nothing.  The last pair's lookahead reads
    // the table for its own last lane again.
#pragma unroll
    for (int ca = 0; ca < CH; ++ca) {
      const int limit = min(kWave, m - kWave * ca);
      uint32_t e0[CH], e1[CH], word0[CH], word1[CH];
      LdsWord* ctr[CH];
      lookup(ca, 0, e0);
      wave_sync();
      for (int l = 0; l < limit; l += 2) {
        increment(ca, l, e0, ctr);
        collect(ca, ctr, word0);
        lookup(ca, l + 1, e1);
        wave_sync();
        const uint32_t p0 = accumulate(ca, e0, word0);
        increment(ca, l + 1, e1, ctr);
        collect(ca, ctr, word1);
        lookup(ca, min(l + 2, limit - 1), e0);
        wave_sync();
        const uint32_t p1 = accumulate(ca, e1, word1);
        // two anchors per reduction: an anchor's sum of counts - 1 is at
        // most (m - 1)(m - 2) < 2^16, so they share a word; its last-slot
        // triples are half of it
        const uint32_t total = wave_sum_dpp(p0 | (p1 << 16));
        own[ca] = write_lane((int)(total >> 17), l + 1, own[ca]);
        own[ca] = write_lane((int)((total & 0xFFFFu) >> 1), l, own[ca]);
      }
    }
    // slot k was an earlier slot of the anchors k+1 .. m-1, and each of them
    // adds its count - 1 (0 where k is alone on their line)
    int sum = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int k = lane + kWave * c;
      if (k < m) {
        const int b = counts[c] + own[c];
        sum += b;
        if (badness) badness[s * m + k] = b;
      }
    }
    sum = wave_sum(sum);
    if (lane == 0) scores[s] = sum / 3;
  }
}

// The least row r_b for which the earlier slot b = a + (dr, dc) (dr <= 0) can
// share its line through the anchor a with another earlier slot; a slot in a
// lower row is alone on that line.  Earlier slots lie in rows <= row(a), so
// the other points of the line that matter are the lattice points strictly
// between a and b and those beyond b.
//  - dr = 0: the other slot of a's own row, the only earlier slot in it: 255.
//  - primitive step (p, q) = (dr, dc) / g with max(|p|, |q|) > (n - 1) / 2:
//    no line of that direction meets three cells of the grid: 255.
//  - g >= 2: lattice points lie between a and b: 0.
//  - g = 1: the first point beyond b is b - (a - b), in row 2 r_b - r_a, so
//    r_b >= r_a - r_b = -dr.
// The zero offset (a lane without an earlier slot) gets 0.
inline uint32_t lone_row_bound(int n, int dr, int dc) {
  if (dr == 0) return dc == 0 ? 0u : 255u;
  const int adr = dr < 0 ? -dr : dr, adc = dc < 0 ? -dc : dc;
  const int g = std::gcd(adr, adc);
  if (2 * std::max(adr / g, adc / g) > n - 1) return 255u;
  return g >= 2 ? 0u : (uint32_t)adr;
}

// Column stride of the table: at least n, and 16 (mod 32).
inline int onesided_stride(int n) { return ((n + 15) / 32) * 32 + 16; }

// Directions with a counter: the primitive (p, q), p > 0, with
// max(p, |q|) <= (n - 1) / 2 (the lines that can meet three cells).
static int onesided_counters(int n) {
  int count = 0;
  for (int p = 1; 2 * p <= n - 1; ++p)
    for (int q = -(n - 1) / 2; q <= (n - 1) / 2; ++q)
      if (std::gcd(p, q < 0 ? -q : q) == 1) ++count;
  return count;
}

// Words of one wave's histogram: the counters, the shared word after them,
// and at least 4n words (the determinant recount's scratch), rounded to 16.
static int onesided_hist_words(int n) {
  int words = ((onesided_counters(n) + 3) & ~3) / 4 + 1;
  if (words < 4 * n) words = 4 * n;
  return (words + 15) / 16 * 16;
}

// Resident one-sided tables per (device, n), built from the direction ids on
// first use.
static int onesided_table(int n, const DirTableInfo& info, OneSidedTable* out) {
  struct Resident {
    int device, n;
    OneSidedTable tab;
  };
  static std::mutex mutex;
  static std::vector<Resident> tables;
  int device = 0;
  int st = hip_status(hipGetDevice(&device), "hipGetDevice");
  if (st) return st;
  std::lock_guard<std::mutex> lock(mutex);
  for (const auto& t : tables) {
    if (t.device == device && t.n == n) {
      *out = t.tab;
      return N3_OK;
    }
  }
  std::vector<uint16_t> ids;
  DirTableInfo built{};
  st = build_dirtable(n, ids, built);
  if (st) return st;
  const int w = info.width, span = n - 1;
  // Counters for the ids [0, counted): the directions whose lines can meet
  // three cells of the grid.  The zero vector and every other direction
  // share the word after them, which no lane increments.
  if (built.counted != onesided_counters(n))
    return set_error(N3_EINVAL, "one-sided table: %d counted directions at n=%d, expected %d", built.counted, n,
                     onesided_counters(n));
  const uint32_t counted = (uint32_t)built.counted, shared_id = (counted + 3u) & ~3u;
  const int stride = onesided_stride(n);
  std::vector<uint32_t> host(((size_t)stride * w + 3) / 4 * 4, 0u);  // padded to 16 bytes
  for (int dr = 0; dr <= span; ++dr)  // rows dr - span = -(n-1) .. 0
    for (int dc = 0; dc < w; ++dc) {
      const uint32_t need = lone_row_bound(n, dr - span, dc - span);
      const uint32_t id = ids[(size_t)dr * w + dc];
      if ((need == 255u) != (id >= counted && !(dr == span && dc == span)))
        return set_error(N3_EINVAL, "one-sided table: lone bound and counted ids disagree at n=%d", n);
      const uint32_t cid = id < counted ? id : shared_id;
      host[(size_t)dr + (size_t)stride * dc] = ((cid & ~3u) << 16) | (need << 8) | ((cid & 3u) << 3);
    }
  OneSidedTable tab{nullptr, (int)(host.size() * 4), stride, onesided_hist_words(n)};
  void* data = nullptr;
  st = hip_status(hipMalloc(&data, host.size() * 4), "hipMalloc(one-sided table)");
  if (st) return st;
  st = hip_status(hipMemcpy(data, host.data(), host.size() * 4, hipMemcpyHostToDevice), "hipMemcpy(one-sided table)");
  if (st) {
    (void)hipFree(data);
    return st;
  }
  tab.data = static_cast<const uint32_t*>(data);
  tables.push_back(Resident{device, n, tab});
  *out = tab;
  return N3_OK;
}

// Up to 16 waves per block; two blocks share a CU where their LDS fits
// (n = 100 stages 87 KB of table and 3 KB per wave; n = 75 46.6 KB and
// 1.75 KB per wave).
template <int CH>
static int launch_onesided(const int32_t* p, const int32_t* q, int n, int64_t count, int32_t* scores, int32_t* badness,
                           const DirTableInfo& info, int* status, hipStream_t stream) {
  int lds_limit = 0, compute_units = 0;
  int st = device_limits(&lds_limit, &compute_units);
  if (st) return st;
  OneSidedTable tab{};
  st = onesided_table(n, info, &tab);
  if (st) return st;
  const int bytes = tab.hist_words * 4;
  int waves = (lds_limit - tab.bytes) / bytes;
  if (waves > 16) waves = 16;
  if (waves < 1) return set_error(N3_EINVAL, "direction table for n=%d does not fit in LDS", n);
  const size_t lds = (size_t)tab.bytes + (size_t)waves * bytes;
  const int per_cu = (2 * lds <= (size_t)lds_limit) ? 2 : 1;
  long long blocks = ((long long)count + waves - 1) / waves;
  const long long max_blocks = (long long)compute_units * per_cu;
  if (blocks > max_blocks) blocks = max_blocks;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL((score_onesided_kernel<CH>), dim3((unsigned)blocks), dim3(waves * kWave), lds, stream, p, q, n,
                     (long long)count, scores, badness, tab, info, status);
  return hip_status(hipGetLastError(), "score_onesided_kernel launch");
}

template <int CH>
static int launch_determinant(const int32_t* p, const int32_t* q, int n, int64_t count, int32_t* scores,
                              int32_t* badness, int* status, hipStream_t stream) {
  long long blocks = ((long long)count + 3) / 4;
  if (blocks > 65535LL * 4) blocks = 65535LL * 4;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(score_determinant_kernel<CH>, dim3((unsigned)blocks), dim3(256), 0, stream, p, q, n,
                     (long long)count, scores, badness, status);
  return hip_status(hipGetLastError(), "score_determinant_kernel launch");
}

// Per-device status word, cleared before and read after a scoring launch.
static int* status_word(int* st_out) {
  static int* words[64] = {nullptr};
  static std::mutex mutex;
  int device = 0;
  (void)hipGetDevice(&device);
  if (device < 0 || device >= 64) {
    *st_out = set_error(N3_ENODEV, "device index %d out of range", device);
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(mutex);
  if (!words[device]) {
    int* w = nullptr;
    *st_out = hip_status(hipMalloc(&w, sizeof(int)), "hipMalloc(status)");
    if (*st_out) return nullptr;
    *st_out = hip_status(hipMemset(w, 0, sizeof(int)), "hipMemset(status)");
    if (*st_out) return nullptr;
    words[device] = w;
  }
  *st_out = N3_OK;
  return words[device];
}

int score_population_impl(const int32_t* p, const int32_t* q, int n, int64_t count, int32_t* scores, int32_t* badness,
                          int method, hipStream_t stream) {
  int st = check_columns_launch_args(n, count);
  if (st) return st;
  if (count == 0) return N3_OK;
  if (!p || !q || !scores) return set_error(N3_EINVAL, "null pointer");
  // the kernels only ever set bits of the status word; n3_score_status
  // reads and clears it, so a launch costs no extra device operation
  int* status = status_word(&st);
  if (st) return st;
  const int m = 2 * n;
  const int ch = (m + kWave - 1) / kWave;
  if (method == N3_SCORE_DETERMINANT || n > kMaxNTable) {
    switch (ch) {
      case 1: st = launch_determinant<1>(p, q, n, count, scores, badness, status, stream); break;
      case 2: st = launch_determinant<2>(p, q, n, count, scores, badness, status, stream); break;
      case 3: st = launch_determinant<3>(p, q, n, count, scores, badness, status, stream); break;
      case 4: st = launch_determinant<4>(p, q, n, count, scores, badness, status, stream); break;
      case 5: st = launch_determinant<5>(p, q, n, count, scores, badness, status, stream); break;
      case 6: st = launch_determinant<6>(p, q, n, count, scores, badness, status, stream); break;
      case 7: st = launch_determinant<7>(p, q, n, count, scores, badness, status, stream); break;
      default: st = launch_determinant<8>(p, q, n, count, scores, badness, status, stream); break;
    }
  } else if (method == N3_SCORE_AUTO || method == N3_SCORE_SLOPE) {
    const uint16_t* table = nullptr;
    DirTableInfo info{};
    st = device_dirtable(n, &table, &info);
    if (st) return st;
    switch (ch) {
      case 1: st = launch_onesided<1>(p, q, n, count, scores, badness, info, status, stream); break;
      case 2: st = launch_onesided<2>(p, q, n, count, scores, badness, info, status, stream); break;
      case 3: st = launch_onesided<3>(p, q, n, count, scores, badness, info, status, stream); break;
      default: st = launch_onesided<4>(p, q, n, count, scores, badness, info, status, stream); break;
    }
  } else {
    return set_error(N3_EINVAL, "unknown scoring method %d", method);
  }
  if (st) return st;
  return N3_OK;
}

// Reads back the status word of the last scoring launch on this device
// (synchronises `stream`).
int score_status(hipStream_t stream) {
  int st = N3_OK;
  int* status = status_word(&st);
  if (st) return st;
  int host = 0;
  st = hip_status(hipMemcpyAsync(&host, status, sizeof(int), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync(status)");
  if (st) return st;
  st = hip_status(hipStreamSynchronize(stream), "hipStreamSynchronize");
  if (st) return st;
  if (host == 0) return N3_OK;
  st = hip_status(hipMemsetAsync(status, 0, sizeof(int), stream), "hipMemsetAsync(status)");
  if (st) return st;
  st = hip_status(hipStreamSynchronize(stream), "hipStreamSynchronize");
  if (st) return st;
  return set_error(N3_EINVAL, "a column lies outside [0, n)");
}

}  // namespace no3

using namespace no3;

extern "C" int n3_generate_states(uint64_t seed, int n, int64_t count, int64_t first_index, int32_t* p, int32_t* q,
                                  void* stream) {
  int st = check_columns_launch_args(n, count);
  if (st) return st;
  if (count == 0) return N3_OK;
  if (!p || !q) return set_error(N3_EINVAL, "null pointer");
  const long long blocks = ((long long)count + kGenThreads - 1) / kGenThreads;
  hipLaunchKernelGGL(generate_kernel, dim3((unsigned)blocks), dim3(kGenThreads), 0, as_stream(stream), seed, n,
                     (long long)count, (long long)first_index, p, q);
  return hip_status(hipGetLastError(), "generate_kernel launch");
}

extern "C" int n3_score_population(const int32_t* p, const int32_t* q, int n, int64_t count, int32_t* scores,
                                   int32_t* badness_or_null, int method, void* stream) {
  return score_population_impl(p, q, n, count, scores, badness_or_null, method, as_stream(stream));
}

extern "C" int n3_score_status(void* stream) { return score_status(as_stream(stream)); }
