// Device search engines built on the scorer:
//  - GA mutation of a parent pool (kernel mutate_population,
//    gpu_device_search.py:27-113): every child copies a parent and, past the
//    elites, applies a random number of random legal 2-switches;
//  - materialisation of enumerated 2-switch children (kernel
//    enumerate_switches, gpu_exhaustive_beam.py:17-57);
//  - the secondary elite-ranking keys of rank_elites
//    (gpu_device_search.py:185-197): max badness and squared badness mass.
//
// One wavefront per child / candidate / state: the 2n columns sit in the
// lanes (slot s = lane + 64 c in chunk c), so copies are coalesced and every
// random draw, legality test and swap is wave-uniform.  Draws come from
// Philox4x32-10 keyed by the seed with counters (child, generation, draw,
// "MUTA"); a draw in [0, k) is mulhi(word, k).
#include "runtime.h"
#include "wave_ops.h"

namespace no3 {

constexpr uint32_t kTagMutate = 0x4D555441u;  // "MUTA"
constexpr int kMutateAttempts = 32;           // gpu_device_search.py:79
constexpr int kMaxStrength = 1 << 16;

template <int CH>
__device__ inline void load_state(const int32_t* __restrict__ p, const int32_t* __restrict__ q, long long index,
                                  int n, int (&col)[CH]) {
  const int lane = lane_id(), m = 2 * n;
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int k = lane + kWave * c;
    col[c] = k < m ? ((k & 1) ? q[index * n + (k >> 1)] : p[index * n + (k >> 1)]) : -1;
  }
}

template <int CH>
__device__ inline void store_state(int32_t* __restrict__ p, int32_t* __restrict__ q, long long index, int n,
                                   const int (&col)[CH]) {
  const int lane = lane_id(), m = 2 * n;
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int k = lane + kWave * c;
    if (k < m) {
      if (k & 1)
        q[index * n + (k >> 1)] = col[c];
      else
        p[index * n + (k >> 1)] = col[c];
    }
  }
}

// The 2-switch of slots u and v if it is legal (different rows, different
// columns, neither row left with one column twice): gpu_exhaustive_beam.py
// :39-54 and the attempt loop of gpu_device_search.py:80-110.
template <int CH>
__device__ inline bool try_switch(int (&col)[CH], int u, int v) {
  if ((u >> 1) == (v >> 1)) return false;
  const int cu = slot_value<CH>(col, u), cv = slot_value<CH>(col, v);
  if (cu == cv) return false;
  const int other_u = slot_value<CH>(col, u ^ 1), other_v = slot_value<CH>(col, v ^ 1);
  if (cv == other_u || cu == other_v) return false;
  const int lane = lane_id();
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int k = lane + kWave * c;
    col[c] = (k == u) ? cv : ((k == v) ? cu : col[c]);
  }
  return true;
}

template <int CH>
__global__ void __launch_bounds__(256) mutate_kernel(const int32_t* __restrict__ parent_p,
                                                     const int32_t* __restrict__ parent_q,
                                                     const int32_t* __restrict__ parent_badness, int elite_count,
                                                     int quality_count, int n, int population, int active_strength,
                                                     int max_strength, int bad_bias, uint32_t seed_lo,
                                                     uint32_t seed_hi, uint32_t generation,
                                                     int32_t* __restrict__ child_p, int32_t* __restrict__ child_q) {
  const int waves = blockDim.x / kWave;
  const int child = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * waves + threadIdx.x / kWave));
  if (child >= population) return;
  const int lane = lane_id(), m = 2 * n;
  uint32_t a[4];
  philox4x32_10((uint32_t)child, generation, 0u, kTagMutate, seed_lo, seed_hi, a);
  int parent = child;
  if (child >= elite_count) {
    // 160 of 256: from the quality pool, else from all elites
    const int pool = (a[0] >> 24) < 160u ? quality_count : elite_count;
    parent = (int)mulhi32(a[1], (uint32_t)pool);
  }
  int col[CH];
  load_state<CH>(parent_p, parent_q, parent, n, col);
  if (child >= elite_count) {
    // one time in 32 the strength may reach max_strength
    const int span = mulhi32(a[2], 32u) == 0u ? max_strength : active_strength;
    const int strength = 1 + (int)mulhi32(a[3], (uint32_t)span);
    // inverse-CDF table of the parent's badness for the biased first slot
    uint32_t cdf[CH];
    uint32_t total = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int k = lane + kWave * c;
      const uint32_t b = k < m ? (uint32_t)parent_badness[(long long)parent * m + k] : 0u;
      const uint32_t x = wave_scan_dpp(b);
      cdf[c] = total + x;
      total += (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
    }
    for (int step = 0; step < strength; ++step) {
      for (int attempt = 0; attempt < kMutateAttempts; ++attempt) {
        uint32_t b[4];
        philox4x32_10((uint32_t)child, generation, 1u + (uint32_t)(step * kMutateAttempts + attempt), kTagMutate,
                      seed_lo, seed_hi, b);
        int u = (int)mulhi32(b[2], (uint32_t)m);
        if ((int)mulhi32(b[0], 100u) < bad_bias && total > 0u) {
          const uint32_t target = mulhi32(b[1], total);
          int below = 0;
#pragma unroll
          for (int c = 0; c < CH; ++c) below += __popcll(lanes(lane + kWave * c < m) & lanes(cdf[c] <= target));
          u = below < m - 1 ? below : m - 1;
        }
        const int v = (int)mulhi32(b[3], (uint32_t)m);
        if (try_switch<CH>(col, u, v)) break;
      }
    }
  }
  store_state<CH>(child_p, child_q, child, n, col);
}

template <int CH>
__global__ void __launch_bounds__(256) enumerate_kernel(const int32_t* __restrict__ parent_p,
                                                        const int32_t* __restrict__ parent_q, int parent_count,
                                                        const int32_t* __restrict__ descriptors, long long count,
                                                        int n, int32_t* __restrict__ child_p,
                                                        int32_t* __restrict__ child_q, uint8_t* __restrict__ valid,
                                                        int* __restrict__ status) {
  const int waves = blockDim.x / kWave;
  const long long cand = (long long)blockIdx.x * waves + threadIdx.x / kWave;
  if (cand >= count) return;
  const int m = 2 * n, lane = lane_id();
  const int parent = descriptors[3 * cand], u = descriptors[3 * cand + 1], v = descriptors[3 * cand + 2];
  if (parent < 0 || parent >= parent_count || u < 0 || u >= m || v < 0 || v >= m) {
    // out-of-range descriptor: no access, child untouched, error bit
    if (lane == 0) {
      valid[cand] = 0;
      atomicOr(status, 1);
    }
    return;
  }
  int col[CH];
  load_state<CH>(parent_p, parent_q, parent, n, col);
  const bool ok = try_switch<CH>(col, u, v);
  store_state<CH>(child_p, child_q, cand, n, col);
  if (lane == 0) valid[cand] = ok ? 1 : 0;
}

__global__ void __launch_bounds__(256) elite_keys_kernel(const int32_t* __restrict__ badness, long long count, int m,
                                                         int32_t* __restrict__ max_badness,
                                                         long long* __restrict__ sum_squares) {
  const int waves = blockDim.x / kWave;
  const long long s = (long long)blockIdx.x * waves + threadIdx.x / kWave;
  if (s >= count) return;
  const int lane = lane_id();
  int mx = INT32_MIN;
  long long sq = 0;
  for (int k = lane; k < m; k += kWave) {
    const int b = badness[s * m + k];
    mx = b > mx ? b : mx;
    sq += (long long)b * b;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const int o = __shfl_xor(mx, off, kWave);
    mx = o > mx ? o : mx;
    sq += __shfl_xor(sq, off, kWave);
  }
  if (lane == 0) {
    max_badness[s] = mx;
    sum_squares[s] = sq;
  }
}

// Per-device status word of n3_enumerate_switches.
static int* enumerate_status(int* st_out) {
  static int* words[64] = {nullptr};
  int device = 0;
  (void)hipGetDevice(&device);
  if (device < 0 || device >= 64) {
    *st_out = set_error(N3_ENODEV, "device index %d out of range", device);
    return nullptr;
  }
  if (!words[device]) {
    int* w = nullptr;
    *st_out = hip_status(hipMalloc(&w, sizeof(int)), "hipMalloc(status)");
    if (*st_out) return nullptr;
    words[device] = w;
  }
  *st_out = N3_OK;
  return words[device];
}

constexpr int kSearchThreads = 256, kSearchWaves = kSearchThreads / kWave;

}  // namespace no3

using namespace no3;

extern "C" int n3_mutate_population(const int32_t* parent_p, const int32_t* parent_q, const int32_t* parent_badness,
                                    int elite_count, int quality_count, int n, int64_t population, int active_strength,
                                    int max_strength, int bad_bias, uint64_t seed, int generation, int32_t* child_p,
                                    int32_t* child_q, void* stream) {
  if (n < 2 || n > kMaxN) return set_error(N3_EINVAL, "n must be in [2, %d] (got %d)", kMaxN, n);
  if (population < 0 || population > INT32_MAX) return set_error(N3_EINVAL, "population must be in [0, 2^31)");
  if (elite_count < 1 || quality_count < 1 || quality_count > elite_count)
    return set_error(N3_EINVAL, "need 1 <= quality_count <= elite_count");
  if (active_strength < 1 || max_strength < 1 || active_strength > kMaxStrength || max_strength > kMaxStrength)
    return set_error(N3_EINVAL, "strengths must be in [1, %d]", kMaxStrength);
  if (bad_bias < 0 || bad_bias > 100) return set_error(N3_EINVAL, "bad_bias must be between zero and 100");
  if (generation < 0) return set_error(N3_EINVAL, "generation must be non-negative");
  if (population == 0) return N3_OK;
  if (!parent_p || !parent_q || !parent_badness || !child_p || !child_q) return set_error(N3_EINVAL, "null pointer");
  const unsigned blocks = (unsigned)((population + kSearchWaves - 1) / kSearchWaves);
  const int ch = (2 * n + kWave - 1) / kWave;
  void (*kernel)(const int32_t*, const int32_t*, const int32_t*, int, int, int, int, int, int, int, uint32_t,
                 uint32_t, uint32_t, int32_t*, int32_t*);
  switch (ch) {
    case 1: kernel = mutate_kernel<1>; break;
    case 2: kernel = mutate_kernel<2>; break;
    case 3: kernel = mutate_kernel<3>; break;
    case 4: kernel = mutate_kernel<4>; break;
    case 5: kernel = mutate_kernel<5>; break;
    case 6: kernel = mutate_kernel<6>; break;
    case 7: kernel = mutate_kernel<7>; break;
    default: kernel = mutate_kernel<8>; break;
  }
  hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kSearchThreads), 0, as_stream(stream), parent_p, parent_q,
                     parent_badness, elite_count, quality_count, n, (int)population, active_strength, max_strength,
                     bad_bias, (uint32_t)seed, (uint32_t)(seed >> 32), (uint32_t)generation, child_p, child_q);
  return hip_status(hipGetLastError(), "mutate_kernel launch");
}

extern "C" int n3_enumerate_switches(const int32_t* parent_p, const int32_t* parent_q, int parent_count,
                                     const int32_t* descriptors, int64_t count, int n, int32_t* child_p,
                                     int32_t* child_q, uint8_t* valid, void* stream) {
  int st = check_columns_launch_args(n, count);
  if (st) return st;
  if (parent_count < 1) return set_error(N3_EINVAL, "parent_count must be positive");
  if (count == 0) return N3_OK;
  if (!parent_p || !parent_q || !descriptors || !child_p || !child_q || !valid)
    return set_error(N3_EINVAL, "null pointer");
  int* status = enumerate_status(&st);
  if (st) return st;
  hipStream_t s = as_stream(stream);
  st = hip_status(hipMemsetAsync(status, 0, sizeof(int), s), "hipMemsetAsync(status)");
  if (st) return st;
  const unsigned blocks = (unsigned)((count + kSearchWaves - 1) / kSearchWaves);
  const int ch = (2 * n + kWave - 1) / kWave;
  void (*kernel)(const int32_t*, const int32_t*, int, const int32_t*, long long, int, int32_t*, int32_t*, uint8_t*,
                 int*);
  switch (ch) {
    case 1: kernel = enumerate_kernel<1>; break;
    case 2: kernel = enumerate_kernel<2>; break;
    case 3: kernel = enumerate_kernel<3>; break;
    case 4: kernel = enumerate_kernel<4>; break;
    case 5: kernel = enumerate_kernel<5>; break;
    case 6: kernel = enumerate_kernel<6>; break;
    case 7: kernel = enumerate_kernel<7>; break;
    default: kernel = enumerate_kernel<8>; break;
  }
  hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kSearchThreads), 0, s, parent_p, parent_q, parent_count, descriptors,
                     (long long)count, n, child_p, child_q, valid, status);
  st = hip_status(hipGetLastError(), "enumerate_kernel launch");
  if (st) return st;
  int host = 0;
  st = hip_status(hipMemcpyAsync(&host, status, sizeof(int), hipMemcpyDeviceToHost, s), "hipMemcpyAsync(status)");
  if (st) return st;
  st = hip_status(hipStreamSynchronize(s), "hipStreamSynchronize");
  if (st) return st;
  if (host) return set_error(N3_EINVAL, "a descriptor names a parent or slot out of range");
  return N3_OK;
}

extern "C" int n3_elite_keys(const int32_t* badness, int64_t count, int n, int32_t* max_badness,
                             int64_t* sum_squares, void* stream) {
  int st = check_columns_launch_args(n, count);
  if (st) return st;
  if (count == 0) return N3_OK;
  if (!badness || !max_badness || !sum_squares) return set_error(N3_EINVAL, "null pointer");
  const unsigned blocks = (unsigned)((count + kSearchWaves - 1) / kSearchWaves);
  hipLaunchKernelGGL(elite_keys_kernel, dim3(blocks), dim3(kSearchThreads), 0, as_stream(stream), badness,
                     (long long)count, 2 * n, max_badness, reinterpret_cast<long long*>(sum_squares));
  return hip_status(hipGetLastError(), "elite_keys_kernel launch");
}
