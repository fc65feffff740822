// Wavefront-level building blocks shared by the move kernels: DPP
// reductions, uniform reads of per-lane slot arrays, and the
// build/query/clear cycle of a byte-packed direction histogram in LDS.
#pragma once

#include "no3_common.h"

namespace no3 {

// Sum over the 64 lanes with DPP (quad perms, row mirrors, row broadcasts);
// the total is read from lane 63 and returned to every lane as a uniform.
// Unsigned, so that packed 16-bit fields may use the top bit.
__device__ inline uint32_t wave_sum_dpp(uint32_t x) {
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x141, 0xF, 0xF, false);  // row_half_mirror
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x140, 0xF, 0xF, false);  // row_mirror
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);  // row_bcast:15
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);  // row_bcast:31
  return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// Inclusive prefix sum over the 64 lanes in lane order, on the VALU:
// row_shr 1, 2, 4, 8 inside each row of 16 (lanes whose source lies before
// the row start add 0), then row 0's total into row 1 and row 2's into row 3
// (row_bcast:15), then the total of rows 0-1 into rows 2 and 3
// (row_bcast:31).
__device__ inline uint32_t wave_scan_dpp(uint32_t x) {
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, false);  // row_shr:1
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, false);  // row_shr:2
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, false);  // row_shr:4
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, false);  // row_shr:8
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);  // row_bcast:15
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);  // row_bcast:31
  return x;
}

// Value of a per-lane register array at a uniform slot (slot = lane + 64 c).
template <int CH>
__device__ inline int slot_value(const int (&arr)[CH], int slot) {
  const int lane = slot & (kWave - 1), chunk = slot >> 6;
  int r = __builtin_amdgcn_readlane(arr[0], lane);
#pragma unroll
  for (int c = 1; c < CH; ++c) {
    const int t = __builtin_amdgcn_readlane(arr[c], lane);
    r = (chunk == c) ? t : r;
  }
  return r;
}

// Replica slot of wavefront `wave` of walk block `block` (`waves` per
// block).  With a ladder that is a multiple of `waves`, block b holds every
// (ladder / waves)-th temperature of one ladder, so that the hot replicas
// (more acceptances, and each acceptance costs more than a rejection) are
// spread over the blocks and SIMDs instead of filling the last blocks of
// every ladder; otherwise slots are consecutive.  The slot keys the random
// stream, so trajectories do not depend on the mapping.
__device__ inline int walk_slot(int block, int wave, int waves, int ladder) {
  if (ladder >= waves && ladder % waves == 0) {
    const int per = ladder / waves;
    return (block / per) * ladder + wave * per + block % per;
  }
  return block * waves + wave;
}

// Entry `index` (uniform, >= 0) of a small per-replica table `global` whose
// first 128 entries the wavefront holds in registers, entry lane + 64 j in
// register j (lo: j = 0, hi: j = 1); later entries are read from memory.
__device__ inline uint32_t register_table(uint32_t lo, uint32_t hi, const uint32_t* global, int index) {
  if (index < kWave) return (uint32_t)__builtin_amdgcn_readlane((int)lo, index);
  if (index < 2 * kWave) return (uint32_t)__builtin_amdgcn_readlane((int)hi, index - kWave);
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)global[index]);
}

__device__ inline unsigned long long register_table64(unsigned long long lo, unsigned long long hi,
                                                      const unsigned long long* global, int index) {
  auto read = [](unsigned long long x, int l) {
    return ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(x >> 32), l) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((int)x, l);
  };
  if (index < kWave) return read(lo, index);
  if (index < 2 * kWave) return read(hi, index - kWave);
  const unsigned long long x = global[index];
  return ((unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((int)(x >> 32)) << 32) |
         (uint32_t)__builtin_amdgcn_readfirstlane((int)x);
}

// Histogram of the directions from `anchor` (a key) to the points keys[c] of
// the lanes with include[c]; afterwards cnt[c] is the number of included
// points on the line (anchor, keys[c]) and id[c] their direction id.  Every
// lane executes every instruction (masked lanes count into the trash word),
// so the LDS reads of all chunks are in flight together.  Every keys[c] must
// be a valid point key.  The histogram is left cleared.
template <int CH>
__device__ inline void direction_counts(const uint16_t* table, uint32_t* hist, int trash, int offset, int anchor,
                                        const int (&keys)[CH], const bool (&include)[CH], int (&cnt)[CH],
                                        int (&id)[CH]) {
  uint32_t word[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) id[c] = table[keys[c] - anchor + offset];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    word[c] = include[c] ? ((uint32_t)id[c] >> 2) : (uint32_t)trash;
    atomicAdd(&hist[word[c]], byte_one((uint32_t)id[c]));
  }
  wave_sync();
#pragma unroll
  for (int c = 0; c < CH; ++c) cnt[c] = (int)byte_counter(hist[word[c]], (uint32_t)id[c]);
  wave_sync();
#pragma unroll
  for (int c = 0; c < CH; ++c) hist[word[c]] = 0u;
  wave_sync();
}

// Same as above when the caller already knows the ids.
template <int CH>
__device__ inline void direction_counts_known(uint32_t* hist, int trash, const int (&id)[CH], const bool (&include)[CH],
                                              int (&cnt)[CH]) {
  uint32_t word[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    word[c] = include[c] ? ((uint32_t)id[c] >> 2) : (uint32_t)trash;
    atomicAdd(&hist[word[c]], byte_one((uint32_t)id[c]));
  }
  wave_sync();
#pragma unroll
  for (int c = 0; c < CH; ++c) cnt[c] = (int)byte_counter(hist[word[c]], (uint32_t)id[c]);
  wave_sync();
#pragma unroll
  for (int c = 0; c < CH; ++c) hist[word[c]] = 0u;
  wave_sync();
}

// As direction_counts_known for a histogram of the counted ids only
// ([0, counted), see DirTableInfo) with a trash word per lane and chunk
// (word trash + 64 c): a lane that is left out, or seen in any other
// direction (alone on its line), counts in its trash word instead, so the
// latter read back 1, and every lane executes every LDS step.
template <int CH>
__device__ inline void direction_counts_counted(uint32_t* hist, int counted, int trash, const int (&id)[CH],
                                                const bool (&include)[CH], int (&cnt)[CH]) {
  uint32_t word[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    word[c] = ((int)include[c] & (int)((uint32_t)id[c] < (uint32_t)counted)) ? ((uint32_t)id[c] >> 2)
                                                                             : (uint32_t)(trash + kWave * c);
    atomicAdd(&hist[word[c]], byte_one((uint32_t)id[c]));
  }
  wave_sync();
#pragma unroll
  for (int c = 0; c < CH; ++c) cnt[c] = (int)byte_counter(hist[word[c]], (uint32_t)id[c]);
  wave_sync();
#pragma unroll
  for (int c = 0; c < CH; ++c) hist[word[c]] = 0u;
  wave_sync();
}

__device__ inline void stage_table(const uint16_t* __restrict__ table_g, uint16_t* table, int entries) {
  const uint2* src = reinterpret_cast<const uint2*>(table_g);
  uint2* dst = reinterpret_cast<uint2*>(table);
  const int words = (entries * 2 + 7) / 8;
  for (int i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
}

}  // namespace no3
