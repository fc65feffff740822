// Shared host/device definitions of the no-three-in-line gfx950 engine.
//
// Geometry conventions (the reference's, accelerated.py:20-24, gpu.py:20-21):
// slot s of a state holds the point (row = s >> 1, column = P[row] if s is
// even, Q[row] if odd); a state has m = 2n slots.
//
// Direction table.  Two points a, b differ by (dr, dc) with |dr|, |dc| <= n-1.
// A point is encoded as key = row * W + column with W = 2n - 1, so
// key_b - key_a = dr * W + dc identifies (dr, dc) uniquely, and
// dir_table[key_b - key_a + (n-1) * W + (n-1)] is the id of the primitive
// undirected direction of (dr, dc) (verifier_b.py:8-19: divide by the gcd,
// flip the sign so that dr > 0 or dr == 0 < dc).  Ids are dense in
// [0, D(n)); the zero vector (coincident points) has id D(n).  The first
// ids go to the directions (p, q), p > 0, with max(p, |q|) <= (n - 1) / 2:
// a line of any other direction meets at most two cells of the grid (a row
// holds only its own two slots), so a point seen in such a direction from
// an anchor is alone on its line, and histograms may leave it out.  Two points b,
// c are collinear with a iff their ids seen from a are equal, so the number
// of collinear triples through a is sum_d C(cnt_d, 2) over the histogram of
// ids seen from a (the counting form of verifier_b.find_collinear_triple).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace no3 {

constexpr int kWave = 64;
constexpr int kMaxN = 256;          // determinant paths: m <= 512
constexpr int kMaxNTable = 100;     // direction-table paths: W*W*2 bytes of LDS
constexpr int kSentinel32 = 1 << 30;           // gpu_delta.py:30
constexpr long long kSentinel64 = 1LL << 60;   // accelerated.py:58,140

// Philox stream tags (counter word 3).
constexpr uint32_t kTagGen = 0x47454E53u;   // state generator
constexpr uint32_t kTagMove = 0x4D4F5645u;  // one tempering proposal
constexpr uint32_t kTagBias = 0x42494153u;  // badness-biased slot draw
constexpr uint32_t kTagExch = 0x45584348u;  // replica exchange

// Trash words at the end of a histogram of the counted ids: one per lane
// and chunk of 64 slots (n <= 100: four chunks).
constexpr int kCountedTrashWords = 4 * kWave;

struct DirTableInfo {
  int n;
  int width;        // W = 2n - 1
  int offset;       // (n-1) * W + (n-1)
  int entries;      // W * W
  int directions;   // D(n); the zero vector has id D(n)
  int hist_words;   // 32-bit words of one byte-packed histogram: the D+1
                    // byte counters, then 64 spare "trash" words, one per
                    // lane, that masked lanes update instead of branching
                    // around the update (one word per lane: no conflicts)
  int counted;      // ids [0, counted): the directions whose lines can meet
                    // three cells of the grid (see build_dirtable)
  int counted_words;  // 32-bit words of a histogram of those ids only and
                      // kCountedTrashWords trash words, a multiple of 16
};

__host__ __device__ inline int trash_word(const DirTableInfo& info, int lane) {
  return info.hist_words - kWave + lane;
}

// Byte-packed histograms keep direction id's 8-bit counter in word id >> 2,
// byte id & 3: its bit offset, and the word increment that adds one to it.
// The offset is masked so that no shift count reaches 32.
__host__ __device__ inline uint32_t byte_shift(uint32_t id) { return (id & 3u) << 3; }
__host__ __device__ inline uint32_t byte_one(uint32_t id) { return 1u << byte_shift(id); }
// Counter id's byte of a histogram word.  v_bfe_u32 and its intrinsic read
// the low five bits of the offset (LLVM's instcombine masks constant offsets
// the same way), so id << 3 needs no mask.
__device__ inline uint32_t byte_counter(uint32_t word, uint32_t id) { return __builtin_amdgcn_ubfe(word, id << 3, 8); }

__host__ __device__ inline uint32_t mulhi32(uint32_t a, uint32_t b) {
  return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__host__ __device__ inline void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                              uint32_t k1, uint32_t out[4]) {
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// Orders LDS traffic of the lanes of one wavefront (LDS executes a wave's
// instructions in issue order; this keeps the compiler from moving them).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// v_writelane_b32: `old` with lane `lane` (uniform, 0..63) replaced by the
// uniform `value`.  Clang has no builtin for it; this binds the LLVM
// intrinsic, so the compiler still tracks the instruction's hazards.
extern "C" __device__ int llvm_amdgcn_writelane(int value, int lane, int old) __asm("llvm.amdgcn.writelane.i32");
__device__ inline int write_lane(int value, int lane, int old) { return llvm_amdgcn_writelane(value, lane, old); }

// Lane mask of `pred`.  __ballot(int) of a compound predicate makes the
// compiler turn the predicate into a 0/1 vector and compare it again (two
// more VALU instructions per call); masks of the parts, AND-ed on the scalar
// unit, avoid that.
__device__ inline uint64_t lanes(bool pred) { return __builtin_amdgcn_ballot_w64(pred); }

__device__ inline int wave_sum(int x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, kWave);
  return x;
}

__device__ inline int lane_id() { return (int)__lane_id(); }

}  // namespace no3
