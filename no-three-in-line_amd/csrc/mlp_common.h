// Shape, parameter layout and bf16 helpers of the move-proposal network,
// shared by the network kernels (mlp.hip) and the learned walks
// (learned.hip).
//
// Parameters are one bf16 vector of kParams entries:
//   W1[128][8] | b1[128] | w2[128] | b2
#pragma once

#include "no3_common.h"

namespace no3 {

constexpr int kIn = 8;
constexpr int kHidden = 128;
constexpr int kW1 = 0;
constexpr int kB1 = kHidden * kIn;
constexpr int kW2 = kB1 + kHidden;
constexpr int kB2 = kW2 + kHidden;
constexpr int kParams = kB2 + 1;  // 1281

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));

__host__ __device__ inline float bf16_to_float(uint16_t h) {
  union {
    uint32_t u;
    float f;
  } v;
  v.u = (uint32_t)h << 16;
  return v.f;
}

// round to nearest even (finite inputs)
__host__ __device__ inline uint16_t float_to_bf16(float f) {
  union {
    uint32_t u;
    float f;
  } v;
  v.f = f;
  const uint32_t lsb = (v.u >> 16) & 1u;
  return (uint16_t)((v.u + 0x7FFFu + lsb) >> 16);
}

// Hidden unit held by accumulator element r of hidden block t in lane half
// `half` for the 32x32 MFMA output layout (row = 8 (r/4) + 4 half + r%4).
__host__ __device__ inline int hidden_of(int t, int r, int half) {
  return 32 * t + 8 * (r >> 2) + 4 * half + (r & 3);
}

// ---- the point network's exact inference --------------------------------
//
// Point features are fixed-point codes k / 128 with |k| <= 255, so every
// feature is exact in bf16.  The trained W1 and b1 are kept on the same grid
// (multiples of 1/128 below 2 in magnitude: the Adam step rounds them onto
// it), so every product W1 x is a multiple of 2^-14 and every partial sum of
// a hidden pre-activation is a multiple of 2^-14 below 2^5 -- exact in fp32,
// whatever order the MFMA adds in.  The output layer is then a fixed-order
// chain of fp32 fused multiply-adds per lane half (hidden units in
// hidden_of order) and one add of the two halves, which a CPU replays bit
// for bit; the logit g picks the integer weight kPointExp[floor(64 g)].
constexpr int kGridScale = 128;
constexpr int kGridMax = 255;
constexpr int kPointQMin = -384;  // g clamped to [-6, 5.5] in steps of 1/64
constexpr int kPointQMax = 352;
constexpr int kPointExpEntries = kPointQMax - kPointQMin + 1;  // 737

// Layout of the walk's float constants: b1 and w2 reordered so that lane
// half h of hidden block t reads its 16 values at [t][h][0..15], then b2.
constexpr int kPointB1 = 0;
constexpr int kPointW2 = kHidden * 1;
constexpr int kPointB2 = kHidden * 2;
constexpr int kPointConsts = kHidden * 2 + 4;  // 260 floats (b2 and padding)

// Feature codes of a point with collinear-pair count b in row `row`,
// column `col`, for a replica whose temperature code is lt_q =
// round(32 ln T) and a state of energy e:
//   b / 32 (b <= 63), min(b, 7) / 4, [b > 0], row / 128, column / 128,
//   ln T / 4, E / 256 (E <= 510), E / 64 (E <= 63).
__host__ __device__ inline void point_codes(int b, int row, int col, int lt_q, int e, int code[kIn]) {
  const int bb = b < 63 ? b : 63, b7 = b < 7 ? b : 7;
  const int e6 = (e < 510 ? e : 510) >> 1, e7 = 2 * (e < 63 ? e : 63);
  code[0] = 4 * bb;
  code[1] = 32 * b7;
  code[2] = b > 0 ? kGridScale : 0;
  code[3] = row;
  code[4] = col;
  code[5] = lt_q;
  code[6] = e6;
  code[7] = e7;
}

// The grid value nearest to w (ties to even), as bf16 bits: k / 128 with k
// clamped to [-255, 255].
__host__ __device__ inline uint16_t grid_bf16(float w) {
  float k = rintf(w * (float)kGridScale);
  k = k < -(float)kGridMax ? -(float)kGridMax : (k > (float)kGridMax ? (float)kGridMax : k);
  return float_to_bf16(k * (1.0f / kGridScale));
}

// bf16 bits of code / 128 (exact for |code| <= 255).
__host__ __device__ inline uint16_t code_bf16(int code) {
  union {
    uint32_t u;
    float f;
  } v;
  v.f = (float)code * (1.0f / kGridScale);
  return (uint16_t)(v.u >> 16);
}

}  // namespace no3
