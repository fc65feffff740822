// Move-proposal network: 8 features -> 128 hidden (ReLU) -> 1 logit, BF16
// parameters, FP32 accumulation, hand-written MFMA kernels for inference,
// training gradients and the tabulated proposal weights.
//
// Parameters are one bf16 vector of kParams entries:
//   W1[128][8] | b1[128] | w2[128] | b2
// Inference on a tile of 32 rows is H^T = W1 X^T with four
// v_mfma_f32_32x32x8_bf16 (M = hidden, N = rows, K = features), so each lane
// ends up holding one row's pre-activations for 16 hidden units per tile and
// the output layer is a per-lane dot product plus one cross-half exchange.
// The weight gradient dW1 = dH X is a second MFMA product with the rows as K.
#include <mutex>

#include "mlp_common.h"
#include "runtime.h"
#include "wave_ops.h"

namespace no3 {

// Features of a candidate point with collinear-pair count b at temperature
// exp(log_t) in a state of energy e: b in four encodings, the temperature,
// the energy in two encodings and one interaction term.
__host__ __device__ inline void point_features(int b, float log_t, int e, float f[kIn]) {
  const float fb = (float)b;
  const float fe = (float)e;
  f[0] = fb * 0.125f;
  f[1] = sqrtf(fb) * (1.0f / 3.0f);
  f[2] = logf(1.0f + fb) * (1.0f / 3.0f);
  f[3] = b > 0 ? 1.0f : 0.0f;
  f[4] = log_t;
  f[5] = fe * (1.0f / 150.0f);
  f[6] = sqrtf(fe) * (1.0f / 12.0f);
  f[7] = fb * 0.125f * log_t;
}

// Pre-activations of the 32 rows whose features this lane supplies:
// lane l provides row (l % 32), features 4 (l / 32) .. +3.
__device__ inline void hidden_tile(const uint16_t* __restrict__ params, bf16x4_t x, floatx16_t acc[4]) {
  const int lane = lane_id();
  const int j = lane & 31, k0 = 4 * (lane >> 5);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    bf16x4_t w;
#pragma unroll
    for (int i = 0; i < 4; ++i) w[i] = (short)params[kW1 + (32 * t + j) * kIn + k0 + i];
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[t][r] = 0.0f;
    acc[t] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(w, x, acc[t], 0, 0, 0);
  }
}

// logit of this lane's row (valid in both halves after the exchange)
__device__ inline float output_row(const float* b1, const float* w2, float b2, const floatx16_t acc[4]) {
  const int half = lane_id() >> 5;
  float s = 0.0f;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int h = hidden_of(t, r, half);
      s += w2[h] * fmaxf(acc[t][r] + b1[h], 0.0f);
    }
  return s + __shfl_xor(s, 32, kWave) + b2;
}

__device__ inline void stage_layer_params(const uint16_t* __restrict__ params, float* b1, float* w2) {
  for (int i = threadIdx.x; i < kHidden; i += blockDim.x) {
    b1[i] = bf16_to_float(params[kB1 + i]);
    w2[i] = bf16_to_float(params[kW2 + i]);
  }
}

__device__ inline bf16x4_t load_row_features(const uint16_t* __restrict__ x, long long row, long long rows) {
  const int k0 = 4 * (lane_id() >> 5);
  bf16x4_t v = {0, 0, 0, 0};
  if (row < rows) {
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = (short)x[row * kIn + k0 + i];
  }
  return v;
}

__global__ void __launch_bounds__(256) mlp_forward_kernel(const uint16_t* __restrict__ params,
                                                          const uint16_t* __restrict__ x, long long rows,
                                                          float* __restrict__ out) {
  __shared__ float b1[kHidden], w2[kHidden];
  stage_layer_params(params, b1, w2);
  __syncthreads();
  const float b2 = bf16_to_float(params[kB2]);
  const int lane = lane_id(), waves = blockDim.x / kWave;
  const long long tiles = (rows + 31) / 32;
  for (long long tile = (long long)blockIdx.x * waves + threadIdx.x / kWave; tile < tiles;
       tile += (long long)gridDim.x * waves) {
    const long long row = tile * 32 + (lane & 31);
    floatx16_t acc[4];
    hidden_tile(params, load_row_features(x, row, rows), acc);
    const float logit = output_row(b1, w2, b2, acc);
    if (lane < 32 && row < rows) out[row] = logit;
  }
}

// Pair-logit BCE gradient: tuple i has rows xu[i], xv[i] and label y[i];
// z = g(xu) + g(xv), loss = mean BCE(sigmoid(z), y).  Every wavefront takes
// tiles of 32 tuples in a grid stride and keeps its gradient in registers;
// the block adds its wavefronts' gradients (and summed losses) in LDS in
// wavefront order and stores them as column `block` of `partial`
// ([kParams + 1][kGradMaxBlocks]); grad_reduce_kernel then adds each
// column's blocks in block order, so the gradient is deterministic (no
// floating-point atomics).
constexpr int kGradWaves = 4;         // wavefronts per block
constexpr int kGradMaxBlocks = 128;   // columns of the partial-gradient buffer
constexpr int kPartial = kParams + 1; // gradient, then the summed loss

__global__ void __launch_bounds__(256) mlp_grad_kernel(const uint16_t* __restrict__ params,
                                                       const uint16_t* __restrict__ xu,
                                                       const uint16_t* __restrict__ xv, const float* __restrict__ y,
                                                       long long pairs, float* __restrict__ partial) {
  __shared__ float b1[kHidden], w2[kHidden];
  __shared__ uint16_t dh_tile[kGradWaves][32 * 32];  // one tile of dH, [hidden 32][rows 32]
  __shared__ uint16_t x_tile[kGradWaves][32 * kIn];  // the tile's features, [row][feature]
  __shared__ float wave_grad[kGradWaves][kPartial];   // the wavefronts' gradients
  stage_layer_params(params, b1, w2);
  __syncthreads();
  const float b2 = bf16_to_float(params[kB2]);
  const int wave = threadIdx.x / kWave, lane = lane_id(), half = lane >> 5, j = lane & 31;
  uint16_t* dh_t = dh_tile[wave];
  uint16_t* xt = x_tile[wave];
  const long long my_wave = (long long)blockIdx.x * kGradWaves + wave;
  const long long all_waves = (long long)gridDim.x * kGradWaves;
  const float inv_pairs = 1.0f / (float)pairs;
  floatx16_t dw1[4];  // dW1 tiles: [hidden 32][feature 32 (8 used)]
  float dw2[4][16], db1[4][16];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      dw1[t][r] = 0.0f;
      dw2[t][r] = 0.0f;
      db1[t][r] = 0.0f;
    }
  float db2 = 0.0f, loss_sum = 0.0f;
  const long long tiles = (pairs + 31) / 32;
  for (long long tile = my_wave; tile < tiles; tile += all_waves) {
    const long long pair = tile * 32 + j;
    const bool live = pair < pairs;
    const bf16x4_t x_u = load_row_features(xu, pair, pairs);
    const bf16x4_t x_v = load_row_features(xv, pair, pairs);
    float z;
    {
      floatx16_t tmp[4];
      hidden_tile(params, x_u, tmp);
      z = output_row(b1, w2, b2, tmp);
      hidden_tile(params, x_v, tmp);
      z += output_row(b1, w2, b2, tmp) - b2;
    }
    // z = g(u) + g(v) counts b2 twice; the pair logit keeps one bias
    const float label = live ? y[pair] : 0.0f;
    const float p = 1.0f / (1.0f + expf(-z));
    const float dz = live ? (p - label) * inv_pairs : 0.0f;
    if (live && half == 0) loss_sum += fmaxf(z, 0.0f) - z * label + logf(1.0f + expf(-fabsf(z)));
    if (half == 0) db2 += dz;
    // both towers: dh = dz * w2 * [pre > 0]; accumulate dw2, db1, and dW1
    for (int side = 0; side < 2; ++side) {
      const bf16x4_t xs = side == 0 ? x_u : x_v;
      // the MFMA B operand (rows x features) of the four K=8 steps, from
      // the tile's features staged in LDS: lane l supplies rows
      // 8 s + 4 half .. +3 of feature l % 32 (zero past the 8 features)
#pragma unroll
      for (int i = 0; i < 4; ++i) xt[j * kIn + 4 * half + i] = (uint16_t)xs[i];
      wave_sync();
      bf16x4_t bv[4];
#pragma unroll
      for (int st = 0; st < 4; ++st)
#pragma unroll
        for (int i = 0; i < 4; ++i) bv[st][i] = j < kIn ? (short)xt[(8 * st + 4 * half + i) * kIn + j] : (short)0;
      floatx16_t a[4];
      hidden_tile(params, xs, a);  // recomputed: cheaper than keeping both sides live
#pragma unroll
      for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int h = hidden_of(t, r, half);
          const float pre = a[t][r] + b1[h];
          const float act = fmaxf(pre, 0.0f);
          const float dh = pre > 0.0f ? dz * w2[h] : 0.0f;
          dw2[t][r] += dz * act;
          db1[t][r] += dh;
          dh_t[(8 * (r >> 2) + 4 * half + (r & 3)) * 32 + j] = float_to_bf16(dh);
        }
        wave_sync();
        // dW1 tile t += dH_t (32 hidden x 32 rows) X (32 rows x 8 features):
        // lane l supplies A[hidden l % 32][rows 8 s + 4 half .. +3]
#pragma unroll
        for (int st = 0; st < 4; ++st) {
          bf16x4_t av4;
#pragma unroll
          for (int i = 0; i < 4; ++i) av4[i] = (short)dh_t[j * 32 + 8 * st + 4 * half + i];
          dw1[t] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av4, bv[st], dw1[t], 0, 0, 0);
        }
        wave_sync();
      }
    }
  }
  float* out = wave_grad[wave];
  // dW1: lane l holds dW1[hidden 32t + row(r, half)][feature l % 32]
  if (j < kIn) {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int r = 0; r < 16; ++r) out[kW1 + hidden_of(t, r, half) * kIn + j] = dw1[t][r];
  }
  // dw2, db1: sum over the 32 rows (lanes) of each half
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      float sa = dw2[t][r], sb = db1[t][r];
#pragma unroll
      for (int off = 16; off >= 1; off >>= 1) {
        sa += __shfl_xor(sa, off, kWave);
        sb += __shfl_xor(sb, off, kWave);
      }
      if (j == 0) {
        out[kW2 + hidden_of(t, r, half)] = sa;
        out[kB1 + hidden_of(t, r, half)] = sb;
      }
    }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    db2 += __shfl_xor(db2, off, kWave);
    loss_sum += __shfl_xor(loss_sum, off, kWave);
  }
  if (lane == 0) {
    out[kB2] = db2;
    out[kParams] = loss_sum;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < kPartial; c += blockDim.x) {
    float sum = wave_grad[0][c];
#pragma unroll
    for (int w = 1; w < kGradWaves; ++w) sum += wave_grad[w][c];
    partial[(long long)c * kGradMaxBlocks + blockIdx.x] = sum;
  }
}

// Sum over the 64 lanes by the DPP pattern of wave_sum_dpp (fixed order).
__device__ inline float wave_sum_f32(float x) {
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xF, 0xF, false));
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x4E, 0xF, 0xF, false));
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x141, 0xF, 0xF, false));
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x140, 0xF, 0xF, false));
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x142, 0xA, 0xF, false));
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x143, 0xC, 0xF, false));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
}

// One wavefront per parameter (and one for the loss): grad[c] += the sum of
// column c's `blocks` entries, lane l adding entries l, l + 64, ... in order.
__global__ void __launch_bounds__(256) grad_reduce_kernel(const float* __restrict__ partial, int blocks,
                                                          float* __restrict__ grad, float* __restrict__ loss) {
  const int c = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (c >= kPartial) return;
  const int lane = lane_id();
  float part = 0.0f;
  for (int b = lane; b < blocks; b += kWave) part += partial[(long long)c * kGradMaxBlocks + b];
  const float sum = wave_sum_f32(part);
  if (lane == 0) {
    if (c < kParams)
      grad[c] += sum;
    else
      loss[0] += sum;
  }
}

// Per-device partial-gradient buffer ([kPartial][kGradMaxBlocks] floats).
static float* grad_partials(int* st_out) {
  static float* buffers[64] = {nullptr};
  int device = 0;
  (void)hipGetDevice(&device);
  if (device < 0 || device >= 64) {
    *st_out = set_error(N3_ENODEV, "device index %d out of range", device);
    return nullptr;
  }
  if (!buffers[device]) {
    float* b = nullptr;
    *st_out = hip_status(hipMalloc(&b, sizeof(float) * kGradMaxBlocks * kPartial), "hipMalloc(partials)");
    if (*st_out) return nullptr;
    buffers[device] = b;
  }
  *st_out = N3_OK;
  return buffers[device];
}

// Both gradient entry points: the partial sums of at most kGradMaxBlocks
// blocks (their grid strides over any number of tiles), then the ordered
// reduction into grad and loss.  Every column the reduction reads is written
// by this call's gradient kernel first, so the buffer needs no clearing.
static int launch_grad(const uint16_t* params, const uint16_t* xu, const uint16_t* xv, const float* y,
                       int64_t pairs, float* grad, float* loss, float* partial, hipStream_t s) {
  const long long tiles = (pairs + 31) / 32;
  long long blocks = (tiles + kGradWaves - 1) / kGradWaves;
  if (blocks > kGradMaxBlocks) blocks = kGradMaxBlocks;
  hipLaunchKernelGGL(mlp_grad_kernel, dim3((unsigned)blocks), dim3(kGradWaves * kWave), 0, s, params, xu, xv, y,
                     (long long)pairs, partial);
  int st = hip_status(hipGetLastError(), "mlp_grad_kernel launch");
  if (st) return st;
  const int reduce_waves = 256 / kWave;
  hipLaunchKernelGGL(grad_reduce_kernel, dim3((kPartial + reduce_waves - 1) / reduce_waves), dim3(256), 0, s, partial,
                     (int)blocks, grad, loss);
  return hip_status(hipGetLastError(), "grad_reduce_kernel launch");
}

// The partial-sum buffer of n3_mlp_grad_noalloc, for callers that cannot
// allocate: a file-scope device array of kPartial x kGradMaxBlocks floats
// (641 KiB), which every GPU the library runs on holds until the process
// ends.  Each GPU has its own copy at its own address; in host code the name
// `scratch` is only a stand-in, so the host passes the address that
// hipGetSymbolAddress reports for the current device (kept per device) and
// never the name itself.  All calls on one GPU share that copy.
static __device__ __align__(256) float scratch[kPartial * kGradMaxBlocks];

static std::mutex scratch_mutex;
static float* scratch_address[64] = {nullptr};

// Looks up (once per device, with `device` current) and returns the device
// address of `scratch`.  Called by n3_device_init, and by the first
// n3_mlp_grad_noalloc on a device that was not initialised that way.
int mlp_scratch_setup(int device, float** address) {
  if (device < 0 || device >= 64) return set_error(N3_ENODEV, "device index %d out of range", device);
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_address[device]) {
    void* ptr = nullptr;
    const int st = hip_status(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (st) return st;
    if (!ptr) return set_error(N3_ENODEV, "no device address for the gradient scratch on device %d", device);
    scratch_address[device] = static_cast<float*>(ptr);
  }
  if (address) *address = scratch_address[device];
  return N3_OK;
}

__global__ void adam_kernel(float* __restrict__ master, uint16_t* __restrict__ params, const float* __restrict__ grad,
                            float* __restrict__ m, float* __restrict__ v, int step, float lr, float beta1, float beta2,
                            float eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kParams) return;
  const float g = grad[i];
  const float mi = beta1 * m[i] + (1.0f - beta1) * g;
  const float vi = beta2 * v[i] + (1.0f - beta2) * g * g;
  m[i] = mi;
  v[i] = vi;
  const float mhat = mi / (1.0f - powf(beta1, (float)step));
  const float vhat = vi / (1.0f - powf(beta2, (float)step));
  const float w = master[i] - lr * mhat / (sqrtf(vhat) + eps);
  master[i] = w;
  // W1 and b1 go onto the fixed-point grid of the point network's exact
  // inference (multiples of 1/128, |k| <= 255); w2 and b2 are plain bf16
  params[i] = i < kW2 ? grid_bf16(w) : float_to_bf16(w);
}

// Proposal weights for every (temperature t, energy e, count b):
// table[(t * (max_energy + 1) + e) * 64 + b] = clamp(round(128 exp(clamp(g,
// -6, 5.5))), 1, 2^15) with g the network's logit (2^15 keeps the learned
// walk's exact acceptance test within 128-bit products).
__global__ void __launch_bounds__(256) weight_table_kernel(const uint16_t* __restrict__ params,
                                                           const float* __restrict__ log_temps, int temps,
                                                           int max_energy, uint32_t* __restrict__ table) {
  __shared__ float b1[kHidden], w2[kHidden];
  stage_layer_params(params, b1, w2);
  __syncthreads();
  const float b2 = bf16_to_float(params[kB2]);
  const int lane = lane_id(), waves = blockDim.x / kWave;
  const long long rows = (long long)temps * (max_energy + 1) * 64;
  const long long tiles = rows / 32;
  for (long long tile = (long long)blockIdx.x * waves + threadIdx.x / kWave; tile < tiles;
       tile += (long long)gridDim.x * waves) {
    const long long row = tile * 32 + (lane & 31);
    const int b = (int)(row & 63);
    const long long te = row >> 6;
    const int e = (int)(te % (max_energy + 1));
    const int t = (int)(te / (max_energy + 1));
    float f[kIn];
    point_features(b, log_temps[t], e, f);
    const int k0 = 4 * (lane >> 5);
    bf16x4_t x;
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] = (short)float_to_bf16(f[k0 + i]);
    floatx16_t acc[4];
    hidden_tile(params, x, acc);
    const float g = fminf(fmaxf(output_row(b1, w2, b2, acc), -6.0f), 5.5f);
    if (lane < 32) {
      const float w = rintf(128.0f * expf(g));
      table[row] = (uint32_t)fminf(fmaxf(w, 1.0f), 32768.0f);
    }
  }
}

}  // namespace no3

using namespace no3;

extern "C" int n3_mlp_params(void) { return kParams; }

extern "C" int n3_mlp_features(int count, float log_temperature, int energy, uint16_t* host_out) {
  if (!host_out || count < 0) return set_error(N3_EINVAL, "bad feature arguments");
  float f[kIn];
  point_features(count, log_temperature, energy, f);
  for (int i = 0; i < kIn; ++i) host_out[i] = float_to_bf16(f[i]);
  return N3_OK;
}

extern "C" int n3_mlp_forward(const uint16_t* params, const uint16_t* x, int64_t rows, float* out, void* stream) {
  if (!params || (!x && rows) || (!out && rows) || rows < 0) return set_error(N3_EINVAL, "bad forward arguments");
  if (rows == 0) return N3_OK;
  long long blocks = (rows + 127) / 128;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), params, x,
                     (long long)rows, out);
  return hip_status(hipGetLastError(), "mlp_forward_kernel launch");
}

extern "C" int n3_mlp_grad(const uint16_t* params, const uint16_t* xu, const uint16_t* xv, const float* y,
                           int64_t pairs, float* grad, float* loss, void* stream) {
  if (!params || !xu || !xv || !y || !grad || !loss || pairs < 1) return set_error(N3_EINVAL, "bad grad arguments");
  int st = N3_OK;
  float* partial = grad_partials(&st);
  if (st) return st;
  return launch_grad(params, xu, xv, y, pairs, grad, loss, partial, as_stream(stream));
}

extern "C" int n3_mlp_grad_noalloc(const uint16_t* params, const uint16_t* xu, const uint16_t* xv, const float* y,
                                   int64_t pairs, float* grad, float* loss, void* stream) {
  if (!params || !xu || !xv || !y || !grad || !loss || pairs < 1) return set_error(N3_EINVAL, "bad grad arguments");
  int device = 0;
  int st = hip_status(hipGetDevice(&device), "hipGetDevice");
  if (st) return st;
  float* partial = nullptr;
  st = mlp_scratch_setup(device, &partial);
  if (st) return st;
  return launch_grad(params, xu, xv, y, pairs, grad, loss, partial, as_stream(stream));
}

extern "C" int n3_mlp_adam(float* master, uint16_t* params, const float* grad, float* m, float* v, int step, float lr,
                           float beta1, float beta2, float eps, void* stream) {
  if (!master || !params || !grad || !m || !v || step < 1) return set_error(N3_EINVAL, "bad adam arguments");
  hipLaunchKernelGGL(adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, as_stream(stream), master, params, grad, m,
                     v, step, lr, beta1, beta2, eps);
  return hip_status(hipGetLastError(), "adam_kernel launch");
}

extern "C" int n3_mlp_weight_table(const uint16_t* params, const float* log_temps, int temps, int max_energy,
                                   uint32_t* table, void* stream) {
  if (!params || !log_temps || !table || temps < 1 || max_energy < 0)
    return set_error(N3_EINVAL, "bad weight-table arguments");
  const long long tiles = (long long)temps * (max_energy + 1) * 64 / 32;
  long long blocks = (tiles + 3) / 4;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(weight_table_kernel, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), params, log_temps,
                     temps, max_energy, table);
  return hip_status(hipGetLastError(), "weight_table_kernel launch");
}
