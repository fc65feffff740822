// Host runtime of the engine: error reporting, device checks, and the
// per-(device, n) direction tables kept resident in HBM.
#include "runtime.h"

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

namespace no3 {

static thread_local char g_error[512] = "";

int set_error(int status, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
  return status;
}

int hip_status(hipError_t e, const char* what) {
  if (e == hipSuccess) return N3_OK;
  return set_error(N3_EHIP - (int)e, "%s: %s", what, hipGetErrorString(e));
}

// verifier_b.py:8-19 without the gcd per query: every primitive undirected
// direction with |dr|, |dc| <= n-1 gets a dense id, in row-major order of
// (dr, dc) with dr >= 0, first those whose lines can meet three cells of the
// grid (dr > 0 and 2 max(dr, |dc|) <= n - 1), then the others.
int build_dirtable(int n, std::vector<uint16_t>& table, DirTableInfo& info) {
  if (n < 2 || n > kMaxNTable) return set_error(N3_EINVAL, "direction table needs 2 <= n <= %d (got %d)", kMaxNTable, n);
  const int w = 2 * n - 1, span = n - 1;
  std::vector<int> id((size_t)w * w, -1);
  int next = 0, counted = 0;
  for (int pass = 0; pass < 2; ++pass) {
    for (int dr = 0; dr <= span; ++dr) {
      for (int dc = -span; dc <= span; ++dc) {
        if (dr == 0 && dc <= 0) continue;
        if (std::gcd(dr, dc < 0 ? -dc : dc) != 1) continue;
        const bool three = dr > 0 && 2 * std::max(dr, dc < 0 ? -dc : dc) <= n - 1;
        if (three == (pass == 0)) id[(size_t)(dr + span) * w + (dc + span)] = next++;
      }
    }
    if (pass == 0) counted = next;
  }
  const int zero_id = next;
  table.assign((size_t)w * w, 0);
  for (int dr = -span; dr <= span; ++dr) {
    for (int dc = -span; dc <= span; ++dc) {
      int value;
      if (dr == 0 && dc == 0) {
        value = zero_id;
      } else {
        const int g = std::gcd(dr < 0 ? -dr : dr, dc < 0 ? -dc : dc);
        int a = dr / g, b = dc / g;
        if (a < 0 || (a == 0 && b < 0)) {
          a = -a;
          b = -b;
        }
        value = id[(size_t)(a + span) * w + (b + span)];
      }
      table[(size_t)(dr + span) * w + (dc + span)] = (uint16_t)value;
    }
  }
  info.n = n;
  info.width = w;
  info.offset = span * w + span;
  info.entries = w * w;
  info.directions = next;
  // byte counters, four per 32-bit word (at least 4n words, the scratch the
  // determinant fallback keeps there), then one trash word per lane, rounded
  // up to 16 words (64 bytes)
  int words = (next + 1 + 3) / 4;
  if (words < 4 * n) words = 4 * n;
  words += kWave;
  info.hist_words = (words + 15) / 16 * 16;
  info.counted = counted;
  info.counted_words = ((counted + 3) / 4 + kCountedTrashWords + 15) / 16 * 16;
  return N3_OK;
}

namespace {
struct DeviceTable {
  int device = -1;
  int n = 0;
  uint16_t* data = nullptr;
  DirTableInfo info{};
};
std::mutex g_table_mutex;
std::vector<DeviceTable> g_tables;
}  // namespace

int device_dirtable(int n, const uint16_t** table, DirTableInfo* info) {
  int device = 0;
  int st = hip_status(hipGetDevice(&device), "hipGetDevice");
  if (st) return st;
  std::lock_guard<std::mutex> lock(g_table_mutex);
  for (const auto& t : g_tables) {
    if (t.device == device && t.n == n) {
      *table = t.data;
      *info = t.info;
      return N3_OK;
    }
  }
  std::vector<uint16_t> host;
  DeviceTable t;
  st = build_dirtable(n, host, t.info);
  if (st) return st;
  // padded to a multiple of 4 entries: kernels copy the table as 64-bit words
  const size_t bytes = (host.size() * sizeof(uint16_t) + 7) / 8 * 8;
  host.resize(bytes / sizeof(uint16_t), 0);
  st = hip_status(hipMalloc(&t.data, bytes), "hipMalloc(direction table)");
  if (st) return st;
  st = hip_status(hipMemcpy(t.data, host.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy(direction table)");
  if (st) {
    (void)hipFree(t.data);
    return st;
  }
  t.device = device;
  t.n = n;
  g_tables.push_back(t);
  *table = t.data;
  *info = t.info;
  return N3_OK;
}

int device_limits(int* lds_limit, int* compute_units) {
  struct Limits {
    std::atomic<int> ready{0};
    int lds = 0, cus = 0;
  };
  static Limits limits[64];
  int device = 0;
  int st = hip_status(hipGetDevice(&device), "hipGetDevice");
  if (st) return st;
  if (device < 0 || device >= 64) return set_error(N3_ENODEV, "device index %d out of range", device);
  Limits& l = limits[device];
  if (!l.ready.load(std::memory_order_acquire)) {
    int lds = 0, cus = 0;
    st = hip_status(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device),
                    "hipDeviceGetAttribute(shared memory)");
    if (st) return st;
    st = hip_status(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device),
                    "hipDeviceGetAttribute(compute units)");
    if (st) return st;
    l.lds = lds > 160 * 1024 ? 160 * 1024 : lds;
    l.cus = cus;
    l.ready.store(1, std::memory_order_release);
  }
  *lds_limit = l.lds;
  *compute_units = l.cus;
  return N3_OK;
}

int check_columns_launch_args(int n, int64_t count) {
  if (n < 2 || n > kMaxN) return set_error(N3_EINVAL, "n must be in [2, %d] (got %d)", kMaxN, n);
  if (count < 0) return set_error(N3_EINVAL, "negative count");
  return N3_OK;
}

}  // namespace no3

using namespace no3;

extern "C" int n3_version(void) { return 10000; }

extern "C" const char* n3_last_error(void) { return g_error; }

extern "C" int64_t n3_struct_size(int which) {
  if (which == 0) return (int64_t)sizeof(n3_chains);
  if (which == 1) return (int64_t)sizeof(n3_group);
  return -1;
}

extern "C" int n3_chains_offsets(int64_t* offsets, int capacity) {
  static const int64_t members[] = {
      offsetof(n3_chains, n),
      offsetof(n3_chains, replicas),
      offsetof(n3_chains, columns),
      offsetof(n3_chains, badness),
      offsetof(n3_chains, scores),
      offsetof(n3_chains, best_columns),
      offsetof(n3_chains, best_scores),
      offsetof(n3_chains, accepted),
      offsetof(n3_chains, reference),
      offsetof(n3_chains, removed_points),
      offsetof(n3_chains, move_thresholds),
      offsetof(n3_chains, k_move),
      offsetof(n3_chains, exchange_thresholds),
      offsetof(n3_chains, exchange_sign),
      offsetof(n3_chains, k_exchange),
      offsetof(n3_chains, counters),
      offsetof(n3_chains, proposals),
      offsetof(n3_chains, block_waves),
      offsetof(n3_chains, log),
      offsetof(n3_chains, log_capacity),
      offsetof(n3_chains, log_count),
      offsetof(n3_chains, centers),
      offsetof(n3_chains, strengths),
      offsetof(n3_chains, inverse_temperatures),
  };
  const int count = (int)(sizeof(members) / sizeof(members[0]));
  for (int i = 0; offsets && i < count && i < capacity; ++i) offsets[i] = members[i];
  return count;
}

extern "C" int n3_device_init(int device) {
  int count = 0;
  int st = hip_status(hipGetDeviceCount(&count), "hipGetDeviceCount");
  if (st) return set_error(N3_ENODEV, "no HIP device: %s", n3_last_error());
  if (device < 0 || device >= count) return set_error(N3_ENODEV, "device %d not present (%d devices)", device, count);
  hipDeviceProp_t prop;
  st = hip_status(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
  if (st) return st;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return set_error(N3_ENODEV, "device %d is %s; this library is built for gfx950 only", device, prop.gcnArchName);
  st = hip_status(hipSetDevice(device), "hipSetDevice");
  if (st) return st;
  return mlp_scratch_setup(device, nullptr);
}

extern "C" int n3_dirtable(int n, uint16_t* host_table, int* directions, int* hist_words) {
  std::vector<uint16_t> table;
  DirTableInfo info{};
  const int st = build_dirtable(n, table, info);
  if (st) return st;
  if (host_table) std::memcpy(host_table, table.data(), table.size() * sizeof(uint16_t));
  if (directions) *directions = info.directions;
  if (hist_words) *hist_words = info.hist_words;
  return N3_OK;
}
