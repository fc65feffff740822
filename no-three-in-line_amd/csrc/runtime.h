// Host-side helpers shared by the engine's translation units.
#pragma once

#include <vector>

#include "../../include/no3.h"
#include "no3_common.h"

namespace no3 {

int set_error(int status, const char* fmt, ...);
int hip_status(hipError_t e, const char* what);
int build_dirtable(int n, std::vector<uint16_t>& table, DirTableInfo& info);
// Resident direction table of size n on the current device (built on first use).
int device_dirtable(int n, const uint16_t** table, DirTableInfo* info);
int check_columns_launch_args(int n, int64_t count);
// LDS bytes per block (at most 160 KiB) and compute-unit count of the current
// device, queried once per device.
int device_limits(int* lds_limit, int* compute_units);

// Device address of the no-allocation gradient's scratch array on `device`
// (current), looked up once per device (csrc/mlp.hip).
int mlp_scratch_setup(int device, float** address);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// LDS bytes of a direction table rounded up to 16 bytes.
__host__ __device__ inline int table_lds_bytes(const DirTableInfo& info) { return (info.entries * 2 + 15) / 16 * 16; }

}  // namespace no3
