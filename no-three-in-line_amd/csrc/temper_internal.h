// Tempering helpers shared between the uniform and the learned walks.
#pragma once

#include "runtime.h"
#include "walk.h"

namespace no3 {

int check_chains(const n3_chains* c);
int check_exchange(const n3_chains* c, int ladder);
WalkArgs walk_args(const n3_chains* c, int steps, int bad_bias, uint64_t seed);
// One exchange round, number counters[1] + round_offset; with `advance` a
// counter kernel advances counters[1] by one after it.
int enqueue_exchange(const n3_chains* c, int ladder, uint64_t seed, hipStream_t stream, uint32_t round_offset = 0,
                     bool advance = true);
// A counter kernel: counters[0] += d_sweep, counters[1] += d_round.
int enqueue_counters(const n3_chains* c, uint32_t d_sweep, uint32_t d_round, hipStream_t stream);

}  // namespace no3
