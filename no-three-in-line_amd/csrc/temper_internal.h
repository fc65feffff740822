// Tempering helpers shared between the uniform and the learned walks.
#pragma once

#include "runtime.h"
#include "walk.h"

namespace no3 {

int check_chains(const n3_chains* c);
int check_exchange(const n3_chains* c, int ladder);
WalkArgs walk_args(const n3_chains* c, int steps, int bad_bias, uint64_t seed);
// One exchange round; without `done` a counter kernel advances the round
// after it, with `done` (a zeroed device word) the exchange kernel does, and
// also the sweep by d_sweep.
int enqueue_exchange(const n3_chains* c, int ladder, uint64_t seed, hipStream_t stream, uint32_t* done = nullptr,
                     uint32_t d_sweep = 0);

}  // namespace no3
