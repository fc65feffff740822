// Arguments of a tempering walk launch (one replica slot per wavefront).
#pragma once

#include <stdint.h>

namespace no3 {

// int32 fields of one proposal-log entry
constexpr int kLogFields = 8;

struct WalkArgs {
  int n;
  int replicas;
  int steps;
  int bad_bias;
  uint32_t seed_lo, seed_hi;
  int32_t* columns;
  int32_t* badness;
  int32_t* scores;
  int32_t* best_columns;
  int32_t* best_scores;
  int32_t* accepted;
  const int32_t* reference;
  int32_t* removed;
  const uint32_t* thr;
  int k_move;
  const uint32_t* counters;
  unsigned long long* proposals;
  int32_t* log;                   // optional proposal log, kLogFields int32 per entry
  long long log_capacity;
  unsigned long long* log_count;
  const float* centers;           // umbrella potentials: null strengths = none
  const float* strengths;
  const double* inverse_temperatures;
  int ladder;                     // > 0: each block takes temperatures spread over one ladder (walk_slot)
  uint32_t sweep_offset;          // the sweep number is counters[0] + sweep_offset (the k-th walk of a graph: k)
};

}  // namespace no3
