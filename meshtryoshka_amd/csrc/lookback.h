// Decoupled look-back over per-tile words, shared by the single-pass scans and the
// single-pass key generator: a tile publishes its aggregate, sums its predecessors'
// published values back to the nearest inclusive prefix, then publishes its own inclusive
// prefix. Tiles take their index from a ticket counter, so every tile a tile waits for
// has started.
#pragma once

#include "common.h"

namespace mt {

// Look-back word per tile: [63:62] state, [61:0] value as 62-bit two's complement. The
// word carries its own value, so publishing and polling need only relaxed agent-scope
// atomics: no release/acquire fences (which on gfx950 write back / invalidate caches) sit
// on the serial chain between neighbouring tiles.
inline constexpr uint64_t kNone = 0;
inline constexpr uint64_t kAggregate = 1ull << 62;
inline constexpr uint64_t kPrefix = 2ull << 62;
inline constexpr uint64_t kStateMask = 3ull << 62;
inline constexpr uint64_t kValueMask = (1ull << 62) - 1;
// Predecessor words loaded at once per look-back round (independent loads in flight).
inline constexpr int kWindow = 64;

struct ScanState {
  uint32_t* ticket;
  uint64_t* words;
};

// Resets the ticket and marks every tile unpublished (a fill kernel, so the sequence can
// be captured in a graph).
__global__ inline void reset_scan_state(uint32_t* ticket, uint64_t* words, int64_t tiles) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i == 0) *ticket = 0;
  if (i < tiles) words[i] = kNone;
}

__device__ __forceinline__ uint64_t load_word(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_word(uint64_t* p, uint64_t state, int64_t value) {
  __hip_atomic_store(p, state | (static_cast<uint64_t>(value) & kValueMask), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int64_t word_value(uint64_t w) {
  return static_cast<int64_t>(w << 2) >> 2;  // sign-extend the 62-bit field
}

// Look-back by the first wave: returns the sum of all tiles before `tile`. Lane l polls
// tile (base - l); the wave consumes the published words up to the nearest inclusive
// prefix, or up to the first unpublished one, and polls again from there.
__device__ inline int64_t look_back(const ScanState& st, int64_t tile) {
  const int lane = __lane_id();
  int64_t running = 0;
  int64_t base = tile - 1;
  while (base >= 0) {
    const int64_t t = base - lane;
    const uint64_t w = t >= 0 ? load_word(&st.words[t]) : kPrefix;  // before tile 0: empty prefix
    const uint64_t state = w & kStateMask;
    const uint64_t pending = __ballot(state == kNone);
    const uint64_t has_prefix = __ballot(state == kPrefix);
    // First lane (nearest predecessor) that is unpublished or an inclusive prefix.
    const uint64_t stops = pending | has_prefix;
    const int first = stops ? __ffsll(static_cast<long long>(stops)) - 1 : kWindow;
    const bool take_first = first < kWindow && ((has_prefix >> first) & 1ull);
    const int take = first + (take_first ? 1 : 0);  // lanes [0, take) are consumed
    int64_t contrib = (lane < take && t >= 0) ? word_value(w) : 0;
    running += wave_sum(contrib);
    if (take_first) break;
    base -= take;
  }
  return running;
}

}  // namespace mt
