// Stable LSD radix sort of (int64 key, int32 value) pairs in signed key order:
// onesweep, 8-bit digits over the key bits that vary, wave64 multi-split ranking,
// decoupled look-back across tiles.
//
//   init       zero the digit histograms, tickets and look-back words
//   mask       one read of all keys -> OR and AND of every key -> the bits that vary
//   histogram  one read of all keys -> counts of every digit of every pass
//   prepare    exclusive digit offsets per pass; a pass whose digit is the same for every
//              key is marked inactive; buffer routing so that the last active pass lands
//              in the output (input -> [tmp <-> out] ... -> out)
//   pass x 8   inactive passes return at once; active ones read 12 B and write 12 B per
//              pair. A tile of 4096 pairs is ranked in registers + LDS, published to the
//              look-back array, reordered in LDS by digit and written as contiguous runs.
//   copy       only if no pass was active (all keys equal): out = in
//
// Constant key bits cannot change the order, so the digits are taken from the varying
// bits only, packed from the least significant one up (pass p sorts varying bits
// 8p .. 8p + 7). Typical fused keys (camera | tile | depth | shell) have ~48 varying bits:
// 6 passes instead of 8. All decisions are taken on the device, so the sequence is
// graph-capturable and never synchronises the host.
#include "common.h"

namespace mt {
namespace {

constexpr int kBlock = 512;
constexpr int kWaves = kBlock / kWave;
constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems;
constexpr int kRadix = 256;
constexpr int kPasses = 8;

// Look-back word: [63:62] state, [61:59] pass tag, [58:0] count.
constexpr uint64_t kStateAggregate = 1ull << 62;
constexpr uint64_t kStatePrefix = 2ull << 62;
constexpr uint64_t kStateMask = 3ull << 62;
constexpr uint64_t kValueMask = (1ull << 59) - 1;
constexpr uint64_t kTagMask = ~kStateMask & ~kValueMask;
// Predecessor look-back words loaded at once per digit (independent loads in flight).
constexpr int kLookback = 8;

struct SortWorkspace {
  unsigned long long* hist;  // [8][256]
  int64_t* digit_base;       // [8][256]
  int32_t* info;             // [0..7] active, [8..15] src, [16..23] dst, [24] num_active
  uint32_t* tickets;         // [8]
  unsigned long long* bits;  // [0] OR of all keys, [1] AND of all keys
  uint64_t* lookback;        // [tiles][256]
  int64_t* tmp_keys;
  int32_t* tmp_values;
};

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

size_t workspace_bytes(int64_t n) {
  const int64_t tiles = (n + kTile - 1) / kTile;
  size_t b = 0;
  b += align_up(kPasses * kRadix * 8, 256);
  b += align_up(kPasses * kRadix * 8, 256);
  b += 256;
  b += 256;
  b += 256;
  b += align_up(size_t(tiles) * kRadix * 8, 256);
  b += align_up(size_t(n) * 8, 256);
  b += align_up(size_t(n) * 4, 256);
  return b;
}

SortWorkspace carve(void* ws, int64_t n) {
  const int64_t tiles = (n + kTile - 1) / kTile;
  char* p = static_cast<char*>(ws);
  SortWorkspace w;
  w.hist = reinterpret_cast<unsigned long long*>(p);
  p += align_up(kPasses * kRadix * 8, 256);
  w.digit_base = reinterpret_cast<int64_t*>(p);
  p += align_up(kPasses * kRadix * 8, 256);
  w.info = reinterpret_cast<int32_t*>(p);
  p += 256;
  w.tickets = reinterpret_cast<uint32_t*>(p);
  p += 256;
  w.bits = reinterpret_cast<unsigned long long*>(p);
  p += 256;
  w.lookback = reinterpret_cast<uint64_t*>(p);
  p += align_up(size_t(tiles) * kRadix * 8, 256);
  w.tmp_keys = reinterpret_cast<int64_t*>(p);
  p += align_up(size_t(n) * 8, 256);
  w.tmp_values = reinterpret_cast<int32_t*>(p);
  return w;
}

// The varying key bits (set = the bit differs between some keys) as runs of contiguous
// bits: run r moves (key >> shift) & (2^width - 1) to bit dst of the compacted key, whose
// bit i is the i-th varying bit counted from the least significant one.
struct Runs {
  int count;
  uint32_t desc[32];  // shift | width << 8 | dst << 16
};

// Written by one thread straight into LDS (dynamic indexing stays out of registers).
__device__ void make_runs(uint64_t v, Runs* r) {
  int count = 0;
  int dst = 0;
  int b = 0;
  while (b < 64) {
    if (!((v >> b) & 1ull)) {
      ++b;
      continue;
    }
    int e = b;
    while (e < 64 && ((v >> e) & 1ull)) ++e;
    r->desc[count++] = uint32_t(b) | (uint32_t(e - b) << 8) | (uint32_t(dst) << 16);
    dst += e - b;
    b = e;
  }
  r->count = count;
}

// Compacted key (varying bits only, packed low). Flipping the sign bit first turns signed
// order into unsigned order. Pass p sorts by bits 8p .. 8p + 7 of this value.
__device__ __forceinline__ uint64_t compact_key(uint64_t key, const Runs& r) {
  const uint64_t k = key ^ (1ull << 63);
  uint64_t out = 0;
  for (int i = 0; i < r.count; ++i) {
    const uint32_t dsc = r.desc[i];
    const uint32_t shift = dsc & 0xFFu, width = (dsc >> 8) & 0xFFu, dst = dsc >> 16;
    const uint64_t m = width >= 64 ? ~0ull : ((1ull << width) - 1ull);
    out |= ((k >> shift) & m) << dst;
  }
  return out;
}

__device__ __forceinline__ uint32_t digit_of(uint64_t key, const Runs& r, int pass) {
  return static_cast<uint32_t>((compact_key(key, r) >> (8 * pass)) & 0xFFull);
}

__device__ __forceinline__ uint64_t varying_bits(const SortWorkspace& w) {
  return w.bits[0] & ~w.bits[1];
}

__device__ __forceinline__ int passes_for(uint64_t v) { return (__popcll(v) + 7) / 8; }

__global__ void init_kernel(SortWorkspace w, int64_t lookback_words) {
  const int64_t tid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = tid; i < kPasses * kRadix; i += stride) w.hist[i] = 0ull;
  if (tid < kPasses) w.tickets[tid] = 0u;
  if (tid == 0) {
    w.bits[0] = 0ull;
    w.bits[1] = ~0ull;
  }
  for (int64_t i = tid; i < lookback_words; i += stride) w.lookback[i] = 0ull;
}

__device__ __forceinline__ int64_t live_count(int64_t n, const int64_t* n_device) {
  if (n_device == nullptr) return n;
  const int64_t d = *n_device;
  return d < n ? (d < 0 ? 0 : d) : n;
}

__global__ void __launch_bounds__(256) mask_kernel(const int64_t* keys, int64_t n, SortWorkspace w,
                                                   const int64_t* n_device) {
  n = live_count(n, n_device);
  uint64_t o = 0ull, a = ~0ull;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t k = static_cast<uint64_t>(keys[i]);
    o |= k;
    a &= k;
  }
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) {
    o |= __shfl_xor(o, d, kWave);
    a &= __shfl_xor(a, d, kWave);
  }
  if (__lane_id() == 0) {
    atomicOr(&w.bits[0], static_cast<unsigned long long>(o));
    atomicAnd(&w.bits[1], static_cast<unsigned long long>(a));
  }
}

__global__ void __launch_bounds__(256) histogram_kernel(const int64_t* keys, int64_t n,
                                                        SortWorkspace w, const int64_t* n_device) {
  n = live_count(n, n_device);
  __shared__ uint32_t s_hist[kPasses][kRadix];
  __shared__ Runs s_runs;
  const uint64_t v = varying_bits(w);
  const int passes = passes_for(v);
  for (int i = threadIdx.x; i < kPasses * kRadix; i += blockDim.x) (&s_hist[0][0])[i] = 0u;
  if (threadIdx.x == 0) make_runs(v, &s_runs);
  __syncthreads();
  // Each thread walks its keys keeping, per pass, the current digit and its run length,
  // and flushes a run when the digit changes: slowly varying digits cost one LDS atomic
  // per run instead of one per key.
  uint32_t cur[kPasses];
  uint32_t run[kPasses];
#pragma unroll
  for (int p = 0; p < kPasses; ++p) {
    cur[p] = 0u;
    run[p] = 0u;
  }
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t ck = compact_key(static_cast<uint64_t>(keys[i]), s_runs);
#pragma unroll
    for (int p = 0; p < kPasses; ++p) {
      if (p < passes) {
        const uint32_t d = static_cast<uint32_t>((ck >> (8 * p)) & 0xFFull);
        if (run[p] != 0u && d != cur[p]) {
          atomicAdd(&s_hist[p][cur[p]], run[p]);
          run[p] = 0u;
        }
        cur[p] = d;
        ++run[p];
      }
    }
  }
#pragma unroll
  for (int p = 0; p < kPasses; ++p)
    if (run[p]) atomicAdd(&s_hist[p][cur[p]], run[p]);
  __syncthreads();
  for (int i = threadIdx.x; i < passes * kRadix; i += blockDim.x) {
    const uint32_t c = (&s_hist[0][0])[i];
    if (c) atomicAdd(&w.hist[i], static_cast<unsigned long long>(c));
  }
}

__global__ void __launch_bounds__(kRadix) prepare_kernel(SortWorkspace w, int64_t n,
                                                         const int64_t* n_device) {
  n = live_count(n, n_device);
  __shared__ int32_t s_active[kPasses];
  __shared__ int64_t s_wave[kRadix / kWave];
  const int passes = passes_for(varying_bits(w));
  const int d = threadIdx.x;
  for (int p = 0; p < kPasses; ++p) {
    const int64_t c = static_cast<int64_t>(w.hist[p * kRadix + d]);
    // Block exclusive scan of the 256 bins.
    const int lane = __lane_id();
    const int wave = d / kWave;
    int64_t inc = wave_inclusive_sum(c);
    if (lane == kWave - 1) s_wave[wave] = inc;
    __syncthreads();
    int64_t before = 0;
    for (int q = 0; q < wave; ++q) before += s_wave[q];
    w.digit_base[p * kRadix + d] = before + inc - c;
    const int full = __syncthreads_or(c == n);
    if (d == 0) s_active[p] = (p < passes && !full) ? 1 : 0;
    __syncthreads();
  }
  if (d == 0) {
    int num_active = 0;
    for (int p = 0; p < kPasses; ++p) num_active += s_active[p];
    // Buffers: 0 = input, 1 = output, 2 = temporary. The last active pass writes 1.
    int remaining = num_active;
    int src = 0;
    for (int p = 0; p < kPasses; ++p) {
      w.info[p] = s_active[p];
      if (!s_active[p]) {
        w.info[8 + p] = -1;
        w.info[16 + p] = -1;
        continue;
      }
      const int dst = (remaining % 2 == 1) ? 1 : 2;
      w.info[8 + p] = src;
      w.info[16 + p] = dst;
      src = dst;
      --remaining;
    }
    w.info[24] = num_active;
  }
}

// A look-back word carries its own count, so relaxed agent-scope atomics suffice: no
// release/acquire fences (cache write-back / invalidation on gfx950) on the tile chain.
__device__ __forceinline__ uint64_t lb_load(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lb_store(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void __launch_bounds__(kBlock) onesweep_kernel(const int64_t* in_keys,
                                                          const int32_t* in_values,
                                                          int64_t* out_keys, int32_t* out_values,
                                                          int64_t n, int pass, SortWorkspace w,
                                                          const int64_t* n_device) {
  if (w.info[pass] == 0) return;
  n = live_count(n, n_device);
  const int src_sel = w.info[8 + pass];
  const int dst_sel = w.info[16 + pass];
  const int64_t* src_k = src_sel == 0 ? in_keys : (src_sel == 1 ? out_keys : w.tmp_keys);
  const int32_t* src_v = src_sel == 0 ? in_values : (src_sel == 1 ? out_values : w.tmp_values);
  int64_t* dst_k = dst_sel == 1 ? out_keys : w.tmp_keys;
  int32_t* dst_v = dst_sel == 1 ? out_values : w.tmp_values;

  __shared__ uint32_t s_wave_count[kWaves][kRadix];
  __shared__ uint32_t s_local_start[kRadix];
  __shared__ int64_t s_global_base[kRadix];
  __shared__ int64_t s_buf[kTile];  // keys, then (as int32) values, in digit order
  __shared__ uint8_t s_dig[kTile];
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_scan_wave[kRadix / kWave];
  __shared__ Runs s_runs;

  for (int i = threadIdx.x; i < kWaves * kRadix; i += kBlock) (&s_wave_count[0][0])[i] = 0u;
  if (threadIdx.x == 0) {
    s_tile = atomicAdd(&w.tickets[pass], 1u);
    make_runs(varying_bits(w), &s_runs);
  }
  __syncthreads();
  const int64_t tile = s_tile;
  const int64_t base = tile * kTile;
  if (base >= n) return;  // beyond the live keys (capacity-sized launch); uniform per block
  const int wave = threadIdx.x / kWave;
  const int lane = __lane_id();
  const uint64_t lt = lanemask_lt();

  // Items: wave w owns [base + w * 512, base + (w + 1) * 512), item i of lane l is at
  // offset i * 64 + l, so (wave, item, lane) order is the input order.
  int64_t keys[kItems];
  int32_t vals[kItems];
  uint32_t digits[kItems];
  uint32_t ranks[kItems];
  const int64_t wbase = base + int64_t(wave) * (kItems * kWave);
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const int64_t idx = wbase + i * kWave + lane;
    const bool valid = idx < n;
    keys[i] = valid ? src_k[idx] : 0;
    vals[i] = valid ? src_v[idx] : 0;
    digits[i] = valid ? digit_of(static_cast<uint64_t>(keys[i]), s_runs, pass) : kRadix;
  }
  // 1. Stable rank within each wave: peers with the same digit via 8 ballots; the first
  //    lane of each peer group adds the group's size to the wave's digit count (one LDS
  //    update per distinct digit, no atomics).
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const bool valid = digits[i] < kRadix;
    const uint32_t dd = valid ? digits[i] : 0u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const bool set = (dd >> bit) & 1u;
      const uint64_t m = __ballot(set);
      peers &= set ? m : ~m;
    }
    const uint32_t before = static_cast<uint32_t>(__popcll(peers & lt));
    const uint32_t old = valid ? s_wave_count[wave][dd] : 0u;
    ranks[i] = old + before;
    if (valid && before == 0) s_wave_count[wave][dd] = old + static_cast<uint32_t>(__popcll(peers));
  }
  __syncthreads();

  // 2. Tile histogram = sum of the wave counts; publish it at once for the successors'
  //    look-back, then turn the wave counts into each wave's offset inside the digit run.
  const int d = threadIdx.x;  // digit owned by threads < kRadix
  const uint64_t tag = uint64_t(pass) << 59;
  uint64_t* lb = w.lookback;
  uint32_t tile_count = 0;
  if (d < kRadix) {
#pragma unroll
    for (int q = 0; q < kWaves; ++q) {
      const uint32_t c = s_wave_count[q][d];
      s_wave_count[q][d] = tile_count;
      tile_count += c;
    }
    lb_store(&lb[tile * kRadix + d], (tile == 0 ? kStatePrefix : kStateAggregate) | tag | tile_count);
  }

  int64_t prefix = 0;
  uint32_t inc = 0;
  if (d < kRadix) {
    // Exclusive scan of the tile counts over digits: start of each digit's run in LDS.
    inc = wave_inclusive_sum(tile_count);
    if (lane == kWave - 1) s_scan_wave[wave] = inc;

    // 3. Look-back over preceding tiles for this digit, kLookback words per round.
    if (tile > 0) {
      int64_t t = tile - 1;
      bool done = false;
      while (!done) {
        uint64_t words[kLookback];
#pragma unroll
        for (int j = 0; j < kLookback; ++j)
          words[j] = (t - j >= 0) ? lb_load(&lb[(t - j) * kRadix + d]) : (kStatePrefix | tag);
        bool stop = false;
        int consumed = 0;
#pragma unroll
        for (int j = 0; j < kLookback; ++j) {
          const uint64_t wd = words[j];
          const bool valid = (wd & kStateMask) != 0 && (wd & kTagMask) == tag;
          if (!stop) {
            if (!valid) {
              stop = true;  // not published yet: poll again from this predecessor
            } else {
              prefix += static_cast<int64_t>(wd & kValueMask);
              ++consumed;
              if ((wd & kStateMask) == kStatePrefix) {
                stop = true;
                done = true;
              }
            }
          }
        }
        t -= consumed;
      }
      lb_store(&lb[tile * kRadix + d], kStatePrefix | tag | uint64_t(prefix + tile_count));
    }
  }
  __syncthreads();
  if (d < kRadix) {
    uint32_t wave_before = 0;
    for (int q = 0; q < wave; ++q) wave_before += s_scan_wave[q];
    s_local_start[d] = wave_before + inc - tile_count;
    s_global_base[d] = w.digit_base[pass * kRadix + d] + prefix;
  }
  __syncthreads();

  // 4. Reorder the keys by digit in LDS and write them as contiguous runs; then the same
  //    for the values through the same buffer.
  const int valid_count = static_cast<int>((n - base) < kTile ? (n - base) : kTile);
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    if (digits[i] < kRadix) {
      const uint32_t dd = digits[i];
      const uint32_t pos = s_local_start[dd] + s_wave_count[wave][dd] + ranks[i];
      s_buf[pos] = keys[i];
      s_dig[pos] = static_cast<uint8_t>(dd);
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < valid_count; j += kBlock) {
    const uint32_t dd = s_dig[j];
    dst_k[s_global_base[dd] + (j - static_cast<int64_t>(s_local_start[dd]))] = s_buf[j];
  }
  __syncthreads();
  int32_t* s_vbuf = reinterpret_cast<int32_t*>(s_buf);
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    if (digits[i] < kRadix) {
      const uint32_t dd = digits[i];
      s_vbuf[s_local_start[dd] + s_wave_count[wave][dd] + ranks[i]] = vals[i];
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < valid_count; j += kBlock) {
    const uint32_t dd = s_dig[j];
    dst_v[s_global_base[dd] + (j - static_cast<int64_t>(s_local_start[dd]))] = s_vbuf[j];
  }
}

__global__ void copy_if_idle_kernel(const int64_t* in_keys, const int32_t* in_values,
                                    int64_t* out_keys, int32_t* out_values, int64_t n,
                                    const int32_t* info, const int64_t* n_device) {
  if (info[24] != 0) return;
  n = live_count(n, n_device);
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    out_keys[i] = in_keys[i];
    out_values[i] = in_values[i];
  }
}

unsigned capped_blocks(int64_t n, int64_t per_block, int64_t cap) {
  int64_t b = (n + per_block - 1) / per_block;
  b = b < 1 ? 1 : (b > cap ? cap : b);
  return static_cast<unsigned>(b);
}

}  // namespace
}  // namespace mt

extern "C" size_t mt_sort_workspace_bytes(int64_t n) { return mt::workspace_bytes(n < 1 ? 1 : n); }

extern "C" int mt_sort_keys(const int64_t* keys_in, const int32_t* values_in, int64_t n,
                            int64_t* keys_out, int32_t* values_out, void* workspace,
                            size_t workspace_bytes, const int64_t* n_device,
                            mt_stream_t stream) {
  using namespace mt;
  MT_REQUIRE(n >= 0, "mt_sort_keys: negative length");
  if (n == 0) return MT_OK;
  MT_REQUIRE(keys_in && values_in && keys_out && values_out, "mt_sort_keys: null pointer");
  MT_REQUIRE(keys_in != keys_out && values_in != values_out, "mt_sort_keys: in-place sort is not supported");
  const size_t need = mt::workspace_bytes(n);
  if (workspace == nullptr || workspace_bytes < need) {
    set_error("mt_sort_keys: workspace of %zu bytes is smaller than the %zu required", workspace_bytes, need);
    return MT_ERR_WORKSPACE_TOO_SMALL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  SortWorkspace w = carve(workspace, n);
  const int64_t tiles = (n + kTile - 1) / kTile;
  const int64_t lb_words = tiles * kRadix;
  init_kernel<<<capped_blocks(lb_words, 256 * 8, 1 << 20), 256, 0, s>>>(w, lb_words);
  mask_kernel<<<capped_blocks(n, 256 * 16, 4096), 256, 0, s>>>(keys_in, n, w, n_device);
  histogram_kernel<<<capped_blocks(n, 256 * 32, 4096), 256, 0, s>>>(keys_in, n, w, n_device);
  prepare_kernel<<<1, kRadix, 0, s>>>(w, n, n_device);
  for (int p = 0; p < kPasses; ++p) {
    onesweep_kernel<<<static_cast<unsigned>(tiles), kBlock, 0, s>>>(keys_in, values_in, keys_out,
                                                                   values_out, n, p, w, n_device);
  }
  copy_if_idle_kernel<<<capped_blocks(n, 256 * 4, 65536), 256, 0, s>>>(
      keys_in, values_in, keys_out, values_out, n, w.info, n_device);
  return check_launch("mt_sort_keys");
}
