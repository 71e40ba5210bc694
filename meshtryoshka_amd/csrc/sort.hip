// Stable LSD radix sort of (int64 key, int32 value) pairs in signed key order:
// reduce-then-scan passes of 8-bit digits over the key bits that vary, wave64 multi-split
// ranking, LDS reorder and contiguous scatter.
//
//   init      the OR / AND accumulators
//   mask      one read of all keys -> the bits that vary (set in some keys, clear in others)
//   prepare   number of passes (8 varying bits each, least significant first) and the buffer
//             routing input -> [tmp <-> out] ... -> out, so the last pass lands in the output
//   per pass  count   : per 4096-key tile, how many keys have each digit (keys read once)
//             scan    : inclusive scan of the counts in (digit, tile) order = every tile's
//                       output position for every digit (no digit base pass needed)
//             scatter : rank each key inside its tile (wave ballots), reorder the tile in LDS
//                       by digit and write each digit's run contiguously
//   copy      only if no bit varies (all keys equal): out = in
//
// Constant key bits cannot change the order, so digits are taken from the varying bits
// only: every pass has at least one varying bit, hence at least two digit values. Typical
// fused keys (camera | tile | depth | shell) have ~48 varying bits: 6 passes instead of 8.
// The passes carry no inter-tile dependency (no look-back): each kernel streams its keys
// at copy speed. All decisions are taken on the device, so the sequence is graph-capturable
// and never synchronises the host.
#include "common.h"

namespace mt {
namespace {

constexpr int kBlock = 512;
constexpr int kWaves = kBlock / kWave;
constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems;
constexpr int kRadix = 256;
constexpr int kPasses = 8;

struct SortWorkspace {
  int32_t* info;             // [0..7] active, [8..15] src, [16..23] dst, [24] passes
  unsigned long long* bits;  // [0] OR of all keys, [1] AND of all keys, [2] bits sorted by
  int32_t* counts;           // [256][tiles] keys per (digit, tile) of the current pass
  int64_t* offsets;          // [256][tiles] inclusive scan of counts
  void* scan_ws;
  size_t scan_ws_bytes;
  int64_t* tmp_keys;
  int32_t* tmp_values;
};

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

int64_t tiles_for(int64_t n) { return (n + kTile - 1) / kTile; }

size_t workspace_bytes(int64_t n) {
  const int64_t cells = tiles_for(n) * kRadix;
  size_t b = 256 + 256;
  b += align_up(size_t(cells) * 4, 256);
  b += align_up(size_t(cells) * 8, 256);
  b += align_up(scan_workspace_bytes_i64(cells), 256);
  b += align_up(size_t(n) * 8, 256);
  b += align_up(size_t(n) * 4, 256);
  return b;
}

SortWorkspace carve(void* ws, int64_t n) {
  const int64_t cells = tiles_for(n) * kRadix;
  char* p = static_cast<char*>(ws);
  SortWorkspace w;
  w.info = reinterpret_cast<int32_t*>(p);
  p += 256;
  w.bits = reinterpret_cast<unsigned long long*>(p);
  p += 256;
  w.counts = reinterpret_cast<int32_t*>(p);
  p += align_up(size_t(cells) * 4, 256);
  w.offsets = reinterpret_cast<int64_t*>(p);
  p += align_up(size_t(cells) * 8, 256);
  w.scan_ws = p;
  w.scan_ws_bytes = align_up(scan_workspace_bytes_i64(cells), 256);
  p += w.scan_ws_bytes;
  w.tmp_keys = reinterpret_cast<int64_t*>(p);
  p += align_up(size_t(n) * 8, 256);
  w.tmp_values = reinterpret_cast<int32_t*>(p);
  return w;
}

// The varying key bits as runs of contiguous bits: run r moves (key >> shift) &
// (2^width - 1) to bit dst of the compacted key, whose bit i is the i-th varying bit
// counted from the least significant one.
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
  return w.bits[0] & ~w.bits[1] & w.bits[2];
}

__device__ __forceinline__ int64_t live_count(int64_t n, const int64_t* n_device) {
  if (n_device == nullptr) return n;
  const int64_t d = *n_device;
  return d < n ? (d < 0 ? 0 : d) : n;
}

__global__ void init_kernel(SortWorkspace w, uint64_t sorted_bits) {
  if (threadIdx.x == 0) {
    w.bits[0] = 0ull;
    w.bits[1] = ~0ull;
    w.bits[2] = sorted_bits;
  }
}

__global__ void __launch_bounds__(256) mask_kernel(const int64_t* keys, int64_t n, SortWorkspace w,
                                                   const int64_t* n_device) {
  n = live_count(n, n_device);
  uint64_t o = 0ull, a = ~0ull;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  // Eight independent loads in flight per thread and round.
  constexpr int kUnroll = 8;
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  for (; i + (kUnroll - 1) * stride < n; i += kUnroll * stride) {
    uint64_t k[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) k[u] = static_cast<uint64_t>(keys[i + u * stride]);
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      o |= k[u];
      a &= k[u];
    }
  }
  for (; i < n; i += stride) {
    const uint64_t k = static_cast<uint64_t>(keys[i]);
    o |= k;
    a &= k;
  }
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) {
    o |= __shfl_xor(o, d, kWave);
    a &= __shfl_xor(a, d, kWave);
  }
  // One atomic pair per workgroup: the pairs on the two shared words serialise.
  __shared__ uint64_t s_o[256 / kWave], s_a[256 / kWave];
  if (__lane_id() == 0) {
    s_o[threadIdx.x / kWave] = o;
    s_a[threadIdx.x / kWave] = a;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 1; q < 256 / kWave; ++q) {
      o |= s_o[q];
      a &= s_a[q];
    }
    atomicOr(&w.bits[0], static_cast<unsigned long long>(o));
    atomicAnd(&w.bits[1], static_cast<unsigned long long>(a));
  }
}

__global__ void prepare_kernel(SortWorkspace w) {
  if (threadIdx.x != 0) return;
  const int passes = (__popcll(varying_bits(w)) + 7) / 8;
  // Buffers: 0 = input, 1 = output, 2 = temporary. The last pass writes 1.
  int src = 0;
  for (int p = 0; p < kPasses; ++p) {
    const int active = p < passes;
    w.info[p] = active;
    w.info[8 + p] = active ? src : -1;
    const int dst = active ? (((passes - p) % 2 == 1) ? 1 : 2) : -1;
    w.info[16 + p] = dst;
    if (active) src = dst;
  }
  w.info[24] = passes;
}

struct Buffers {
  const int64_t* src_k;
  const int32_t* src_v;
  int64_t* dst_k;
  int32_t* dst_v;
};

__device__ __forceinline__ Buffers route(const SortWorkspace& w, int pass, const int64_t* in_k,
                                         const int32_t* in_v, int64_t* out_k, int32_t* out_v) {
  const int src = w.info[8 + pass], dst = w.info[16 + pass];
  Buffers b;
  b.src_k = src == 0 ? in_k : (src == 1 ? out_k : w.tmp_keys);
  b.src_v = src == 0 ? in_v : (src == 1 ? out_v : w.tmp_values);
  b.dst_k = dst == 1 ? out_k : w.tmp_keys;
  b.dst_v = dst == 1 ? out_v : w.tmp_values;
  return b;
}

// Loads one tile's digits (item i of lane l of wave w is key base + w * 512 + i * 64 + l)
// and ranks them inside each wave: peers with the same digit via 8 ballots, ranks[i] =
// earlier occurrences of the digit in the wave, s_wave_count[w][d] = occurrences per wave.
// Digits of keys past n are kRadix (no digit).
__device__ __forceinline__ void rank_tile(const uint32_t* digits, uint32_t* ranks,
                                          uint32_t (*s_wave_count)[kRadix]) {
  const int wave = threadIdx.x / kWave;
  const uint64_t lt = lanemask_lt();
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
}

// Pass step 1: keys per (digit, tile), digit-major. Tiles past the live keys write zeros.
// A plain LDS histogram (one sub-histogram per wave): counting needs no ranks, and every
// load instruction reads 64 consecutive keys.
__global__ void __launch_bounds__(kBlock) count_kernel(const int64_t* in_k, const int32_t* in_v,
                                                       int64_t* out_k, int32_t* out_v, int64_t n,
                                                       int pass, SortWorkspace w,
                                                       const int64_t* n_device) {
  if (w.info[pass] == 0) return;
  n = live_count(n, n_device);
  __shared__ uint32_t s_hist[kWaves][kRadix];
  __shared__ Runs s_runs;
  for (int i = threadIdx.x; i < kWaves * kRadix; i += kBlock) (&s_hist[0][0])[i] = 0u;
  if (threadIdx.x == 0) make_runs(varying_bits(w), &s_runs);
  __syncthreads();
  const int64_t tile = xcd_swizzle(blockIdx.x, gridDim.x);
  const int64_t tiles = gridDim.x;
  const int64_t base = tile * kTile;
  if (base < n) {  // uniform per block
    const Buffers bf = route(w, pass, in_k, in_v, out_k, out_v);
    const int wave = threadIdx.x / kWave;
    uint64_t keys[kItems];
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const int64_t idx = base + i * kBlock + threadIdx.x;
      keys[i] = idx < n ? static_cast<uint64_t>(bf.src_k[idx]) : 0ull;
    }
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const int64_t idx = base + i * kBlock + threadIdx.x;
      if (idx < n) atomicAdd(&s_hist[wave][digit_of(keys[i], s_runs, pass)], 1u);
    }
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kRadix; d += kBlock) {
    uint32_t total = 0;
#pragma unroll
    for (int q = 0; q < kWaves; ++q) total += s_hist[q][d];
    w.counts[int64_t(d) * tiles + tile] = static_cast<int32_t>(total);
  }
}

// Pass step 3: scatter. Global position of digit d in this tile = inclusive scan value of
// the previous (digit, tile) cell.
__global__ void __launch_bounds__(kBlock) scatter_kernel(const int64_t* in_k, const int32_t* in_v,
                                                         int64_t* out_k, int32_t* out_v, int64_t n,
                                                         int pass, SortWorkspace w,
                                                         const int64_t* n_device) {
  if (w.info[pass] == 0) return;
  n = live_count(n, n_device);
  // Tiles in XCD-contiguous ranges: consecutive tiles write adjacent pieces of every
  // digit's output run, so those cache lines are completed in one XCD's L2.
  const int64_t tile = xcd_swizzle(blockIdx.x, gridDim.x);
  const int64_t tiles = gridDim.x;
  const int64_t base = tile * kTile;
  if (base >= n) return;  // beyond the live keys (capacity-sized launch); uniform per block
  const Buffers bf = route(w, pass, in_k, in_v, out_k, out_v);

  __shared__ uint32_t s_wave_count[kWaves][kRadix];
  __shared__ uint32_t s_local_start[kRadix];
  __shared__ int64_t s_global_base[kRadix];
  __shared__ int64_t s_buf[kTile];  // keys, then (as int32) values, in digit order
  __shared__ uint8_t s_dig[kTile];
  __shared__ uint32_t s_scan_wave[kRadix / kWave];
  __shared__ Runs s_runs;

  for (int i = threadIdx.x; i < kWaves * kRadix; i += kBlock) (&s_wave_count[0][0])[i] = 0u;
  if (threadIdx.x == 0) make_runs(varying_bits(w), &s_runs);
  __syncthreads();
  const int wave = threadIdx.x / kWave;
  const int lane = __lane_id();

  int64_t keys[kItems];
  int32_t vals[kItems];
  uint32_t digits[kItems];
  uint32_t ranks[kItems];
  const int64_t wbase = base + int64_t(wave) * (kItems * kWave);
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const int64_t idx = wbase + i * kWave + lane;
    const bool valid = idx < n;
    keys[i] = valid ? bf.src_k[idx] : 0;
    vals[i] = valid ? bf.src_v[idx] : 0;
    digits[i] = valid ? digit_of(static_cast<uint64_t>(keys[i]), s_runs, pass) : kRadix;
  }
  rank_tile(digits, ranks, s_wave_count);
  __syncthreads();

  const int d = threadIdx.x;  // digit owned by threads < kRadix
  uint32_t tile_count = 0, inc = 0;
  if (d < kRadix) {
    // Offsets of each wave inside the tile's run of the digit, then the run starts.
#pragma unroll
    for (int q = 0; q < kWaves; ++q) {
      const uint32_t c = s_wave_count[q][d];
      s_wave_count[q][d] = tile_count;
      tile_count += c;
    }
    inc = wave_inclusive_sum(tile_count);
    if (lane == kWave - 1) s_scan_wave[wave] = inc;
    const int64_t cell = int64_t(d) * tiles + tile;
    s_global_base[d] = cell == 0 ? 0 : w.offsets[cell - 1];
  }
  __syncthreads();
  if (d < kRadix) {
    uint32_t wave_before = 0;
    for (int q = 0; q < wave; ++q) wave_before += s_scan_wave[q];
    s_local_start[d] = wave_before + inc - tile_count;
  }
  __syncthreads();

  // Reorder the keys by digit in LDS and write them as contiguous runs; then the same for
  // the values through the same buffer.
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
    bf.dst_k[s_global_base[dd] + (j - static_cast<int64_t>(s_local_start[dd]))] = s_buf[j];
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
    bf.dst_v[s_global_base[dd] + (j - static_cast<int64_t>(s_local_start[dd]))] = s_vbuf[j];
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

namespace mt {
namespace {

// ------------------------------------------------------------------------------------
// Sorting ranges in place by their low key bits (the second half of a tile-bucketed
// sort: mt_sort_keys_high orders the keys by their high bits, e.g. camera | tile, a
// delineation gives each tile's range, and every range is then sorted by the remaining
// bits, e.g. depth | shell, inside one workgroup). Stable, so the result is the stable
// sort by the whole key.
// ------------------------------------------------------------------------------------

// Per pass of a block sort: the items of the block are in position order (item i of lane l
// of wave w is position w kIt 64 + i 64 + l). Each item's digit is ranked among its wave's
// lanes by ballots (rank_tile's multi-split, generalised to kIt items) and the wave's
// running count of the digit in LDS; the waves' counts are scanned per digit and over the
// digits, and pos receives each item's place in the stable order by this digit.
template <int kThreads, int kIt>
struct BlockRank {
  static constexpr int kW = kThreads / kWave;
  uint32_t wave_count[kW][kRadix];
  uint32_t digit_start[kRadix];
  uint32_t digit_total[kRadix];
  uint32_t scan_wave[kRadix / kWave];
};

// live: the wave's items at or past this index hold no key (a bound the wave agrees on), so
// their ranking is skipped. Ends with a barrier.
template <int kThreads, int kIt>
__device__ __forceinline__ void block_rank(const uint32_t (&digits)[kIt], uint32_t (&pos)[kIt],
                                           BlockRank<kThreads, kIt>& R, int live = kIt) {
  constexpr int kW = kThreads / kWave;
  for (int i = threadIdx.x; i < kW * kRadix; i += kThreads) (&R.wave_count[0][0])[i] = 0u;
  __syncthreads();
  const int wave = threadIdx.x / kWave;
  const int lane = __lane_id();
  const uint64_t lt = lanemask_lt();
  uint32_t ranks[kIt];
#pragma unroll
  for (int i = 0; i < kIt; ++i) {
    ranks[i] = 0u;
    if (i >= live) continue;
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
    const uint32_t old = valid ? R.wave_count[wave][dd] : 0u;
    ranks[i] = old + before;
    if (valid && before == 0) R.wave_count[wave][dd] = old + static_cast<uint32_t>(__popcll(peers));
  }
  __syncthreads();
  // Thread t < kOwners owns digits [t kDpt, (t + 1) kDpt): the waves' offsets inside each
  // digit's run, then the digits' starts by a scan over the owners (in digit order).
  constexpr int kDpt = kRadix / kThreads > 0 ? kRadix / kThreads : 1;
  constexpr int kOwners = kRadix / kDpt;  // 256 or 64: whole waves
  const int t = threadIdx.x;
  uint32_t tot[kDpt];
  uint32_t local = 0, inc = 0;
  if (t < kOwners) {
#pragma unroll
    for (int j = 0; j < kDpt; ++j) {
      const int d = t * kDpt + j;
      uint32_t total = 0;
#pragma unroll
      for (int q = 0; q < kW; ++q) {
        const uint32_t c = R.wave_count[q][d];
        R.wave_count[q][d] = total;
        total += c;
      }
      tot[j] = total;
      local += total;
    }
    inc = wave_inclusive_sum(local);
    if (lane == kWave - 1) R.scan_wave[t / kWave] = inc;
  }
  __syncthreads();
  if (t < kOwners) {
    uint32_t run = inc - local;
    for (int q = 0; q < t / kWave; ++q) run += R.scan_wave[q];
#pragma unroll
    for (int j = 0; j < kDpt; ++j) {
      const int d = t * kDpt + j;
      R.digit_start[d] = run;
      R.digit_total[d] = tot[j];
      run += tot[j];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kIt; ++i)
    pos[i] = (i < live && digits[i] < kRadix)
                 ? R.digit_start[digits[i]] + R.wave_count[wave][digits[i]] + ranks[i]
                 : 0xffffffffu;
}

// Varying bits of the range's keys below low_mask, and the runs that compact them.
template <int kThreads>
__device__ __forceinline__ void range_runs(unsigned long long o, unsigned long long a,
                                           unsigned long long* s_oa, Runs* runs, int* passes,
                                           uint64_t low_mask) {
#pragma unroll
  for (int dd = kWave / 2; dd >= 1; dd >>= 1) {
    o |= __shfl_xor(o, dd, kWave);
    a &= __shfl_xor(a, dd, kWave);
  }
  if (threadIdx.x == 0) {
    s_oa[0] = 0ull;
    s_oa[1] = ~0ull;
  }
  __syncthreads();
  if (__lane_id() == 0) {
    atomicOr(&s_oa[0], o);
    atomicAnd(&s_oa[1], a);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint64_t v = s_oa[0] & ~s_oa[1] & low_mask;
    make_runs(v, runs);
    *passes = (__popcll(v) + 7) / 8;
  }
  __syncthreads();
}

template <int kThreads, int kIt>
struct RangeSortLds {
  int64_t buf[kThreads * kIt];  // keys, then (as int32) values, in the pass's new order
  BlockRank<kThreads, kIt> rank;
  unsigned long long oa[2];
  Runs runs;
  int passes;
};

// One workgroup per range of at most kThreads * kIt keys (ranges of other sizes are left
// to the other kernels): the keys and values stay in registers and LDS for every pass.
template <int kThreads, int kIt>
__global__ void __launch_bounds__(kThreads) range_sort_kernel(int64_t* keys, int32_t* values,
                                                              const int64_t* ranges,
                                                              const int32_t* list,
                                                              const int32_t* list_count, int64_t n,
                                                              int64_t min_size, uint64_t low_mask) {
  constexpr int kCap = kThreads * kIt;
  constexpr int kPerWave = kIt * kWave;
  extern __shared__ __align__(16) char s_dyn[];
  auto& L = *reinterpret_cast<RangeSortLds<kThreads, kIt>*>(s_dyn);
  const int wave = threadIdx.x / kWave;
  const int lane = __lane_id();
  const int32_t listed = *list_count;
  for (int32_t li = blockIdx.x; li < listed; li += gridDim.x) {
    const int64_t r = list[li];
    int64_t start = ranges[2 * r], end = ranges[2 * r + 1];
    start = start < 0 ? 0 : (start > n ? n : start);
    end = end < start ? start : (end > n ? n : end);
    const int64_t m = end - start;
    if (m <= min_size || m > kCap) continue;  // uniform over the block
    int64_t k[kIt];
    int32_t v[kIt];
    unsigned long long o = 0ull, a = ~0ull;
#pragma unroll
    for (int i = 0; i < kIt; ++i) {
      const int p = wave * kPerWave + i * kWave + lane;
      const bool valid = p < m;
      k[i] = valid ? keys[start + p] : 0;
      v[i] = valid ? values[start + p] : 0;
      if (valid) {
        o |= static_cast<uint64_t>(k[i]);
        a &= static_cast<uint64_t>(k[i]);
      }
    }
    range_runs<kThreads>(o, a, L.oa, &L.runs, &L.passes, low_mask);
    const int passes = L.passes;
    // Items of this wave that hold some key (the same for every lane of the wave).
    const int64_t in_wave = m - int64_t(wave) * kPerWave;
    const int live = in_wave <= 0 ? 0 : int(in_wave >= kPerWave ? kIt : (in_wave + kWave - 1) / kWave);
    for (int pass = 0; pass < passes; ++pass) {
      uint32_t dg[kIt], pos[kIt];
#pragma unroll
      for (int i = 0; i < kIt; ++i) {
        const int p = wave * kPerWave + i * kWave + lane;
        dg[i] = p < m ? digit_of(static_cast<uint64_t>(k[i]), L.runs, pass) : kRadix;
      }
      block_rank<kThreads, kIt>(dg, pos, L.rank, live);
#pragma unroll
      for (int i = 0; i < kIt; ++i)
        if (pos[i] != 0xffffffffu) L.buf[pos[i]] = k[i];
      __syncthreads();
#pragma unroll
      for (int i = 0; i < kIt; ++i) {
        const int p = wave * kPerWave + i * kWave + lane;
        if (p < m) k[i] = L.buf[p];
      }
      __syncthreads();
      int32_t* vb = reinterpret_cast<int32_t*>(L.buf);
#pragma unroll
      for (int i = 0; i < kIt; ++i)
        if (pos[i] != 0xffffffffu) vb[pos[i]] = v[i];
      __syncthreads();
#pragma unroll
      for (int i = 0; i < kIt; ++i) {
        const int p = wave * kPerWave + i * kWave + lane;
        if (p < m) v[i] = vb[p];
      }
      __syncthreads();
    }
    if (passes > 0) {
#pragma unroll
      for (int i = 0; i < kIt; ++i) {
        const int p = wave * kPerWave + i * kWave + lane;
        if (p < m) {
          keys[start + p] = k[i];
          values[start + p] = v[i];
        }
      }
    }
    __syncthreads();  // the next range reuses the LDS
  }
}

// Ranges longer than kThreads * kIt: one workgroup per range runs the passes through
// global memory (the range in keys / values and the same places of tmp, alternately),
// chunk by chunk in order: a histogram of the whole range, then each chunk's items go to
// their digit's next places. Stable like the in-LDS passes.
template <int kThreads, int kIt>
__global__ void __launch_bounds__(kThreads) long_range_sort_kernel(
    int64_t* keys, int32_t* values, int64_t* tmp_keys, int32_t* tmp_values,
    const int64_t* ranges, const int32_t* list, const int32_t* list_count, int64_t n,
    int64_t min_size, uint64_t low_mask) {
  constexpr int kCap = kThreads * kIt;
  constexpr int kPerWave = kIt * kWave;
  __shared__ BlockRank<kThreads, kIt> R;
  __shared__ uint32_t s_hist[kRadix];
  __shared__ int64_t s_next[kRadix];
  __shared__ unsigned long long s_oa[2];
  __shared__ Runs s_runs;
  __shared__ int s_passes;
  const int wave = threadIdx.x / kWave;
  const int lane = __lane_id();
  const int32_t listed = *list_count;
  for (int32_t li = blockIdx.x; li < listed; li += gridDim.x) {
    const int64_t r = list[li];
    int64_t start = ranges[2 * r], end = ranges[2 * r + 1];
    start = start < 0 ? 0 : (start > n ? n : start);
    end = end < start ? start : (end > n ? n : end);
    const int64_t m = end - start;
    if (m <= min_size) continue;  // uniform over the block
    unsigned long long o = 0ull, a = ~0ull;
    for (int64_t j = threadIdx.x; j < m; j += kThreads) {
      const uint64_t kk = static_cast<uint64_t>(keys[start + j]);
      o |= kk;
      a &= kk;
    }
    range_runs<kThreads>(o, a, s_oa, &s_runs, &s_passes, low_mask);
    const int passes = s_passes;
    for (int pass = 0; pass < passes; ++pass) {
      const bool from_tmp = (pass & 1) != 0;
      const int64_t* src_k = (from_tmp ? tmp_keys : keys) + start;
      const int32_t* src_v = (from_tmp ? tmp_values : values) + start;
      int64_t* dst_k = (from_tmp ? keys : tmp_keys) + start;
      int32_t* dst_v = (from_tmp ? values : tmp_values) + start;
      for (int d = threadIdx.x; d < kRadix; d += kThreads) s_hist[d] = 0u;
      __syncthreads();
      for (int64_t j = threadIdx.x; j < m; j += kThreads)
        atomicAdd(&s_hist[digit_of(static_cast<uint64_t>(src_k[j]), s_runs, pass)], 1u);
      __syncthreads();
      if (threadIdx.x < kRadix) {  // exclusive scan of the histogram (one wave per 64 digits)
        const uint32_t c = s_hist[threadIdx.x];
        const uint32_t inc = wave_inclusive_sum(c);
        if (lane == kWave - 1) R.scan_wave[threadIdx.x / kWave] = inc;
        s_next[threadIdx.x] = inc - c;
      }
      __syncthreads();
      if (threadIdx.x < kRadix) {
        int64_t before = 0;
        for (int q = 0; q < int(threadIdx.x) / kWave; ++q) before += R.scan_wave[q];
        s_next[threadIdx.x] += before;
      }
      __syncthreads();
      for (int64_t c0 = 0; c0 < m; c0 += kCap) {
        const int64_t cm = (m - c0) < kCap ? (m - c0) : kCap;
        int64_t k[kIt];
        int32_t v[kIt];
        uint32_t dg[kIt], pos[kIt];
#pragma unroll
        for (int i = 0; i < kIt; ++i) {
          const int p = wave * kPerWave + i * kWave + lane;
          const bool valid = p < cm;
          k[i] = valid ? src_k[c0 + p] : 0;
          v[i] = valid ? src_v[c0 + p] : 0;
          dg[i] = valid ? digit_of(static_cast<uint64_t>(k[i]), s_runs, pass) : kRadix;
        }
        block_rank<kThreads, kIt>(dg, pos, R);
#pragma unroll
        for (int i = 0; i < kIt; ++i) {
          if (pos[i] == 0xffffffffu) continue;
          // pos[i] - digit_start = the item's place among the chunk's items of its digit.
          const int64_t at = s_next[dg[i]] + (pos[i] - R.digit_start[dg[i]]);
          dst_k[at] = k[i];
          dst_v[at] = v[i];
        }
        __syncthreads();
        if (threadIdx.x < kRadix) s_next[threadIdx.x] += R.digit_total[threadIdx.x];
        __syncthreads();
      }
    }
    if (passes & 1) {  // the last pass wrote tmp: copy the range back
      for (int64_t j = threadIdx.x; j < m; j += kThreads) {
        keys[start + j] = tmp_keys[start + j];
        values[start + j] = tmp_values[start + j];
      }
    }
    __syncthreads();  // the next range reuses the LDS
  }
}



// Size classes of the range sort (keys per range): up to 1024, 2048, 4096, 8192, 16384,
// longer.
constexpr int kRangeClasses = 6;
constexpr int64_t kClassMax[kRangeClasses - 1] = {1024, 2048, 4096, 8192, 16384};

// Lists every range of two or more keys in its size class (one atomic per wave and class),
// so that each class's kernel visits only its own ranges.
__global__ void __launch_bounds__(256) classify_ranges_kernel(const int64_t* ranges,
                                                              int64_t num_ranges, int64_t n,
                                                              int32_t* lists, int32_t* counts) {
  const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int cls = -1;
  if (r < num_ranges) {
    int64_t start = ranges[2 * r], end = ranges[2 * r + 1];
    start = start < 0 ? 0 : (start > n ? n : start);
    end = end < start ? start : (end > n ? n : end);
    const int64_t m = end - start;
    if (m >= 2) {
      cls = kRangeClasses - 1;
      for (int c = kRangeClasses - 2; c >= 0; --c)
        if (m <= kClassMax[c]) cls = c;
    }
  }
  const int lane = __lane_id();
#pragma unroll
  for (int c = 0; c < kRangeClasses; ++c) {
    const uint64_t m = __ballot(cls == c);
    if (m == 0) continue;
    int at = 0;
    if (lane == 0) at = atomicAdd(&counts[c], __popcll(m));
    at = __shfl(at, 0, kWave);
    if (cls == c) lists[c * num_ranges + at + __popcll(m & lanemask_lt())] = static_cast<int32_t>(r);
  }
}

__global__ void zero_class_counts_kernel(int32_t* counts) {
  if (threadIdx.x < kRangeClasses) counts[threadIdx.x] = 0;
}

}  // namespace
}  // namespace mt

extern "C" size_t mt_sort_key_ranges_workspace_bytes(int64_t num_ranges) {
  const size_t r = static_cast<size_t>(num_ranges < 1 ? 1 : num_ranges);
  return 256 + (r * 4 * mt::kRangeClasses + 255) / 256 * 256;
}

extern "C" int mt_sort_key_ranges(int64_t* keys, int32_t* values, int64_t n,
                                  const int64_t* ranges, int64_t num_ranges, int32_t low_bits,
                                  int64_t* tmp_keys, int32_t* tmp_values, void* workspace,
                                  size_t workspace_bytes, mt_stream_t stream) {
  using namespace mt;
  MT_REQUIRE(n >= 0 && num_ranges >= 0, "mt_sort_key_ranges: negative size");
  MT_REQUIRE(num_ranges < (int64_t(1) << 31), "mt_sort_key_ranges: too many ranges");
  MT_REQUIRE(low_bits >= 0 && low_bits <= 63, "mt_sort_key_ranges: low_bits %d (0..63)", low_bits);
  if (n == 0 || num_ranges == 0 || low_bits == 0) return MT_OK;
  MT_REQUIRE(keys && values && ranges && tmp_keys && tmp_values, "mt_sort_key_ranges: null pointer");
  MT_REQUIRE(tmp_keys != keys && tmp_values != values,
             "mt_sort_key_ranges: the temporary arrays must be separate from the keys and values");
  const size_t need = mt_sort_key_ranges_workspace_bytes(num_ranges);
  if (workspace == nullptr || workspace_bytes < need) {
    set_error("mt_sort_key_ranges: workspace of %zu bytes is smaller than the %zu required",
              workspace_bytes, need);
    return MT_ERR_WORKSPACE_TOO_SMALL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint64_t low_mask = (1ull << low_bits) - 1ull;
  int32_t* counts = static_cast<int32_t*>(workspace);
  int32_t* lists = reinterpret_cast<int32_t*>(static_cast<char*>(workspace) + 256);
  zero_class_counts_kernel<<<1, 64, 0, s>>>(counts);
  classify_ranges_kernel<<<blocks_for(num_ranges, 256), 256, 0, s>>>(ranges, num_ranges, n, lists,
                                                                      counts);
  // (threads, items) per class: (64, 16), (128, 16), (256, 16), (512, 16) (75 KB of LDS, two
  // workgroups per CU), (1024, 16) (146 KB); longer ranges through the temporary arrays.
  const size_t lds_a = sizeof(RangeSortLds<64, 16>), lds_a2 = sizeof(RangeSortLds<128, 16>);
  const size_t lds_b = sizeof(RangeSortLds<256, 16>);
  const size_t lds_c = sizeof(RangeSortLds<512, 16>), lds_d = sizeof(RangeSortLds<1024, 16>);
  static bool attributes_set = false;
  if (!attributes_set) {
    const hipError_t e1 =
        hipFuncSetAttribute(reinterpret_cast<const void*>(&range_sort_kernel<512, 16>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds_c));
    const hipError_t e2 =
        hipFuncSetAttribute(reinterpret_cast<const void*>(&range_sort_kernel<1024, 16>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds_d));
    if (e1 != hipSuccess || e2 != hipSuccess) {
      set_error("mt_sort_key_ranges: cannot allow %zu bytes of LDS: %s", lds_d,
                hipGetErrorString(e1 != hipSuccess ? e1 : e2));
      return MT_ERR_HIP;
    }
    attributes_set = true;
  }
  // Each class's workgroups loop over its list (its length is read on the device); the
  // grids cover the GPU a few times over.
  const auto grid = [&](int64_t cap) {
    return static_cast<unsigned>(num_ranges < cap ? num_ranges : cap);
  };
  const int64_t stride = num_ranges;
  const int32_t* list = lists;
  range_sort_kernel<64, 16><<<grid(8192), 64, lds_a, s>>>(keys, values, ranges, list, counts + 0, n,
                                                           1, low_mask);
  range_sort_kernel<128, 16><<<grid(4096), 128, lds_a2, s>>>(keys, values, ranges, list + stride,
                                                             counts + 1, n, kClassMax[0], low_mask);
  range_sort_kernel<256, 16><<<grid(2048), 256, lds_b, s>>>(keys, values, ranges, list + 2 * stride,
                                                            counts + 2, n, kClassMax[1], low_mask);
  range_sort_kernel<512, 16><<<grid(1024), 512, lds_c, s>>>(keys, values, ranges, list + 3 * stride,
                                                            counts + 3, n, kClassMax[2], low_mask);
  range_sort_kernel<1024, 16><<<grid(512), 1024, lds_d, s>>>(keys, values, ranges, list + 4 * stride,
                                                             counts + 4, n, kClassMax[3], low_mask);
  long_range_sort_kernel<1024, 8><<<grid(512), 1024, 0, s>>>(
      keys, values, tmp_keys, tmp_values, ranges, list + 5 * stride, counts + 5, n, kClassMax[4],
      low_mask);
  return check_launch("mt_sort_key_ranges");
}

extern "C" size_t mt_sort_workspace_bytes(int64_t n) { return mt::workspace_bytes(n < 1 ? 1 : n); }

namespace mt {
namespace {

// The LSD sort over the key bits in sorted_bits (all 64 for mt_sort_keys).
int sort_impl(const int64_t* keys_in, const int32_t* values_in, int64_t n, int64_t* keys_out,
              int32_t* values_out, void* workspace, size_t workspace_bytes,
              const int64_t* n_device, uint64_t sorted_bits, hipStream_t s, const char* name) {
  MT_REQUIRE(n >= 0, "%s: negative length", name);
  if (n == 0) return MT_OK;
  MT_REQUIRE(keys_in && values_in && keys_out && values_out, "%s: null pointer", name);
  MT_REQUIRE(keys_in != keys_out && values_in != values_out, "%s: in-place sort is not supported", name);
  const int64_t tiles = tiles_for(n);
  MT_REQUIRE(tiles < (int64_t(1) << 31), "%s: too many keys", name);
  const size_t need = mt::workspace_bytes(n);
  if (workspace == nullptr || workspace_bytes < need) {
    set_error("%s: workspace of %zu bytes is smaller than the %zu required", name, workspace_bytes, need);
    return MT_ERR_WORKSPACE_TOO_SMALL;
  }
  SortWorkspace w = carve(workspace, n);
  init_kernel<<<1, 64, 0, s>>>(w, sorted_bits);
  mask_kernel<<<capped_blocks(n, 256 * 16, 4096), 256, 0, s>>>(keys_in, n, w, n_device);
  prepare_kernel<<<1, 64, 0, s>>>(w);
  const unsigned grid = static_cast<unsigned>(tiles);
  // The varying bits are a subset of sorted_bits: passes past those bits are never active,
  // so they are not launched (the device decides among the rest).
  const int max_passes = (__builtin_popcountll(sorted_bits) + 7) / 8;
  for (int p = 0; p < max_passes && p < kPasses; ++p) {
    count_kernel<<<grid, kBlock, 0, s>>>(keys_in, values_in, keys_out, values_out, n, p, w, n_device);
    // Inactive passes leave the counts as they are; scanning them is harmless.
    if (int rc = inclusive_scan_i32_to_i64(w.counts, w.offsets, tiles * kRadix, w.scan_ws,
                                           w.scan_ws_bytes, s))
      return rc;
    scatter_kernel<<<grid, kBlock, 0, s>>>(keys_in, values_in, keys_out, values_out, n, p, w, n_device);
  }
  copy_if_idle_kernel<<<capped_blocks(n, 256 * 4, 65536), 256, 0, s>>>(
      keys_in, values_in, keys_out, values_out, n, w.info, n_device);
  return check_launch(name);
}

}  // namespace
}  // namespace mt

extern "C" int mt_sort_keys(const int64_t* keys_in, const int32_t* values_in, int64_t n,
                            int64_t* keys_out, int32_t* values_out, void* workspace,
                            size_t workspace_bytes, const int64_t* n_device,
                            mt_stream_t stream) {
  return mt::sort_impl(keys_in, values_in, n, keys_out, values_out, workspace, workspace_bytes,
                       n_device, ~0ull, static_cast<hipStream_t>(stream), "mt_sort_keys");
}

extern "C" int mt_sort_keys_high(const int64_t* keys_in, const int32_t* values_in, int64_t n,
                                 int64_t* keys_out, int32_t* values_out, void* workspace,
                                 size_t workspace_bytes, const int64_t* n_device,
                                 int32_t low_bits, mt_stream_t stream) {
  MT_REQUIRE(low_bits >= 0 && low_bits <= 64, "mt_sort_keys_high: low_bits %d", low_bits);
  const uint64_t high = low_bits >= 64 ? 0ull : (~0ull << low_bits);
  return mt::sort_impl(keys_in, values_in, n, keys_out, values_out, workspace, workspace_bytes,
                       n_device, high, static_cast<hipStream_t>(stream), "mt_sort_keys_high");
}
