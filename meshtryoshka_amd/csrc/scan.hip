// Single-pass prefix sums with decoupled look-back.
//
// Each block takes a tile of kTile = 256 x 16 elements in ticket order, scans it in
// registers (16 consecutive items per thread, wave64 shuffles across threads), publishes
// its aggregate, then one wave walks back over the predecessors' published aggregates /
// inclusive prefixes up to 64 tiles at a time. Every element is read once and written once:
// 8 bytes per element for the int32 scan, 12 for int32 -> int64.
#include "common.h"
#include "lookback.h"

namespace mt {
namespace {

constexpr int kBlock = 256;
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

size_t scan_state_bytes(int64_t n) {
  const int64_t tiles = (n + kTile - 1) / kTile;
  return 256 + align_up(tiles * sizeof(uint64_t), 256);
}

ScanState scan_state(void* ws) {
  char* p = static_cast<char*>(ws);
  ScanState s;
  s.ticket = reinterpret_cast<uint32_t*>(p);
  s.words = reinterpret_cast<uint64_t*>(p + 256);
  return s;
}


// Block-wide exclusive scan of one value per thread; returns the exclusive prefix and
// writes the block total into *total (all threads).
template <typename T>
__device__ __forceinline__ T block_exclusive(T x, T* s_wave, T* total) {
  const int wave = threadIdx.x / kWave;
  const int lane = __lane_id();
  T inc = wave_inclusive_sum(x);
  if (lane == kWave - 1) s_wave[wave] = inc;
  __syncthreads();
  T before = 0, sum = 0;
#pragma unroll
  for (int w = 0; w < kBlock / kWave; ++w) {
    const T v = s_wave[w];
    if (w < wave) before += v;
    sum += v;
  }
  *total = sum;
  return before + inc - x;
}

// LDS index with one pad slot per 16 elements: thread t's 16 consecutive items start at
// 17 t, so a wave's simultaneous accesses fall in distinct banks.
__device__ __forceinline__ int pad(int i) { return i + (i >> 4); }
constexpr int kPadded = kTile + kTile / 16;

template <typename In, typename Out, bool kInclusive>
__global__ void __launch_bounds__(kBlock) scan_kernel(const In* in, Out* out, int64_t n,
                                                      ScanState st) {
  __shared__ union {
    In in[kPadded];
    Out out[kPadded];
  } s_buf;
  __shared__ Out s_wave[kBlock / kWave];
  __shared__ int64_t s_tile;
  __shared__ Out s_prefix;

  if (threadIdx.x == 0) s_tile = atomicAdd(st.ticket, 1u);
  __syncthreads();
  const int64_t tile = s_tile;
  const int64_t base = tile * kTile;
  const int64_t count = (n - base) < kTile ? (n - base) : kTile;

  // Coalesced striped load into LDS.
  for (int i = threadIdx.x; i < kTile; i += kBlock) s_buf.in[pad(i)] = (i < count) ? in[base + i] : In(0);
  __syncthreads();
  Out items[kItems];
  Out local = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    items[k] = static_cast<Out>(s_buf.in[pad(threadIdx.x * kItems + k)]);
    local += items[k];
  }
  Out block_total;
  const Out thread_prefix = block_exclusive<Out>(local, s_wave, &block_total);

  if (threadIdx.x < kWave) {
    int64_t prefix = 0;
    if (tile == 0) {
      if (threadIdx.x == 0) store_word(&st.words[0], kPrefix, block_total);
    } else {
      if (threadIdx.x == 0) store_word(&st.words[tile], kAggregate, block_total);
      prefix = look_back(st, tile);
      if (threadIdx.x == 0) store_word(&st.words[tile], kPrefix, prefix + int64_t(block_total));
    }
    if (threadIdx.x == 0) s_prefix = static_cast<Out>(prefix);
  }
  __syncthreads();  // also orders the LDS reads above before the writes below
  Out run = s_prefix + thread_prefix;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    if (kInclusive) {
      run += items[k];
      s_buf.out[pad(threadIdx.x * kItems + k)] = run;
    } else {
      s_buf.out[pad(threadIdx.x * kItems + k)] = run;
      run += items[k];
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < count; i += kBlock) out[base + i] = s_buf.out[pad(i)];
}

template <typename In, typename Out, bool kInclusive>
int run_scan(const In* in, Out* out, int64_t n, void* ws, size_t ws_bytes, hipStream_t stream,
             const char* name) {
  MT_REQUIRE(n >= 0, "%s: negative length", name);
  if (n == 0) return MT_OK;
  MT_REQUIRE(in != nullptr && out != nullptr, "%s: null data pointer", name);
  const size_t need = scan_state_bytes(n);
  if (ws == nullptr || ws_bytes < need) {
    set_error("%s: workspace of %zu bytes is smaller than the %zu required", name, ws_bytes, need);
    return MT_ERR_WORKSPACE_TOO_SMALL;
  }
  const int64_t tiles = (n + kTile - 1) / kTile;
  ScanState st = scan_state(ws);
  reset_scan_state<<<blocks_for(tiles, 256), 256, 0, stream>>>(st.ticket, st.words, tiles);
  scan_kernel<In, Out, kInclusive><<<static_cast<unsigned>(tiles), kBlock, 0, stream>>>(in, out, n, st);
  return check_launch(name);
}

}  // namespace

size_t scan_workspace_bytes_i64(int64_t n) { return scan_state_bytes(n); }

int exclusive_scan_i32(int32_t* data, int64_t n, void* ws, size_t ws_bytes, hipStream_t stream) {
  return run_scan<int32_t, int32_t, false>(data, data, n, ws, ws_bytes, stream,
                                           "mt_exclusive_scan_i32");
}

int inclusive_scan_i32_to_i64(const int32_t* in, int64_t* out, int64_t n, void* ws,
                              size_t ws_bytes, hipStream_t stream) {
  return run_scan<int32_t, int64_t, true>(in, out, n, ws, ws_bytes, stream,
                                          "mt_inclusive_scan_i32_to_i64");
}

}  // namespace mt

extern "C" size_t mt_scan_workspace_bytes(int64_t n) {
  // Large enough for either output type.
  return mt::scan_workspace_bytes_i64(n < 1 ? 1 : n);
}

extern "C" int mt_exclusive_scan_i32(int32_t* data, int64_t n, void* workspace,
                                     size_t workspace_bytes, mt_stream_t stream) {
  return mt::exclusive_scan_i32(data, n, workspace, workspace_bytes,
                                static_cast<hipStream_t>(stream));
}

extern "C" int mt_inclusive_scan_i32_to_i64(const int32_t* in, int64_t* out, int64_t n,
                                            void* workspace, size_t workspace_bytes,
                                            mt_stream_t stream) {
  return mt::inclusive_scan_i32_to_i64(in, out, n, workspace, workspace_bytes,
                                       static_cast<hipStream_t>(stream));
}
