// Single-pass prefix sums with decoupled look-back.
//
// Each block takes a tile of kTile = 256 x 16 elements in ticket order, scans it in
// registers (16 consecutive items per thread, wave64 shuffles across threads), publishes
// its aggregate, then one wave walks back over the predecessors' published aggregates /
// inclusive prefixes 64 tiles at a time. Every element is read once and written once:
// 8 bytes per element for the int32 scan, 12 for int32 -> int64.
#include "common.h"

namespace mt {
namespace {

constexpr int kBlock = 256;
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;

enum : uint32_t { kNone = 0, kAggregate = 1, kPrefix = 2 };

template <typename Out>
struct ScanState {
  uint32_t* ticket;
  uint32_t* flags;
  Out* aggregate;
  Out* prefix;
};

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

template <typename Out>
size_t scan_state_bytes(int64_t n) {
  const int64_t tiles = (n + kTile - 1) / kTile;
  return 256 + align_up(tiles * sizeof(uint32_t), 256) + 2 * align_up(tiles * sizeof(Out), 256);
}

template <typename Out>
ScanState<Out> scan_state(void* ws, int64_t n) {
  const int64_t tiles = (n + kTile - 1) / kTile;
  char* p = static_cast<char*>(ws);
  ScanState<Out> s;
  s.ticket = reinterpret_cast<uint32_t*>(p);
  p += 256;
  s.flags = reinterpret_cast<uint32_t*>(p);
  p += align_up(tiles * sizeof(uint32_t), 256);
  s.aggregate = reinterpret_cast<Out*>(p);
  p += align_up(tiles * sizeof(Out), 256);
  s.prefix = reinterpret_cast<Out*>(p);
  return s;
}

__global__ void reset_scan_state(uint32_t* ticket, uint32_t* flags, int64_t tiles) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i == 0) *ticket = 0;
  if (i < tiles) flags[i] = kNone;
}

__device__ __forceinline__ uint32_t load_flag(const uint32_t* f) {
  return __hip_atomic_load(f, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_flag(uint32_t* f, uint32_t v) {
  __hip_atomic_store(f, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
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

// Look-back by the first wave: returns the sum of all tiles before `tile`.
template <typename Out>
__device__ Out look_back(const ScanState<Out>& st, int64_t tile) {
  const int lane = __lane_id();
  Out running = 0;
  int64_t base = tile - 1;
  while (base >= 0) {
    const int64_t t = base - lane;
    uint32_t f = kPrefix;  // lanes before tile 0 count as an empty inclusive prefix
    Out v = 0;
    if (t >= 0) {
      do {
        f = load_flag(&st.flags[t]);
      } while (f == kNone);
      v = (f == kPrefix) ? st.prefix[t] : st.aggregate[t];
    }
    const uint64_t has_prefix = __ballot(f == kPrefix);
    // Lanes up to and including the nearest inclusive prefix contribute.
    const int stop = has_prefix ? __ffsll(static_cast<long long>(has_prefix)) - 1 : kWave - 1;
    Out contrib = (lane <= stop) ? v : Out(0);
    contrib = wave_sum(contrib);
    running += contrib;
    if (has_prefix) break;
    base -= kWave;
  }
  return running;
}

// LDS index with one pad slot per 16 elements: thread t's 16 consecutive items start at
// 17 t, so a wave's simultaneous accesses fall in distinct banks.
__device__ __forceinline__ int pad(int i) { return i + (i >> 4); }
constexpr int kPadded = kTile + kTile / 16;

template <typename In, typename Out, bool kInclusive>
__global__ void __launch_bounds__(kBlock) scan_kernel(const In* in, Out* out, int64_t n,
                                                      ScanState<Out> st) {
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
    Out prefix = 0;
    if (tile == 0) {
      if (threadIdx.x == 0) {
        st.prefix[0] = block_total;
        store_flag(&st.flags[0], kPrefix);
      }
    } else {
      if (threadIdx.x == 0) {
        st.aggregate[tile] = block_total;
        store_flag(&st.flags[tile], kAggregate);
      }
      prefix = look_back(st, tile);
      if (threadIdx.x == 0) {
        st.prefix[tile] = prefix + block_total;
        store_flag(&st.flags[tile], kPrefix);
      }
    }
    if (threadIdx.x == 0) s_prefix = prefix;
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
  const size_t need = scan_state_bytes<Out>(n);
  if (ws == nullptr || ws_bytes < need) {
    set_error("%s: workspace of %zu bytes is smaller than the %zu required", name, ws_bytes, need);
    return MT_ERR_WORKSPACE_TOO_SMALL;
  }
  const int64_t tiles = (n + kTile - 1) / kTile;
  ScanState<Out> st = scan_state<Out>(ws, n);
  reset_scan_state<<<blocks_for(tiles, 256), 256, 0, stream>>>(st.ticket, st.flags, tiles);
  scan_kernel<In, Out, kInclusive><<<static_cast<unsigned>(tiles), kBlock, 0, stream>>>(in, out, n, st);
  return check_launch(name);
}

}  // namespace

size_t scan_workspace_bytes_i64(int64_t n) { return scan_state_bytes<int64_t>(n); }

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
