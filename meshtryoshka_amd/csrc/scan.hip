// Single-pass prefix sums with decoupled look-back.
//
// Each block takes a tile of kTile = 1024 x 16 elements in ticket order, scans it in
// registers (16 consecutive items per thread, wave64 shuffles across threads), publishes
// its aggregate, then one wave walks back over the predecessors' published aggregates /
// inclusive prefixes up to 64 tiles at a time. Every element is read once and written once:
// 8 bytes per element for the int32 scan, 12 for int32 -> int64.
#include <mutex>

#include "common.h"
#include "lookback.h"

namespace mt {
namespace {

// 16384-element tiles (139 KB of LDS, one workgroup of 16 waves per CU): the look-back
// chain has a quarter of the links of 4096-element tiles (C1 scans 0.68 -> 0.49 ms and
// 0.85 -> 0.71 ms).
constexpr int kBlock = 1024;
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

// Look-back state of mt_exclusive_scan_i32_noalloc: a ticket word and one look-back word
// per tile, in a file-scope device array instead of a caller workspace. Every GPU the
// library runs on holds its own copy (128 KiB + 256 B) until the process ends.
constexpr int64_t kNoAllocMaxTiles = MT_SCAN_NOALLOC_MAX_ELEMENTS / kTile;
static_assert(kNoAllocMaxTiles * kTile == MT_SCAN_NOALLOC_MAX_ELEMENTS, "whole tiles");
constexpr size_t kScratchBytes = 256 + kNoAllocMaxTiles * sizeof(uint64_t);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// In host code `scratch` names a host-side stand-in, not device memory: kernels must get
// the device address, which differs per GPU. It is looked up once per device, with that
// device current, and kept here.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_address[kMaxDevices] = {};

int scratch_address(void** out, const char* name) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error("%s: hipGetDevice failed: %s", name, hipGetErrorString(e));
    return MT_ERR_HIP;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("%s: device %d is beyond the %d devices this library tracks", name, dev, kMaxDevices);
    return MT_ERR_INVALID_ARGUMENT;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_address[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      set_error("%s: hipGetSymbolAddress of the scan scratch array failed on device %d: %s", name,
                dev, hipGetErrorString(e));
      return MT_ERR_HIP;
    }
    g_scratch_address[dev] = p;
  }
  *out = g_scratch_address[dev];
  return MT_OK;
}

}  // namespace

size_t scan_workspace_bytes_i64(int64_t n) { return scan_state_bytes(n); }

int exclusive_scan_i32(int32_t* data, int64_t n, void* ws, size_t ws_bytes, hipStream_t stream) {
  return run_scan<int32_t, int32_t, false>(data, data, n, ws, ws_bytes, stream,
                                           "mt_exclusive_scan_i32");
}

int exclusive_scan_i32_state(int32_t* data, int64_t n, void* state, size_t state_bytes,
                             hipStream_t stream, const char* name) {
  return run_scan<int32_t, int32_t, false>(data, data, n, state, state_bytes, stream, name);
}

size_t scratch_bytes() { return kScratchBytes; }

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

extern "C" int mt_exclusive_scan_i32_noalloc(int32_t* data, int64_t n, mt_stream_t stream) {
  const char* name = "mt_exclusive_scan_i32_noalloc";
  MT_REQUIRE(n >= 0, "%s: negative length", name);
  if (n > MT_SCAN_NOALLOC_MAX_ELEMENTS) {
    mt::set_error("%s: %lld elements exceed the no-allocation limit of %lld; use "
                  "mt_exclusive_scan_i32 with a workspace", name, static_cast<long long>(n),
                  static_cast<long long>(MT_SCAN_NOALLOC_MAX_ELEMENTS));
    return MT_ERR_INVALID_ARGUMENT;
  }
  if (n == 0) return MT_OK;
  MT_REQUIRE(data != nullptr, "%s: null data pointer", name);
  void* state = nullptr;
  const int rc = mt::scratch_address(&state, name);
  if (rc != MT_OK) return rc;
  // The same kernels as mt_exclusive_scan_i32; the reset kernel clears the ticket and the
  // look-back words of this call's tiles before the scan, so a previous call's contents
  // never leak into this one.
  return mt::exclusive_scan_i32_state(data, n, state, mt::scratch_bytes(),
                                      static_cast<hipStream_t>(stream), name);
}

extern "C" int mt_inclusive_scan_i32_to_i64(const int32_t* in, int64_t* out, int64_t n,
                                            void* workspace, size_t workspace_bytes,
                                            mt_stream_t stream) {
  return mt::inclusive_scan_i32_to_i64(in, out, n, workspace, workspace_bytes,
                                       static_cast<hipStream_t>(stream));
}
