// Vertex welding (index_vertices.py:6-51): every point is given the input index of the first
// point, in input order, at the same quantised position. Each point is quantised and packed
// into a non-negative int64 key, the (key, input index) pairs are sorted stably by the
// library's radix sort (sort.hip), and every sorted position then takes the input index
// held at the head of its run of equal keys. A stable sort keeps each run in input order, so
// the head of a run holds the run's smallest input index.
#include "common.h"

namespace mt {
namespace {

constexpr int kBlock = 256;
constexpr int kItems = 8;                   // sorted positions per thread in the run pass
constexpr int kTile = kBlock * kItems;      // sorted positions per block
constexpr uint64_t kHuge = uint64_t(1) << 62;  // a magnitude that no bit budget admits

// The quantised coordinate torch.round(p / tol) of a float32 tensor on the device: ATen
// divides by a host scalar as a float32 multiplication by the scalar's reciprocal (taken in
// double, then rounded to float32), and rounds half to even.
__device__ __forceinline__ float quantise(float p, float inv_tol) { return rintf(p * inv_tol); }

// |q| as an integer; infinities, NaN and magnitudes of 2^62 or more map to kHuge.
__device__ __forceinline__ uint64_t magnitude(float q) {
  const float a = fabsf(q);
  return a < 0x1p62f ? uint64_t(a) : kHuge;
}

__global__ void zero_u64_kernel(unsigned long long* x) { *x = 0ull; }

__global__ void __launch_bounds__(kBlock) max_magnitude_kernel(const float* coords, int64_t n,
                                                               float inv_tol,
                                                               unsigned long long* out) {
  uint64_t m = 0;
  auto larger = [](uint64_t a, uint64_t b) { return a > b ? a : b; };
  const int64_t stride = int64_t(gridDim.x) * kBlock;
  for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride)
    m = larger(m, magnitude(quantise(coords[i], inv_tol)));
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1)
    m = larger(m, uint64_t(__shfl_xor((unsigned long long)m, d, kWave)));
  __shared__ uint64_t part[kBlock / kWave];
  if (__lane_id() == 0) part[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / kWave; ++w) m = larger(m, part[w]);
    atomicMax(out, (unsigned long long)m);
  }
}

// key = (qx + bias) << 2 bits | (qy + bias) << bits | (qz + bias): with |q| <= bias and
// 2 bias < 2^bits every field is a non-negative integer below 2^bits, so distinct quantised
// positions give distinct keys.
__global__ void __launch_bounds__(kBlock) pack_kernel(const float* points, int64_t n,
                                                      float inv_tol, int64_t bias, int bits,
                                                      int64_t* keys, int32_t* index) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const float* p = points + 3 * i;
  const int64_t x = int64_t(quantise(p[0], inv_tol)) + bias;
  const int64_t y = int64_t(quantise(p[1], inv_tol)) + bias;
  const int64_t z = int64_t(quantise(p[2], inv_tol)) + bias;
  keys[i] = (x << (2 * bits)) | (y << bits) | z;
  index[i] = int32_t(i);
}

__global__ void __launch_bounds__(kBlock) iota_kernel(int64_t n, int32_t* index) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) index[i] = int32_t(i);
}

// First position t <= s with keys[t] == keys[s] (keys ascending): steps back by doubling
// strides until a smaller key (or the start) is passed, then bisects. Runs are short in
// practice, so the first probe usually ends the search.
__device__ int64_t run_head(const int64_t* keys, int64_t s) {
  const int64_t key = keys[s];
  int64_t hi = s;  // keys[hi] == key
  int64_t lo;      // keys[lo] < key, or lo == -1
  for (int64_t step = 1;; step <<= 1) {
    lo = hi - step;
    if (lo < 0) {
      lo = -1;
      break;
    }
    if (keys[lo] != key) break;
    hi = lo;
  }
  while (hi - lo > 1) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (keys[mid] == key) hi = mid;
    else lo = mid;
  }
  return hi;
}

// out[index[s]] = index[head(s)] for every sorted position s, head(s) being the first
// position of s's run of equal keys. Thread t of a block owns the kTile-tile's positions
// [t kItems, (t + 1) kItems); the last run head at or before each position comes from a
// block-wide max-scan of the threads' last heads; positions before the tile's first head
// belong to a run that started in an earlier tile, found by run_head.
__global__ void __launch_bounds__(kBlock) first_of_run_kernel(const int64_t* keys,
                                                              const int32_t* index, int64_t n,
                                                              int32_t* out) {
  __shared__ int64_t s_key[kTile + 1];  // s_key[j + 1] = keys[b0 + j]; s_key[0] = keys[b0 - 1]
  __shared__ int32_t s_index[kTile];
  __shared__ int32_t s_wave_last[kBlock / kWave];
  __shared__ int32_t s_carry;           // input index at the head of the run holding b0
  const int64_t b0 = int64_t(blockIdx.x) * kTile;
  const int count = n - b0 < kTile ? int(n - b0) : kTile;
  for (int j = threadIdx.x; j < count; j += kBlock) {
    s_key[j + 1] = keys[b0 + j];
    s_index[j] = index[b0 + j];
  }
  if (threadIdx.x == 0) {
    s_key[0] = b0 > 0 ? keys[b0 - 1] : 0;
    // Only read when position b0 continues a run from the previous tile.
    s_carry = (b0 > 0 && keys[b0 - 1] == keys[b0]) ? index[run_head(keys, b0)] : -1;
  }
  __syncthreads();
  const int lo = threadIdx.x * kItems;
  int last = -1;  // last head among this thread's positions (tile-local), -1 if none
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const int j = lo + k;
    if (j < count && (b0 + j == 0 || s_key[j + 1] != s_key[j])) last = j;
  }
  // Inclusive max-scan of `last` over the wave, then over the waves before this one.
  const int lane = __lane_id(), wave = threadIdx.x / kWave;
  int incl = last;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int y = __shfl_up(incl, d, kWave);
    if (lane >= d && y > incl) incl = y;
  }
  if (lane == kWave - 1) s_wave_last[wave] = incl;
  __syncthreads();
  int head = __shfl_up(incl, 1, kWave);
  if (lane == 0) head = -1;
  for (int w = 0; w < wave; ++w) head = s_wave_last[w] > head ? s_wave_last[w] : head;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const int j = lo + k;
    if (j >= count) break;
    if (b0 + j == 0 || s_key[j + 1] != s_key[j]) head = j;
    out[s_index[j]] = head >= 0 ? s_index[head] : s_carry;
  }
}

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

struct IndexWorkspace {
  int64_t* keys_in;
  int64_t* keys_sorted;
  int32_t* index_in;
  int32_t* index_sorted;
  void* sort_ws;
  size_t sort_ws_bytes;
};

size_t index_workspace_bytes(int64_t n) {
  const size_t m = size_t(n < 1 ? 1 : n);
  return 2 * align_up(8 * m) + 2 * align_up(4 * m) + mt_sort_workspace_bytes(int64_t(m));
}

IndexWorkspace carve(void* ws, int64_t n) {
  const size_t m = size_t(n < 1 ? 1 : n);
  char* p = static_cast<char*>(ws);
  IndexWorkspace w;
  w.keys_in = reinterpret_cast<int64_t*>(p);
  p += align_up(8 * m);
  w.keys_sorted = reinterpret_cast<int64_t*>(p);
  p += align_up(8 * m);
  w.index_in = reinterpret_cast<int32_t*>(p);
  p += align_up(4 * m);
  w.index_sorted = reinterpret_cast<int32_t*>(p);
  p += align_up(4 * m);
  w.sort_ws = p;
  w.sort_ws_bytes = mt_sort_workspace_bytes(int64_t(m));
  return w;
}

// Sorts the packed pairs of `w` and writes every input index's first-of-run index to out.
int weld_sorted(const IndexWorkspace& w, int64_t n, int32_t* out, hipStream_t s,
                const char* what) {
  if (int rc = mt_sort_keys(w.keys_in, w.index_in, n, w.keys_sorted, w.index_sorted, w.sort_ws,
                            w.sort_ws_bytes, nullptr, s))
    return rc;
  first_of_run_kernel<<<unsigned((n + kTile - 1) / kTile), kBlock, 0, s>>>(
      w.keys_sorted, w.index_sorted, n, out);
  return check_launch(what);
}

int bit_length(uint64_t x) { return x == 0 ? 0 : 64 - __builtin_clzll(x); }

}  // namespace
}  // namespace mt

extern "C" int mt_quantized_max_abs(const float* points, int64_t n, double tol,
                                    int64_t* max_abs, mt_stream_t stream) {
  using namespace mt;
  MT_REQUIRE(n >= 0, "mt_quantized_max_abs: negative length");
  MT_REQUIRE(max_abs != nullptr, "mt_quantized_max_abs: null output");
  MT_REQUIRE(n == 0 || points != nullptr, "mt_quantized_max_abs: null points");
  MT_REQUIRE(n <= (int64_t(1) << 40), "mt_quantized_max_abs: too many points");
  const float inv_tol = float(1.0 / tol);
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto* out = reinterpret_cast<unsigned long long*>(max_abs);
  zero_u64_kernel<<<1, 1, 0, s>>>(out);
  if (n > 0) {
    const int64_t coords = 3 * n;
    const unsigned want = blocks_for(coords, kBlock * 16);
    const unsigned grid = want > 4096u ? 4096u : want;
    max_magnitude_kernel<<<grid, kBlock, 0, s>>>(points, coords, inv_tol, out);
  }
  return check_launch("mt_quantized_max_abs");
}

extern "C" size_t mt_index_vertices_workspace_bytes(int64_t n) {
  return mt::index_workspace_bytes(n);
}

extern "C" int mt_index_vertices(const float* points, int64_t n, double tol, int64_t max_abs,
                                 int32_t* out, void* workspace, size_t workspace_bytes,
                                 mt_stream_t stream) {
  using namespace mt;
  MT_REQUIRE(n >= 0, "mt_index_vertices: negative length");
  MT_REQUIRE(n < (int64_t(1) << 31), "mt_index_vertices: %lld points do not fit int32 indices",
             (long long)n);
  MT_REQUIRE(max_abs >= 0, "mt_index_vertices: negative max_abs");
  const int bits = bit_length(uint64_t(max_abs)) + 1;
  MT_REQUIRE(bits * 3 <= 63, "need %d bits, exceed 63 bits total", bits * 3);
  if (n == 0) return MT_OK;
  MT_REQUIRE(points && out, "mt_index_vertices: null pointer");
  if (workspace == nullptr || workspace_bytes < index_workspace_bytes(n)) {
    set_error("mt_index_vertices: workspace of %zu bytes is smaller than the %zu required",
              workspace_bytes, index_workspace_bytes(n));
    return MT_ERR_WORKSPACE_TOO_SMALL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const IndexWorkspace w = carve(workspace, n);
  pack_kernel<<<blocks_for(n, kBlock), kBlock, 0, s>>>(points, n, float(1.0 / tol), max_abs,
                                                       bits, w.keys_in, w.index_in);
  return weld_sorted(w, n, out, s, "mt_index_vertices");
}

extern "C" int mt_index_keys(const int64_t* keys, int64_t n, int32_t* out, void* workspace,
                             size_t workspace_bytes, mt_stream_t stream) {
  using namespace mt;
  MT_REQUIRE(n >= 0, "mt_index_keys: negative length");
  MT_REQUIRE(n < (int64_t(1) << 31), "mt_index_keys: %lld keys do not fit int32 indices",
             (long long)n);
  if (n == 0) return MT_OK;
  MT_REQUIRE(keys && out, "mt_index_keys: null pointer");
  if (workspace == nullptr || workspace_bytes < index_workspace_bytes(n)) {
    set_error("mt_index_keys: workspace of %zu bytes is smaller than the %zu required",
              workspace_bytes, index_workspace_bytes(n));
    return MT_ERR_WORKSPACE_TOO_SMALL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  IndexWorkspace w = carve(workspace, n);
  iota_kernel<<<blocks_for(n, kBlock), kBlock, 0, s>>>(n, w.index_in);
  // The sort reads the caller's keys directly; it leaves its input unmodified.
  w.keys_in = const_cast<int64_t*>(keys);
  return weld_sorted(w, n, out, s, "mt_index_keys");
}
