// Coarse-to-fine subdivision of the sparse grid: marking occupied cells, dilating and
// upscaling bit-packed occupancy grids, and resampling the per-vertex SDF and SH onto the
// finer grid (write_occupancy.slang, dilate_occupancy.slang, upscale_occupancy.slang,
// upscale_parameters.slang). Packed grids are int32[I, J, K / 32], bit b of word
// (i, j, kp) being cell k = 32 kp + b. All integer work is exact; the resampling keeps the
// reference's floating-point operation order (no contraction), so it is bit-reproducible.
#include "common.h"

namespace mt {
namespace {

constexpr int kBlock = 256;

__global__ void __launch_bounds__(kBlock) write_occupancy_kernel(const int32_t* voxel_indices,
                                                                 int64_t n, uint32_t* occupancy,
                                                                 int64_t num_words) {
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int32_t v = voxel_indices[t];
  if (v < 0 || (int64_t(v) >> 5) >= num_words) return;  // not a cell of this grid
  // Many triangles share a cell: an atomic only where the bit is not yet visible.
  const uint32_t bit = 1u << (v & 31);
  uint32_t* word = occupancy + (v >> 5);
  if ((__hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bit) == 0)
    atomicOr(word, bit);
}

// OR of the (2d + 1)^2 words around (i, j) in word column kp (dilate_occupancy.slang:28-36).
__device__ __forceinline__ uint32_t window_or(const uint32_t* occ, int I, int J, int KP, int i,
                                              int j, int kp, int d) {
  uint32_t r = 0;
  for (int di = -d; di <= d; ++di) {
    const int ii = i + di;
    if (ii < 0 || ii >= I) continue;
    for (int dj = -d; dj <= d; ++dj) {
      const int jj = j + dj;
      if (jj < 0 || jj >= J) continue;
      r |= occ[(int64_t(ii) * J + jj) * KP + kp];
    }
  }
  return r;
}

// Box dilation by d cells along all three axes (dilate_occupancy.slang:9-94): the window
// OR of the word and of its two k neighbours, shifted by up to d bits across the words.
__global__ void __launch_bounds__(kBlock) dilate_occupancy_kernel(const uint32_t* occ, int I,
                                                                  int J, int KP, int d,
                                                                  uint32_t* out) {
  const int64_t w = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (w >= int64_t(I) * J * KP) return;
  const int kp = static_cast<int>(w % KP);
  const int j = static_cast<int>((w / KP) % J);
  const int i = static_cast<int>(w / (int64_t(KP) * J));
  uint32_t result = 0;
  if (kp > 0) {  // the previous word's top d bits reach this word's low d bits
    uint32_t prev = window_or(occ, I, J, KP, i, j, kp - 1, d) >> (32 - d);
    for (int b = 0; b < d; ++b) {
      result |= prev;
      prev >>= 1;
    }
  }
  const uint32_t cur = window_or(occ, I, J, KP, i, j, kp, d);
  result |= cur;
  uint32_t left = cur, right = cur;
  for (int b = 0; b < d; ++b) {
    right >>= 1;
    left <<= 1;
    result |= left | right;
  }
  if (kp < KP - 1) {  // the next word's low d bits reach this word's top d bits
    uint32_t next = window_or(occ, I, J, KP, i, j, kp + 1, d) << (32 - d);
    for (int b = 0; b < d; ++b) {
      result |= next;
      next <<= 1;
    }
  }
  out[w] = result;
}

// A finer cell (i, j, k) is occupied if a coarse cell under any of its corners is
// (upscale_occupancy.slang:1-74; integer corner mapping, no rounding).
__global__ void __launch_bounds__(kBlock) upscale_occupancy_kernel(const uint32_t* occ, int I,
                                                                   int J, int KP, uint32_t* out,
                                                                   int I2, int J2, int KP2) {
  const int64_t w = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (w >= int64_t(I2) * J2 * KP2) return;
  const int kp = static_cast<int>(w % KP2);
  const int j = static_cast<int>((w / KP2) % J2);
  const int i = static_cast<int>(w / (int64_t(KP2) * J2));
  const int K = 32 * KP, K2 = 32 * KP2;
  // Coarse rows / columns under the low and high corners in i and j.
  int is[2], js[2];
  for (int s = 0; s < 2; ++s) {
    is[s] = ((i + s) * I - s) / I2;
    js[s] = ((j + s) * J - s) / J2;
  }
  uint32_t result = 0;
  for (int b = 0; b < 32; ++b) {
    const int k = 32 * kp + b;
    bool value = false;
    for (int sk = 0; sk < 2; ++sk) {
      const int ks = ((k + sk) * K - sk) / K2;
      const uint32_t bit = 1u << (ks & 31);
      for (int si = 0; si < 2; ++si)
        for (int sj = 0; sj < 2; ++sj)
          value = value || (occ[(int64_t(is[si]) * J + js[sj]) * KP + (ks >> 5)] & bit) != 0;
    }
    result |= uint32_t(value) << b;
  }
  out[w] = result;
}

// Vertex (i, j, k) of a vertex-occupancy grid: its index, or -1 where it is outside the
// grid or not occupied (upscale_parameters.slang:7-52; the index is the word's offset plus
// the occupied bits below).
__device__ __forceinline__ int64_t vertex_index(const uint32_t* vocc, const int32_t* voff,
                                                int VI, int VJ, int VKP, int i, int j, int k) {
  if (i < 0 || j < 0 || k < 0) return -1;
  const int kp = k >> 5;
  if (i >= VI || j >= VJ || kp >= VKP) return -1;
  const int64_t w = (int64_t(i) * VJ + j) * VKP + kp;
  const uint32_t chunk = vocc[w];
  const int bit = k & 31;
  if (((chunk >> bit) & 1u) == 0) return -1;
  return int64_t(voff[w]) + __popc(chunk & ((1u << bit) - 1u));
}

// One thread per bit of the new vertex-occupancy grid (neighbouring lanes, neighbouring
// vertices along k): every occupied new vertex gets the occupancy-weighted trilinear
// interpolation of the old vertices around its position (upscale_parameters.slang:54-198),
// weights and sums in the reference's order.
template <int kNsh>
__global__ void __launch_bounds__(kBlock) upscale_parameters_kernel(
    const uint32_t* old_vocc, const int32_t* old_voff, int OI, int OJ, int OKP,
    const float* old_sdf, const float* old_sh, int64_t old_v, const uint32_t* new_vocc,
    const int32_t* new_voff, int NI, int NJ, int NKP, float* new_sdf, float* new_sh,
    int64_t new_v) {
  const int64_t tid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t w = tid >> 5;
  const int b = static_cast<int>(tid & 31);
  if (w >= int64_t(NI) * NJ * NKP) return;
  const uint32_t chunk = new_vocc[w];
  if (((chunk >> b) & 1u) == 0) return;
  const int kp = static_cast<int>(w % NKP);
  const int j = static_cast<int>((w / NKP) % NJ);
  const int i = static_cast<int>(w / (int64_t(NKP) * NJ));
  const float inv_i = 1.0f / float(NI - 1);
  const float inv_j = 1.0f / float(NJ - 1);
  const float inv_k = 1.0f / float(32 * (NKP - 1));
  const float scale_x = float((OKP - 1) * 32), scale_y = float(OJ - 1), scale_z = float(OI - 1);
  const int64_t out = int64_t(new_voff[w]) + __popc(chunk & ((1u << b) - 1u));
  {
    const int k = 32 * kp + b;
    const float x = float(k) * inv_k, y = float(j) * inv_j, z = float(i) * inv_i;
    const float xu = x * scale_x, yu = y * scale_y, zu = z * scale_z;
    const int ci = int(zu), cj = int(yu), ck = int(xu);
    const float ti = fmodf(zu, 1.0f), tj = fmodf(yu, 1.0f), tk = fmodf(xu, 1.0f);
    float sh[kNsh][3];
#pragma unroll
    for (int s = 0; s < kNsh; ++s) sh[s][0] = sh[s][1] = sh[s][2] = 0.0f;
    float sd = 0.0f, den = 0.0f;
    for (int di = 0; di < 2; ++di)
      for (int dj = 0; dj < 2; ++dj)
        for (int dk = 0; dk < 2; ++dk) {
          const int64_t v = vertex_index(old_vocc, old_voff, OI, OJ, OKP, ci + di, cj + dj, ck + dk);
          if (v < 0 || v >= old_v) continue;
          const float wgt = ((di == 0 ? 1.0f - ti : ti) * (dj == 0 ? 1.0f - tj : tj)) *
                            (dk == 0 ? 1.0f - tk : tk);
#pragma unroll
          for (int s = 0; s < kNsh; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c) sh[s][c] += wgt * old_sh[(int64_t(s) * old_v + v) * 3 + c];
          sd += wgt * old_sdf[v];
          den += wgt;
        }
    const float m = 1.0f / den;
    if (out < new_v) {
#pragma unroll
      for (int s = 0; s < kNsh; ++s)
#pragma unroll
        for (int c = 0; c < 3; ++c) new_sh[(int64_t(s) * new_v + out) * 3 + c] = sh[s][c] * m;
      new_sdf[out] = sd * m;
    }
  }
}

}  // namespace
}  // namespace mt

extern "C" int mt_write_occupancy(const int32_t* voxel_indices, int64_t n, int32_t* occupancy,
                                  int64_t num_words, int32_t min_shell, int32_t max_shell,
                                  mt_stream_t stream) {
  (void)min_shell;  // accepted and unused, as in write_occupancy.slang:3-20
  (void)max_shell;
  MT_REQUIRE(n >= 0 && num_words >= 0, "mt_write_occupancy: negative size");
  if (n == 0) return MT_OK;
  mt::write_occupancy_kernel<<<mt::blocks_for(n, mt::kBlock), mt::kBlock, 0,
                               static_cast<hipStream_t>(stream)>>>(
      voxel_indices, n, reinterpret_cast<uint32_t*>(occupancy), num_words);
  return mt::check_launch("mt_write_occupancy");
}

extern "C" int mt_dilate_occupancy(const int32_t* occupancy, int32_t I, int32_t J, int32_t KP,
                                   int32_t dilation, int32_t* out, mt_stream_t stream) {
  MT_REQUIRE(I >= 0 && J >= 0 && KP >= 0, "mt_dilate_occupancy: negative size");
  MT_REQUIRE(dilation > 0 && dilation < 32, "mt_dilate_occupancy: dilation %d outside (0, 32)",
             dilation);
  const int64_t words = int64_t(I) * J * KP;
  if (words == 0) return MT_OK;
  mt::dilate_occupancy_kernel<<<mt::blocks_for(words, mt::kBlock), mt::kBlock, 0,
                                static_cast<hipStream_t>(stream)>>>(
      reinterpret_cast<const uint32_t*>(occupancy), I, J, KP, dilation,
      reinterpret_cast<uint32_t*>(out));
  return mt::check_launch("mt_dilate_occupancy");
}

extern "C" int mt_upscale_occupancy(const int32_t* occupancy, int32_t I, int32_t J, int32_t KP,
                                    int32_t* out, int32_t I2, int32_t J2, int32_t KP2,
                                    mt_stream_t stream) {
  MT_REQUIRE(I > 0 && J > 0 && KP > 0 && I2 >= I && J2 >= J && KP2 >= KP,
             "mt_upscale_occupancy: the target grid must be at least as large as the source");
  MT_REQUIRE(int64_t(I2 + 1) * I < (int64_t(1) << 31) && int64_t(J2 + 1) * J < (int64_t(1) << 31) &&
                 int64_t(32) * (KP2 + 1) * 32 * KP < (int64_t(1) << 31),
             "mt_upscale_occupancy: grid too large for the int32 corner mapping");
  const int64_t words = int64_t(I2) * J2 * KP2;
  mt::upscale_occupancy_kernel<<<mt::blocks_for(words, mt::kBlock), mt::kBlock, 0,
                                 static_cast<hipStream_t>(stream)>>>(
      reinterpret_cast<const uint32_t*>(occupancy), I, J, KP, reinterpret_cast<uint32_t*>(out),
      I2, J2, KP2);
  return mt::check_launch("mt_upscale_occupancy");
}

extern "C" int mt_upscale_parameters(const int32_t* old_vertex_occupancy,
                                     const int32_t* old_vertex_offsets, int32_t OI, int32_t OJ,
                                     int32_t OKP, const float* old_sdf, const float* old_sh,
                                     int64_t old_num_vertices, int32_t num_sh,
                                     const int32_t* new_vertex_occupancy,
                                     const int32_t* new_vertex_offsets, int32_t NI, int32_t NJ,
                                     int32_t NKP, float* new_sdf, float* new_sh,
                                     int64_t new_num_vertices, mt_stream_t stream) {
  MT_REQUIRE(OI > 1 && OJ > 1 && OKP > 1 && NI > 1 && NJ > 1 && NKP > 1,
             "mt_upscale_parameters: vertex grids need at least two vertices per axis");
  MT_REQUIRE(num_sh == 1 || num_sh == 4 || num_sh == 9 || num_sh == 16,
             "mt_upscale_parameters: %d SH coefficients (need 1, 4, 9 or 16)", num_sh);
  const int64_t words = int64_t(NI) * NJ * NKP;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const auto* ov = reinterpret_cast<const uint32_t*>(old_vertex_occupancy);
  const auto* nv = reinterpret_cast<const uint32_t*>(new_vertex_occupancy);
#define MT_UPS(N)                                                                              \
  mt::upscale_parameters_kernel<N><<<mt::blocks_for(32 * words, mt::kBlock), mt::kBlock, 0, st>>>( \
      ov, old_vertex_offsets, OI, OJ, OKP, old_sdf, old_sh, old_num_vertices, nv,             \
      new_vertex_offsets, NI, NJ, NKP, new_sdf, new_sh, new_num_vertices)
  switch (num_sh) {
    case 1: MT_UPS(1); break;
    case 4: MT_UPS(4); break;
    case 9: MT_UPS(9); break;
    default: MT_UPS(16); break;
  }
#undef MT_UPS
  return mt::check_launch("mt_upscale_parameters");
}
