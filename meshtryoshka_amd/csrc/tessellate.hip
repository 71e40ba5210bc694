// Multi-level-set marching cubes over the sparse column structure.
//
// Column c owns the edges from its origin vertex lower[0][c] along +x/+y/+z to
// lower[1..3][c] (negative = no edge). For level set s an edge crosses iff exactly one
// end is below the iso value; the triangle vertex is placed at
// t = clamp((iso - d0) / (d1 - d0), 0, 1) and written at
// vertex_offsets[s * N + c] + (number of crossing axes before this one).
#include "common.h"
#include "sh.h"

#define MT_MC_TABLE_QUALIFIER static __device__ const
#include "mc_tables.h"

namespace mt {
namespace {

constexpr int kBlock = 256;

struct ColumnSdf {
  float d0;
  float d[3];
  int32_t v0;
  int32_t v[3];
};

__device__ __forceinline__ ColumnSdf load_column(const float* sdf, const int32_t* lower,
                                                 int64_t n, int64_t c) {
  ColumnSdf col;
  col.v0 = lower[c];
  col.d0 = sdf[col.v0];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    col.v[a] = lower[(a + 1) * n + c];
    col.d[a] = (col.v[a] >= 0) ? sdf[col.v[a]] : col.d0;
  }
  return col;
}

__device__ __forceinline__ bool crosses(float d0, float d1, float iso) {
  return (d0 < iso) != (d1 < iso);
}

__device__ __forceinline__ float edge_t(float d0, float d1, float iso) {
  if (d0 == d1) return 0.5f;
  const float t = (iso - d0) / (d1 - d0);
  return fminf(fmaxf(t, 0.0f), 1.0f);
}

// axes (may be NULL): bit a of axes[s n + c] set when the column's axis-a edge crosses
// level set s, i.e. when the column has a vertex of type a there (vertices are written in
// axis order, so its index is the column's first vertex plus the set bits below a).
__global__ void __launch_bounds__(kBlock) count_vertices_kernel(const float* sdf, const int32_t* lower,
                                                                int64_t n, const float* level_sets,
                                                                int S, int32_t* counts,
                                                                uint8_t* axes) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const ColumnSdf col = load_column(sdf, lower, n, c);
  for (int s = 0; s < S; ++s) {
    const float iso = level_sets[s];
    const int x0 = crosses(col.d0, col.d[0], iso), x1 = crosses(col.d0, col.d[1], iso);
    const int x2 = crosses(col.d0, col.d[2], iso);
    counts[s * n + c] = x0 + x1 + x2;
    if (axes) axes[s * n + c] = static_cast<uint8_t>(x0 | (x1 << 1) | (x2 << 2));
  }
}

// A block owns 256 consecutive columns. For each level set their vertices occupy one
// contiguous output range (the offsets are a scan in (level set, column) order), so the
// block first lists its crossings in LDS in output order, then writes positions, SDF,
// types and every SH plane of that range with consecutive threads on consecutive words.
__global__ void __launch_bounds__(kBlock) create_vertices_kernel(
    const float* gv, const float* gsdf, const float* gsh, int64_t ng, int nsh, const int32_t* lower,
    int64_t n, const float* level_sets, int S, const int32_t* offsets, int64_t nt, float* ov,
    float* od, float* osh, uint8_t* otype, int32_t* overflow) {
  __shared__ int32_t s_v0[3 * kBlock], s_v1[3 * kBlock];
  __shared__ float s_t[3 * kBlock], s_d0[3 * kBlock], s_d1[3 * kBlock];
  __shared__ uint8_t s_axis[3 * kBlock];
  __shared__ int s_total;
  const int64_t c_first = int64_t(blockIdx.x) * kBlock;
  const int64_t c = c_first + threadIdx.x;
  const int64_t c_last = (c_first + kBlock < n ? c_first + kBlock : n) - 1;
  const bool valid = c < n;
  ColumnSdf col{};
  if (valid) col = load_column(gsdf, lower, n, c);
  for (int s = 0; s < S; ++s) {
    const float iso = level_sets[s];
    const int64_t base = offsets[int64_t(s) * n + c_first];
    int k = 0;
    if (valid) {
      int pos = static_cast<int>(offsets[int64_t(s) * n + c] - base);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (!crosses(col.d0, col.d[a], iso)) continue;
        s_v0[pos] = col.v0;
        s_v1[pos] = col.v[a];
        s_t[pos] = edge_t(col.d0, col.d[a], iso);
        s_d0[pos] = col.d0;
        s_d1[pos] = col.d[a];
        s_axis[pos] = static_cast<uint8_t>(a);
        ++pos;
        ++k;
      }
      if (c == c_last) s_total = pos;
    }
    __syncthreads();
    const int total = s_total;
    if (base + total > nt && threadIdx.x == 0 && overflow) atomicOr(overflow, 1);
    const int live = static_cast<int>(base >= nt ? 0 : (base + total <= nt ? total : nt - base));
    for (int v = threadIdx.x; v < live; v += kBlock) {
      const int64_t idx = base + v;
      const int64_t g0 = s_v0[v], g1 = s_v1[v];
      const float t = s_t[v];
      const float x0 = gv[3 * g0 + 0], y0 = gv[3 * g0 + 1], z0 = gv[3 * g0 + 2];
      const float x1 = gv[3 * g1 + 0], y1 = gv[3 * g1 + 1], z1 = gv[3 * g1 + 2];
      ov[3 * idx + 0] = x0 + t * (x1 - x0);
      ov[3 * idx + 1] = y0 + t * (y1 - y0);
      ov[3 * idx + 2] = z0 + t * (z1 - z0);
      od[idx] = s_d0[v] + t * (s_d1[v] - s_d0[v]);
      otype[idx] = s_axis[v];
    }
    for (int q = 0; q < nsh; ++q) {
      const float* plane = gsh + int64_t(q) * ng * 3;
      float* out = osh + (int64_t(q) * nt + base) * 3;
      for (int i = threadIdx.x; i < 3 * live; i += kBlock) {
        const int v = i / 3, ch = i - 3 * v;
        const float t = s_t[v];
        const float a = plane[3 * int64_t(s_v0[v]) + ch];
        const float b = plane[3 * int64_t(s_v1[v]) + ch];
        out[i] = a + t * (b - a);
      }
    }
    __syncthreads();  // the next level set reuses the lists
  }
}

__global__ void __launch_bounds__(kBlock) create_vertices_backward_kernel(
    const float* gsdf, int64_t ng, int nsh, const int32_t* lower, int64_t n,
    const float* level_sets, int S, const int32_t* offsets, int64_t nt, const float* g_v,
    const float* g_d, const float* g_sh, float* gg_v, float* gg_d, float* gg_sh,
    const int32_t* skip_unless_flag) {
  if (skip_unless_flag != nullptr && *skip_unless_flag == 0) return;
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const ColumnSdf col = load_column(gsdf, lower, n, c);
  const bool want_v = g_v != nullptr && gg_v != nullptr;
  const bool want_d = g_d != nullptr && gg_d != nullptr;
  const bool want_sh = g_sh != nullptr && gg_sh != nullptr;
  // The origin's gradient is accumulated over all of the column's crossings, then
  // added once; each axis end receives the sum over level sets.
  float acc0_v[3] = {0.f, 0.f, 0.f};
  float acc0_d = 0.f;
  bool any = false;
  for (int a = 0; a < 3; ++a) {
    if (col.v[a] < 0) continue;
    const int64_t g1 = col.v[a];
    float acc1_v[3] = {0.f, 0.f, 0.f};
    float acc1_d = 0.f;
    bool axis_any = false;
    for (int s = 0; s < S; ++s) {
      const float iso = level_sets[s];
      if (!crosses(col.d0, col.d[a], iso)) continue;
      int64_t idx = offsets[s * n + c];
      for (int b = 0; b < a; ++b) idx += crosses(col.d0, col.d[b], iso) ? 1 : 0;
      if (idx >= nt) continue;
      const float t = edge_t(col.d0, col.d[a], iso);
      const float w0 = 1.0f - t;
      axis_any = true;
      if (want_v) {
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
          const float g = g_v[3 * idx + ch];
          acc0_v[ch] += w0 * g;
          acc1_v[ch] += t * g;
        }
      }
      if (want_d) {
        const float g = g_d[idx];
        acc0_d += w0 * g;
        acc1_d += t * g;
      }
      if (want_sh) {
        for (int k = 0; k < nsh; ++k) {
          const float* gs = g_sh + (int64_t(k) * nt + idx) * 3;
          float* o0 = gg_sh + (int64_t(k) * ng + col.v0) * 3;
          float* o1 = gg_sh + (int64_t(k) * ng + g1) * 3;
#pragma unroll
          for (int ch = 0; ch < 3; ++ch) {
            const float g = gs[ch];
            atomicAdd(o0 + ch, w0 * g);
            atomicAdd(o1 + ch, t * g);
          }
        }
      }
    }
    if (!axis_any) continue;
    any = true;
    if (want_v) {
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) atomicAdd(gg_v + 3 * g1 + ch, acc1_v[ch]);
    }
    if (want_d) atomicAdd(gg_d + g1, acc1_d);
  }
  if (!any) return;
  if (want_v) {
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) atomicAdd(gg_v + 3 * int64_t(col.v0) + ch, acc0_v[ch]);
  }
  if (want_d) atomicAdd(gg_d + col.v0, acc0_d);
}

// ------------------------------------------------------------------------------------
// Atomic-free adjoint. Each grid vertex g gathers from the (at most four) columns that
// reference it: the column whose origin is g, and the columns whose +x / +y / +z end is g.
// The inverse maps are built per call by a scatter over columns; if two columns claim the
// same slot (a structure whose origins are not a permutation of the grid vertices) a flag
// is raised and the atomic-scatter kernel above runs instead (decided on the device).
// ------------------------------------------------------------------------------------

__global__ void fill_maps_kernel(int32_t* maps, int64_t n_words) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n_words) maps[i] = -1;
  if (i == 0) maps[n_words] = 0;  // collision flag
}

__global__ void __launch_bounds__(kBlock) build_maps_kernel(const int32_t* lower, int64_t n,
                                                            int64_t ng, int32_t* maps) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  int32_t* flag = maps + 4 * ng;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int32_t g = lower[a * n + c];
    if (g < 0 || g >= ng) {
      if (a == 0) atomicOr(flag, 1);  // an origin outside the grid: use the scatter path
      continue;
    }
    if (atomicCAS(maps + a * ng + g, -1, static_cast<int32_t>(c)) != -1) atomicOr(flag, 1);
  }
}

template <int kNsh>
__global__ void __launch_bounds__(kBlock) gather_vertices_backward_kernel(
    const float* gsdf, int64_t ng, const int32_t* lower, int64_t n, const float* level_sets,
    int S, const int32_t* offsets, int64_t nt, const float* g_v, const float* g_d,
    const float* g_sh, float* gg_v, float* gg_d, float* gg_sh, const int32_t* maps) {
  if (maps[4 * ng] != 0) return;  // not injective: the scatter kernel handles it
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  float acc_v[3] = {0.f, 0.f, 0.f};
  float acc_d = 0.f;
  float acc_sh[kNsh][3];
#pragma unroll
  for (int k = 0; k < kNsh; ++k) acc_sh[k][0] = acc_sh[k][1] = acc_sh[k][2] = 0.f;
  // Role 0: g is the origin (weight 1 - t) of every crossing of its column; roles 1..3:
  // g is the +x / +y / +z end (weight t) of that axis' crossings of another column.
  for (int role = 0; role < 4; ++role) {
    const int32_t c = maps[role * ng + g];
    if (c < 0) continue;
    const ColumnSdf col = load_column(gsdf, lower, n, c);
    for (int s = 0; s < S; ++s) {
      const float iso = level_sets[s];
      bool cr[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) cr[a] = crosses(col.d0, col.d[a], iso);
      int64_t idx = offsets[s * n + c];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (!cr[a]) continue;
        const int64_t here = idx++;
        if (role != 0 && role != a + 1) continue;
        if (here >= nt) continue;
        const float t = edge_t(col.d0, col.d[a], iso);
        const float wgt = (role == 0) ? 1.0f - t : t;
        if (g_v) {
#pragma unroll
          for (int ch = 0; ch < 3; ++ch) acc_v[ch] += wgt * g_v[3 * here + ch];
        }
        if (g_d) acc_d += wgt * g_d[here];
        if (g_sh) {
#pragma unroll
          for (int k = 0; k < kNsh; ++k) {
            const float* gs = g_sh + (int64_t(k) * nt + here) * 3;
            acc_sh[k][0] += wgt * gs[0];
            acc_sh[k][1] += wgt * gs[1];
            acc_sh[k][2] += wgt * gs[2];
          }
        }
      }
    }
  }
  if (gg_v && g_v) {
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) gg_v[3 * g + ch] = acc_v[ch];
  }
  if (gg_d && g_d) gg_d[g] = acc_d;
  if (gg_sh && g_sh) {
#pragma unroll
    for (int k = 0; k < kNsh; ++k) {
      float* o = gg_sh + (int64_t(k) * ng + g) * 3;
      o[0] = acc_sh[k][0];
      o[1] = acc_sh[k][1];
      o[2] = acc_sh[k][2];
    }
  }
}

// ------------------------------------------------------------------------------------
// Batched training step: triangle vertices fused with their colours, and the adjoint of
// both. The per-vertex SH array [NSH, V, 3] is never stored: the forward interpolates each
// vertex's coefficients from the grid while it has the vertex in registers and writes only
// the colours of the B cameras; the adjoint recomputes the SH gradient of each triangle
// vertex from its colour gradient (sh_backward_kernel's arithmetic) and distributes it to
// the two grid vertices of its edge. Colours are bit-identical to the separate kernels;
// gradients equal them up to float summation order.
// ------------------------------------------------------------------------------------

// The block's 256 column origins' SH rows are staged once in LDS (read by every crossing
// of the column, over all level sets, and by the +x crossing of the previous column);
// only the other axis ends' rows are read from global memory per vertex.
// 128 columns per block: the staged origin rows take 25 KB of LDS, so four or five blocks
// share a CU (256 columns: two).
constexpr int kVcBlock = 128;

// SH rows: a grid vertex's (deg + 1)^2 coefficients x 3 channels contiguous (coefficient-
// major, then channel), zero-padded to whole float4s. The fused forward gathers one row
// per axis-end vertex (one or two cache lines) instead of one 12-byte entry from each of
// the (deg + 1)^2 coefficient planes of grid_sh [NSH, V, 3].
__host__ __device__ constexpr int sh_row_floats(int deg) { return (3 * (deg + 1) * (deg + 1) + 3) / 4 * 4; }

constexpr int kPackVerts = 64;               // grid vertices per block
constexpr int kPackBlock = 3 * kPackVerts;  // one thread per float of a plane's range

// grid_sh [NSH, V, 3] -> rows [V, sh_row_floats(deg)] through LDS: each coefficient plane's
// 64 x 3 floats are read contiguously, the block's 64 rows written as contiguous float4s.
// The staged rows are one float longer than the rows written, so that the plane reads'
// LDS stores spread over the banks.
template <int kDeg>
__global__ void __launch_bounds__(kPackBlock) pack_sh_rows_kernel(const float* gsh, int64_t ng,
                                                                 float* rows) {
  constexpr int kN = (kDeg + 1) * (kDeg + 1);
  constexpr int kR = sh_row_floats(kDeg);
  constexpr int kS = kR + 1;
  __shared__ float s[kPackVerts * kS];
  const int64_t g0 = int64_t(blockIdx.x) * kPackVerts;
  const int nv = static_cast<int>(ng - g0 < kPackVerts ? ng - g0 : kPackVerts);
  const int r = threadIdx.x;  // float r of the block's range of each plane
#pragma unroll
  for (int k = 0; k < kN; ++k)
    if (r < 3 * nv) s[(r / 3) * kS + 3 * k + r % 3] = gsh[(int64_t(k) * ng + g0) * 3 + r];
  __syncthreads();
  float4* out = reinterpret_cast<float4*>(rows + g0 * kR);
  for (int q = threadIdx.x; q < nv * kR / 4; q += kPackBlock) {
    const int row = (4 * q) / kR, col = 4 * q - row * kR;  // a float4 never spans two rows
    const float* src = s + row * kS + col;
    float4 x;
    x.x = col + 0 < 3 * kN ? src[0] : 0.0f;
    x.y = col + 1 < 3 * kN ? src[1] : 0.0f;
    x.z = col + 2 < 3 * kN ? src[2] : 0.0f;
    x.w = col + 3 < 3 * kN ? src[3] : 0.0f;
    out[q] = x;
  }
}

// kRows: gsh holds SH rows (pack_sh_rows_kernel's layout) instead of [NSH, V, 3] planes.
template <int kDeg, bool kRows>
__global__ void __launch_bounds__(kVcBlock) create_vertices_colors_kernel(
    const float* gv, const float* gsdf, const float* gsh, int64_t ng, const int32_t* lower,
    int64_t n, const float* level_sets, int S, const int32_t* offsets, int64_t nt,
    const float* cams, int B, float* ov, float* od, uint8_t* otype, float* colors,
    int32_t* overflow, int parts, int cstride, uint32_t* words) {
  constexpr int kN = (kDeg + 1) * (kDeg + 1);
  constexpr int kRow = 3 * kN + 1;  // padded: consecutive rows start in different banks
  constexpr int kR = sh_row_floats(kDeg);
  const bool write_vertices = (parts & 1) != 0;
  const bool write_colors = (parts & 2) != 0 && B > 0;
  __shared__ float s_sh[kVcBlock * kRow];
  __shared__ int32_t s_g0[kVcBlock];
  __shared__ float s_c0[kVcBlock];  // the column's origin SDF
  __shared__ int32_t s_v1[3 * kVcBlock];
  __shared__ float s_t[3 * kVcBlock], s_d1[3 * kVcBlock];
  __shared__ uint8_t s_axis[3 * kVcBlock], s_col[3 * kVcBlock];
  __shared__ int s_total;
  const int64_t c_first = int64_t(blockIdx.x) * kVcBlock;
  const int64_t c = c_first + threadIdx.x;
  const int64_t c_last = (c_first + kVcBlock < n ? c_first + kVcBlock : n) - 1;
  const bool valid = c < n;
  ColumnSdf col{};
  if (valid) col = load_column(gsdf, lower, n, c);
  if (valid && write_colors) {
    if constexpr (kRows) {
      const float4* src = reinterpret_cast<const float4*>(gsh + int64_t(col.v0) * kR);
#pragma unroll
      for (int q = 0; q < kR / 4; ++q) {
        const float4 x = src[q];
        const float e[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (4 * q + i < 3 * kN) s_sh[threadIdx.x * kRow + 4 * q + i] = e[i];
      }
    } else {
#pragma unroll
      for (int k = 0; k < kN; ++k) {
        const float* plane = gsh + int64_t(k) * ng * 3;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) s_sh[threadIdx.x * kRow + 3 * k + ch] = plane[3 * int64_t(col.v0) + ch];
      }
    }
  }
  s_g0[threadIdx.x] = valid ? col.v0 : -1;
  s_c0[threadIdx.x] = col.d0;
  for (int s = 0; s < S; ++s) {
    const float iso = level_sets[s];
    const int64_t base = offsets[int64_t(s) * n + c_first];
    if (valid) {
      const int32_t first = offsets[int64_t(s) * n + c];
      int pos = static_cast<int>(first - base);
      uint32_t axes = 0;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (!crosses(col.d0, col.d[a], iso)) continue;
        axes |= 1u << a;
        s_v1[pos] = col.v[a];
        s_t[pos] = edge_t(col.d0, col.d[a], iso);
        s_d1[pos] = col.d[a];
        s_axis[pos] = static_cast<uint8_t>(a);
        s_col[pos] = static_cast<uint8_t>(threadIdx.x);
        ++pos;
      }
      if (c == c_last) s_total = pos;
      if (words != nullptr) {  // offsets past the word's 29 bits only occur on overflow
        constexpr int32_t kMaxFirst = int32_t(MT_COLUMN_WORD_MAX_VERTICES - 1);
        const int32_t f = first < 0 || first > kMaxFirst ? kMaxFirst : first;
        words[int64_t(s) * n + c] = (static_cast<uint32_t>(f) << 3) | axes;
      }
    }
    __syncthreads();
    const int total = s_total;
    if (base + total > nt && threadIdx.x == 0 && overflow) atomicOr(overflow, 1);
    const int live = static_cast<int>(base >= nt ? 0 : (base + total <= nt ? total : nt - base));
    for (int v = threadIdx.x; v < live; v += kVcBlock) {
      const int64_t idx = base + v;
      const int lc = s_col[v];
      const int64_t g0 = s_g0[lc], g1 = s_v1[v];
      const float t = s_t[v];
      const float x0 = gv[3 * g0 + 0], y0 = gv[3 * g0 + 1], z0 = gv[3 * g0 + 2];
      const float x1 = gv[3 * g1 + 0], y1 = gv[3 * g1 + 1], z1 = gv[3 * g1 + 2];
      const float3 p = make_float3(x0 + t * (x1 - x0), y0 + t * (y1 - y0), z0 + t * (z1 - z0));
      if (write_vertices) {
        ov[3 * idx + 0] = p.x;
        ov[3 * idx + 1] = p.y;
        ov[3 * idx + 2] = p.z;
        const float d0 = s_c0[lc];
        od[idx] = d0 + t * (s_d1[v] - d0);
        otype[idx] = s_axis[v];
      }
      if (!write_colors) continue;
      // The axis end's row: the next column's staged origin when that is the end vertex.
      const int nb = (lc + 1 < kVcBlock && s_g0[lc + 1] == g1) ? lc + 1 : -1;
      float coef[kN][3];
      if constexpr (kRows) {
        float e[kR];
        if (nb >= 0) {
#pragma unroll
          for (int i = 0; i < 3 * kN; ++i) e[i] = s_sh[nb * kRow + i];
        } else {
          const float4* src = reinterpret_cast<const float4*>(gsh + g1 * kR);
#pragma unroll
          for (int q = 0; q < kR / 4; ++q) {
            const float4 x = src[q];
            e[4 * q + 0] = x.x;
            e[4 * q + 1] = x.y;
            e[4 * q + 2] = x.z;
            e[4 * q + 3] = x.w;
          }
        }
#pragma unroll
        for (int k = 0; k < kN; ++k) {
#pragma unroll
          for (int ch = 0; ch < 3; ++ch) {
            const float a = s_sh[lc * kRow + 3 * k + ch];
            coef[k][ch] = a + t * (e[3 * k + ch] - a);
          }
        }
      } else {
#pragma unroll
        for (int k = 0; k < kN; ++k) {
#pragma unroll
          for (int ch = 0; ch < 3; ++ch) {
            const float a = s_sh[lc * kRow + 3 * k + ch];
            const float b = nb >= 0 ? s_sh[nb * kRow + 3 * k + ch] : gsh[(int64_t(k) * ng + g1) * 3 + ch];
            coef[k][ch] = a + t * (b - a);
          }
        }
      }
      // Camera b's colour at (b nt + idx) 3 (cstride 0) or (idx cstride + b) 3.
      float* cout = colors + 3 * (cstride > 0 ? idx * cstride : idx);
      const int64_t cstep = cstride > 0 ? 3 : 3 * nt;
      for (int b = 0; b < B; ++b, cout += cstep) {
        const float3 d = view_dir(p, cams + 3 * b);
        float basis[16];
        sh_basis(d.x, d.y, d.z, kDeg, basis);
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) cout[ch] = fmaxf(sh_channel_pre<kN>(basis, coef, ch), 0.0f);
      }
    }
    __syncthreads();  // the next level set reuses the lists
  }
}

// Rows of the colour adjoint's workspace: [SDF sum, SH sums (coefficient-major, then
// channel)], padded to whole float4s. A live column's four rows (axis ends x, y, z, then
// its origin) are contiguous at its position in the live-column list.
__host__ __device__ constexpr int adjoint_row_floats(int n_coef) { return (1 + 3 * n_coef + 3) / 4 * 4; }

template <int kN>
__device__ __forceinline__ void store_row(float* row, float d, const float (&acc)[kN][3]) {
  constexpr int kPad = adjoint_row_floats(kN);
  float r[kPad];
  r[0] = d;
#pragma unroll
  for (int k = 0; k < kN; ++k)
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) r[1 + 3 * k + ch] = acc[k][ch];
#pragma unroll
  for (int j = 1 + 3 * kN; j < kPad; ++j) r[j] = 0.f;
  float4* r4 = reinterpret_cast<float4*>(row);
#pragma unroll
  for (int j = 0; j < kPad / 4; ++j) r4[j] = make_float4(r[4 * j], r[4 * j + 1], r[4 * j + 2], r[4 * j + 3]);
}

template <int kN>
__device__ __forceinline__ void add_row(const float* row, float& d, float (&acc)[kN][3]) {
  constexpr int kPad = adjoint_row_floats(kN);
  float r[kPad];
  const float4* r4 = reinterpret_cast<const float4*>(row);
#pragma unroll
  for (int j = 0; j < kPad / 4; ++j) {
    const float4 q = r4[j];
    r[4 * j] = q.x;
    r[4 * j + 1] = q.y;
    r[4 * j + 2] = q.z;
    r[4 * j + 3] = q.w;
  }
  d += r[0];
#pragma unroll
  for (int k = 0; k < kN; ++k)
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) acc[k][ch] += r[1 + 3 * k + ch];
}

// Phase 0, one thread per column: bit a of col_live[c] is set when some vertex of the
// column's axis-a edge takes part (below the vertex count and, with touched flags, flagged);
// bit 3 s + a of col_slots[c] when the one of level set s does (S <= kSlotLevels). The
// columns with any set bit are appended to list (one atomic per wave).
constexpr int kSlotLevels = 10;

// Phase 0 for column c: bits of col_live[c] and col_slots[c] (written here); returns the live bits.
__device__ __forceinline__ uint32_t adjoint_column_mask(
    const float* gsdf, const int32_t* lower, int64_t n, const float* level_sets, int S,
    const int32_t* offsets, int64_t nt, const uint8_t* touched, uint8_t* col_live,
    uint32_t* col_slots, int64_t c) {
  uint32_t bits = 0, slots = 0;
  if (c < n) {
    const ColumnSdf col = load_column(gsdf, lower, n, c);
    int s0 = 0;
    if (S <= kSlotLevels && touched != nullptr) {
      // All level sets' first vertices, then every flag, are loaded before any is used.
      int32_t first[kSlotLevels];
#pragma unroll
      for (int s = 0; s < kSlotLevels; ++s) first[s] = s < S ? offsets[s * n + c] : 0;
      uint8_t flag[kSlotLevels][3];
#pragma unroll
      for (int s = 0; s < kSlotLevels; ++s) {
        const float iso = s < S ? level_sets[s] : 0.f;
        int64_t here = first[s];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          flag[s][a] = 0;
          if (s >= S || !crosses(col.d0, col.d[a], iso)) continue;
          if (col.v[a] >= 0 && here < nt) flag[s][a] = touched[here];
          ++here;
        }
      }
#pragma unroll
      for (int s = 0; s < kSlotLevels; ++s)
#pragma unroll
        for (int a = 0; a < 3; ++a)
          if (flag[s][a]) {
            bits |= 1u << a;
            slots |= 1u << (3 * s + a);
          }
      s0 = S;
    }
    for (int s = s0; s < S; ++s) {
      const float iso = level_sets[s];
      int64_t here = offsets[s * n + c];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (!crosses(col.d0, col.d[a], iso)) continue;
        if (col.v[a] >= 0 && here < nt && (touched == nullptr || touched[here])) {
          bits |= 1u << a;
          if (s < kSlotLevels) slots |= 1u << (3 * s + a);
        }
        ++here;
      }
    }
    col_live[c] = static_cast<uint8_t>(bits);
    col_slots[c] = slots;
  }
  return bits;
}

// kMaskRounds columns per thread (consecutive runs of kBlock), so that a workgroup appends
// kMaskRounds x kBlock columns' live ones with one atomic on the shared list counter (the
// counter's atomics serialise).
constexpr int kMaskRounds = 4;

__global__ void __launch_bounds__(kBlock) colors_adjoint_mask_kernel(
    const float* gsdf, int64_t ng, const int32_t* lower, int64_t n, const float* level_sets,
    int S, const int32_t* offsets, int64_t nt, const uint8_t* touched, const int32_t* maps,
    uint8_t* col_live, uint32_t* col_slots, int32_t* list, int32_t* col_pos, int32_t* count) {
  (void)ng;
  (void)maps;
  const int64_t c0 = int64_t(blockIdx.x) * (kBlock * kMaskRounds) + threadIdx.x;
  uint32_t live[kMaskRounds];
#pragma unroll
  for (int r = 0; r < kMaskRounds; ++r) {
    const int64_t c = c0 + int64_t(r) * kBlock;
    live[r] = c < n ? adjoint_column_mask(gsdf, lower, n, level_sets, S, offsets, nt, touched,
                                          col_live, col_slots, c)
                    : 0u;
  }
  __shared__ int s_count[kMaskRounds][kBlock / kWave];
  __shared__ int s_base;
  const int lane = __lane_id();
  const int wave = threadIdx.x / kWave;
  uint64_t m[kMaskRounds];
#pragma unroll
  for (int r = 0; r < kMaskRounds; ++r) {
    m[r] = __ballot(live[r] != 0);
    if (lane == 0) s_count[r][wave] = __popcll(m[r]);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int total = 0;
#pragma unroll
    for (int r = 0; r < kMaskRounds; ++r)
#pragma unroll
      for (int w = 0; w < kBlock / kWave; ++w) total += s_count[r][w];
    s_base = total > 0 ? atomicAdd(count, total) : 0;
  }
  __syncthreads();
  int at = s_base;  // rounds in order, waves in order inside a round
#pragma unroll
  for (int r = 0; r < kMaskRounds; ++r) {
    int before = 0, round_total = 0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) {
      before += w < wave ? s_count[r][w] : 0;
      round_total += s_count[r][w];
    }
    if (live[r] != 0) {
      const int64_t c = c0 + int64_t(r) * kBlock;
      const int32_t i = at + before + __popcll(m[r] & lanemask_lt());
      list[i] = static_cast<int32_t>(c);
      col_pos[c] = i;
    }
    at += round_total;
  }
}

__global__ void zero_counts_kernel(int32_t* a, int32_t* b) {
  *a = 0;
  *b = 0;
}

// Phase 1, one thread per live column: the column's crossings are visited once (axis by
// axis, level sets in order); (1 - t) g is summed into the column's origin row and t g into
// the row of the axis end (rows: [SDF, SH coefficient-major then channel]) for phase 2. Each camera's term
// basis_k * gc is added to both sums directly (no per-vertex gradient array: 48 fewer
// registers, 3 waves per SIMD instead of 2).
template <int kNsh, int kDeg>
__global__ void __launch_bounds__(kBlock) colors_adjoint_columns_kernel(
    const float* gsdf, int64_t ng, const int32_t* lower, int64_t n, const float* level_sets,
    int S, const int32_t* offsets, int64_t nt, const float* tv, const float* colors,
    const float* cams, int B, float* g_d, float* g_col, float* part, const int32_t* list,
    const int32_t* count, const uint8_t* col_live, const uint32_t* col_slots, const int32_t* maps,
    const uint8_t* touched, bool clear, int cstride) {
  constexpr int kN = (kDeg + 1) * (kDeg + 1);
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= *count) return;
  const int64_t c = list[i];  // a column with some contributing vertex
  const uint32_t axes = col_live[c], slots = col_slots[c];
  const ColumnSdf col = load_column(gsdf, lower, n, c);
  float acc0_d = 0.f;
  float acc0[kN][3];
#pragma unroll
  for (int k = 0; k < kN; ++k) acc0[k][0] = acc0[k][1] = acc0[k][2] = 0.f;
  bool live0 = false;  // some vertex of the column takes part
  for (int a = 0; a < 3; ++a) {
    if (col.v[a] < 0 || !((axes >> a) & 1u)) continue;
    bool live = false;  // some vertex of this edge takes part
    float acce_d = 0.f;
    float acce[kN][3];
#pragma unroll
    for (int k = 0; k < kN; ++k) acce[k][0] = acce[k][1] = acce[k][2] = 0.f;
    for (int s = 0; s < S; ++s) {
      if (s < kSlotLevels && !((slots >> (3 * s + a)) & 1u)) continue;  // nothing to add
      const float iso = level_sets[s];
      if (!crosses(col.d0, col.d[a], iso)) continue;
      int64_t here = offsets[s * n + c];
      for (int q = 0; q < a; ++q) here += crosses(col.d0, col.d[q], iso) ? 1 : 0;
      if (here >= nt) continue;
      if (touched && !touched[here]) continue;  // no gradient reached this vertex
      live = true;
      const float t = edge_t(col.d0, col.d[a], iso);
      const float w0 = 1.0f - t;
      const float gd = g_d[here];
      if (clear) g_d[here] = 0.f;
      acc0_d += w0 * gd;
      acce_d += t * gd;
      const float3 p = vertex_at(tv, here);
      // Camera b's colour of this vertex: at (b nt + here) 3 (cstride 0) or (here cstride + b) 3.
      const int64_t first = 3 * (cstride > 0 ? here * cstride : here);
      const float* cb = colors + first;
      float* gb = g_col + first;
      const int64_t stride = cstride > 0 ? 3 : 3 * nt;
      for (int b = 0; b < B; ++b, cb += stride, gb += stride) {
        // The view direction with the hardware reciprocal square root (~1 ulp): the
        // gradient tolerates it, and it saves the correctly rounded sqrt and division that
        // the forward needs for bit-exact colours.
        const float dx = p.x - cams[3 * b + 0], dy = p.y - cams[3 * b + 1], dz = p.z - cams[3 * b + 2];
        const float inv = __builtin_amdgcn_rsqf(dx * dx + dy * dy + dz * dz);
        float basis[16];
        sh_basis(dx * inv, dy * inv, dz * inv, kDeg, basis);
        const float gr[3] = {gb[0], gb[1], gb[2]};
        if (clear) gb[0] = gb[1] = gb[2] = 0.f;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
          const float gc = cb[ch] > 0.0f ? gr[ch] : 0.0f;
          const float g0 = w0 * gc, g1 = t * gc;
          // Fused multiply-adds (the build keeps separate roundings elsewhere, for bit-exact
          // forward values; a gradient sum only needs the tolerance).
#pragma unroll
          for (int k = 0; k < kN; ++k) {
            acc0[k][ch] = __builtin_fmaf(basis[k], g0, acc0[k][ch]);
            acce[k][ch] = __builtin_fmaf(basis[k], g1, acce[k][ch]);
          }
        }
      }
    }
    // Rows of edges without a contributing vertex (col_live bit clear) stay unwritten.
    if (!live) continue;
    live0 = true;
    store_row<kN>(part + (i * 4 + a) * adjoint_row_floats(kN), acce_d, acce);
  }
  store_row<kN>(part + (i * 4 + 3) * adjoint_row_floats(kN), acc0_d, acc0);  // the origin's
}

// For accumulating calls: the grid vertices that some live row ends at (one atomic per
// block), so that phase 2 runs over them only.
__global__ void __launch_bounds__(kBlock) colors_adjoint_ends_mask_kernel(
    int64_t ng, const uint8_t* col_live, const int32_t* maps, int32_t* list, int32_t* count) {
  if (maps[4 * ng] != 0) return;
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  bool live = false;
  if (g < ng) {
    const int32_t own = maps[g];
    live = own >= 0 && col_live[own] != 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int32_t e = maps[(a + 1) * ng + g];
      live = live || (e >= 0 && ((col_live[e] >> a) & 1u));
    }
  }
  __shared__ int s_count[kBlock / kWave];
  __shared__ int s_base;
  const uint64_t m = __ballot(live);
  const int lane = __lane_id();
  const int wave = threadIdx.x / kWave;
  if (lane == 0) s_count[wave] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int total = 0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) total += s_count[w];
    s_base = total > 0 ? atomicAdd(count, total) : 0;
  }
  __syncthreads();
  int at = s_base;
#pragma unroll
  for (int w = 0; w < kBlock / kWave; ++w) at += w < wave ? s_count[w] : 0;
  if (live) list[at + __popcll(m & lanemask_lt())] = static_cast<int32_t>(g);
}

// Phase 2, one thread per grid vertex: add the axis-end sums of the (at most three)
// columns that end at it; a grid vertex that is no column's origin starts from zero.
template <int kNsh, int kDeg>
__global__ void __launch_bounds__(kBlock) colors_adjoint_ends_kernel(
    int64_t ng, int64_t n, const float* part, const uint8_t* col_live, const int32_t* col_pos,
    float* gg_d, float* gg_sh, const int32_t* maps, bool accumulate, int32_t* status,
    const int32_t* list, const int32_t* count) {
  constexpr int kN = (kDeg + 1) * (kDeg + 1);
  constexpr int kPad = adjoint_row_floats(kN);
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  (void)status;
  if (maps[4 * ng] != 0) return;  // not one-to-one: colors_adjoint_scatter_kernel adds the rows
  // With a list (accumulating calls): only the listed grid vertices, which have live ends.
  if (list != nullptr && i >= *count) return;
  const int64_t g = list != nullptr ? int64_t(list[i]) : i;
  if (g >= ng) return;
  // The sum: the origin row of the grid vertex's own column, then the axis-end rows of the
  // columns that end at it (axis order); rows of dead columns and edges are zero.
  float sd = 0.f;
  float v[kN][3];
#pragma unroll
  for (int k = 0; k < kN; ++k) v[k][0] = v[k][1] = v[k][2] = 0.f;
  const int32_t own = maps[g];
  if (own >= 0 && col_live[own] != 0)
    add_row<kN>(part + (int64_t(col_pos[own]) * 4 + 3) * kPad, sd, v);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int32_t e = maps[(a + 1) * ng + g];
    if (e >= 0 && ((col_live[e] >> a) & 1u))
      add_row<kN>(part + (int64_t(col_pos[e]) * 4 + a) * kPad, sd, v);
  }
  gg_d[g] = accumulate ? gg_d[g] + sd : sd;
  struct F3 {
    float x, y, z;
  };
#pragma unroll
  for (int k = 0; k < kNsh; ++k) {
    F3* o = reinterpret_cast<F3*>(gg_sh + (int64_t(k) * ng + g) * 3);
    F3 x{0.f, 0.f, 0.f};
    if (k < kN) x = F3{v[k < kN ? k : 0][0], v[k < kN ? k : 0][1], v[k < kN ? k : 0][2]};
    if (accumulate) {  // read-modify-write of the three channels
      const F3 y = *o;
      x = F3{y.x + x.x, y.y + x.y, y.z + x.z};
    }
    *o = x;  // one 12-byte store per plane
  }
}

// Phase 2 when some grid vertex is the origin or the +axis end of more than one column (a
// multi-component grid whose coincident vertices were welded): one thread per live column
// adds its rows to their grid vertices with atomics. A non-accumulating call clears the
// grid gradients first (clear_grid_gradients_kernel, also decided on the device).
__global__ void __launch_bounds__(kBlock) clear_grid_gradients_kernel(const int32_t* maps,
                                                                      int64_t ng, int num_sh,
                                                                      float* gg_d, float* gg_sh) {
  if (maps[4 * ng] == 0) return;
  const int64_t total = ng * (1 + 3 * int64_t(num_sh));
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < total; j += stride) {
    if (j < ng) gg_d[j] = 0.f;
    else gg_sh[j - ng] = 0.f;
  }
}

template <int kNsh, int kDeg>
__global__ void __launch_bounds__(kBlock) colors_adjoint_scatter_kernel(
    int64_t ng, const int32_t* lower, int64_t n, const float* part, const int32_t* list,
    const int32_t* count, const uint8_t* col_live, float* gg_d, float* gg_sh,
    const int32_t* maps) {
  constexpr int kN = (kDeg + 1) * (kDeg + 1);
  constexpr int kPad = adjoint_row_floats(kN);
  if (maps[4 * ng] == 0) return;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= *count) return;
  const int64_t c = list[i];
  const uint32_t axes = col_live[c];
  for (int r = 0; r < 4; ++r) {  // rows 0..2: the +x / +y / +z ends; row 3: the origin
    if (r < 3 && !((axes >> r) & 1u)) continue;  // unwritten row
    const int32_t g = r < 3 ? lower[(r + 1) * n + c] : lower[c];
    if (g < 0 || g >= ng) continue;
    float sd = 0.f;
    float v[kN][3];
#pragma unroll
    for (int k = 0; k < kN; ++k) v[k][0] = v[k][1] = v[k][2] = 0.f;
    add_row<kN>(part + (i * 4 + r) * kPad, sd, v);
    atomicAdd(gg_d + g, sd);
#pragma unroll
    for (int k = 0; k < kN; ++k)
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) atomicAdd(gg_sh + (int64_t(k) * ng + g) * 3 + ch, v[k][ch]);
  }
}

__global__ void __launch_bounds__(kBlock) count_faces_kernel(const float* sdf, const int32_t* lower,
                                                             int64_t ncol, const int32_t* upper,
                                                             int64_t n, const float* level_sets,
                                                             int S, int32_t* counts,
                                                             uint8_t* codes) {
  const int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= n) return;
  // Corner order (000, 100, 110, 010, 001, 101, 111, 011) in (x, y, z).
  float cs[8];
  cs[0] = sdf[lower[0 * ncol + v]];
  cs[1] = sdf[lower[1 * ncol + v]];
  cs[2] = sdf[upper[3 * n + v]];
  cs[3] = sdf[lower[2 * ncol + v]];
  cs[4] = sdf[lower[3 * ncol + v]];
  cs[5] = sdf[upper[2 * n + v]];
  cs[6] = sdf[upper[0 * n + v]];
  cs[7] = sdf[upper[1 * n + v]];
  for (int s = 0; s < S; ++s) {
    const float iso = level_sets[s];
    int code = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) code |= (cs[q] >= iso) ? (1 << q) : 0;
    codes[s * n + v] = static_cast<uint8_t>(code);
    counts[s * n + v] = MT_MC_CASE_START[code + 1] - MT_MC_CASE_START[code];
  }
}

// A block owns 256 consecutive voxels. For each level set their face corners occupy one
// contiguous output range (face offsets are a scan in (level set, voxel) order), so the
// corners are spread evenly over the block's threads: each thread finds its corner's
// voxel by binary search over the voxels' corner offsets in LDS, resolves the corner's
// triangle vertex, and consecutive threads write consecutive corners.
// Face offsets are int32 (the reference's layout) or int64 (more than 2^31 - 1 face
// corners); within one block they span at most 256 voxels x 15 corners.
// kMode 1 (kAxesMode): a corner's vertex is found from the owner column's crossing-axis
// bits (vtypes holds them, as count_vertices_kernel writes them; one byte beside the
// offsets) instead of by reading the types of its vertices, which saves one dependent load
// per corner. kMode 2 (kWordsMode): voffsets holds one word per (level set, column),
// (first vertex << 3) | crossing axes, so a corner costs one 4-byte load. voxel_indices may
// be NULL (not written).
constexpr int kTypesMode = 0, kAxesMode = 1, kWordsMode = 2;
template <typename FOff, int kMode>
__global__ void __launch_bounds__(kBlock) create_faces_kernel(
    const int32_t* neighbors, const int32_t* indices, int64_t n, int64_t ncol,
    const int32_t* voffsets, const uint8_t* vtypes, const uint8_t* codes, int S,
    const FOff* foffsets, int32_t* faces, int32_t* voxel_indices, int64_t fcap,
    int64_t vcap, int32_t* overflow) {
  constexpr int kMaxTri = 5 * kBlock;  // a marching-cubes case has at most 5 triangles
  __shared__ int32_t s_start[kBlock];
  __shared__ uint8_t s_code[kBlock];
  __shared__ uint8_t s_owner[kMaxTri];  // the block-local voxel of each of the block's triangles
  // The voxels' neighbour columns (slots 0..5: +x, +y, +z, +xy, +yz, +xz), read once for all
  // level sets, so that a corner's owner column costs no dependent global load.
  __shared__ int32_t s_nbr[6][kBlock];
  __shared__ int s_total;
  const int64_t v_first = int64_t(blockIdx.x) * kBlock;
  const int last = static_cast<int>((v_first + kBlock < n ? v_first + kBlock : n) - 1 - v_first);
  const int64_t v = v_first + threadIdx.x;
  const bool valid = int(threadIdx.x) <= last;
  if (valid) {
#pragma unroll
    for (int q = 0; q < 6; ++q) s_nbr[q][threadIdx.x] = neighbors[q * n + v];
  }
  for (int s = 0; s < S; ++s) {
    const int64_t base = foffsets[int64_t(s) * n + v_first];
    if (valid) {
      const int code = codes[int64_t(s) * n + v];
      const int start = static_cast<int>(foffsets[int64_t(s) * n + v] - base);
      const int corners = MT_MC_CASE_START[code + 1] - MT_MC_CASE_START[code];
      s_code[threadIdx.x] = static_cast<uint8_t>(code);
      s_start[threadIdx.x] = start;
      for (int t = start / 3; t < (start + corners) / 3; ++t)
        if (t < kMaxTri) s_owner[t] = static_cast<uint8_t>(threadIdx.x);
      if (int(threadIdx.x) == last) s_total = start + corners;
    }
    __syncthreads();
    const int total = s_total;
    // One thread per triangle: its three corners are consecutive corners of one voxel (the
    // block's range starts on a triangle, as every voxel contributes 3 per triangle).
    for (int t = threadIdx.x; 3 * t < total; t += kBlock) {
      const int i = 3 * t;
      int lo;  // the voxel whose corners hold corner i
      if (t < kMaxTri) {
        lo = s_owner[t];
      } else {  // never taken (at most 5 triangles per voxel): binary search
        lo = 0;
        int hi = last;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (s_start[mid] <= i) lo = mid; else hi = mid - 1;
        }
      }
      const int q0 = i - s_start[lo];
      const int code = s_code[lo];
      const int64_t vv = v_first + lo;
      int32_t target[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int e = MT_MC_CASE_EDGES[MT_MC_CASE_START[code] + q0 + c];
        const int slot = MT_MC_EDGE_OWNER[e][0];
        const int axis = MT_MC_EDGE_OWNER[e][1];
        const int64_t owner = (slot < 0) ? vv : s_nbr[slot][lo];
        const int64_t vo = s * ncol + owner;
        target[c] = -1;
        if constexpr (kMode != kTypesMode) {
          int32_t start;
          uint32_t m;  // the owner's crossing axes at this level set
          if constexpr (kMode == kWordsMode) {
            const uint32_t w = reinterpret_cast<const uint32_t*>(voffsets)[vo];
            start = static_cast<int32_t>(w >> 3);
            m = w & 7u;
          } else {
            start = voffsets[vo];
            m = vtypes[vo];
          }
          if (int64_t(start) + __popc(m) > vcap) {  // never written
            if (overflow) atomicOr(overflow, 1);
          } else if ((m >> axis) & 1u) {
            target[c] = start + __popc(m & ((1u << axis) - 1u));
          }
          continue;
        }
        const int32_t start = voffsets[vo];
        int32_t cnt = voffsets[vo + 1] - start;
        if (int64_t(start) + cnt > vcap) {  // vertices past the capacity were never written
          if (overflow) atomicOr(overflow, 1);
          cnt = 0;
        }
        // A column has at most one vertex per axis, written in axis order: its (at most
        // three) types are read together, so the three corners' lookups overlap instead of
        // waiting on each other load by load. Same result as a first-match search.
        const int t0 = cnt > 0 ? vtypes[start + 0] : -1;
        const int t1 = cnt > 1 ? vtypes[start + 1] : -1;
        const int t2 = cnt > 2 ? vtypes[start + 2] : -1;
        target[c] = t0 == axis ? start : t1 == axis ? start + 1 : t2 == axis ? start + 2 : -1;
        for (int k = 3; k < cnt && target[c] < 0; ++k)  // never taken: counts are <= 3
          if (vtypes[start + k] == axis) target[c] = start + k;
      }
      const int64_t tri = (base + i) / 3;
      if (tri >= fcap) {
        if (overflow) atomicOr(overflow, 1);
        continue;
      }
      // Corner c of the case lands at position 2 - c (the reference's reversed winding).
      faces[3 * tri + 0] = target[2];
      faces[3 * tri + 1] = target[1];
      faces[3 * tri + 2] = target[0];
      if (voxel_indices) voxel_indices[tri] = indices[vv];
    }
    __syncthreads();  // the next level set reuses the LDS lists
  }
}

}  // namespace
}  // namespace mt

extern "C" int mt_count_triangle_vertices(const float* sdf, const int32_t* lower_corners,
                                          int64_t num_columns, const float* level_sets, int32_t S,
                                          int32_t* counts, mt_stream_t stream) {
  MT_REQUIRE(num_columns >= 0 && S >= 0, "mt_count_triangle_vertices: negative size");
  MT_REQUIRE(int64_t(S) * num_columns < (int64_t(1) << 31),
             "mt_count_triangle_vertices: S * columns exceeds int32 offsets");
  if (num_columns == 0 || S == 0) return MT_OK;
  mt::count_vertices_kernel<<<mt::blocks_for(num_columns, mt::kBlock), mt::kBlock, 0,
                              static_cast<hipStream_t>(stream)>>>(sdf, lower_corners, num_columns,
                                                                  level_sets, S, counts, nullptr);
  return mt::check_launch("mt_count_triangle_vertices");
}

extern "C" int mt_count_triangle_vertices_axes(const float* sdf, const int32_t* lower_corners,
                                               int64_t num_columns, const float* level_sets,
                                               int32_t S, int32_t* counts, uint8_t* axes,
                                               mt_stream_t stream) {
  MT_REQUIRE(num_columns >= 0 && S >= 0, "mt_count_triangle_vertices_axes: negative size");
  MT_REQUIRE(int64_t(S) * num_columns < (int64_t(1) << 31),
             "mt_count_triangle_vertices_axes: S * columns exceeds int32 offsets");
  MT_REQUIRE(axes != nullptr || num_columns == 0 || S == 0,
             "mt_count_triangle_vertices_axes: null axes");
  if (num_columns == 0 || S == 0) return MT_OK;
  mt::count_vertices_kernel<<<mt::blocks_for(num_columns, mt::kBlock), mt::kBlock, 0,
                              static_cast<hipStream_t>(stream)>>>(sdf, lower_corners, num_columns,
                                                                  level_sets, S, counts, axes);
  return mt::check_launch("mt_count_triangle_vertices_axes");
}

extern "C" int mt_create_triangle_vertices_forward(
    const float* grid_vertices, const float* grid_sdf, const float* grid_sh,
    int64_t num_grid_vertices, int32_t num_sh, const int32_t* lower_corners, int64_t num_columns,
    const float* level_sets, int32_t S, const int32_t* vertex_offsets,
    int64_t num_triangle_vertices, float* out_vertices, float* out_sdf, float* out_sh,
    uint8_t* out_types, int32_t* overflow, mt_stream_t stream) {
  MT_REQUIRE(num_sh >= 0 && S >= 0 && num_columns >= 0,
             "mt_create_triangle_vertices_forward: negative size");
  if (num_columns == 0 || S == 0 || num_triangle_vertices == 0) return MT_OK;
  mt::create_vertices_kernel<<<mt::blocks_for(num_columns, mt::kBlock), mt::kBlock, 0,
                               static_cast<hipStream_t>(stream)>>>(
      grid_vertices, grid_sdf, grid_sh, num_grid_vertices, num_sh, lower_corners, num_columns,
      level_sets, S, vertex_offsets, num_triangle_vertices, out_vertices, out_sdf, out_sh,
      out_types, overflow);
  return mt::check_launch("mt_create_triangle_vertices_forward");
}

extern "C" size_t mt_create_triangle_vertices_backward_workspace_bytes(int64_t num_grid_vertices) {
  return static_cast<size_t>(4 * num_grid_vertices + 1) * sizeof(int32_t);
}

extern "C" int mt_create_triangle_vertices_backward(
    const float* grid_sdf, int64_t num_grid_vertices, int32_t num_sh, const int32_t* lower_corners,
    int64_t num_columns, const float* level_sets, int32_t S, const int32_t* vertex_offsets,
    int64_t num_triangle_vertices, const float* grad_vertices, const float* grad_sdf,
    const float* grad_sh, float* grid_vertices_grad, float* grid_sdf_grad, float* grid_sh_grad,
    void* workspace, size_t workspace_bytes, mt_stream_t stream) {
  MT_REQUIRE(num_sh >= 0 && S >= 0 && num_columns >= 0,
             "mt_create_triangle_vertices_backward: negative size");
  if (num_columns == 0 || S == 0 || num_triangle_vertices == 0) return MT_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const float* gv = grid_vertices_grad ? grad_vertices : nullptr;
  const float* gd = grid_sdf_grad ? grad_sdf : nullptr;
  const float* gs = grid_sh_grad ? grad_sh : nullptr;
  const size_t need = mt_create_triangle_vertices_backward_workspace_bytes(num_grid_vertices);
  const bool gather = workspace != nullptr && workspace_bytes >= need &&
                      (num_sh == 1 || num_sh == 4 || num_sh == 9 || num_sh == 16);
  if (!gather) {
    mt::create_vertices_backward_kernel<<<mt::blocks_for(num_columns, mt::kBlock), mt::kBlock, 0, st>>>(
        grid_sdf, num_grid_vertices, num_sh, lower_corners, num_columns, level_sets, S,
        vertex_offsets, num_triangle_vertices, gv, gd, gs, grid_vertices_grad, grid_sdf_grad,
        grid_sh_grad, nullptr);
    return mt::check_launch("mt_create_triangle_vertices_backward");
  }
  int32_t* maps = static_cast<int32_t*>(workspace);
  const int64_t words = 4 * num_grid_vertices;
  mt::fill_maps_kernel<<<mt::blocks_for(words, 256), 256, 0, st>>>(maps, words);
  mt::build_maps_kernel<<<mt::blocks_for(num_columns, mt::kBlock), mt::kBlock, 0, st>>>(
      lower_corners, num_columns, num_grid_vertices, maps);
  const dim3 ggrid(mt::blocks_for(num_grid_vertices, mt::kBlock));
#define MT_GATHER(N)                                                                          \
  mt::gather_vertices_backward_kernel<N><<<ggrid, mt::kBlock, 0, st>>>(                       \
      grid_sdf, num_grid_vertices, lower_corners, num_columns, level_sets, S, vertex_offsets, \
      num_triangle_vertices, gv, gd, gs, grid_vertices_grad, grid_sdf_grad, grid_sh_grad, maps)
  switch (num_sh) {
    case 1: MT_GATHER(1); break;
    case 4: MT_GATHER(4); break;
    case 9: MT_GATHER(9); break;
    default: MT_GATHER(16); break;
  }
#undef MT_GATHER
  // Runs only if the maps found a collision (device-side decision).
  mt::create_vertices_backward_kernel<<<mt::blocks_for(num_columns, mt::kBlock), mt::kBlock, 0, st>>>(
      grid_sdf, num_grid_vertices, num_sh, lower_corners, num_columns, level_sets, S,
      vertex_offsets, num_triangle_vertices, gv, gd, gs, grid_vertices_grad, grid_sdf_grad,
      grid_sh_grad, maps + words);
  return mt::check_launch("mt_create_triangle_vertices_backward");
}

extern "C" int mt_count_triangle_faces(const float* sdf, const int32_t* lower_corners,
                                       int64_t num_columns, const int32_t* upper_corners,
                                       int64_t num_voxels, const float* level_sets, int32_t S,
                                       int32_t* face_counts, uint8_t* cell_codes,
                                       mt_stream_t stream) {
  MT_REQUIRE(num_voxels >= 0 && num_voxels <= num_columns && S >= 0,
             "mt_count_triangle_faces: bad sizes");
  MT_REQUIRE(int64_t(S) * num_voxels < (int64_t(1) << 31),
             "mt_count_triangle_faces: S * voxels exceeds int32 offsets");
  if (num_voxels == 0 || S == 0) return MT_OK;
  mt::count_faces_kernel<<<mt::blocks_for(num_voxels, mt::kBlock), mt::kBlock, 0,
                           static_cast<hipStream_t>(stream)>>>(
      sdf, lower_corners, num_columns, upper_corners, num_voxels, level_sets, S, face_counts,
      cell_codes);
  return mt::check_launch("mt_count_triangle_faces");
}

extern "C" int mt_create_triangle_faces(const int32_t* neighbors, const int32_t* indices,
                                        int64_t num_voxels, int64_t num_columns,
                                        const int32_t* vertex_offsets, const uint8_t* vertex_types,
                                        const uint8_t* cell_codes, int32_t S,
                                        const int32_t* face_offsets, int32_t* faces,
                                        int32_t* voxel_indices, int64_t face_capacity,
                                        int64_t vertex_capacity, int32_t* overflow,
                                        mt_stream_t stream) {
  MT_REQUIRE(num_voxels >= 0 && S >= 0 && face_capacity >= 0, "mt_create_triangle_faces: bad sizes");
  if (num_voxels == 0 || S == 0) return MT_OK;
  mt::create_faces_kernel<int32_t, mt::kTypesMode><<<mt::blocks_for(num_voxels, mt::kBlock), mt::kBlock, 0,
                                     static_cast<hipStream_t>(stream)>>>(
      neighbors, indices, num_voxels, num_columns, vertex_offsets, vertex_types, cell_codes, S,
      face_offsets, faces, voxel_indices, face_capacity, vertex_capacity, overflow);
  return mt::check_launch("mt_create_triangle_faces");
}

extern "C" int mt_create_triangle_faces_i64(const int32_t* neighbors, const int32_t* indices,
                                            int64_t num_voxels, int64_t num_columns,
                                            const int32_t* vertex_offsets,
                                            const uint8_t* vertex_types, const uint8_t* cell_codes,
                                            int32_t S, const int64_t* face_offsets, int32_t* faces,
                                            int32_t* voxel_indices, int64_t face_capacity,
                                            int64_t vertex_capacity, int32_t* overflow,
                                            mt_stream_t stream) {
  MT_REQUIRE(num_voxels >= 0 && S >= 0 && face_capacity >= 0,
             "mt_create_triangle_faces_i64: bad sizes");
  MT_REQUIRE(face_capacity < (int64_t(1) << 31), "mt_create_triangle_faces_i64: face ids are int32");
  if (num_voxels == 0 || S == 0) return MT_OK;
  mt::create_faces_kernel<int64_t, mt::kTypesMode><<<mt::blocks_for(num_voxels, mt::kBlock), mt::kBlock, 0,
                                     static_cast<hipStream_t>(stream)>>>(
      neighbors, indices, num_voxels, num_columns, vertex_offsets, vertex_types, cell_codes, S,
      face_offsets, faces, voxel_indices, face_capacity, vertex_capacity, overflow);
  return mt::check_launch("mt_create_triangle_faces_i64");
}

extern "C" int mt_create_triangle_faces_axes_i64(const int32_t* neighbors, int64_t num_voxels,
                                                 int64_t num_columns, const int32_t* vertex_offsets,
                                                 const uint8_t* column_axes,
                                                 const uint8_t* cell_codes, int32_t S,
                                                 const int64_t* face_offsets, int32_t* faces,
                                                 int64_t face_capacity, int64_t vertex_capacity,
                                                 int32_t* overflow, mt_stream_t stream) {
  MT_REQUIRE(num_voxels >= 0 && S >= 0 && face_capacity >= 0,
             "mt_create_triangle_faces_axes_i64: bad sizes");
  MT_REQUIRE(face_capacity < (int64_t(1) << 31), "mt_create_triangle_faces_axes_i64: face ids are int32");
  if (num_voxels == 0 || S == 0) return MT_OK;
  mt::create_faces_kernel<int64_t, mt::kAxesMode><<<mt::blocks_for(num_voxels, mt::kBlock), mt::kBlock, 0,
                                           static_cast<hipStream_t>(stream)>>>(
      neighbors, nullptr, num_voxels, num_columns, vertex_offsets, column_axes, cell_codes, S,
      face_offsets, faces, nullptr, face_capacity, vertex_capacity, overflow);
  return mt::check_launch("mt_create_triangle_faces_axes_i64");
}

extern "C" int mt_create_triangle_faces_words_i64(const int32_t* neighbors, int64_t num_voxels,
                                                  int64_t num_columns, const uint32_t* column_words,
                                                  const uint8_t* cell_codes, int32_t S,
                                                  const int64_t* face_offsets, int32_t* faces,
                                                  int64_t face_capacity, int64_t vertex_capacity,
                                                  int32_t* overflow, mt_stream_t stream) {
  MT_REQUIRE(num_voxels >= 0 && S >= 0 && face_capacity >= 0,
             "mt_create_triangle_faces_words_i64: bad sizes");
  MT_REQUIRE(face_capacity < (int64_t(1) << 31), "mt_create_triangle_faces_words_i64: face ids are int32");
  MT_REQUIRE(vertex_capacity <= MT_COLUMN_WORD_MAX_VERTICES,
             "mt_create_triangle_faces_words_i64: %lld vertices exceed the 2^29 a column word holds",
             static_cast<long long>(vertex_capacity));
  if (num_voxels == 0 || S == 0) return MT_OK;
  mt::create_faces_kernel<int64_t, mt::kWordsMode><<<mt::blocks_for(num_voxels, mt::kBlock), mt::kBlock, 0,
                                                     static_cast<hipStream_t>(stream)>>>(
      neighbors, nullptr, num_voxels, num_columns, reinterpret_cast<const int32_t*>(column_words),
      nullptr, cell_codes, S, face_offsets, faces, nullptr, face_capacity, vertex_capacity, overflow);
  return mt::check_launch("mt_create_triangle_faces_words_i64");
}

extern "C" int mt_create_triangle_vertices_colors_forward(
    const float* grid_vertices, const float* grid_sdf, const float* grid_sh,
    int64_t num_grid_vertices, int32_t num_sh, int32_t sh_degree, const int32_t* lower_corners,
    int64_t num_columns, const float* level_sets, int32_t S, const int32_t* vertex_offsets,
    int64_t num_triangle_vertices, const float* camera_positions, int32_t B, float* out_vertices,
    float* out_sdf, uint8_t* out_types, float* out_colors, int32_t* overflow, int32_t parts,
    int32_t colors_stride, uint32_t* column_words, mt_stream_t stream) {
  MT_REQUIRE((parts & 3) != 0 && (parts & ~7) == 0,
             "mt_create_triangle_vertices_colors_forward: parts %d", parts);
  MT_REQUIRE(column_words == nullptr || num_triangle_vertices <= MT_COLUMN_WORD_MAX_VERTICES,
             "mt_create_triangle_vertices_colors_forward: %lld vertices exceed the 2^29 a column word holds",
             static_cast<long long>(num_triangle_vertices));
  MT_REQUIRE(colors_stride == 0 || colors_stride >= B,
             "mt_create_triangle_vertices_colors_forward: colors_stride must be 0 or at least B");
  MT_REQUIRE(num_sh >= 0 && S >= 0 && num_columns >= 0 && B >= 0,
             "mt_create_triangle_vertices_colors_forward: negative size");
  MT_REQUIRE(sh_degree >= 0 && sh_degree <= 3 && (sh_degree + 1) * (sh_degree + 1) <= num_sh,
             "mt_create_triangle_vertices_colors_forward: degree %d needs (deg+1)^2 <= %d coefficients",
             sh_degree, num_sh);
  MT_REQUIRE((parts & 4) == 0 || reinterpret_cast<uintptr_t>(grid_sh) % 16 == 0,
             "mt_create_triangle_vertices_colors_forward: SH rows not 16-byte aligned");
  if (num_columns == 0 || S == 0 || num_triangle_vertices == 0) return MT_OK;
  const dim3 grid(mt::blocks_for(num_columns, mt::kVcBlock));
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool rows = (parts & 4) != 0;
#define MT_VCOL(D)                                                                              \
  do {                                                                                          \
    if (rows)                                                                                   \
      mt::create_vertices_colors_kernel<D, true><<<grid, mt::kVcBlock, 0, st>>>(                \
          grid_vertices, grid_sdf, grid_sh, num_grid_vertices, lower_corners, num_columns,     \
          level_sets, S, vertex_offsets, num_triangle_vertices, camera_positions, B,           \
          out_vertices, out_sdf, out_types, out_colors, overflow, parts, colors_stride,        \
          column_words);                                                                        \
    else                                                                                        \
      mt::create_vertices_colors_kernel<D, false><<<grid, mt::kVcBlock, 0, st>>>(               \
          grid_vertices, grid_sdf, grid_sh, num_grid_vertices, lower_corners, num_columns,     \
          level_sets, S, vertex_offsets, num_triangle_vertices, camera_positions, B,           \
          out_vertices, out_sdf, out_types, out_colors, overflow, parts, colors_stride,        \
          column_words);                                                                        \
  } while (0)
  // Without colours the degree-0 instance stages no SH rows (less LDS, more blocks per CU).
  switch ((parts & 2) != 0 && B > 0 ? sh_degree : 0) {
    case 0: MT_VCOL(0); break;
    case 1: MT_VCOL(1); break;
    case 2: MT_VCOL(2); break;
    default: MT_VCOL(3); break;
  }
#undef MT_VCOL
  return mt::check_launch("mt_create_triangle_vertices_colors_forward");
}

extern "C" int64_t mt_sh_row_floats(int32_t sh_degree) {
  return sh_degree < 0 || sh_degree > 3 ? -1 : mt::sh_row_floats(sh_degree);
}

extern "C" int mt_pack_sh_rows(const float* grid_sh, int64_t num_grid_vertices, int32_t num_sh,
                               int32_t sh_degree, float* rows, mt_stream_t stream) {
  MT_REQUIRE(num_grid_vertices >= 0, "mt_pack_sh_rows: negative size");
  MT_REQUIRE(sh_degree >= 0 && sh_degree <= 3 && (sh_degree + 1) * (sh_degree + 1) <= num_sh,
             "mt_pack_sh_rows: degree %d needs (deg+1)^2 <= %d coefficients", sh_degree, num_sh);
  MT_REQUIRE(reinterpret_cast<uintptr_t>(rows) % 16 == 0, "mt_pack_sh_rows: rows not 16-byte aligned");
  if (num_grid_vertices == 0) return MT_OK;
  const dim3 grid(mt::blocks_for(num_grid_vertices, mt::kPackVerts));
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (sh_degree) {
    case 0: mt::pack_sh_rows_kernel<0><<<grid, mt::kPackBlock, 0, st>>>(grid_sh, num_grid_vertices, rows); break;
    case 1: mt::pack_sh_rows_kernel<1><<<grid, mt::kPackBlock, 0, st>>>(grid_sh, num_grid_vertices, rows); break;
    case 2: mt::pack_sh_rows_kernel<2><<<grid, mt::kPackBlock, 0, st>>>(grid_sh, num_grid_vertices, rows); break;
    default: mt::pack_sh_rows_kernel<3><<<grid, mt::kPackBlock, 0, st>>>(grid_sh, num_grid_vertices, rows); break;
  }
  return mt::check_launch("mt_pack_sh_rows");
}

extern "C" size_t mt_create_triangle_vertices_colors_backward_workspace_bytes(
    int64_t num_grid_vertices, int64_t num_columns, int32_t sh_degree) {
  const int64_t row = mt::adjoint_row_floats((sh_degree + 1) * (sh_degree + 1));
  const size_t maps = static_cast<size_t>(4 * num_grid_vertices + 1) * sizeof(int32_t);
  const size_t part = static_cast<size_t>(4 * row * num_columns) * sizeof(float);
  const size_t list = static_cast<size_t>(num_columns + 1) * sizeof(int32_t);
  const size_t slots = static_cast<size_t>(num_columns) * sizeof(uint32_t);
  const size_t live = static_cast<size_t>(num_columns);
  const size_t ends = static_cast<size_t>(num_grid_vertices + 1) * sizeof(int32_t);
  const size_t pos = static_cast<size_t>(num_columns) * sizeof(int32_t);
  auto up = [](size_t x) { return (x + 255) / 256 * 256; };
  return up(maps) + up(part) + up(list) + up(slots) + up(live) + up(ends) + pos;
}

extern "C" int mt_create_triangle_vertices_colors_backward(
    const float* grid_sdf, int64_t num_grid_vertices, int32_t num_sh, int32_t sh_degree,
    const int32_t* lower_corners, int64_t num_columns, const float* level_sets, int32_t S,
    const int32_t* vertex_offsets, int64_t num_triangle_vertices, const float* vertices,
    const float* colors, const float* camera_positions, int32_t B, float* grad_sdf,
    float* grad_colors, float* grid_sdf_grad, float* grid_sh_grad, void* workspace,
    size_t workspace_bytes, int32_t accumulate, const uint8_t* vertex_touched,
    int32_t clear_gradients, int32_t maps_ready, int32_t* status, int32_t colors_stride,
    mt_stream_t stream) {
  MT_REQUIRE(num_sh >= 0 && S >= 0 && num_columns >= 0 && B >= 0,
             "mt_create_triangle_vertices_colors_backward: negative size");
  MT_REQUIRE(colors_stride == 0 || colors_stride >= B,
             "mt_create_triangle_vertices_colors_backward: colors_stride must be 0 or at least B");
  MT_REQUIRE(num_sh == 1 || num_sh == 4 || num_sh == 9 || num_sh == 16,
             "mt_create_triangle_vertices_colors_backward: %d SH coefficients (need 1, 4, 9 or 16)",
             num_sh);
  MT_REQUIRE(sh_degree >= 0 && (sh_degree + 1) * (sh_degree + 1) <= num_sh,
             "mt_create_triangle_vertices_colors_backward: degree %d needs (deg+1)^2 <= %d",
             sh_degree, num_sh);
  const size_t need = mt_create_triangle_vertices_colors_backward_workspace_bytes(
      num_grid_vertices, num_columns, sh_degree);
  if (workspace == nullptr || workspace_bytes < need) {
    mt::set_error("mt_create_triangle_vertices_colors_backward: workspace of %zu bytes is smaller than the %zu required",
                  workspace_bytes, need);
    return MT_ERR_WORKSPACE_TOO_SMALL;
  }
  if (num_grid_vertices == 0) return MT_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  int32_t* maps = static_cast<int32_t*>(workspace);
  const int64_t words = 4 * num_grid_vertices;
  float* part = reinterpret_cast<float*>(static_cast<char*>(workspace) +
                                         (static_cast<size_t>(words + 1) * 4 + 255) / 256 * 256);
  const int64_t row = mt::adjoint_row_floats((sh_degree + 1) * (sh_degree + 1));
  int32_t* list = reinterpret_cast<int32_t*>(
      reinterpret_cast<char*>(part) +
      (static_cast<size_t>(4 * row * num_columns) * sizeof(float) + 255) / 256 * 256);
  int32_t* live_count = list + num_columns;
  uint32_t* col_slots = reinterpret_cast<uint32_t*>(
      reinterpret_cast<char*>(list) +
      (static_cast<size_t>(num_columns + 1) * sizeof(int32_t) + 255) / 256 * 256);
  uint8_t* col_live = reinterpret_cast<uint8_t*>(col_slots) +
                      (static_cast<size_t>(num_columns) * sizeof(uint32_t) + 255) / 256 * 256;
  int32_t* ends_list = reinterpret_cast<int32_t*>(
      col_live + (static_cast<size_t>(num_columns) + 255) / 256 * 256);
  int32_t* ends_count = ends_list + num_grid_vertices;
  int32_t* col_pos = reinterpret_cast<int32_t*>(
      reinterpret_cast<char*>(ends_list) +
      (static_cast<size_t>(num_grid_vertices + 1) * sizeof(int32_t) + 255) / 256 * 256);
  // The grid-vertex maps depend only on the columns: a caller whose workspace holds them
  // from an earlier call over the same columns skips rebuilding them.
  if (maps_ready == 0) {
    mt::fill_maps_kernel<<<mt::blocks_for(words, 256), 256, 0, st>>>(maps, words);
    if (num_columns > 0)
      mt::build_maps_kernel<<<mt::blocks_for(num_columns, mt::kBlock), mt::kBlock, 0, st>>>(
          lower_corners, num_columns, num_grid_vertices, maps);
  }
  mt::zero_counts_kernel<<<1, 1, 0, st>>>(live_count, ends_count);
  if (num_columns > 0)
    mt::colors_adjoint_mask_kernel<<<mt::blocks_for(num_columns, mt::kBlock * mt::kMaskRounds),
                                     mt::kBlock, 0, st>>>(
        grid_sdf, num_grid_vertices, lower_corners, num_columns, level_sets, S, vertex_offsets,
        num_triangle_vertices, vertex_touched, maps, col_live, col_slots, list, col_pos,
        live_count);
  const dim3 cgrid(mt::blocks_for(num_columns, mt::kBlock));
  const dim3 ggrid(mt::blocks_for(num_grid_vertices, mt::kBlock));
  // An accumulating call leaves the grid vertices without live ends alone: phase 2 runs
  // over the list of the others.
  if (accumulate != 0)
    mt::colors_adjoint_ends_mask_kernel<<<ggrid, mt::kBlock, 0, st>>>(
        num_grid_vertices, col_live, maps, ends_list, ends_count);
#define MT_ADJ(N, D)                                                                             \
  do {                                                                                           \
    if (num_columns > 0)                                                                         \
      mt::colors_adjoint_columns_kernel<N, D><<<cgrid, mt::kBlock, 0, st>>>(                     \
          grid_sdf, num_grid_vertices, lower_corners, num_columns, level_sets, S,                \
          vertex_offsets, num_triangle_vertices, vertices, colors, camera_positions, B, grad_sdf, \
          grad_colors, part, list, live_count, col_live, col_slots, maps, vertex_touched,       \
          clear_gradients != 0, colors_stride);                                                 \
    mt::colors_adjoint_ends_kernel<N, D><<<ggrid, mt::kBlock, 0, st>>>(                          \
        num_grid_vertices, num_columns, part, col_live, col_pos, grid_sdf_grad, grid_sh_grad,     \
        maps, accumulate != 0, status, accumulate != 0 ? ends_list : nullptr, ends_count);        \
    if (accumulate == 0)                                                                         \
      mt::clear_grid_gradients_kernel<<<ggrid.x < 4096u ? ggrid.x : 4096u, mt::kBlock, 0, st>>>(  \
          maps, num_grid_vertices, num_sh, grid_sdf_grad, grid_sh_grad);                          \
    if (num_columns > 0)                                                                         \
      mt::colors_adjoint_scatter_kernel<N, D><<<cgrid, mt::kBlock, 0, st>>>(                     \
          num_grid_vertices, lower_corners, num_columns, part, list, live_count, col_live,        \
          grid_sdf_grad, grid_sh_grad, maps);                                                     \
  } while (0)
  const int key = num_sh * 4 + sh_degree;
  switch (key) {
    case 1 * 4 + 0: MT_ADJ(1, 0); break;
    case 4 * 4 + 0: MT_ADJ(4, 0); break;
    case 4 * 4 + 1: MT_ADJ(4, 1); break;
    case 9 * 4 + 0: MT_ADJ(9, 0); break;
    case 9 * 4 + 1: MT_ADJ(9, 1); break;
    case 9 * 4 + 2: MT_ADJ(9, 2); break;
    case 16 * 4 + 0: MT_ADJ(16, 0); break;
    case 16 * 4 + 1: MT_ADJ(16, 1); break;
    case 16 * 4 + 2: MT_ADJ(16, 2); break;
    default: MT_ADJ(16, 3); break;
  }
#undef MT_ADJ
  return mt::check_launch("mt_create_triangle_vertices_colors_backward");
}
