// Tile compositing: fused intersect-and-composite (forward and adjoint) and the explicit
// mode's intersect / interpolate / composite.
//
// One workgroup per (camera, tile), one lane per pixel (wave64, so a 16x16 tile is four
// waves). The tile's depth-sorted keys are staged through LDS in chunks; for each staged
// triangle the per-triangle terms of the barycentric solve are computed once at staging,
// so a pixel spends ~25 VALU ops per (pixel, triangle) test. The three IEEE divisions
// of the test are only evaluated when a cheap reciprocal estimate cannot already prove
// a miss (see bary_test): the hit decisions stay bit-identical to the reference formula.
#include "common.h"

namespace mt {
namespace {

constexpr float kEpsBary = 1e-8f;   // intersect_and_composite.slang:99
constexpr float kSumLo = 0.999f;    // :132
constexpr float kSumHi = 1.001f;
constexpr float kAlphaEps = 1e-5f;  // :161
constexpr float kSuffixEps = 1e-6f; // :3
constexpr float kDummySdf = 1e10f;  // :318

// Staged triangle (7 x float4 = 112 bytes).
struct Staged {
  float4 g0;  // p1.x, p1.y, p2.x, p2.y
  float4 g1;  // e12.x, e12.y, e13.x, e13.y
  float4 g2;  // d_13_13, d_12_13, d_12_12, d_12_23
  float4 g3;  // d_13_23, D = denom + eps, 1 / D (estimate), shell (bits)
};
struct StagedColor {
  float4 a, b, c;  // rgb + sdf of the three corners
};

__device__ __forceinline__ Staged stage_geometry(float2 p1, float2 p2, float2 p3, int shell) {
  const float e12x = p2.x - p1.x, e12y = p2.y - p1.y;
  const float e23x = p3.x - p2.x, e23y = p3.y - p2.y;
  const float e13x = p3.x - p1.x, e13y = p3.y - p1.y;
  const float d_12_12 = e12x * e12x + e12y * e12y;
  const float d_12_23 = e12x * e23x + e12y * e23y;
  const float d_12_13 = e12x * e13x + e12y * e13y;
  const float d_13_13 = e13x * e13x + e13y * e13y;
  const float d_13_23 = e13x * e23x + e13y * e23y;
  const float d_23_23 = e23x * e23x + e23y * e23y;
  const float a = fabsf(d_12_13), b = fabsf(d_12_23), c = fabsf(d_13_23);
  float denom;
  if (a <= b && a <= c) {
    denom = d_12_12 * d_13_13 - d_12_13 * d_12_13;
  } else if (b <= a && b <= c) {
    denom = d_12_12 * d_23_23 - d_12_23 * d_12_23;
  } else if (c <= a && c <= b) {
    denom = d_13_13 * d_23_23 - d_13_23 * d_13_23;
  } else {
    denom = __builtin_nanf("");  // only reachable with NaN inputs; every test then misses
  }
  const float D = denom + kEpsBary;
  // The estimate is only trusted where its error bound holds (normal, moderate |D|).
  const float ad = fabsf(D);
  const float r = (ad >= 0x1p-60f && ad <= 0x1p60f) ? __builtin_amdgcn_rcpf(D) : 0.0f;
  Staged s;
  s.g0 = make_float4(p1.x, p1.y, p2.x, p2.y);
  s.g1 = make_float4(e12x, e12y, e13x, e13y);
  s.g2 = make_float4(d_13_13, d_12_13, d_12_12, d_12_23);
  s.g3 = make_float4(d_13_23, D, r, __int_as_float(shell));
  return s;
}

struct Bary {
  float u, v, w;
  bool hit;
};

// The reference's barycentric test (intersect_and_composite.slang:65-145) at pixel centre
// (px, py). Miss proof without division: with r = rcp(D) (relative error <= 2^-22) the
// product q * r is within a factor (1 +- 2^-21) of q / D. If it is < -2^-100 the exactly
// rounded q / D is a nonzero negative number; if it is > 1 + 2^-19 the rounded quotient
// exceeds 1. Either way the reference rejects the pixel, so the division is skipped.
// r == 0 (D outside [2^-60, 2^60] or NaN) disables the shortcut.
__device__ __forceinline__ Bary bary_test(const Staged& s, float px, float py) {
  Bary out;
  const float a1x = px - s.g0.x, a1y = py - s.g0.y;
  const float a2x = px - s.g0.z, a2y = py - s.g0.w;
  const float d_1x_12 = a1x * s.g1.x + a1y * s.g1.y;
  const float d_1x_13 = a1x * s.g1.z + a1y * s.g1.w;
  const float d_2x_12 = a2x * s.g1.x + a2y * s.g1.y;
  const float d_2x_13 = a2x * s.g1.z + a2y * s.g1.w;
  const float vn = s.g2.x * d_1x_12 - s.g2.y * d_1x_13;
  const float wn = s.g2.z * d_1x_13 - s.g2.y * d_1x_12;
  const float un = s.g2.w * d_2x_13 - s.g3.x * d_2x_12;
  const float r = s.g3.z;
  if (r != 0.0f) {
    const float ua = un * r, va = vn * r, wa = wn * r;
    constexpr float kNeg = -0x1p-100f;
    constexpr float kBig = 1.0f + 0x1p-19f;
    if (ua < kNeg || va < kNeg || wa < kNeg || ua > kBig || va > kBig || wa > kBig) {
      out.hit = false;
      out.u = out.v = out.w = 0.0f;
      return out;
    }
  }
  const float D = s.g3.y;
  const float v = vn / D;
  const float w = wn / D;
  const float u = un / D;
  const float sum = (u + v) + w;
  const bool miss = isnan(u) || isnan(v) || isnan(w) || u < 0.0f || u > 1.0f || v < 0.0f ||
                    v > 1.0f || w < 0.0f || w > 1.0f || sum < kSumLo || sum > kSumHi;
  out.u = u;
  out.v = v;
  out.w = w;
  out.hit = !miss;
  return out;
}

__device__ __forceinline__ float sdf_to_cdf(float sdf, float k) { return 1.0f / (1.0f + expf(-sdf * k)); }

__device__ __forceinline__ int shell_of(const Staged& s) { return __float_as_int(s.g3.w); }

// Fused kernels stage up to kMaxChunk triangles per round in dynamic LDS
// (64 + 48 + 4 bytes each: 58 KB at 512); the explicit intersect kernel stages up to
// kMaxIntersectChunk (geometry + face only).
constexpr int kMaxChunk = 512;
constexpr int kMaxIntersectChunk = 256;
constexpr int kStagedBytes = sizeof(Staged) + sizeof(StagedColor) + sizeof(int32_t);

struct StageBuffers {
  Staged* geo;
  StagedColor* col;
  int32_t* face;
};

__device__ __forceinline__ StageBuffers carve_stage(char* smem, int chunk) {
  StageBuffers b;
  b.geo = reinterpret_cast<Staged*>(smem);
  b.col = reinterpret_cast<StagedColor*>(smem + sizeof(Staged) * chunk);
  b.face = reinterpret_cast<int32_t*>(smem + (sizeof(Staged) + sizeof(StagedColor)) * chunk);
  return b;
}

extern __shared__ __align__(16) char g_dynamic_lds[];

struct TileCoord {
  int cam, tile, x, y;
  bool in_image;
};

__device__ __forceinline__ TileCoord tile_coord(const mt_composite_args& a) {
  const int T = a.grid_rows * a.grid_cols;
  const uint32_t bid = xcd_swizzle(blockIdx.x, gridDim.x);
  TileCoord c;
  c.cam = bid / T;
  c.tile = bid % T;
  const int ty = c.tile / a.grid_cols, tx = c.tile % a.grid_cols;
  const int ly = threadIdx.x / a.tile_w, lx = threadIdx.x % a.tile_w;
  c.x = tx * a.tile_w + lx;
  c.y = ty * a.tile_h + ly;
  c.in_image = c.x < a.width && c.y < a.height;
  return c;
}

__device__ __forceinline__ void stage_chunk(const mt_composite_args& a, int cam, int64_t first,
                                            int count, Staged* sg, StagedColor* sc, int32_t* sf,
                                            bool want_color) {
  const float2* xy = reinterpret_cast<const float2*>(a.xy) + int64_t(cam) * a.num_vertices;
  const float* col = a.colors + int64_t(cam) * a.num_vertices * 3;
  for (int j = threadIdx.x; j < count; j += blockDim.x) {
    const int64_t key_index = first + j;
    const int32_t face = a.sorted_values[key_index];
    const uint64_t key = static_cast<uint64_t>(a.sorted_keys[key_index]);
    const int shell = static_cast<int>(key_field(key, a.layout.shell_offset, a.layout.shell_bits));
    const int64_t va = a.faces[3 * int64_t(face) + 0];
    const int64_t vb = a.faces[3 * int64_t(face) + 1];
    const int64_t vc = a.faces[3 * int64_t(face) + 2];
    sg[j] = stage_geometry(xy[va], xy[vb], xy[vc], shell);
    sf[j] = face;
    if (want_color) {
      sc[j].a = make_float4(col[3 * va + 0], col[3 * va + 1], col[3 * va + 2], a.sdf[va]);
      sc[j].b = make_float4(col[3 * vb + 0], col[3 * vb + 1], col[3 * vb + 2], a.sdf[vb]);
      sc[j].c = make_float4(col[3 * vc + 0], col[3 * vc + 1], col[3 * vc + 2], a.sdf[vc]);
    }
  }
}

// Block-reduced statistics counters (one atomic pair per block).
__device__ void add_stats(unsigned long long* stats, uint32_t tests, uint32_t hits) {
  __shared__ unsigned long long s_t[1024 / kWave], s_h[1024 / kWave];
  unsigned long long t = wave_sum<unsigned long long>(tests);
  unsigned long long h = wave_sum<unsigned long long>(hits);
  if (__lane_id() == 0) {
    s_t[threadIdx.x / kWave] = t;
    s_h[threadIdx.x / kWave] = h;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long st = 0, sh = 0;
    for (int w = 0; w < (int(blockDim.x) + kWave - 1) / kWave; ++w) {
      st += s_t[w];
      sh += s_h[w];
    }
    atomicAdd(stats + 0, st);
    atomicAdd(stats + 1, sh);
  }
}

__global__ void __launch_bounds__(1024) fused_forward_kernel(mt_composite_args a, float* out) {
  const StageBuffers sb = carve_stage(g_dynamic_lds, a.chunk);
  Staged* s_geo = sb.geo;
  StagedColor* s_col = sb.col;
  int32_t* s_face = sb.face;
  const TileCoord tc = tile_coord(a);
  const int T = a.grid_rows * a.grid_cols;
  const int64_t start = a.tile_ranges[2 * (int64_t(tc.cam) * T + tc.tile) + 0];
  const int64_t end = a.tile_ranges[2 * (int64_t(tc.cam) * T + tc.tile) + 1];
  const float k = a.sharpness[0];
  const float px = float(tc.x) + 0.5f, py = float(tc.y) + 0.5f;
  const float dummy_cdf = sdf_to_cdf(kDummySdf, k);

  float r = 0.f, g = 0.f, bl = 0.f, trans = 1.f;
  float outer_cdf = dummy_cdf;
  int last_shell = -1;
  bool active = tc.in_image;
  const int chunk = a.chunk;
  uint32_t n_tests = 0, n_hits = 0;
  for (int64_t first = start; first < end; first += chunk) {
    const int count = static_cast<int>((end - first) < chunk ? (end - first) : chunk);
    __syncthreads();
    stage_chunk(a, tc.cam, first, count, s_geo, s_col, s_face, true);
    __syncthreads();
    if (active) {
      for (int i = 0; i < count; ++i) {
        const Staged st = s_geo[i];
        const int shell = shell_of(st);
        if (shell <= last_shell) continue;
        ++n_tests;
        const Bary bc = bary_test(st, px, py);
        if (!bc.hit) continue;
        ++n_hits;
        const StagedColor c = s_col[i];
        const float sdf = (bc.u * c.a.w + bc.v * c.b.w) + bc.w * c.c.w;
        const float cr = (bc.u * c.a.x + bc.v * c.b.x) + bc.w * c.c.x;
        const float cg = (bc.u * c.a.y + bc.v * c.b.y) + bc.w * c.c.y;
        const float cb = (bc.u * c.a.z + bc.v * c.b.z) + bc.w * c.c.z;
        const float inner_cdf = sdf_to_cdf(sdf, k);
        float alpha = ((outer_cdf - inner_cdf) + kAlphaEps) / (outer_cdf + kAlphaEps);
        alpha = fminf(fmaxf(alpha, 0.0f), 1.0f);
        r = r + (cr * trans) * alpha;
        g = g + (cg * trans) * alpha;
        bl = bl + (cb * trans) * alpha;
        trans = trans * (1.0f - alpha);
        outer_cdf = inner_cdf;
        last_shell = shell;
        if (shell == a.num_shells - 1) {
          active = false;
          break;
        }
      }
    }
    if (!__syncthreads_or(active)) break;
  }
  if (tc.in_image) {
    const int64_t plane = int64_t(a.height) * a.width;
    float* o = out + int64_t(tc.cam) * 4 * plane + int64_t(tc.y) * a.width + tc.x;
    o[0] = r;
    o[plane] = g;
    o[2 * plane] = bl;
    o[3 * plane] = trans;
  }
  if (a.stats) add_stats(a.stats, n_tests, n_hits);
}

// Scatter one sample's gradient (colour and SDF) to its triangle's corners.
__device__ __forceinline__ void scatter_sample(const mt_composite_args& a, int cam, int32_t face,
                                               float bu, float bv, float bw, float dr, float dg,
                                               float db, float dsdf, float* colors_grad,
                                               float* sdf_grad) {
  const int64_t v[3] = {a.faces[3 * int64_t(face) + 0], a.faces[3 * int64_t(face) + 1],
                        a.faces[3 * int64_t(face) + 2]};
  const float wgt[3] = {bu, bv, bw};
  float* cgrad = colors_grad + int64_t(cam) * a.num_vertices * 3;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    atomicAdd(cgrad + 3 * v[q] + 0, wgt[q] * dr);
    atomicAdd(cgrad + 3 * v[q] + 1, wgt[q] * dg);
    atomicAdd(cgrad + 3 * v[q] + 2, wgt[q] * db);
    atomicAdd(sdf_grad + v[q], wgt[q] * dsdf);
  }
}

__global__ void __launch_bounds__(1024) fused_backward_kernel(mt_composite_args a, const float* out,
                                                              const float* gout, float* colors_grad,
                                                              float* sdf_grad, float* sharp_grad,
                                                              int quirk) {
  const StageBuffers sb = carve_stage(g_dynamic_lds, a.chunk);
  Staged* s_geo = sb.geo;
  StagedColor* s_col = sb.col;
  int32_t* s_face = sb.face;
  __shared__ float s_red[1024 / kWave];
  const TileCoord tc = tile_coord(a);
  const int T = a.grid_rows * a.grid_cols;
  const int64_t start = a.tile_ranges[2 * (int64_t(tc.cam) * T + tc.tile) + 0];
  const int64_t end = a.tile_ranges[2 * (int64_t(tc.cam) * T + tc.tile) + 1];
  const float k = a.sharpness[0];
  const float px = float(tc.x) + 0.5f, py = float(tc.y) + 0.5f;
  const float dummy_cdf = sdf_to_cdf(kDummySdf, k);

  float fr = 0.f, fg = 0.f, fb = 0.f, ft = 0.f;    // final accumulator
  float gr = 0.f, gg = 0.f, gb = 0.f, gt = 0.f;    // its gradient
  if (tc.in_image) {
    const int64_t plane = int64_t(a.height) * a.width;
    const int64_t o = int64_t(tc.cam) * 4 * plane + int64_t(tc.y) * a.width + tc.x;
    fr = out[o];
    fg = out[o + plane];
    fb = out[o + 2 * plane];
    ft = out[o + 3 * plane];
    gr = gout[o];
    gg = gout[o + plane];
    gb = gout[o + 2 * plane];
    gt = gout[o + 3 * plane];
  }
  float ar = 0.f, ag = 0.f, ab = 0.f, trans = 1.f;
  float outer_cdf = dummy_cdf, outer_sdf = kDummySdf;
  int last_shell = -1;
  bool active = tc.in_image;
  float dk = 0.f;
  // Sample waiting for its outer-role SDF gradient (set by the next accepted hit).
  uint32_t n_tests = 0, n_hits = 0;
  bool pending = false;
  int32_t p_face = 0;
  float p_u = 0.f, p_v = 0.f, p_w = 0.f, p_dr = 0.f, p_dg = 0.f, p_db = 0.f, p_dsdf = 0.f;

  const int chunk = a.chunk;
  for (int64_t first = start; first < end; first += chunk) {
    const int count = static_cast<int>((end - first) < chunk ? (end - first) : chunk);
    __syncthreads();
    stage_chunk(a, tc.cam, first, count, s_geo, s_col, s_face, true);
    __syncthreads();
    if (active) {
      for (int i = 0; i < count; ++i) {
        const Staged st = s_geo[i];
        const int shell = shell_of(st);
        if (shell <= last_shell) continue;
        ++n_tests;
        const Bary bc = bary_test(st, px, py);
        if (!bc.hit) continue;
        ++n_hits;
        const StagedColor c = s_col[i];
        const float sdf = (bc.u * c.a.w + bc.v * c.b.w) + bc.w * c.c.w;
        const float cr = (bc.u * c.a.x + bc.v * c.b.x) + bc.w * c.c.x;
        const float cg = (bc.u * c.a.y + bc.v * c.b.y) + bc.w * c.c.y;
        const float cb = (bc.u * c.a.z + bc.v * c.b.z) + bc.w * c.c.z;
        const float ci = sdf_to_cdf(sdf, k);
        const float co = outer_cdf;
        const float den = co + kAlphaEps;
        const float x = ((co - ci) + kAlphaEps) / den;
        const float alpha = fminf(fmaxf(x, 0.0f), 1.0f);
        // Suffix trick (slang:460-473).
        const float weight = alpha * trans;
        ar += weight * cr;
        ag += weight * cg;
        ab += weight * cb;
        const float inv = 1.0f / (1.0f - alpha + kSuffixEps);
        const float d_alpha = gr * (trans * cr - (fr - ar) * inv) + gg * (trans * cg - (fg - ag) * inv) +
                              gb * (trans * cb - (fb - ab) * inv) - gt * ft * inv;
        float d_in = 0.f, d_out = 0.f;
        if (x > 0.0f && x < 1.0f) {
          const float dci = ci * (1.0f - ci);
          const float dco = co * (1.0f - co);
          const float dx_di = -1.0f / den;
          const float dx_do = ci / (den * den);
          d_in = d_alpha * dx_di * k * dci;
          d_out = d_alpha * dx_do * k * dco;
          dk += d_alpha * (dx_di * sdf * dci + dx_do * outer_sdf * dco);
        }
        if (pending) {
          p_dsdf += d_out;
          const bool q = quirk != 0;
          scatter_sample(a, tc.cam, p_face, q ? bc.u : p_u, q ? bc.v : p_v, q ? bc.w : p_w, p_dr,
                         p_dg, p_db, p_dsdf, colors_grad, sdf_grad);
        }
        pending = true;
        p_face = s_face[i];
        p_u = bc.u;
        p_v = bc.v;
        p_w = bc.w;
        p_dr = weight * gr;
        p_dg = weight * gg;
        p_db = weight * gb;
        p_dsdf = d_in;
        trans *= 1.0f - alpha;
        outer_cdf = ci;
        outer_sdf = sdf;
        last_shell = shell;
        if (shell == a.num_shells - 1) {
          active = false;
          break;
        }
      }
    }
    if (!__syncthreads_or(active)) break;
  }
  if (pending) {
    scatter_sample(a, tc.cam, p_face, p_u, p_v, p_w, p_dr, p_dg, p_db, p_dsdf, colors_grad,
                   sdf_grad);
  }
  // Sharpness gradient: block reduction, one atomic per block.
  float s = wave_sum_f(dk);
  if (__lane_id() == 0) s_red[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int w = 0; w < (int(blockDim.x) + kWave - 1) / kWave; ++w) t += s_red[w];
    if (t != 0.0f) atomicAdd(sharp_grad, t);
  }
  if (a.stats) add_stats(a.stats, n_tests, n_hits);
}

// ------------------------------------------------------------------------------------
// Explicit mode
// ------------------------------------------------------------------------------------

__global__ void __launch_bounds__(1024) intersect_kernel(const int32_t* values, const int64_t* ranges,
                                                         const float* xy, int64_t V,
                                                         const int32_t* faces, int B, int S,
                                                         int grid_rows, int grid_cols, int th,
                                                         int tw, int H, int W, int chunk,
                                                         float* out_uv, int32_t* out_index) {
  __shared__ Staged s_geo[kMaxIntersectChunk];
  __shared__ int32_t s_face[kMaxIntersectChunk];
  const int T = grid_rows * grid_cols;
  const uint32_t bid = xcd_swizzle(blockIdx.x, gridDim.x);
  const int tile = bid % T;
  const int bs = bid / T;  // camera * S + shell
  const int cam = bs / S;
  const int ty = tile / grid_cols, tx = tile % grid_cols;
  const int x = tx * tw + threadIdx.x % tw;
  const int y = ty * th + threadIdx.x / tw;
  const bool in_image = x < W && y < H;
  const int64_t start = ranges[2 * (int64_t(bs) * T + tile) + 0];
  const int64_t end = ranges[2 * (int64_t(bs) * T + tile) + 1];
  const float px = float(x) + 0.5f, py = float(y) + 0.5f;
  const float2* pxy = reinterpret_cast<const float2*>(xy) + int64_t(cam) * V;
  bool active = in_image;
  float hu = 0.f, hv = 0.f;
  int32_t hidx = -1;
  for (int64_t first = start; first < end; first += chunk) {
    const int count = static_cast<int>((end - first) < chunk ? (end - first) : chunk);
    __syncthreads();
    for (int j = threadIdx.x; j < count; j += blockDim.x) {
      const int32_t face = values[first + j];
      const int64_t va = faces[3 * int64_t(face)], vb = faces[3 * int64_t(face) + 1],
                    vc = faces[3 * int64_t(face) + 2];
      s_geo[j] = stage_geometry(pxy[va], pxy[vb], pxy[vc], 0);
      s_face[j] = face;
    }
    __syncthreads();
    if (active) {
      for (int i = 0; i < count; ++i) {
        const Bary bc = bary_test(s_geo[i], px, py);
        if (!bc.hit) continue;
        hu = bc.u;
        hv = bc.v;
        hidx = s_face[i];
        active = false;
        break;
      }
    }
    if (!__syncthreads_or(active)) break;
  }
  if (in_image) {
    const int64_t p = (int64_t(bs) * H + y) * W + x;
    reinterpret_cast<float2*>(out_uv)[p] = make_float2(hu, hv);
    out_index[p] = hidx;
  }
}

__global__ void __launch_bounds__(256) interpolate_kernel(const float* uv, const int32_t* index,
                                                          const int32_t* faces, const float* sdf,
                                                          const float* colors, int64_t V, int S,
                                                          int64_t HW, int64_t n, float* out_sdf,
                                                          float* out_col) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int32_t f = index[p];
  if (f < 0) {
    out_sdf[p] = 0.f;
    out_col[3 * p + 0] = 0.f;
    out_col[3 * p + 1] = 0.f;
    out_col[3 * p + 2] = 0.f;
    return;
  }
  const int cam = static_cast<int>((p / HW) / S);
  const float2 q = reinterpret_cast<const float2*>(uv)[p];
  const float u = q.x, v = q.y, w = (1.0f - u) - v;
  const int64_t va = faces[3 * int64_t(f)], vb = faces[3 * int64_t(f) + 1], vc = faces[3 * int64_t(f) + 2];
  const float* col = colors + int64_t(cam) * V * 3;
  out_sdf[p] = (u * sdf[va] + v * sdf[vb]) + w * sdf[vc];
#pragma unroll
  for (int c = 0; c < 3; ++c) out_col[3 * p + c] = (u * col[3 * va + c] + v * col[3 * vb + c]) + w * col[3 * vc + c];
}

__global__ void __launch_bounds__(256) interpolate_backward_kernel(
    const float* uv, const int32_t* index, const int32_t* faces, int64_t V, int S, int64_t HW,
    int64_t n, const float* g_sdf, const float* g_col, float* sdf_grad, float* col_grad) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int32_t f = index[p];
  if (f < 0) return;
  const int cam = static_cast<int>((p / HW) / S);
  const float2 q = reinterpret_cast<const float2*>(uv)[p];
  const float wgt[3] = {q.x, q.y, (1.0f - q.x) - q.y};
  const float gs = g_sdf[p];
  const float gc[3] = {g_col[3 * p], g_col[3 * p + 1], g_col[3 * p + 2]};
  float* cg = col_grad + int64_t(cam) * V * 3;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int64_t vi = faces[3 * int64_t(f) + k];
    atomicAdd(sdf_grad + vi, wgt[k] * gs);
#pragma unroll
    for (int c = 0; c < 3; ++c) atomicAdd(cg + 3 * vi + c, wgt[k] * gc[c]);
  }
}

__global__ void __launch_bounds__(256) composite_forward_kernel(const float* sdf, const float* col,
                                                                const uint8_t* mask, int64_t P, int S,
                                                                const float* sharpness,
                                                                float* out_col, float* out_t) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const float k = sharpness[0];
  float r = 0.f, g = 0.f, b = 0.f, trans = 1.f;
  float outer_cdf = sdf_to_cdf(kDummySdf, k);
  for (int s = 0; s < S; ++s) {
    const int64_t q = p * S + s;
    if (!mask[q]) continue;
    const float ci = sdf_to_cdf(sdf[q], k);
    float alpha = ((outer_cdf - ci) + kAlphaEps) / (outer_cdf + kAlphaEps);
    alpha = fminf(fmaxf(alpha, 0.0f), 1.0f);
    r = r + (col[3 * q + 0] * trans) * alpha;
    g = g + (col[3 * q + 1] * trans) * alpha;
    b = b + (col[3 * q + 2] * trans) * alpha;
    trans = trans * (1.0f - alpha);
    outer_cdf = ci;
  }
  out_col[3 * p + 0] = r;
  out_col[3 * p + 1] = g;
  out_col[3 * p + 2] = b;
  out_t[p] = trans;
}

__global__ void __launch_bounds__(256) composite_backward_kernel(
    const float* sdf, const float* col, const uint8_t* mask, int64_t P, int S,
    const float* sharpness, const float* out_col, const float* out_t, const float* g_col,
    const float* g_t, float* sdf_grad, float* col_grad, float* sharp_grad, int quirk) {
  __shared__ float s_red[256 / kWave];
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  float dk = 0.f;
  if (p < P) {
    const float k = sharpness[0];
    const float fr = out_col[3 * p], fg = out_col[3 * p + 1], fb = out_col[3 * p + 2], ft = out_t[p];
    const float gr = g_col[3 * p], gg = g_col[3 * p + 1], gb = g_col[3 * p + 2], gt = g_t[p];
    float ar = 0.f, ag = 0.f, ab = 0.f, trans = 1.f;
    float outer_cdf = sdf_to_cdf(kDummySdf, k), outer_sdf = kDummySdf;
    int prev = -1;
    float partial = 0.f;
    for (int s = 0; s < S; ++s) {
      const int64_t q = p * S + s;
      if (!mask[q]) continue;
      const float si = sdf[q];
      const float cr = col[3 * q], cg = col[3 * q + 1], cb = col[3 * q + 2];
      const float ci = sdf_to_cdf(si, k);
      const float co = outer_cdf;
      const float den = co + kAlphaEps;
      const float x = ((co - ci) + kAlphaEps) / den;
      const float alpha = fminf(fmaxf(x, 0.0f), 1.0f);
      const float weight = alpha * trans;
      ar += weight * cr;
      ag += weight * cg;
      ab += weight * cb;
      const float inv = 1.0f / (1.0f - alpha + kSuffixEps);
      col_grad[3 * q + 0] += weight * gr;
      col_grad[3 * q + 1] += weight * gg;
      col_grad[3 * q + 2] += weight * gb;
      const float d_alpha = gr * (trans * cr - (fr - ar) * inv) + gg * (trans * cg - (fg - ag) * inv) +
                            gb * (trans * cb - (fb - ab) * inv) - gt * ft * inv;
      float d_in = 0.f, d_out = 0.f;
      if (x > 0.0f && x < 1.0f) {
        const float dci = ci * (1.0f - ci);
        const float dco = co * (1.0f - co);
        const float dx_di = -1.0f / den;
        const float dx_do = ci / (den * den);
        d_in = d_alpha * dx_di * k * dci;
        d_out = d_alpha * dx_do * k * dco;
        dk += d_alpha * (dx_di * si * dci + dx_do * outer_sdf * dco);
      }
      if (prev >= 0) {
        const int target = quirk ? s - 1 : prev;
        sdf_grad[p * S + target] += partial + d_out;
      }
      partial = d_in;
      trans *= 1.0f - alpha;
      outer_cdf = ci;
      outer_sdf = si;
      prev = s;
    }
    if (prev >= 0) {
      const int target = quirk ? S - 1 : prev;
      sdf_grad[p * S + target] += partial;
    }
  }
  float s = wave_sum_f(dk);
  if (__lane_id() == 0) s_red[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float t = s_red[0] + s_red[1] + s_red[2] + s_red[3];
    if (t != 0.0f) atomicAdd(sharp_grad, t);
  }
}

// Mean squared error of the background-composited, [0, 1]-clipped image against a target
// and its gradient with respect to the (rgb, transmittance) accumulators. The clip passes
// gradient where 0 <= x <= 1, like torch.clamp.
__global__ void __launch_bounds__(256) image_l2_loss_kernel(const float* out, const float* bg,
                                                            const float* target, int64_t B,
                                                            int64_t HW, float inv_n, float* loss,
                                                            float* grad) {
  __shared__ float s_red[256 / kWave];
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  float acc = 0.f;
  if (p < B * HW) {
    const int64_t b = p / HW, q = p % HW;
    const float* o = out + b * 4 * HW + q;
    const float* t = target + b * 3 * HW + q;
    float* g = grad + b * 4 * HW + q;
    const float trans = o[3 * HW];
    float gt = 0.f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float pre = o[c * HW] + trans * bg[c];
      const float img = fminf(fmaxf(pre, 0.0f), 1.0f);
      const float diff = img - t[c * HW];
      acc += diff * diff;
      const float gc = (pre >= 0.0f && pre <= 1.0f) ? 2.0f * diff * inv_n : 0.0f;
      g[c * HW] = gc;
      gt += gc * bg[c];
    }
    g[3 * HW] = gt;
  }
  const float s = wave_sum_f(acc);
  if (__lane_id() == 0) s_red[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, (s_red[0] + s_red[1] + s_red[2] + s_red[3]) * inv_n);
}

int check_composite_args(const mt_composite_args* a, const char* name) {
  MT_REQUIRE(a != nullptr, "%s: null args", name);
  MT_REQUIRE(a->tile_h > 0 && a->tile_w > 0 && a->tile_h * a->tile_w <= 1024,
             "%s: tile of %dx%d pixels must hold 1..1024 pixels", name, a->tile_h, a->tile_w);
  MT_REQUIRE(a->chunk >= 0 && a->chunk <= kMaxChunk, "%s: chunk must be in [0, %d]", name, kMaxChunk);
  MT_REQUIRE(a->num_cameras >= 0 && a->grid_rows >= 0 && a->grid_cols >= 0, "%s: bad grid", name);
  MT_REQUIRE(int64_t(a->num_cameras) * a->grid_rows * a->grid_cols < (int64_t(1) << 31),
             "%s: too many tiles", name);
  return MT_OK;
}

}  // namespace
}  // namespace mt

static int resolve_chunk(const mt_composite_args* a) {
  const int pix = a->tile_h * a->tile_w;
  int c = a->chunk > 0 ? a->chunk : pix;
  return c > mt::kMaxChunk ? mt::kMaxChunk : c;
}

extern "C" int mt_intersect_and_composite_forward(const mt_composite_args* args, float* out,
                                                  mt_stream_t stream) {
  if (int rc = mt::check_composite_args(args, "mt_intersect_and_composite_forward")) return rc;
  mt_composite_args a = *args;
  a.chunk = resolve_chunk(args);
  const int64_t blocks = int64_t(a.num_cameras) * a.grid_rows * a.grid_cols;
  if (blocks == 0) return MT_OK;
  mt::fused_forward_kernel<<<static_cast<unsigned>(blocks), a.tile_h * a.tile_w,
                             static_cast<size_t>(a.chunk) * mt::kStagedBytes,
                             static_cast<hipStream_t>(stream)>>>(a, out);
  return mt::check_launch("mt_intersect_and_composite_forward");
}

extern "C" int mt_intersect_and_composite_backward(const mt_composite_args* args, const float* out,
                                                   const float* grad_out, float* colors_grad,
                                                   float* sdf_grad, float* sharpness_grad,
                                                   int32_t reference_quirk, mt_stream_t stream) {
  if (int rc = mt::check_composite_args(args, "mt_intersect_and_composite_backward")) return rc;
  mt_composite_args a = *args;
  a.chunk = resolve_chunk(args);
  const int64_t blocks = int64_t(a.num_cameras) * a.grid_rows * a.grid_cols;
  if (blocks == 0) return MT_OK;
  mt::fused_backward_kernel<<<static_cast<unsigned>(blocks), a.tile_h * a.tile_w,
                              static_cast<size_t>(a.chunk) * mt::kStagedBytes,
                              static_cast<hipStream_t>(stream)>>>(
      a, out, grad_out, colors_grad, sdf_grad, sharpness_grad, reference_quirk);
  return mt::check_launch("mt_intersect_and_composite_backward");
}

extern "C" int mt_image_l2_loss(const float* out, const float* background, const float* target,
                                int32_t B, int32_t H, int32_t W, float* loss, float* grad_out,
                                mt_stream_t stream) {
  MT_REQUIRE(B >= 0 && H >= 0 && W >= 0, "mt_image_l2_loss: negative size");
  const int64_t n = int64_t(B) * H * W;
  if (n == 0) return MT_OK;
  mt::image_l2_loss_kernel<<<mt::blocks_for(n, 256), 256, 0, static_cast<hipStream_t>(stream)>>>(
      out, background, target, B, int64_t(H) * W, 1.0f / float(3 * n), loss, grad_out);
  return mt::check_launch("mt_image_l2_loss");
}

extern "C" int mt_intersect(const int32_t* sorted_values, const int64_t* tile_ranges, const float* xy,
                            int64_t V, const int32_t* faces, int32_t B, int32_t S,
                            int32_t grid_rows, int32_t grid_cols, int32_t tile_h, int32_t tile_w,
                            int32_t height, int32_t width, float* out_uv, int32_t* out_index,
                            mt_stream_t stream) {
  MT_REQUIRE(tile_h > 0 && tile_w > 0 && tile_h * tile_w <= 1024, "mt_intersect: bad tile shape");
  const int64_t blocks = int64_t(B) * S * grid_rows * grid_cols;
  MT_REQUIRE(blocks < (int64_t(1) << 31), "mt_intersect: too many tiles");
  if (blocks == 0) return MT_OK;
  int chunk = tile_h * tile_w;
  if (chunk > mt::kMaxIntersectChunk) chunk = mt::kMaxIntersectChunk;
  mt::intersect_kernel<<<static_cast<unsigned>(blocks), tile_h * tile_w, 0,
                         static_cast<hipStream_t>(stream)>>>(sorted_values, tile_ranges, xy, V,
                                                              faces, B, S, grid_rows, grid_cols,
                                                              tile_h, tile_w, height, width, chunk,
                                                              out_uv, out_index);
  return mt::check_launch("mt_intersect");
}

extern "C" int mt_interpolate_forward(const float* uv, const int32_t* index, const int32_t* faces,
                                      const float* sdf, const float* colors, int64_t V, int32_t B,
                                      int32_t S, int32_t H, int32_t W, float* out_sdf,
                                      float* out_colors, mt_stream_t stream) {
  const int64_t n = int64_t(B) * S * H * W;
  if (n == 0) return MT_OK;
  mt::interpolate_kernel<<<mt::blocks_for(n, 256), 256, 0, static_cast<hipStream_t>(stream)>>>(
      uv, index, faces, sdf, colors, V, S, int64_t(H) * W, n, out_sdf, out_colors);
  return mt::check_launch("mt_interpolate_forward");
}

extern "C" int mt_interpolate_backward(const float* uv, const int32_t* index, const int32_t* faces,
                                       int64_t V, int32_t B, int32_t S, int32_t H, int32_t W,
                                       const float* grad_sdf, const float* grad_colors,
                                       float* sdf_grad, float* colors_grad, mt_stream_t stream) {
  const int64_t n = int64_t(B) * S * H * W;
  if (n == 0) return MT_OK;
  mt::interpolate_backward_kernel<<<mt::blocks_for(n, 256), 256, 0, static_cast<hipStream_t>(stream)>>>(
      uv, index, faces, V, S, int64_t(H) * W, n, grad_sdf, grad_colors, sdf_grad, colors_grad);
  return mt::check_launch("mt_interpolate_backward");
}

extern "C" int mt_composite_forward(const float* sdf, const float* colors, const uint8_t* mask,
                                    int64_t P, int32_t S, const float* sharpness, float* out_colors,
                                    float* out_trans, mt_stream_t stream) {
  if (P == 0) return MT_OK;
  mt::composite_forward_kernel<<<mt::blocks_for(P, 256), 256, 0, static_cast<hipStream_t>(stream)>>>(
      sdf, colors, mask, P, S, sharpness, out_colors, out_trans);
  return mt::check_launch("mt_composite_forward");
}

extern "C" int mt_composite_backward(const float* sdf, const float* colors, const uint8_t* mask,
                                     int64_t P, int32_t S, const float* sharpness,
                                     const float* out_colors, const float* out_trans,
                                     const float* grad_colors, const float* grad_trans,
                                     float* sdf_grad, float* colors_grad, float* sharpness_grad,
                                     int32_t reference_quirk, mt_stream_t stream) {
  if (P == 0) return MT_OK;
  mt::composite_backward_kernel<<<mt::blocks_for(P, 256), 256, 0, static_cast<hipStream_t>(stream)>>>(
      sdf, colors, mask, P, S, sharpness, out_colors, out_trans, grad_colors, grad_trans, sdf_grad,
      colors_grad, sharpness_grad, reference_quirk);
  return mt::check_launch("mt_composite_backward");
}
