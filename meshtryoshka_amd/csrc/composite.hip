// Tile compositing: fused intersect-and-composite (forward and adjoint) and the explicit
// mode's intersect / interpolate / composite.
//
// One workgroup per (camera, tile) with one thread per pixel (rounded up to whole
// waves). The tile's depth-sorted keys are processed in chunks staged in LDS:
//
//   stage    one thread per key: the triangle's barycentric terms (TriTerms) and the
//            rectangle of tile pixels it can possibly hit (candidate_rect, proven
//            conservative), then a block scan of the rectangle areas;
//   test     (pixel, triangle) candidate pairs spread evenly over all threads; each
//            accepted pair sets bit i of its pixel's hit mask (atomicOr in LDS);
//   composite one thread per pixel walks its mask bits in key order and applies the
//            reference's per-pixel compositing rule to the accepted hits, recomputing
//            their barycentrics (the same IEEE operations, so the same values).
//
// A pixel therefore tests only the triangles whose rectangle covers it instead of every
// key of the tile, and the per-pixel traversal order, hits and arithmetic are exactly
// those of the reference's sequential loop.
#include "common.h"
#include "geometry.h"

namespace mt {
namespace {

constexpr float kAlphaEps = 1e-5f;  // intersect_and_composite.slang:161
constexpr float kSuffixEps = 1e-6f; // :3
constexpr float kDummySdf = 1e10f;  // :318

// Staged triangle (4 x float4 = 64 bytes).
struct Staged {
  float4 g0;  // p1.x, p1.y, p2.x, p2.y
  float4 g1;  // e12.x, e12.y, e13.x, e13.y
  float4 g2;  // d_13_13, d_12_13, d_12_12, d_12_23
  float4 g3;  // d_13_23, D = denom + eps, 1 / D (estimate), shell (bits)
};

__device__ __forceinline__ Staged to_staged(const TriTerms& t, int shell) {
  // The estimate is only trusted where its error bound holds (normal, moderate |D|).
  const float ad = fabsf(t.D);
  const float r = (ad >= 0x1p-60f && ad <= 0x1p60f) ? __builtin_amdgcn_rcpf(t.D) : 0.0f;
  Staged s;
  s.g0 = make_float4(t.p1x, t.p1y, t.p2x, t.p2y);
  s.g1 = make_float4(t.e12x, t.e12y, t.e13x, t.e13y);
  s.g2 = make_float4(t.d1313, t.d1213, t.d1212, t.d1223);
  s.g3 = make_float4(t.d1323, t.D, r, __int_as_float(shell));
  return s;
}

__device__ __forceinline__ TriTerms to_terms(const Staged& s) {
  TriTerms t;
  t.p1x = s.g0.x;
  t.p1y = s.g0.y;
  t.p2x = s.g0.z;
  t.p2y = s.g0.w;
  t.e12x = s.g1.x;
  t.e12y = s.g1.y;
  t.e13x = s.g1.z;
  t.e13y = s.g1.w;
  t.d1313 = s.g2.x;
  t.d1213 = s.g2.y;
  t.d1212 = s.g2.z;
  t.d1223 = s.g2.w;
  t.d1323 = s.g3.x;
  t.D = s.g3.y;
  return t;
}

struct Bary {
  float u, v, w;
  bool hit;
};

// The reference's barycentric test (intersect_and_composite.slang:65-145) at pixel centre
// (px, py), with the divisions skipped where rcp(D) already proves a miss (geometry.h;
// s.g3.z = rcp(D), or 0 when D is outside [2^-60, 2^60] or NaN).
__device__ __forceinline__ Bary bary_test(const Staged& s, float px, float py) {
  Bary out;
  const BaryNum n = bary_numerators(to_terms(s), px, py);
  if (rcp_proves_miss(n, s.g3.z)) {
    out.hit = false;
    out.u = out.v = out.w = 0.0f;
    return out;
  }
  const float D = s.g3.y;
  out.v = n.vn / D;
  out.w = n.wn / D;
  out.u = n.un / D;
  out.hit = bary_accept(out.u, out.v, out.w);
  return out;
}

__device__ __forceinline__ float sdf_to_cdf(float sdf, float k) { return 1.0f / (1.0f + expf(-sdf * k)); }

__device__ __forceinline__ int shell_of(const Staged& s) { return __float_as_int(s.g3.w); }

// ------------------------------------------------------------------------------------
// Chunk staging, candidate pairs and per-pixel hit masks (LDS)
// ------------------------------------------------------------------------------------

constexpr int kMaxChunk = 512;
constexpr int kMaxLds = 64 * 1024;
constexpr int kMaxTilePixels = 512;  // one thread per pixel, at most 8 waves per workgroup

struct ChunkLds {
  Staged* geo;       // [chunk]
  uint2* rect;       // [chunk] (x | w << 16, y | h << 16), tile-local pixels
  int32_t* face;     // [chunk]
  int32_t* prefix;   // [chunk + 1] exclusive prefix of rectangle areas
  uint32_t* mask;    // [words][npix]: bit i of word i / 32 = triangle i hits the pixel
  uint8_t* live;     // [npix] pixel still compositing
  int words, npix;
};

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

__host__ __device__ inline size_t chunk_lds_bytes(int chunk, int npix) {
  const int words = (chunk + 31) / 32;
  return align16(sizeof(Staged) * chunk) + align16(sizeof(uint2) * chunk) +
         align16(4 * size_t(chunk)) + align16(4 * size_t(chunk + 1)) +
         align16(4 * size_t(words) * npix) + align16(npix);
}

extern __shared__ __align__(16) char g_dynamic_lds[];

__device__ __forceinline__ ChunkLds carve_chunk(int chunk, int npix) {
  ChunkLds L;
  char* p = g_dynamic_lds;
  L.words = (chunk + 31) / 32;
  L.npix = npix;
  L.geo = reinterpret_cast<Staged*>(p);
  p += align16(sizeof(Staged) * chunk);
  L.rect = reinterpret_cast<uint2*>(p);
  p += align16(sizeof(uint2) * chunk);
  L.face = reinterpret_cast<int32_t*>(p);
  p += align16(4 * size_t(chunk));
  L.prefix = reinterpret_cast<int32_t*>(p);
  p += align16(4 * size_t(chunk + 1));
  L.mask = reinterpret_cast<uint32_t*>(p);
  p += align16(4 * size_t(L.words) * npix);
  L.live = reinterpret_cast<uint8_t*>(p);
  return L;
}

// Exclusive block prefix sum (blockDim.x a multiple of 64, at most 512); *total gets the
// block sum. Ends with a barrier, so s_wave may be reused at once.
__device__ __forceinline__ int block_exclusive_sum(int x, int* s_wave, int* total) {
  const int wave = threadIdx.x / kWave;
  const int lane = __lane_id();
  const int inc = wave_inclusive_sum(x);
  if (lane == kWave - 1) s_wave[wave] = inc;
  __syncthreads();
  int before = 0, sum = 0;
  const int nw = blockDim.x / kWave;
  for (int w = 0; w < nw; ++w) {
    const int v = s_wave[w];
    if (w < wave) before += v;
    sum += v;
  }
  *total = sum;
  __syncthreads();
  return before + inc - x;
}

struct TileCoord {
  int cam, tile, x0, y0;  // tile origin in pixels
  int pix;                // this thread's pixel (tile-local), -1 for padding threads
  int x, y;
  bool in_image;
  int height, width;      // the camera's image
  int64_t slot;           // the tile's pair in tile_ranges
  int64_t pixel_base;     // the camera's first pixel (planes of height x width)
};

__device__ __forceinline__ TileCoord tile_coord(int bid_linear, int grid_cols, int T, int th,
                                                int tw, int height, int width) {
  TileCoord c;
  c.cam = bid_linear / T;
  c.tile = bid_linear % T;
  c.slot = bid_linear;
  c.height = height;
  c.width = width;
  c.pixel_base = int64_t(c.cam) * height * width;
  const int ty = c.tile / grid_cols, tx = c.tile % grid_cols;
  c.x0 = tx * tw;
  c.y0 = ty * th;
  const int npix = th * tw;
  c.pix = int(threadIdx.x) < npix ? int(threadIdx.x) : -1;
  const int ly = c.pix >= 0 ? c.pix / tw : 0, lx = c.pix >= 0 ? c.pix % tw : 0;
  c.x = c.x0 + lx;
  c.y = c.y0 + ly;
  c.in_image = c.pix >= 0 && c.x < width && c.y < height;
  return c;
}

// Workgroup bid of a fused launch: its camera and tile, with the camera table when the
// cameras have different resolutions (block ids run over every camera's tiles in order).
__device__ __forceinline__ TileCoord locate_tile(const mt_composite_args& a, int64_t bid) {
  if (a.tile_base == nullptr)
    return tile_coord(int(bid), a.grid_cols, a.grid_rows * a.grid_cols, a.tile_h, a.tile_w,
                      a.height, a.width);
  int lo = 0, hi = a.num_cameras - 1;  // last camera whose first tile is <= bid
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a.tile_base[mid] <= bid) lo = mid; else hi = mid - 1;
  }
  const int h = a.cam_dims[4 * lo + 0], w = a.cam_dims[4 * lo + 1];
  const int rows = a.cam_dims[4 * lo + 2], cols = a.cam_dims[4 * lo + 3];
  const int64_t local = bid - a.tile_base[lo];
  TileCoord c = tile_coord(int(local), cols, rows * cols, a.tile_h, a.tile_w, h, w);
  c.cam = lo;
  c.slot = bid;
  c.pixel_base = a.pixel_base[lo];
  return c;
}

// Stage keys [first, first + count) of one tile: terms, candidate rectangle, face; clear
// the hit masks. Returns nothing; callers barrier afterwards.
// Largest candidate box taken from the bounding box of a certified triangle (larger boxes
// use the candidate rectangle, which is tighter for thin diagonal triangles).
constexpr int kStageCentrePixels = 16;

__device__ __forceinline__ void stage_chunk(const ChunkLds& L, const float2* xy,
                                            const int32_t* faces, const int32_t* values,
                                            const int64_t* keys, int shell_offset,
                                            int shell_bits, int64_t first, int count, int x0,
                                            int y0, int tw, int th) {
  for (int j = threadIdx.x; j < count; j += blockDim.x) {
    const int32_t face = values[first + j];
    const int shell =
        keys ? static_cast<int>(key_field(static_cast<uint64_t>(keys[first + j]), shell_offset,
                                          shell_bits))
             : 0;
    const int64_t va = faces[3 * int64_t(face) + 0];
    const int64_t vb = faces[3 * int64_t(face) + 1];
    const int64_t vc = faces[3 * int64_t(face) + 2];
    const float2 p1 = xy[va], p2 = xy[vb], p3 = xy[vc];
    const TriTerms t = tri_terms(p1.x, p1.y, p2.x, p2.y, p3.x, p3.y);
    // A small, well-shaped triangle's candidates are the pixel centres near its bounding
    // box (geometry.h, centre_range: float arithmetic under a proven shape certificate);
    // the others get the candidate rectangle (double precision).
    const CentreRange cen = centre_range(p1.x, p1.y, p2.x, p2.y, p3.x, p3.y, x0, y0, tw, th,
                                         kStageCentrePixels);
    const Rect r = cen.state == kCentresNone  ? Rect{0, 0, 0, 0}
                   : cen.state == kCentresFew ? Rect{cen.x, cen.y, cen.w, cen.h}
                                              : candidate_rect(t, x0, y0, tw, th);
    L.geo[j] = to_staged(t, shell);
    L.rect[j] = make_uint2(uint32_t(r.x) | (uint32_t(r.w) << 16), uint32_t(r.y) | (uint32_t(r.h) << 16));
    L.face[j] = face;
  }
  const int mask_words = L.words * L.npix;
  for (int i = threadIdx.x; i < mask_words; i += blockDim.x) L.mask[i] = 0u;
}

// Exclusive prefix of the staged rectangle areas into L.prefix[0..count]; returns the
// number of candidate pairs. Each thread scans a contiguous run of entries.
__device__ __forceinline__ int scan_areas(const ChunkLds& L, int count, int* s_wave) {
  const int per = (count + int(blockDim.x) - 1) / int(blockDim.x);
  const int begin = int(threadIdx.x) * per;
  int local = 0;
  for (int k = 0; k < per; ++k) {
    const int j = begin + k;
    if (j < count) {
      const uint2 rc = L.rect[j];
      local += int(rc.x >> 16) * int(rc.y >> 16);
    }
  }
  int total;
  int run = block_exclusive_sum(local, s_wave, &total);
  for (int k = 0; k < per; ++k) {
    const int j = begin + k;
    if (j < count) {
      L.prefix[j] = run;
      const uint2 rc = L.rect[j];
      run += int(rc.x >> 16) * int(rc.y >> 16);
    }
  }
  if (threadIdx.x == 0) L.prefix[count] = total;
  return total;
}

// Test every candidate pair of the chunk; set mask bits of accepted hits. Pairs of pixels
// that no longer composite are skipped. Returns the number of tests run by this thread.
__device__ __forceinline__ uint32_t test_pairs(const ChunkLds& L, int count, int total, int x0,
                                               int y0, int tw) {
  uint32_t tests = 0;
  for (int q = threadIdx.x; q < total; q += blockDim.x) {
    int lo = 0, hi = count - 1;  // last entry whose prefix is <= q (it has a nonzero area)
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (L.prefix[mid] <= q) lo = mid; else hi = mid - 1;
    }
    const uint2 rc = L.rect[lo];
    const int w = int(rc.x >> 16);
    const int local = q - L.prefix[lo];
    const int lx = int(rc.x & 0xFFFFu) + local % w;
    const int ly = int(rc.y & 0xFFFFu) + local / w;
    const int pix = ly * tw + lx;
    if (!L.live[pix]) continue;
    ++tests;
    const Bary bc = bary_test(L.geo[lo], float(x0 + lx) + 0.5f, float(y0 + ly) + 0.5f);
    if (bc.hit) atomicOr(&L.mask[(lo >> 5) * L.npix + pix], 1u << (lo & 31));
  }
  return tests;
}

// Block-reduced statistics counters (one atomic pair per block).
__device__ void add_stats(unsigned long long* stats, uint32_t tests, uint32_t hits) {
  __shared__ unsigned long long s_t[kMaxTilePixels / kWave], s_h[kMaxTilePixels / kWave];
  unsigned long long t = wave_sum<unsigned long long>(tests);
  unsigned long long h = wave_sum<unsigned long long>(hits);
  if (__lane_id() == 0) {
    s_t[threadIdx.x / kWave] = t;
    s_h[threadIdx.x / kWave] = h;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long st = 0, sh = 0;
    for (int w = 0; w < int(blockDim.x) / kWave; ++w) {
      st += s_t[w];
      sh += s_h[w];
    }
    atomicAdd(stats + 0, st);
    atomicAdd(stats + 1, sh);
  }
}

struct Corner {
  float r, g, b, sdf;
};

// Barycentric interpolation of colour and SDF of one hit (intersect_and_composite.slang:
// 150-158); the three corners are read from global memory (hits are rare per key).
struct Sample {
  float r, g, b, sdf;
  int64_t v[3];
};

// Camera cam's part of a colour array (colours or their gradients): its first float and the
// floats between consecutive vertices. colors_stride 0: f32[B, V, 3] (camera-major);
// s > 0: f32[V, s, 3] (each vertex's cameras together).
__device__ __forceinline__ int64_t camera_colors_offset(int stride, int cam, int64_t V) {
  return stride > 0 ? 3 * int64_t(cam) : 3 * int64_t(cam) * V;
}
__device__ __forceinline__ int64_t vertex_colors_step(int stride) {
  return stride > 0 ? 3 * int64_t(stride) : 3;
}

__device__ __forceinline__ Sample interpolate_hit(const mt_composite_args& a, const float* col,
                                                  int32_t face, const Bary& bc) {
  Sample s;
  s.v[0] = a.faces[3 * int64_t(face) + 0];
  s.v[1] = a.faces[3 * int64_t(face) + 1];
  s.v[2] = a.faces[3 * int64_t(face) + 2];
  const int64_t vs = vertex_colors_step(a.colors_stride);
  const float* ca = col + vs * s.v[0];
  const float* cb = col + vs * s.v[1];
  const float* cc = col + vs * s.v[2];
  s.sdf = (bc.u * a.sdf[s.v[0]] + bc.v * a.sdf[s.v[1]]) + bc.w * a.sdf[s.v[2]];
  s.r = (bc.u * ca[0] + bc.v * cb[0]) + bc.w * cc[0];
  s.g = (bc.u * ca[1] + bc.v * cb[1]) + bc.w * cc[1];
  s.b = (bc.u * ca[2] + bc.v * cb[2]) + bc.w * cc[2];
  return s;
}

// ------------------------------------------------------------------------------------
// Fused forward (intersect_and_composite.slang:224-339)
// ------------------------------------------------------------------------------------

__global__ void __launch_bounds__(512) fused_forward_kernel(mt_composite_args a, float* out) {
  const int npix = a.tile_h * a.tile_w;
  const ChunkLds L = carve_chunk(a.chunk, npix);
  __shared__ int s_wave[kMaxTilePixels / kWave];
  const TileCoord tc = locate_tile(a, xcd_swizzle(blockIdx.x, gridDim.x));
  const int64_t start = a.tile_ranges[2 * tc.slot + 0];
  const int64_t end = a.tile_ranges[2 * tc.slot + 1];
  const float k = a.sharpness[0];
  const float px = float(tc.x) + 0.5f, py = float(tc.y) + 0.5f;
  const float2* xy = reinterpret_cast<const float2*>(a.xy) + int64_t(tc.cam) * a.num_vertices;
  const float* col = a.colors + camera_colors_offset(a.colors_stride, tc.cam, a.num_vertices);

  float r = 0.f, g = 0.f, bl = 0.f, trans = 1.f;
  float outer_cdf = sdf_to_cdf(kDummySdf, k);
  int last_shell = -1;
  bool active = tc.in_image;
  if (tc.pix >= 0) L.live[tc.pix] = active ? 1 : 0;
  uint32_t n_tests = 0, n_hits = 0;
  const int chunk = a.chunk;
  for (int64_t first = start; first < end; first += chunk) {
    const int count = static_cast<int>((end - first) < chunk ? (end - first) : chunk);
    __syncthreads();
    stage_chunk(L, xy, a.faces, a.sorted_values, a.sorted_keys, a.layout.shell_offset,
                a.layout.shell_bits, first, count, tc.x0, tc.y0, a.tile_w, a.tile_h);
    __syncthreads();
    const int total = scan_areas(L, count, s_wave);
    __syncthreads();
    n_tests += test_pairs(L, count, total, tc.x0, tc.y0, a.tile_w);
    __syncthreads();
    if (active) {
      const int words = (count + 31) / 32;
      for (int wd = 0; wd < words && active; ++wd) {
        uint32_t bits = L.mask[wd * npix + tc.pix];
        while (bits) {
          const int i = wd * 32 + __builtin_ctz(bits);
          bits &= bits - 1u;
          const Staged st = L.geo[i];
          const int shell = shell_of(st);
          if (shell <= last_shell) continue;
          const Bary bc = bary_test(st, px, py);
          if (!bc.hit) continue;
          ++n_hits;
          const Sample sm = interpolate_hit(a, col, L.face[i], bc);
          const float inner_cdf = sdf_to_cdf(sm.sdf, k);
          float alpha = ((outer_cdf - inner_cdf) + kAlphaEps) / (outer_cdf + kAlphaEps);
          alpha = fminf(fmaxf(alpha, 0.0f), 1.0f);
          r = r + (sm.r * trans) * alpha;
          g = g + (sm.g * trans) * alpha;
          bl = bl + (sm.b * trans) * alpha;
          trans = trans * (1.0f - alpha);
          outer_cdf = inner_cdf;
          last_shell = shell;
          if (shell == a.num_shells - 1) {
            active = false;
            break;
          }
        }
      }
      L.live[tc.pix] = active ? 1 : 0;
    }
    if (!__syncthreads_or(active)) break;
  }
  if (tc.in_image) {
    const int64_t plane = int64_t(tc.height) * tc.width;
    float* o = out + 4 * tc.pixel_base + int64_t(tc.y) * tc.width + tc.x;
    o[0] = r;
    o[plane] = g;
    o[2 * plane] = bl;
    o[3 * plane] = trans;
  }
  if (a.stats) add_stats(a.stats, n_tests, n_hits);
}

// Scatter one sample's gradient (colour and SDF) to its triangle's corners.
__device__ __forceinline__ void scatter_sample(const mt_composite_args& a, int cam,
                                               const int64_t* v, float bu, float bv, float bw,
                                               float dr, float dg, float db, float dsdf,
                                               float* colors_grad, float* sdf_grad) {
  const float wgt[3] = {bu, bv, bw};
  float* cgrad = colors_grad + camera_colors_offset(a.colors_stride, cam, a.num_vertices);
  const int64_t vs = vertex_colors_step(a.colors_stride);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (a.vertex_touched) a.vertex_touched[v[q]] = 1;
    atomicAdd(cgrad + vs * v[q] + 0, wgt[q] * dr);
    atomicAdd(cgrad + vs * v[q] + 1, wgt[q] * dg);
    atomicAdd(cgrad + vs * v[q] + 2, wgt[q] * db);
    atomicAdd(sdf_grad + v[q], wgt[q] * dsdf);
  }
}

// Per staged triangle of the current chunk, its three vertices' colour (3) and SDF (1)
// gradient sums in LDS: the pixels of a tile that hit one triangle add into LDS, and one
// set of global atomics per triangle and chunk replaces one per hit.
constexpr int kGradSlots = 12;

__host__ __device__ inline size_t grad_lds_bytes(int chunk) {
  return align16(sizeof(float) * kGradSlots * size_t(chunk));
}

__device__ __forceinline__ void scatter_sample_lds(float* acc, float bu, float bv, float bw,
                                                   float dr, float dg, float db, float dsdf) {
  const float wgt[3] = {bu, bv, bw};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    atomicAdd(acc + 4 * q + 0, wgt[q] * dr);
    atomicAdd(acc + 4 * q + 1, wgt[q] * dg);
    atomicAdd(acc + 4 * q + 2, wgt[q] * db);
    atomicAdd(acc + 4 * q + 3, wgt[q] * dsdf);
  }
}

// Adds the chunk's nonzero per-triangle sums to the global gradients and clears them for
// the next chunk (the sums start at zero and only hit triangles' sums become nonzero).
__device__ __forceinline__ void flush_chunk_grads(const ChunkLds& L, float* acc, int count,
                                                  const int32_t* faces, int64_t num_vertices,
                                                  int cam, int colors_stride, float* colors_grad,
                                                  float* sdf_grad, uint8_t* touched) {
  float* cgrad = colors_grad + camera_colors_offset(colors_stride, cam, num_vertices);
  const int64_t vs = vertex_colors_step(colors_stride);
  for (int j = threadIdx.x; j < count; j += blockDim.x) {
    float* g = acc + kGradSlots * j;
    float e[kGradSlots];
    bool any = false;
#pragma unroll
    for (int i = 0; i < kGradSlots; ++i) {
      e[i] = g[i];
      any = any || e[i] != 0.0f;
    }
    if (!any) continue;
    const int64_t face = L.face[j];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int64_t v = faces[3 * face + q];
      if (touched) touched[v] = 1;
#pragma unroll
      for (int c = 0; c < 3; ++c)
        if (e[4 * q + c] != 0.0f) atomicAdd(cgrad + vs * v + c, e[4 * q + c]);
      if (e[4 * q + 3] != 0.0f) atomicAdd(sdf_grad + v, e[4 * q + 3]);
    }
#pragma unroll
    for (int i = 0; i < kGradSlots; ++i) g[i] = 0.0f;
  }
}

// ------------------------------------------------------------------------------------
// Fused backward (intersect_and_composite.slang:341-564): the same traversal, recomputed,
// with the suffix trick for the accumulated colour. A sample's gradient is added to its
// triangle's LDS sums when the triangle is still staged (the usual case), otherwise
// straight to global memory.
// ------------------------------------------------------------------------------------

__global__ void __launch_bounds__(512) fused_backward_kernel(mt_composite_args a, const float* out,
                                                              const float* gout, float* colors_grad,
                                                              float* sdf_grad, float* sharp_grad,
                                                              int quirk) {
  const int npix = a.tile_h * a.tile_w;
  const ChunkLds L = carve_chunk(a.chunk, npix);
  float* gacc = reinterpret_cast<float*>(g_dynamic_lds + chunk_lds_bytes(a.chunk, npix));
  for (int j = threadIdx.x; j < kGradSlots * a.chunk; j += blockDim.x) gacc[j] = 0.0f;
  __shared__ int s_wave[kMaxTilePixels / kWave];
  __shared__ float s_red[kMaxTilePixels / kWave];
  const TileCoord tc = locate_tile(a, xcd_swizzle(blockIdx.x, gridDim.x));
  const int64_t start = a.tile_ranges[2 * tc.slot + 0];
  const int64_t end = a.tile_ranges[2 * tc.slot + 1];
  const float k = a.sharpness[0];
  const float px = float(tc.x) + 0.5f, py = float(tc.y) + 0.5f;
  const float2* xy = reinterpret_cast<const float2*>(a.xy) + int64_t(tc.cam) * a.num_vertices;
  const float* col = a.colors + camera_colors_offset(a.colors_stride, tc.cam, a.num_vertices);

  float fr = 0.f, fg = 0.f, fb = 0.f, ft = 0.f;    // final accumulator
  float gr = 0.f, gg = 0.f, gb = 0.f, gt = 0.f;    // its gradient
  if (tc.in_image) {
    const int64_t plane = int64_t(tc.height) * tc.width;
    const int64_t o = 4 * tc.pixel_base + int64_t(tc.y) * tc.width + tc.x;
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
  float outer_cdf = sdf_to_cdf(kDummySdf, k), outer_sdf = kDummySdf;
  int last_shell = -1;
  bool active = tc.in_image;
  if (tc.pix >= 0) L.live[tc.pix] = active ? 1 : 0;
  float dk = 0.f;
  uint32_t n_tests = 0, n_hits = 0;
  // Sample waiting for its outer-role SDF gradient (set by the next accepted hit).
  bool pending = false;
  int p_slot = -1;  // the pending sample's triangle in the staged chunk, -1 once restaged
  int64_t p_v[3] = {0, 0, 0};
  float p_u = 0.f, p_v_ = 0.f, p_w = 0.f, p_dr = 0.f, p_dg = 0.f, p_db = 0.f, p_dsdf = 0.f;

  const int chunk = a.chunk;
  for (int64_t first = start; first < end; first += chunk) {
    const int count = static_cast<int>((end - first) < chunk ? (end - first) : chunk);
    __syncthreads();
    stage_chunk(L, xy, a.faces, a.sorted_values, a.sorted_keys, a.layout.shell_offset,
                a.layout.shell_bits, first, count, tc.x0, tc.y0, a.tile_w, a.tile_h);
    p_slot = -1;
    __syncthreads();
    const int total = scan_areas(L, count, s_wave);
    __syncthreads();
    n_tests += test_pairs(L, count, total, tc.x0, tc.y0, a.tile_w);
    __syncthreads();
    if (active) {
      const int words = (count + 31) / 32;
      for (int wd = 0; wd < words && active; ++wd) {
        uint32_t bits = L.mask[wd * npix + tc.pix];
        while (bits) {
          const int i = wd * 32 + __builtin_ctz(bits);
          bits &= bits - 1u;
          const Staged st = L.geo[i];
          const int shell = shell_of(st);
          if (shell <= last_shell) continue;
          const Bary bc = bary_test(st, px, py);
          if (!bc.hit) continue;
          ++n_hits;
          const Sample sm = interpolate_hit(a, col, L.face[i], bc);
          const float sdf = sm.sdf;
          const float ci = sdf_to_cdf(sdf, k);
          const float co = outer_cdf;
          const float den = co + kAlphaEps;
          const float x = ((co - ci) + kAlphaEps) / den;
          const float alpha = fminf(fmaxf(x, 0.0f), 1.0f);
          // Suffix trick (slang:460-473).
          const float weight = alpha * trans;
          ar += weight * sm.r;
          ag += weight * sm.g;
          ab += weight * sm.b;
          const float inv = 1.0f / (1.0f - alpha + kSuffixEps);
          const float d_alpha = gr * (trans * sm.r - (fr - ar) * inv) +
                                gg * (trans * sm.g - (fg - ag) * inv) +
                                gb * (trans * sm.b - (fb - ab) * inv) - gt * ft * inv;
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
            const float su = q ? bc.u : p_u, sv = q ? bc.v : p_v_, sw = q ? bc.w : p_w;
            if (p_slot >= 0)
              scatter_sample_lds(gacc + kGradSlots * p_slot, su, sv, sw, p_dr, p_dg, p_db, p_dsdf);
            else
              scatter_sample(a, tc.cam, p_v, su, sv, sw, p_dr, p_dg, p_db, p_dsdf, colors_grad,
                             sdf_grad);
          }
          pending = true;
          p_slot = i;
          p_v[0] = sm.v[0];
          p_v[1] = sm.v[1];
          p_v[2] = sm.v[2];
          p_u = bc.u;
          p_v_ = bc.v;
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
      L.live[tc.pix] = active ? 1 : 0;
    }
    const bool more = __syncthreads_or(active);  // also: every LDS sum of the chunk is in
    flush_chunk_grads(L, gacc, count, a.faces, a.num_vertices, tc.cam, a.colors_stride,
                      colors_grad, sdf_grad, a.vertex_touched);
    if (!more) break;
  }
  if (pending) {  // its triangle's chunk has been flushed
    scatter_sample(a, tc.cam, p_v, p_u, p_v_, p_w, p_dr, p_dg, p_db, p_dsdf, colors_grad,
                   sdf_grad);
  }
  // Sharpness gradient: block reduction, one atomic per block.
  float s = wave_sum_f(dk);
  if (__lane_id() == 0) s_red[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int w = 0; w < int(blockDim.x) / kWave; ++w) t += s_red[w];
    if (t != 0.0f) atomicAdd(sharp_grad, t);
  }
  if (a.stats) add_stats(a.stats, n_tests, n_hits);
}

// ------------------------------------------------------------------------------------
// Explicit mode: first hit per (camera, shell, pixel) (intersect.slang:85-182)
// ------------------------------------------------------------------------------------

constexpr int kIntersectChunk = 256;

__global__ void __launch_bounds__(512) intersect_kernel(const int32_t* values, const int64_t* ranges,
                                                         const float* xy, int64_t V,
                                                         const int32_t* faces, int B, int S,
                                                         int grid_rows, int grid_cols, int th,
                                                         int tw, int H, int W, int chunk,
                                                         float* out_uv, int32_t* out_index) {
  const int npix = th * tw;
  const ChunkLds L = carve_chunk(chunk, npix);
  __shared__ int s_wave[kMaxTilePixels / kWave];
  const int T = grid_rows * grid_cols;
  // Block = (camera * S + shell, tile).
  const TileCoord tc = tile_coord(int(xcd_swizzle(blockIdx.x, gridDim.x)), grid_cols, T, th, tw,
                                  H, W);
  const int bs = tc.cam;
  const int cam = bs / S;
  const int64_t start = ranges[2 * (int64_t(bs) * T + tc.tile) + 0];
  const int64_t end = ranges[2 * (int64_t(bs) * T + tc.tile) + 1];
  const float px = float(tc.x) + 0.5f, py = float(tc.y) + 0.5f;
  const float2* pxy = reinterpret_cast<const float2*>(xy) + int64_t(cam) * V;
  bool active = tc.in_image;
  if (tc.pix >= 0) L.live[tc.pix] = active ? 1 : 0;
  float hu = 0.f, hv = 0.f;
  int32_t hidx = -1;
  for (int64_t first = start; first < end; first += chunk) {
    const int count = static_cast<int>((end - first) < chunk ? (end - first) : chunk);
    __syncthreads();
    stage_chunk(L, pxy, faces, values, nullptr, 0, 0, first, count, tc.x0, tc.y0, tw, th);
    __syncthreads();
    const int total = scan_areas(L, count, s_wave);
    __syncthreads();
    test_pairs(L, count, total, tc.x0, tc.y0, tw);
    __syncthreads();
    if (active) {
      const int words = (count + 31) / 32;
      for (int wd = 0; wd < words && active; ++wd) {
        uint32_t bits = L.mask[wd * npix + tc.pix];
        while (bits) {
          const int i = wd * 32 + __builtin_ctz(bits);
          bits &= bits - 1u;
          const Bary bc = bary_test(L.geo[i], px, py);
          if (!bc.hit) continue;
          hu = bc.u;
          hv = bc.v;
          hidx = L.face[i];
          active = false;
          break;
        }
      }
      L.live[tc.pix] = active ? 1 : 0;
    }
    if (!__syncthreads_or(active)) break;
  }
  if (tc.in_image) {
    const int64_t p = (int64_t(bs) * H + tc.y) * W + tc.x;
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
                                                            int64_t HW_uniform,
                                                            const int64_t* pixel_base,
                                                            int64_t total, float inv_n,
                                                            double* loss, float* grad) {
  __shared__ float s_red[256 / kWave];
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  float acc = 0.f;
  if (p < total) {
    int64_t first, HW;  // the camera's first pixel and pixel count
    if (pixel_base) {
      int lo = 0, hi = int(B) - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pixel_base[mid] <= p) lo = mid; else hi = mid - 1;
      }
      first = pixel_base[lo];
      HW = pixel_base[lo + 1] - first;
    } else {
      HW = HW_uniform;
      first = (p / HW) * HW;
    }
    const int64_t q = p - first;
    const float* o = out + 4 * first + q;
    const float* t = target + 3 * first + q;
    float* g = grad + 4 * first + q;
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
  // Block partials are summed in float64: tens of thousands of FP32 atomic additions to
  // one float would leave the mean ~1e-6 (relative) from the exact sum of the same terms.
  if (threadIdx.x == 0) {
    const double part = double(s_red[0]) + double(s_red[1]) + double(s_red[2]) + double(s_red[3]);
    atomicAdd(loss, part * double(inv_n));
  }
}

int check_composite_args(const mt_composite_args* a, const char* name) {
  MT_REQUIRE(a != nullptr, "%s: null args", name);
  MT_REQUIRE(a->tile_h > 0 && a->tile_w > 0 && a->tile_h * a->tile_w <= kMaxTilePixels,
             "%s: tile of %dx%d pixels must hold 1..%d pixels", name, a->tile_h, a->tile_w,
             kMaxTilePixels);
  MT_REQUIRE(a->chunk >= 0 && a->chunk <= kMaxChunk, "%s: chunk must be in [0, %d]", name, kMaxChunk);
  MT_REQUIRE(a->num_cameras >= 0 && a->grid_rows >= 0 && a->grid_cols >= 0, "%s: bad grid", name);
  MT_REQUIRE(a->colors_stride == 0 || a->colors_stride >= a->num_cameras,
             "%s: colors_stride must be 0 or at least num_cameras", name);
  MT_REQUIRE(int64_t(a->num_cameras) * a->grid_rows * a->grid_cols < (int64_t(1) << 31),
             "%s: too many tiles", name);
  MT_REQUIRE((a->cam_dims == nullptr) == (a->tile_base == nullptr) &&
                 (a->tile_base == nullptr) == (a->pixel_base == nullptr),
             "%s: cam_dims, tile_base and pixel_base go together", name);
  MT_REQUIRE(a->tile_base == nullptr || (a->total_tiles >= 0 && a->total_tiles < (int64_t(1) << 31)),
             "%s: bad total_tiles", name);
  return MT_OK;
}

}  // namespace
}  // namespace mt

// Chunk actually staged: the requested one (0 = one key per pixel), at most kMaxChunk and
// small enough that the stage buffers and hit masks fit the LDS budget.
static int resolve_chunk(int requested, int npix, bool grads = false) {
  int c = requested > 0 ? requested : npix;
  if (c > mt::kMaxChunk) c = mt::kMaxChunk;
  auto bytes = [&](int cc) {
    return mt::chunk_lds_bytes(cc, npix) + (grads ? mt::grad_lds_bytes(cc) : size_t(0));
  };
  while (c > 32 && bytes(c) > size_t(mt::kMaxLds)) c -= 32;
  return c;
}

static unsigned block_threads(int npix) { return static_cast<unsigned>((npix + 63) / 64 * 64); }

extern "C" int mt_intersect_and_composite_forward(const mt_composite_args* args, float* out,
                                                  mt_stream_t stream) {
  if (int rc = mt::check_composite_args(args, "mt_intersect_and_composite_forward")) return rc;
  mt_composite_args a = *args;
  const int npix = a.tile_h * a.tile_w;
  a.chunk = resolve_chunk(args->chunk, npix);
  const int64_t blocks =
      a.tile_base ? a.total_tiles : int64_t(a.num_cameras) * a.grid_rows * a.grid_cols;
  if (blocks == 0) return MT_OK;
  mt::fused_forward_kernel<<<static_cast<unsigned>(blocks), block_threads(npix),
                             mt::chunk_lds_bytes(a.chunk, npix), static_cast<hipStream_t>(stream)>>>(
      a, out);
  return mt::check_launch("mt_intersect_and_composite_forward");
}

extern "C" int mt_intersect_and_composite_backward(const mt_composite_args* args, const float* out,
                                                   const float* grad_out, float* colors_grad,
                                                   float* sdf_grad, float* sharpness_grad,
                                                   int32_t reference_quirk, mt_stream_t stream) {
  if (int rc = mt::check_composite_args(args, "mt_intersect_and_composite_backward")) return rc;
  mt_composite_args a = *args;
  const int npix = a.tile_h * a.tile_w;
  a.chunk = resolve_chunk(args->chunk, npix, true);
  const int64_t blocks =
      a.tile_base ? a.total_tiles : int64_t(a.num_cameras) * a.grid_rows * a.grid_cols;
  if (blocks == 0) return MT_OK;
  mt::fused_backward_kernel<<<static_cast<unsigned>(blocks), block_threads(npix),
                              mt::chunk_lds_bytes(a.chunk, npix) + mt::grad_lds_bytes(a.chunk),
                              static_cast<hipStream_t>(stream)>>>(
      a, out, grad_out, colors_grad, sdf_grad, sharpness_grad, reference_quirk);
  return mt::check_launch("mt_intersect_and_composite_backward");
}

extern "C" int mt_image_l2_loss(const float* out, const float* background, const float* target,
                                int32_t B, int32_t H, int32_t W, const int64_t* pixel_base,
                                int64_t total_pixels, int64_t mean_pixels, double* loss,
                                float* grad_out, mt_stream_t stream) {
  MT_REQUIRE(B >= 0 && H >= 0 && W >= 0 && total_pixels >= 0 && mean_pixels >= 0,
             "mt_image_l2_loss: negative size");
  const int64_t n = pixel_base ? total_pixels : int64_t(B) * H * W;
  if (n == 0) return MT_OK;
  const int64_t m = mean_pixels > 0 ? mean_pixels : n;
  mt::image_l2_loss_kernel<<<mt::blocks_for(n, 256), 256, 0, static_cast<hipStream_t>(stream)>>>(
      out, background, target, B, int64_t(H) * W, pixel_base, n, 1.0f / float(3 * m), loss,
      grad_out);
  return mt::check_launch("mt_image_l2_loss");
}

extern "C" int mt_intersect(const int32_t* sorted_values, const int64_t* tile_ranges, const float* xy,
                            int64_t V, const int32_t* faces, int32_t B, int32_t S,
                            int32_t grid_rows, int32_t grid_cols, int32_t tile_h, int32_t tile_w,
                            int32_t height, int32_t width, float* out_uv, int32_t* out_index,
                            mt_stream_t stream) {
  MT_REQUIRE(tile_h > 0 && tile_w > 0 && tile_h * tile_w <= mt::kMaxTilePixels,
             "mt_intersect: tile of %dx%d pixels must hold 1..%d pixels", tile_h, tile_w,
             mt::kMaxTilePixels);
  const int64_t blocks = int64_t(B) * S * grid_rows * grid_cols;
  MT_REQUIRE(blocks < (int64_t(1) << 31), "mt_intersect: too many tiles");
  if (blocks == 0) return MT_OK;
  const int npix = tile_h * tile_w;
  const int chunk = resolve_chunk(mt::kIntersectChunk, npix);
  mt::intersect_kernel<<<static_cast<unsigned>(blocks), block_threads(npix),
                         mt::chunk_lds_bytes(chunk, npix), static_cast<hipStream_t>(stream)>>>(sorted_values, tile_ranges, xy, V,
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
