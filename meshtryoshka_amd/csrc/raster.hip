// Projection, triangle bounding, key generation, tile delineation and spherical
// harmonics. Floating-point expressions keep the reference's operation order; the
// library is compiled without FP contraction so that keys (which embed depth bits) and
// tile bounds are reproducible bit for bit.
#include "common.h"
#include "geometry.h"
#include "sh.h"

namespace mt {
namespace {

constexpr int kBlock = 256;

// ------------------------------------------------------------------------------------
// Projection (project_vertices.slang:51-92)
// ------------------------------------------------------------------------------------

// One world vertex through one camera (project_vertices.slang:51-69): p = E [v, 1] summed
// left to right, depth = p.z, xy = (K p).xy / depth. e: the 4 x 4 world-to-camera matrix
// (row-major), k: the 3 x 3 intrinsics.
struct Projected {
  float x, y, depth;
};

__device__ __forceinline__ Projected project_point(const float* e, const float* k, float x,
                                                   float y, float z) {
  float cs[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) cs[r] = ((e[4 * r + 0] * x + e[4 * r + 1] * y) + e[4 * r + 2] * z) + e[4 * r + 3];
  const float depth = cs[2];
  const float px = (k[0] * cs[0] + k[1] * cs[1]) + k[2] * cs[2];
  const float py = (k[3] * cs[0] + k[4] * cs[1]) + k[5] * cs[2];
  return Projected{px / depth, py / depth, depth};
}

// Four consecutive vertices per thread, every camera: the vertices are read once (three
// 16-byte loads), and each camera's xy and depth of the four are written with 16-byte
// stores where the output offset allows (8-byte or 4-byte stores otherwise; vec_in /
// vec_out: the vertex / output arrays are 16-byte aligned). Vertices at or past *num_live
// (when given) are neither read nor written.
__global__ void __launch_bounds__(kBlock) project_kernel(const float* verts, int64_t V,
                                                         const int32_t* num_live, const float* E,
                                                         const float* K, int B, float* out_xy,
                                                         float* out_d, bool vec_in, bool vec_out) {
  const int64_t v0 = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) * 4;
  int64_t n = V;
  if (num_live != nullptr) {
    const int64_t live = *num_live;
    n = live < n ? (live < 0 ? 0 : live) : n;
  }
  if (v0 >= n) return;
  const bool full = v0 + 4 <= n;
  float c[12];
  if (full && vec_in) {
    const float4* src = reinterpret_cast<const float4*>(verts + 3 * v0);  // 48 * (v0 / 4) bytes
    const float4 a = src[0], bb = src[1], cc = src[2];
    const float e[12] = {a.x, a.y, a.z, a.w, bb.x, bb.y, bb.z, bb.w, cc.x, cc.y, cc.z, cc.w};
#pragma unroll
    for (int i = 0; i < 12; ++i) c[i] = e[i];
  } else {
#pragma unroll
    for (int i = 0; i < 12; ++i) c[i] = v0 + i / 3 < n ? verts[3 * v0 + i] : 0.0f;
  }
  for (int b = 0; b < B; ++b) {
    Projected p[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      p[i] = project_point(E + 16 * b, K + 9 * b, c[3 * i], c[3 * i + 1], c[3 * i + 2]);
    const int64_t o = int64_t(b) * V + v0;
    if (full && vec_out && (o & 1) == 0) {  // 16-byte aligned xy
      float4* xy4 = reinterpret_cast<float4*>(out_xy + 2 * o);
      xy4[0] = make_float4(p[0].x, p[0].y, p[1].x, p[1].y);
      xy4[1] = make_float4(p[2].x, p[2].y, p[3].x, p[3].y);
      if (out_d == nullptr) continue;
      if ((o & 3) == 0) {
        *reinterpret_cast<float4*>(out_d + o) =
            make_float4(p[0].depth, p[1].depth, p[2].depth, p[3].depth);
      } else {
        *reinterpret_cast<float2*>(out_d + o) = make_float2(p[0].depth, p[1].depth);
        *reinterpret_cast<float2*>(out_d + o + 2) = make_float2(p[2].depth, p[3].depth);
      }
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (v0 + i >= n) break;
        reinterpret_cast<float2*>(out_xy)[o + i] = make_float2(p[i].x, p[i].y);
        if (out_d != nullptr) out_d[o + i] = p[i].depth;
      }
    }
  }
}

// ------------------------------------------------------------------------------------
// Bounds (bound_triangles.slang:62-192)
// ------------------------------------------------------------------------------------

// float -> int32 rounding toward zero, saturating, NaN -> 0: exactly what v_cvt_i32_f32
// does (checked on gfx950 against the branchy form -- NaN -> 0, x >= 2^31 -> INT_MAX,
// x <= -2^31 -> INT_MIN, else truncation -- for the edge cases and 2^20 random bit patterns).
__device__ __forceinline__ int32_t f2i_rz(float x) {
  int32_t r;
  asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// Tile rows [y0, y1) x columns [x0, x1) overlapped by one face in one camera
// (bound_triangles.slang:144-192): empty for back faces (negative signed area) and faces
// with a corner in front of the near plane.
struct TileRect {
  int y0, y1, x0, x1;
};

// Tile-side arithmetic: truncating integer quotients and correctly rounded float quotients
// by the tile height (rows) and width (columns). With power-of-two sides (kPow2) they are
// shifts and exact scalings: x * 2^-k is the same correctly rounded number as x / 2^k
// (f32 denormals are preserved), and the truncated quotient of a negative x is minus the
// shifted |x|.
template <bool kPow2>
struct TileSize {
  int th, tw, lh, lw;
  float ih, iw;  // 2^-lh, 2^-lw

  __device__ __forceinline__ TileSize(int h, int w) : th(h), tw(w), lh(0), lw(0), ih(1.f), iw(1.f) {
    if (kPow2) {
      lh = __builtin_ctz(unsigned(h));
      lw = __builtin_ctz(unsigned(w));
      ih = __int_as_float((127 - lh) << 23);
      iw = __int_as_float((127 - lw) << 23);
    }
  }
  __device__ __forceinline__ static int shift_div(int x, int l) {
    return x >= 0 ? (x >> l) : -static_cast<int>((-static_cast<int64_t>(x)) >> l);
  }
  __device__ __forceinline__ int row(int x) const { return kPow2 ? shift_div(x, lh) : x / th; }
  __device__ __forceinline__ int col(int x) const { return kPow2 ? shift_div(x, lw) : x / tw; }
  __device__ __forceinline__ float rowf(float x) const { return kPow2 ? x * ih : x / float(th); }
  __device__ __forceinline__ float colf(float x) const { return kPow2 ? x * iw : x / float(tw); }
};

__host__ __forceinline__ bool pow2_tiles(int th, int tw) {
  return th > 0 && tw > 0 && (th & (th - 1)) == 0 && (tw & (tw - 1)) == 0;
}

template <bool kPow2>
__device__ __forceinline__ TileRect corner_tile_rect(float2 a, float2 bb, float2 c, float da,
                                                     float db, float dc, const TileSize<kPow2>& ts,
                                                     int rmin, int cmin, int rmax, int cmax,
                                                     float near_plane) {
  TileRect r{0, 0, 0, 0};
  const float check = (bb.x - a.x) * (c.y - a.y) - (bb.y - a.y) * (c.x - a.x);
  const float min_depth = fminf(fminf(da, db), dc);
  if (!(check < 0.0f) && !(min_depth < near_plane)) {
    const float ylo = fminf(fminf(a.y, bb.y), c.y), yhi = fmaxf(fmaxf(a.y, bb.y), c.y);
    const float xlo = fminf(fminf(a.x, bb.x), c.x), xhi = fmaxf(fmaxf(a.x, bb.x), c.x);
    r.y0 = clampi(ts.row(f2i_rz(ylo)), rmin, rmax);
    r.y1 = clampi(f2i_rz(ceilf(ts.rowf(yhi))), rmin, rmax);
    r.x0 = clampi(ts.col(f2i_rz(xlo)), cmin, cmax);
    r.x1 = clampi(f2i_rz(ceilf(ts.colf(xhi))), cmin, cmax);
  }
  return r;
}

template <bool kPow2>
__device__ __forceinline__ TileRect face_tile_rect(const float* xy, const float* depths, int64_t o,
                                                   int64_t ia, int64_t ib, int64_t ic,
                                                   const TileSize<kPow2>& ts, int rmin, int cmin,
                                                   int rmax, int cmax, float near_plane) {
  const float2* pxy = reinterpret_cast<const float2*>(xy);
  return corner_tile_rect(pxy[o + ia], pxy[o + ib], pxy[o + ic], depths[o + ia], depths[o + ib],
                          depths[o + ic], ts, rmin, cmin, rmax, cmax, near_plane);
}

// Narrows a face's tile rectangle to the tiles where some pixel can pass the barycentric
// test. For a small, well-shaped face the pixel centres near its bounding box are the only
// candidates (geometry.h, centre_range; most sub-pixel faces have none and get no key);
// otherwise the candidate pixel rectangle over the whole tile-aligned box (geometry.h, the
// bound the composites use per tile) excludes the other pixels. When at most kExactPixels
// candidates remain, each of them is put through the test itself (the reference's float
// operations, as the composites evaluate them), and only tiles with an accepted pixel
// keep a key. Keys dropped this way cannot change any pixel or gradient.
constexpr int kExactPixels = 16;

// The pixel tests of cull_tiles (below) for a candidate rectangle q of at most kExactPixels
// pixels of the face (a, bb, c): the tiles of its tile rectangle r that hold an accepted pixel.
template <bool kPow2>
__device__ __forceinline__ TileRect exact_tiles_at(TileRect r, Rect q, float2 a, float2 bb,
                                                   float2 c, const TileSize<kPow2>& ts) {
  const int th = ts.th, tw = ts.tw;
  TileRect n;
  n.y0 = n.y1 = r.y0;
  n.x0 = n.x1 = r.x0;
  const int px0 = r.x0 * tw, py0 = r.y0 * th;
  const TriTerms t = tri_terms(a.x, a.y, bb.x, bb.y, c.x, c.y);
  const float ad = fabsf(t.D);
  const float rd = (ad >= 0x1p-60f && ad <= 0x1p60f) ? __builtin_amdgcn_rcpf(t.D) : 0.0f;
  int ax0 = 0x7fffffff, ay0 = 0x7fffffff, ax1 = -1, ay1 = -1;  // accepted pixels' box
  for (int j = 0; j < q.h; ++j) {
    for (int i = 0; i < q.w; ++i) {
      const int X = px0 + q.x + i, Y = py0 + q.y + j;
      const BaryNum m = bary_numerators(t, float(X) + 0.5f, float(Y) + 0.5f);
      if (rcp_proves_miss(m, rd)) continue;
      const float v = m.vn / t.D, w = m.wn / t.D, u = m.un / t.D;
      if (bary_accept(u, v, w)) {
        ax0 = X < ax0 ? X : ax0;
        ax1 = X > ax1 ? X : ax1;
        ay0 = Y < ay0 ? Y : ay0;
        ay1 = Y > ay1 ? Y : ay1;
      }
    }
  }
  if (ax1 < 0) return n;
  n.y0 = ts.row(ay0);
  n.y1 = ts.row(ay1) + 1;
  n.x0 = ts.col(ax0);
  n.x1 = ts.col(ax1) + 1;
  return n;
}


// The pixel centres of face (a, b, c) in its tile-aligned box r that can hit
// (geometry.h, centre_range): none, a few, or unknown.
__device__ __forceinline__ CentreRange box_centres(float2 a, float2 bb, float2 c, TileRect r,
                                                   int th, int tw) {
  return centre_range(a.x, a.y, bb.x, bb.y, c.x, c.y, r.x0 * tw, r.y0 * th, (r.x1 - r.x0) * tw,
                      (r.y1 - r.y0) * th, kExactPixels);
}

// Face (a, bb, c) with tile rectangle r. Small, well-shaped faces take their candidate pixels
// from the bounding box (box_centres); the others from the candidate rectangle. reload(a,
// bb, c) may load the corners again before the pixel tests, so that they need not stay in
// registers across the candidate rectangle.
template <bool kPow2, typename Reload>
__device__ __forceinline__ TileRect cull_tiles_at(TileRect r, float2 a, float2 bb, float2 c,
                                                  const TileSize<kPow2>& ts, Reload reload) {
  if (r.y1 <= r.y0 || r.x1 <= r.x0) return r;
  const int th = ts.th, tw = ts.tw;
  TileRect n;
  n.y0 = n.y1 = r.y0;
  n.x0 = n.x1 = r.x0;
  const CentreRange cen = box_centres(a, bb, c, r, th, tw);
  if (cen.state == kCentresNone) return n;
  Rect q{cen.x, cen.y, cen.w, cen.h};
  if (cen.state != kCentresFew) {
    const TriTerms t = tri_terms(a.x, a.y, bb.x, bb.y, c.x, c.y);
    q = candidate_rect(t, r.x0 * tw, r.y0 * th, (r.x1 - r.x0) * tw, (r.y1 - r.y0) * th);
    if (q.w <= 0 || q.h <= 0) return n;
    if (q.w * q.h > kExactPixels) {
      n.y0 = r.y0 + ts.row(q.y);
      n.y1 = r.y0 + ts.row(q.y + q.h - 1) + 1;
      n.x0 = r.x0 + ts.col(q.x);
      n.x1 = r.x0 + ts.col(q.x + q.w - 1) + 1;
      return n;
    }
    reload(a, bb, c);
  }
  return exact_tiles_at(r, q, a, bb, c, ts);
}

// cull_tiles_at for face f (whose three corners the caller has checked), camera offset o
// into xy.
template <bool kPow2>
__device__ __forceinline__ TileRect cull_tiles(TileRect r, const float* xy, const int32_t* faces,
                                               int64_t f, int64_t o, const TileSize<kPow2>& ts) {
  if (r.y1 <= r.y0 || r.x1 <= r.x0) return r;
  const float2* pxy = reinterpret_cast<const float2*>(xy);
  auto load = [&](float2& a, float2& bb, float2& c) {
    a = pxy[o + faces[3 * f + 0]];
    bb = pxy[o + faces[3 * f + 1]];
    c = pxy[o + faces[3 * f + 2]];
  };
  float2 a, bb, c;
  load(a, bb, c);
  return cull_tiles_at(r, a, bb, c, ts, [&](float2& a2, float2& b2, float2& c2) {
    asm volatile("" ::: "memory");
    load(a2, b2, c2);
  });
}

// Face whose keys are generated: live (below the device-side face count) and with all
// three corners inside the vertex array.
__device__ __forceinline__ bool keyed_face(const int32_t* faces, int64_t f, int64_t V,
                                           const int64_t* live_x3, int64_t& ia, int64_t& ib,
                                           int64_t& ic) {
  ia = faces[3 * f + 0];
  ib = faces[3 * f + 1];
  ic = faces[3 * f + 2];
  const bool bad_corner = ia < 0 || ib < 0 || ic < 0 || ia >= V || ib >= V || ic >= V;
  return !((live_x3 != nullptr && f >= int64_t(*live_x3) / 3) || bad_corner);
}

__global__ void __launch_bounds__(kBlock) bound_kernel(const float* xy, const float* depths,
                                                       int64_t V, const int32_t* faces, int64_t F,
                                                       int B, int th, int tw, int rmin, int cmin,
                                                       int rmax, int cmax, float near_plane,
                                                       int32_t* tmin, int32_t* tmax,
                                                       int32_t* counts, const int64_t* live_x3) {
  const int64_t f = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= F) return;
  const int64_t ia = faces[3 * f + 0], ib = faces[3 * f + 1], ic = faces[3 * f + 2];
  const bool bad_corner = ia < 0 || ib < 0 || ic < 0 || ia >= V || ib >= V || ic >= V;
  if ((live_x3 != nullptr && f >= int64_t(*live_x3) / 3) || bad_corner) {
    for (int b = 0; b < B; ++b) {
      const int64_t q = int64_t(b) * F + f;
      reinterpret_cast<int2*>(tmin)[q] = make_int2(0, 0);
      reinterpret_cast<int2*>(tmax)[q] = make_int2(0, 0);
      counts[q] = 0;
    }
    return;
  }
  for (int b = 0; b < B; ++b) {
    const TileRect r = face_tile_rect(xy, depths, int64_t(b) * V, ia, ib, ic,
                                      TileSize<false>(th, tw), rmin, cmin, rmax, cmax, near_plane);
    const int64_t q = int64_t(b) * F + f;
    reinterpret_cast<int2*>(tmin)[q] = make_int2(r.y0, r.x0);
    reinterpret_cast<int2*>(tmax)[q] = make_int2(r.y1, r.x1);
    counts[q] = (r.y1 - r.y0) * (r.x1 - r.x0);
  }
}

// ------------------------------------------------------------------------------------
// Keys (generate_keys.slang:4-68). Each block owns 256 consecutive faces of one camera;
// their keys form one contiguous output range, which the block writes cooperatively
// (coalesced) after a block scan of the per-face tile counts.
// ------------------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) keys_kernel(const float* depths, int64_t V,
                                                      const int32_t* faces, int64_t F,
                                                      const int32_t* boundaries, int S,
                                                      const int32_t* tmin, const int32_t* tmax,
                                                      const int64_t* offsets, int grid_cols,
                                                      KeyFields kf, int64_t* keys,
                                                      int32_t* values, int64_t kcap,
                                                      int32_t* overflow) {
  __shared__ int32_t s_prefix[kBlock + 1];
  __shared__ int32_t s_r0[kBlock], s_c0[kBlock], s_w[kBlock];
  __shared__ uint64_t s_key_hi[kBlock];  // camera | shell | depth bits (tile added later)
  __shared__ int32_t s_wave[kBlock / kWave];
  const int b = blockIdx.y;
  const int64_t f = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  int32_t count = 0;
  int64_t block_start = 0;
  if (f < F) {
    const int64_t q = int64_t(b) * F + f;
    const int2 lo = reinterpret_cast<const int2*>(tmin)[q];
    const int2 hi = reinterpret_cast<const int2*>(tmax)[q];
    count = (hi.x - lo.x) * (hi.y - lo.y);
    s_r0[threadIdx.x] = lo.x;
    s_c0[threadIdx.x] = lo.y;
    s_w[threadIdx.x] = hi.y - lo.y;
    if (count > 0) {
      uint32_t shell = 0;
      while (shell < uint32_t(S) && !(f < boundaries[shell])) ++shell;
      const int64_t o = int64_t(b) * V;
      const float d0 = depths[o + faces[3 * f + 0]];
      const float d1 = depths[o + faces[3 * f + 1]];
      const float d2 = depths[o + faces[3 * f + 2]];
      const float mean_depth = ((d0 + d1) + d2) / 3.f;
      s_key_hi[threadIdx.x] = pack_key(kf, uint32_t(b), shell, 0u, mean_depth);
    }
  }
  // Block-exclusive scan of counts.
  const int lane = __lane_id();
  const int wave = threadIdx.x / kWave;
  int32_t inc = wave_inclusive_sum(count);
  if (lane == kWave - 1) s_wave[wave] = inc;
  __syncthreads();
  int32_t before = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kBlock / kWave; ++w) {
    if (w < wave) before += s_wave[w];
    total += s_wave[w];
  }
  s_prefix[threadIdx.x] = before + inc - count;
  if (threadIdx.x == 0) s_prefix[kBlock] = total;
  // Output position of the block's first key (the inclusive offset of the face before).
  __shared__ int64_t s_block_start;
  if (threadIdx.x == 0) {
    const int64_t q0 = int64_t(b) * F + int64_t(blockIdx.x) * kBlock;
    block_start = (q0 == 0) ? 0 : offsets[q0 - 1];
    s_block_start = block_start;
  }
  __syncthreads();
  const int64_t base = s_block_start;
  const uint64_t tile_mask = field_mask(kf.tile_bits);
  for (int32_t j = threadIdx.x; j < total; j += kBlock) {
    // Owner: the last face whose prefix is <= j (binary search over 256 entries).
    int lo = 0, hi = kBlock - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_prefix[mid] <= j) lo = mid; else hi = mid - 1;
    }
    const int32_t local = j - s_prefix[lo];
    const int32_t w = s_w[lo];
    const uint32_t row = uint32_t(s_r0[lo] + local / w);
    const uint32_t col = uint32_t(s_c0[lo] + local % w);
    const uint32_t tile = row * uint32_t(grid_cols) + col;
    const uint64_t key = s_key_hi[lo] | ((uint64_t(tile) & tile_mask) << kf.tile_offset);
    if (base + j >= kcap) {
      if (overflow) atomicOr(overflow, 1);
      continue;
    }
    keys[base + j] = static_cast<int64_t>(key);
    values[base + j] = static_cast<int32_t>(int64_t(blockIdx.x) * kBlock + lo);
  }
}

// ------------------------------------------------------------------------------------
// Batched step: keys straight from the projected geometry. The per-(camera, face) tile
// bounds, counts and the inclusive scan over all of them are never stored: one kernel
// counts the keys of each (camera, 256-face block), a scan over the block counts gives
// each block its output position, and the key kernel recomputes each face's tile
// rectangle (the same face_tile_rect arithmetic) while writing the block's keys.
// ------------------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) key_block_counts_kernel(
    const float* xy, const float* depths, int64_t V, const int32_t* faces, int64_t F, int B,
    int th, int tw, int rmin, int cmin, int rmax, int cmax, float near_plane,
    const int32_t* cam_dims, int cull, const int64_t* live_x3, int32_t* block_counts) {
  __shared__ int32_t s_wave[kBlock / kWave];
  const int64_t f = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  const int64_t nblk = gridDim.x;
  int64_t ia = 0, ib = 0, ic = 0;
  const bool keyed = f < F && keyed_face(faces, f, V, live_x3, ia, ib, ic);
  for (int b = 0; b < B; ++b) {
    int32_t count = 0;
    if (keyed) {
      const int r0 = cam_dims ? 0 : rmin, c0 = cam_dims ? 0 : cmin;
      const int r1 = cam_dims ? cam_dims[4 * b + 2] : rmax, c1 = cam_dims ? cam_dims[4 * b + 3] : cmax;
      const TileSize<false> ts(th, tw);
      TileRect r = face_tile_rect(xy, depths, int64_t(b) * V, ia, ib, ic, ts, r0, c0, r1, c1,
                                  near_plane);
      if (cull) r = cull_tiles(r, xy, faces, f, int64_t(b) * V, ts);
      count = (r.y1 - r.y0) * (r.x1 - r.x0);
    }
    const int32_t w = wave_sum(count);
    if (__lane_id() == 0) s_wave[threadIdx.x / kWave] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
      int32_t total = 0;
#pragma unroll
      for (int q = 0; q < kBlock / kWave; ++q) total += s_wave[q];
      block_counts[int64_t(b) * nblk + blockIdx.x] = total;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kBlock) keys_from_geometry_kernel(
    const float* xy, const float* depths, int64_t V, const int32_t* faces, int64_t F,
    const int64_t* live_x3, const int32_t* boundaries, int S, int th, int tw, int rmin, int cmin,
    int rmax, int cmax, float near_plane, const int32_t* cam_dims, int cull,
    const int64_t* block_incl,
    int grid_cols, KeyFields kf, int64_t* keys, int32_t* values, int64_t kcap, int32_t* overflow) {
  __shared__ int32_t s_prefix[kBlock + 1];
  __shared__ int32_t s_r0[kBlock], s_c0[kBlock], s_w[kBlock];
  __shared__ uint64_t s_key_hi[kBlock];
  __shared__ int32_t s_wave[kBlock / kWave];
  __shared__ int64_t s_block_start;
  const int b = blockIdx.y;
  if (cam_dims) {  // this camera's own tile grid
    rmin = 0;
    cmin = 0;
    rmax = cam_dims[4 * b + 2];
    cmax = cam_dims[4 * b + 3];
    grid_cols = cmax;
  }
  const int64_t f = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  int32_t count = 0;
  int64_t ia = 0, ib = 0, ic = 0;
  if (f < F && keyed_face(faces, f, V, live_x3, ia, ib, ic)) {
    const int64_t o = int64_t(b) * V;
    const TileSize<false> ts(th, tw);
    TileRect r = face_tile_rect(xy, depths, o, ia, ib, ic, ts, rmin, cmin, rmax, cmax,
                                near_plane);
    if (cull) r = cull_tiles(r, xy, faces, f, o, ts);
    count = (r.y1 - r.y0) * (r.x1 - r.x0);
    s_r0[threadIdx.x] = r.y0;
    s_c0[threadIdx.x] = r.x0;
    s_w[threadIdx.x] = r.x1 - r.x0;
    if (count > 0) {
      uint32_t shell = 0;
      while (shell < uint32_t(S) && !(f < boundaries[shell])) ++shell;
      const float mean_depth = ((depths[o + ia] + depths[o + ib]) + depths[o + ic]) / 3.f;
      s_key_hi[threadIdx.x] = pack_key(kf, uint32_t(b), shell, 0u, mean_depth);
    }
  }
  const int lane = __lane_id();
  const int wave = threadIdx.x / kWave;
  const int32_t inc = wave_inclusive_sum(count);
  if (lane == kWave - 1) s_wave[wave] = inc;
  if (threadIdx.x == 0) {
    const int64_t q = int64_t(b) * gridDim.x + blockIdx.x;
    s_block_start = q == 0 ? 0 : block_incl[q - 1];
  }
  __syncthreads();
  int32_t before = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kBlock / kWave; ++w) {
    if (w < wave) before += s_wave[w];
    total += s_wave[w];
  }
  s_prefix[threadIdx.x] = before + inc - count;
  __syncthreads();
  const int64_t base = s_block_start;
  const uint64_t tile_mask = field_mask(kf.tile_bits);
  for (int32_t j = threadIdx.x; j < total; j += kBlock) {
    int lo = 0, hi = kBlock - 1;  // last face whose prefix is <= j (it has keys)
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_prefix[mid] <= j) lo = mid; else hi = mid - 1;
    }
    const int32_t local = j - s_prefix[lo];
    const int32_t w = s_w[lo];
    const uint32_t row = uint32_t(s_r0[lo] + local / w);
    const uint32_t col = uint32_t(s_c0[lo] + local % w);
    const uint32_t tile = row * uint32_t(grid_cols) + col;
    const uint64_t key = s_key_hi[lo] | ((uint64_t(tile) & tile_mask) << kf.tile_offset);
    if (base + j >= kcap) {
      if (overflow) atomicOr(overflow, 1);
      continue;
    }
    keys[base + j] = static_cast<int64_t>(key);
    values[base + j] = static_cast<int32_t>(int64_t(blockIdx.x) * kBlock + lo);
  }
}

// Single pass: the keys of every (camera, face) pair without per-pair arrays. A block owns
// one unit: a camera slice (kPassFaces / 2^lfpb cameras) of a face block (2^lfpb faces;
// pairs are camera-major inside the unit). Units are numbered slice-major; the blocks of
// one XCD (equal blockIdx mod 8, the hardware's round-robin) take consecutive units, so
// that the corners a unit shares with its neighbours stay in that XCD's L2. A block
//   1. bounds each pair (culling, tile-aligned box), drops the pairs whose bounding box
//      reaches no pixel centre (box_centres), and lists the others for the pixel-level
//      narrowing (cull_tiles) in LDS, grouped by the route they will take,
//   2. narrows the listed pairs with every thread busy (back-facing, clipped and keyless
//      pairs take no thread),
//   3. scans the key counts in pair order, reserves its output range with one atomic add on
//      a key counter (no block waits for another), records the unit's count and range, and
//      writes its keys there.
// reorder_keys_kernel then moves every unit's keys to their place in unit order (the scan
// of the unit counts). A face's corner indices are read from memory once per camera slice.
// The keys of one camera come out in (face, tile) order, the order of
// mt_generate_keys_from_geometry; the cameras of a slice are interleaved unit by unit. Keys
// of different cameras differ in their camera field, so a stable sort orders both
// arrangements identically.
constexpr int kPassThreads = 256;  // one wave per SIMD per block
constexpr int kPassPer = 4;
constexpr int kPassFaces = kPassPer * kPassThreads;
constexpr int kMaxPassShells = 64;
constexpr int kStaged = 256;

template <bool kPow2>
__global__ void __launch_bounds__(kPassThreads) __attribute__((amdgpu_waves_per_eu(kPow2 ? 6 : 5)))
keys_single_pass_kernel(
    const float* xy, const float* depths, int64_t V, const int32_t* faces, int64_t F,
    const int64_t* live_x3, const int32_t* boundaries, int S, int B, int lfpb, int nslices,
    int th, int tw,
    int rmin, int cmin, int rmax, int cmax, float near_plane, const int32_t* cam_dims, int cull,
    int grid_cols, KeyFields kf, int64_t* keys, int32_t* values, int64_t kcap, int64_t* n_keys,
    int32_t* overflow, unsigned long long* key_counter, int32_t* unit_count, int64_t* unit_at) {
  __shared__ int32_t s_prefix[kPassFaces];  // key count, then its exclusive scan
  // Tile rectangles packed 16 + 16 bits: origin (row << 16 | column), extent (height << 16 |
  // width); grids have at most 65535 tile rows and columns.
  __shared__ uint32_t s_org[kPassFaces], s_ext[kPassFaces];
  // Corner depths' sum (the reference's order, bits of a float) of the pairs with keys: the
  // key's depth field is its third, made when the keys are written.
  __shared__ uint32_t s_zsum[kPassFaces];
  __shared__ int16_t s_work[kPassFaces];
  // The corners of the first kStaged pairs of route 0, so that step 2 reads nothing from
  // memory for them.
  __shared__ float2 s_corner[kStaged * 3];
  __shared__ int32_t s_wave[kPassPer][kPassThreads / kWave];
  __shared__ int32_t s_nwork[2];  // pairs listed for each route of step 2
  __shared__ int64_t s_block_start;
  __shared__ int64_t s_bounds[kMaxPassShells];  // the shells' face boundaries
  if (int(threadIdx.x) < S) s_bounds[threadIdx.x] = boundaries[threadIdx.x];
  if (threadIdx.x == 0) s_nwork[0] = s_nwork[1] = 0;
  // The host keeps the unit count below 2^31, so unit indices are 32-bit.
  const int fpb = 1 << lfpb;
  const uint32_t nfb = uint32_t((F + fpb - 1) >> lfpb);  // face blocks
  const uint32_t nblk = nfb * uint32_t(nslices);         // units
  const uint32_t per_xcd = (nblk + 7u) / 8u;
  const uint32_t q = (blockIdx.x & 7u) * per_xcd + (blockIdx.x >> 3);  // this block's unit
  if (q >= nblk) return;  // the grid is rounded up to a multiple of 8
  __syncthreads();
  const uint32_t slice = q / nfb;
  const int64_t f0 = int64_t(q - slice * nfb) << lfpb;  // first face of the block
  const int bfirst = int(slice) * (kPassFaces >> lfpb);  // first camera of the slice
  const TileSize<kPow2> ts(th, tw);
  const int lane = __lane_id();
  const int wave = threadIdx.x / kWave;
  const int64_t live = live_x3 != nullptr ? int64_t(*live_x3) / 3 : F;
  const int64_t fend = live < F ? live : F;  // faces below it are keyed if their corners are valid
  // Pair l: camera bfirst + (l >> lfpb), face f0 + (l & (fpb - 1)).
  // The first shell s with f < boundaries[s] (S if none), from the block's LDS copy.
  auto shell_of = [&](int64_t f) {
    uint32_t shell = 0;
    while (shell < uint32_t(S) && !(f < s_bounds[shell])) ++shell;
    return shell;
  };

  // 1. Bounds. A pair slot i (pairs i * kPassThreads + threadIdx.x) holds one camera for the
  // whole block (lfpb >= log2(kPassThreads)), so its camera's arrays are uniform; corner
  // indices are checked without branches and used as 32-bit offsets.
  const float2* pxy = reinterpret_cast<const float2*>(xy);
  const int64_t nlive64 = fend - f0;
  const int nlive = nlive64 <= 0 ? 0 : (nlive64 >= fpb ? fpb : int(nlive64));  // live faces here
  const int32_t* fblock = faces + 3 * f0;
  const uint32_t vlim = V < 0x80000000ll ? uint32_t(V) : 0x80000000u;
  // Corner indices of slot i's face (-1 when the face is not live), loaded before use.
  auto face_corners = [&](int i, int32_t& ia, int32_t& ib, int32_t& ic) {
    const int local = (i * kPassThreads + int(threadIdx.x)) & (fpb - 1);
    ia = ib = ic = -1;
    if (local < nlive) {
      ia = fblock[3 * local + 0];
      ib = fblock[3 * local + 1];
      ic = fblock[3 * local + 2];
    }
  };
  // Bounds slot i's pair (face corners ia, ib, ic) and lists it for step 2.
  auto bound_slot = [&](int i, int32_t ia, int32_t ib, int32_t ic, float2 pa, float2 pb, float2 pc,
                        float za, float zb, float zc, bool valid) {
    const int l = i * kPassThreads + threadIdx.x;
    const int b = bfirst + (l >> lfpb);
    int32_t count = 0;
    int route = -1;  // step 2: 0 = the pixel centres near the box, 1 = candidate rectangle
    if (valid) {
      int r0 = rmin, c0 = cmin, r1 = rmax, c1 = cmax;
      if (cam_dims) {  // this camera's own tile grid
        r0 = 0;
        c0 = 0;
        r1 = cam_dims[4 * b + 2];
        c1 = cam_dims[4 * b + 3];
      }
      const TileRect r = corner_tile_rect(pa, pb, pc, za, zb, zc, ts, r0, c0, r1, c1, near_plane);
      s_org[l] = (uint32_t(r.y0) << 16) | uint32_t(r.x0);
      s_ext[l] = (uint32_t(r.y1 - r.y0) << 16) | uint32_t(r.x1 - r.x0);
      count = (r.y1 - r.y0) * (r.x1 - r.x0);
      if (count > 0 && cull) {
        // Faces whose bounding box reaches no pixel centre have no keys; the others are
        // narrowed in step 2.
        // Only the certificate and the box's emptiness here: the candidate centres of the
        // others are found in step 2 (cull_tiles_at), which keeps this part short.
        const bool cert = centre_certificate(pa.x, pa.y, pb.x, pb.y, pc.x, pc.y);
        const bool none = cert && centre_box_empty(pa.x, pa.y, pb.x, pb.y, pc.x, pc.y);
        route = none ? -1 : (cert ? 0 : 1);
        if (route >= 0) s_zsum[l] = __float_as_uint((za + zb) + zc);
        count = 0;  // set by step 2
      } else if (count > 0) {
        s_zsum[l] = __float_as_uint((za + zb) + zc);
      }
    }
    s_prefix[l] = count;
    // Append the listed pairs, one LDS atomic per wave and route: route 0 from the front
    // of s_work, route 1 from the back, so that the waves of step 2 take one route each.
    const uint64_t m0 = __ballot(route == 0), m1 = __ballot(route == 1);
    int at0 = 0, at1 = 0;
    if (lane == 0) {
      if (m0 != 0) at0 = atomicAdd(&s_nwork[0], __popcll(m0));
      if (m1 != 0) at1 = atomicAdd(&s_nwork[1], __popcll(m1));
    }
    at0 = __shfl(at0, 0, kWave);
    at1 = __shfl(at1, 0, kWave);
    if (route == 0) {
      const int pos = at0 + __popcll(m0 & lanemask_lt());
      s_work[pos] = static_cast<int16_t>(l);
      if (pos < kStaged) {
        s_corner[3 * pos + 0] = pa;
        s_corner[3 * pos + 1] = pb;
        s_corner[3 * pos + 2] = pc;
      }
    }
    if (route == 1)
      s_work[kPassFaces - 1 - (at1 + __popcll(m1 & lanemask_lt()))] = static_cast<int16_t>(l);
  };
  // kHoist slots at a time: their corners are all loaded before any is used, so that the
  // loads overlap. With lfpb == log2(kPassThreads) every slot holds the same face (the
  // slice's cameras), whose corner indices are then read once.
  constexpr int kHoist = 2;
  const bool shared_face = (fpb == kPassThreads);
  int32_t sa = -1, sb = -1, sc = -1;
  if (shared_face) face_corners(0, sa, sb, sc);
#pragma unroll
  for (int i0 = 0; i0 < kPassPer; i0 += kHoist) {
    int32_t ca[kHoist], cb[kHoist], cc[kHoist];
    bool valid[kHoist];
    float2 pa[kHoist], pb[kHoist], pc[kHoist];
    float za[kHoist], zb[kHoist], zc[kHoist];
#pragma unroll
    for (int h = 0; h < kHoist; ++h) {
      const int i = i0 + h;
      if (shared_face) {
        ca[h] = sa;
        cb[h] = sb;
        cc[h] = sc;
      } else {
        face_corners(i, ca[h], cb[h], cc[h]);
      }
    }
#pragma unroll
    for (int h = 0; h < kHoist; ++h) {
      const int i = i0 + h;
      const int b = bfirst + ((i * kPassThreads) >> lfpb);  // uniform
      // Negative indices are >= 2^31 as unsigned numbers, so one comparison each.
      valid[h] = (b < B) & (uint32_t(ca[h]) < vlim) & (uint32_t(cb[h]) < vlim) &
                 (uint32_t(cc[h]) < vlim);
      const float2* cxy = pxy + int64_t(b) * V;
      const float* cd = depths + int64_t(b) * V;
      if (valid[h]) {
        pa[h] = cxy[uint32_t(ca[h])];
        pb[h] = cxy[uint32_t(cb[h])];
        pc[h] = cxy[uint32_t(cc[h])];
        za[h] = cd[uint32_t(ca[h])];
        zb[h] = cd[uint32_t(cb[h])];
        zc[h] = cd[uint32_t(cc[h])];
      }
    }
#pragma unroll
    for (int h = 0; h < kHoist; ++h)
      bound_slot(i0 + h, ca[h], cb[h], cc[h], pa[h], pb[h], pc[h], za[h], zb[h], zc[h], valid[h]);
  }
  __syncthreads();
  // 2. Pixel-level narrowing of the listed pairs: route 0 (positions [0, n0) of s_work)
  // then route 1 (the last n1 positions); a wave mixes the routes only where they meet.
  const int n0 = s_nwork[0], nwork = n0 + s_nwork[1];
  for (int k = threadIdx.x; k < nwork; k += kPassThreads) {
    const int l = s_work[k < n0 ? k : kPassFaces - 1 - (k - n0)];
    const int b = bfirst + (l >> lfpb);
    const int64_t f = f0 + (l & (fpb - 1));
    TileRect r;
    r.y0 = int(s_org[l] >> 16);
    r.x0 = int(s_org[l] & 0xffffu);
    r.y1 = r.y0 + int(s_ext[l] >> 16);
    r.x1 = r.x0 + int(s_ext[l] & 0xffffu);
    if (k < n0 && k < kStaged) {  // corners staged by step 1
      r = cull_tiles_at(r, s_corner[3 * k + 0], s_corner[3 * k + 1], s_corner[3 * k + 2], ts,
                        [](float2&, float2&, float2&) {});
    } else {
      r = cull_tiles(r, xy, faces, f, int64_t(b) * V, ts);
    }
    const int32_t count = (r.y1 - r.y0) * (r.x1 - r.x0);
    s_org[l] = (uint32_t(r.y0) << 16) | uint32_t(r.x0);
    s_ext[l] = (uint32_t(r.y1 - r.y0) << 16) | uint32_t(r.x1 - r.x0);
    s_prefix[l] = count;
  }
  __syncthreads();
  // 3. Scan in pair order (slot-major), output range, keys. The slots are scanned together:
  // wave sums of all of them, one barrier, then each slot's offset from the earlier slots'
  // totals.
  int32_t cnt[kPassPer], inc[kPassPer];
#pragma unroll
  for (int i = 0; i < kPassPer; ++i) {
    cnt[i] = s_prefix[i * kPassThreads + threadIdx.x];
    inc[i] = wave_inclusive_sum(cnt[i]);
    if (lane == kWave - 1) s_wave[i][wave] = inc[i];
  }
  __syncthreads();
  int32_t carry = 0;
#pragma unroll
  for (int i = 0; i < kPassPer; ++i) {
    int32_t before = 0, slot_total = 0;
#pragma unroll
    for (int w = 0; w < kPassThreads / kWave; ++w) {
      const int32_t x = s_wave[i][w];
      before += w < wave ? x : 0;
      slot_total += x;
    }
    s_prefix[i * kPassThreads + threadIdx.x] = carry + before + inc[i] - cnt[i];
    carry += slot_total;
  }
  const int32_t total = carry;
  if (threadIdx.x == 0) {
    const int64_t at =
        total > 0 ? static_cast<int64_t>(atomicAdd(key_counter, static_cast<unsigned long long>(total)))
                  : 0;
    s_block_start = at;
    unit_count[q] = total;
    unit_at[q] = at;
  }
  __syncthreads();
  const int64_t base = s_block_start;
  const uint64_t tile_mask = field_mask(kf.tile_bits);
  for (int32_t j = threadIdx.x; j < total; j += kPassThreads) {
    int lo = 0, hi = kPassFaces - 1;  // last pair whose prefix is <= j (it has keys)
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_prefix[mid] <= j) lo = mid; else hi = mid - 1;
    }
    const int32_t local = j - s_prefix[lo];
    const int32_t w = int32_t(s_ext[lo] & 0xffffu);
    const int b = bfirst + (lo >> lfpb);
    const int cols = cam_dims ? cam_dims[4 * b + 3] : grid_cols;
    const uint32_t row = (s_org[lo] >> 16) + uint32_t(local / w);
    const uint32_t col = (s_org[lo] & 0xffffu) + uint32_t(local % w);
    const uint32_t tile = row * uint32_t(cols) + col;
    const int64_t f = f0 + (lo & (fpb - 1));
    const float mean_depth = __uint_as_float(s_zsum[lo]) / 3.f;
    const uint64_t key = pack_key(kf, uint32_t(b), shell_of(f), 0u, mean_depth) |
                         ((uint64_t(tile) & tile_mask) << kf.tile_offset);
    if (base + j < 0 || base + j >= kcap) {
      if (overflow) atomicOr(overflow, 1);
      continue;
    }
    keys[base + j] = static_cast<int64_t>(key);
    values[base + j] = static_cast<int32_t>(f);
  }
}

// log2 of the faces per block of the single-pass kernel: the largest power of two with
// (faces x B) <= kPassFaces, at least kPassThreads faces (then the cameras come in slices of
// kPassFaces / kPassThreads).
__host__ inline int single_pass_lfpb(int B) {
  int l = 0;
  while ((int64_t(2) << l) * (B > 0 ? B : 1) <= kPassFaces) ++l;
  constexpr int kMin = __builtin_ctz(kPassThreads);
  return l > kMin ? l : kMin;
}

__host__ inline int single_pass_slices(int B, int lfpb) {
  const int per = kPassFaces >> lfpb;
  return (B + per - 1) / per;
}

__global__ void fill_i64_kernel(int64_t* p, int64_t n, int64_t value) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

// One wave per unit: moves the unit's keys from the range it reserved (src) to its place in
// unit order (dst, after the keys of every earlier unit: incl is the inclusive scan of the
// unit counts). The last unit's wave stores the total key count. When the total exceeds the
// capacity (the key kernel raised the overflow flag), a unit's keys whose reserved range
// fell past the capacity were never written: their places below the capacity get key 0
// with value 0 (camera 0, tile 0, shell 0, face 0), so that every key below the capacity
// indexes valid tile ranges and faces in the stages after the sort.
constexpr int kReorderWaves = 4;
__global__ void __launch_bounds__(kReorderWaves * kWave) reorder_keys_kernel(
    const int64_t* src_keys, const int32_t* src_values, const int32_t* unit_count,
    const int64_t* unit_at, const int64_t* incl, int64_t units, int64_t kcap, int64_t* keys,
    int32_t* values, int64_t* n_keys) {
  const int64_t u = int64_t(blockIdx.x) * kReorderWaves + threadIdx.x / kWave;
  if (u >= units) return;
  const int lane = __lane_id();
  const int32_t n = unit_count[u];
  const int64_t from = unit_at[u], to = incl[u] - n;
  for (int32_t j = lane; j < n; j += kWave) {
    if (to + j < kcap) {
      const bool kept = from + j < kcap;
      keys[to + j] = kept ? src_keys[from + j] : 0;
      values[to + j] = kept ? src_values[from + j] : 0;
    }
  }
  if (u == units - 1 && lane == 0 && n_keys) *n_keys = incl[u];
}

int64_t single_pass_blocks(int64_t F, int B) {
  if (B <= 0) return 0;
  const int lfpb = single_pass_lfpb(B);
  return ((F + (int64_t(1) << lfpb) - 1) >> lfpb) * single_pass_slices(B, lfpb);
}

// Workspace of the single pass: the key counter (256 bytes), each unit's key count (int32)
// and reserved range start (int64), the inclusive scan of the counts (int64), and that
// scan's own workspace.
struct SinglePassState {
  unsigned long long* counter;
  int32_t* unit_count;
  int64_t* unit_at;
  int64_t* incl;
  void* scan;
  size_t scan_bytes;
  size_t total;
};

inline size_t pad256(size_t x) { return (x + 255) / 256 * 256; }

SinglePassState single_pass_state(void* ws, int64_t F, int B) {
  const size_t units = static_cast<size_t>(single_pass_blocks(F, B));
  const size_t o_count = 256, o_at = o_count + pad256(units * 4), o_incl = o_at + pad256(units * 8);
  const size_t o_scan = o_incl + pad256(units * 8);
  SinglePassState s{};
  s.scan_bytes = mt_scan_workspace_bytes(static_cast<int64_t>(units));
  s.total = o_scan + s.scan_bytes;
  if (ws != nullptr) {
    char* p = static_cast<char*>(ws);
    s.counter = reinterpret_cast<unsigned long long*>(p);
    s.unit_count = reinterpret_cast<int32_t*>(p + o_count);
    s.unit_at = reinterpret_cast<int64_t*>(p + o_at);
    s.incl = reinterpret_cast<int64_t*>(p + o_incl);
    s.scan = p + o_scan;
  }
  return s;
}

size_t single_pass_state_bytes(int64_t F, int B) { return single_pass_state(nullptr, F, B).total; }

// ------------------------------------------------------------------------------------
// Delineation (delineate_tiles.slang:12-58)
// ------------------------------------------------------------------------------------

__device__ __forceinline__ void run_index(uint64_t key, const KeyFields& kf, bool separate,
                                          uint32_t& cam, uint32_t& shell, uint32_t& tile) {
  cam = key_field(key, kf.camera_offset, kf.camera_bits);
  shell = separate ? key_field(key, kf.shell_offset, kf.shell_bits) : 0u;
  tile = key_field(key, kf.tile_offset, kf.tile_bits);
}

__global__ void __launch_bounds__(kBlock) delineate_kernel(const int64_t* keys, int64_t n,
                                                           KeyFields kf, int separate, int B,
                                                           int S_out, int T, int64_t* ranges,
                                                           const int64_t* n_device,
                                                           const int64_t* tile_base) {
  if (n_device != nullptr && *n_device < n) n = *n_device;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t c, s, t;
  run_index(static_cast<uint64_t>(keys[i]), kf, separate != 0, c, s, t);
  auto slot = [&](uint32_t cc, uint32_t ss, uint32_t tt) -> int64_t {
    if (tile_base) return (tile_base[cc] + tt) * 2;  // per-camera tile grids, S' = 1
    return ((int64_t(cc) * S_out + ss) * T + tt) * 2;
  };
  if (i > 0) {
    uint32_t pc, ps, pt;
    run_index(static_cast<uint64_t>(keys[i - 1]), kf, separate != 0, pc, ps, pt);
    if (pc != c || ps != s || pt != t) {
      ranges[slot(pc, ps, pt) + 1] = i;
      ranges[slot(c, s, t) + 0] = i;
    }
  }
  if (i == n - 1) ranges[slot(c, s, t) + 1] = n;
}

// ------------------------------------------------------------------------------------
// Camera positions: inverse(E)[:3, 3] in double precision, from the adjugate: column 3 of
// the inverse is the cofactors of E's row 3 over det(E). Every index is a compile-time
// constant, so the matrix stays in registers (no private memory).
// ------------------------------------------------------------------------------------

__device__ __forceinline__ double det3_rows012(const double (&m)[4][4], int c0, int c1, int c2) {
  return m[0][c0] * (m[1][c1] * m[2][c2] - m[1][c2] * m[2][c1]) -
         m[0][c1] * (m[1][c0] * m[2][c2] - m[1][c2] * m[2][c0]) +
         m[0][c2] * (m[1][c0] * m[2][c1] - m[1][c1] * m[2][c0]);
}

__global__ void camera_positions_kernel(const float* E, int B, float* out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double m[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) m[r][c] = double(E[16 * b + 4 * r + c]);
  // Cofactor C(3, c) = (-1)^(3 + c) * minor(3, c), the minor over rows 0..2 without column c.
  double cof[4];
  cof[0] = -det3_rows012(m, 1, 2, 3);
  cof[1] = det3_rows012(m, 0, 2, 3);
  cof[2] = -det3_rows012(m, 0, 1, 3);
  cof[3] = det3_rows012(m, 0, 1, 2);
  const double det = m[3][0] * cof[0] + m[3][1] * cof[1] + m[3][2] * cof[2] + m[3][3] * cof[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) out[3 * b + r] = static_cast<float>(cof[r] / det);
}

// ------------------------------------------------------------------------------------
// Spherical harmonics (evaluate_spherical_harmonics.slang:124-270)
// ------------------------------------------------------------------------------------

template <int kDeg>
__global__ void __launch_bounds__(kBlock) sh_forward_kernel(const float* verts, int64_t V,
                                                            const float* sh, const float* cams,
                                                            int B, float* colors) {
  constexpr int kN = (kDeg + 1) * (kDeg + 1);
  const int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= V) return;
  float coef[kN][3];
#pragma unroll
  for (int k = 0; k < kN; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) coef[k][c] = sh[(int64_t(k) * V + v) * 3 + c];
  for (int b = 0; b < B; ++b) {
    const float3 d = view_dir(vertex_at(verts, v), cams + 3 * b);
    float basis[16];
    sh_basis(d.x, d.y, d.z, kDeg, basis);
#pragma unroll
    for (int c = 0; c < 3; ++c)
      colors[(int64_t(b) * V + v) * 3 + c] = fmaxf(sh_channel_pre<kN>(basis, coef, c), 0.0f);
  }
}

template <int kDeg>
__global__ void __launch_bounds__(kBlock) sh_backward_kernel(const float* verts, int64_t V,
                                                             const float* sh, const float* cams,
                                                             int B, const float* gcol,
                                                             float* gsh) {
  constexpr int kN = (kDeg + 1) * (kDeg + 1);
  const int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= V) return;
  float coef[kN][3];
  float grad[kN][3];
#pragma unroll
  for (int k = 0; k < kN; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      coef[k][c] = sh[(int64_t(k) * V + v) * 3 + c];
      grad[k][c] = 0.f;
    }
  for (int b = 0; b < B; ++b) {
    const float3 d = view_dir(vertex_at(verts, v), cams + 3 * b);
    float basis[16];
    sh_basis(d.x, d.y, d.z, kDeg, basis);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      // d max(x, 0) / dx passes only for x > 0.
      const float g = (sh_channel_pre<kN>(basis, coef, c) > 0.0f) ? gcol[(int64_t(b) * V + v) * 3 + c] : 0.0f;
#pragma unroll
      for (int k = 0; k < kN; ++k) grad[k][c] += basis[k] * g;
    }
  }
#pragma unroll
  for (int k = 0; k < kN; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) gsh[(int64_t(k) * V + v) * 3 + c] += grad[k][c];
}

}  // namespace
}  // namespace mt

namespace mt {
namespace {
int project_vertices(const char* name, const float* vertices, int64_t V, const int32_t* num_live,
                     const float* extrinsics, const float* intrinsics, int32_t B, float* out_xy,
                     float* out_depth, mt_stream_t stream) {
  MT_REQUIRE(V >= 0 && B >= 0, "%s: negative size", name);
  if (V == 0 || B == 0) return MT_OK;
  MT_REQUIRE(vertices && extrinsics && intrinsics && out_xy, "%s: null pointer", name);
  const bool vec_in = (reinterpret_cast<uintptr_t>(vertices) & 15) == 0;
  const bool vec_out = (reinterpret_cast<uintptr_t>(out_xy) & 15) == 0 &&
                       (reinterpret_cast<uintptr_t>(out_depth) & 15) == 0;
  const int64_t blocks = ((V + 3) / 4 + kBlock - 1) / kBlock;
  MT_REQUIRE(blocks <= INT32_MAX, "%s: %lld vertices are too many for one launch", name,
             static_cast<long long>(V));
  project_kernel<<<static_cast<unsigned>(blocks), kBlock, 0, static_cast<hipStream_t>(stream)>>>(
      vertices, V, num_live, extrinsics, intrinsics, B, out_xy, out_depth, vec_in, vec_out);
  return check_launch(name);
}
}  // namespace
}  // namespace mt

extern "C" int mt_project_vertices(const float* vertices, int64_t V, const float* extrinsics,
                                   const float* intrinsics, int32_t B, float* out_xy,
                                   float* out_depth, mt_stream_t stream) {
  return mt::project_vertices("mt_project_vertices", vertices, V, nullptr, extrinsics, intrinsics,
                              B, out_xy, out_depth, stream);
}

extern "C" int mt_project_vertices_live(const float* vertices, int64_t V, const int32_t* num_live,
                                        const float* extrinsics, const float* intrinsics,
                                        int32_t B, float* out_xy, float* out_depth,
                                        mt_stream_t stream) {
  MT_REQUIRE(num_live != nullptr, "mt_project_vertices_live: num_live is NULL");
  return mt::project_vertices("mt_project_vertices_live", vertices, V, num_live, extrinsics,
                              intrinsics, B, out_xy, out_depth, stream);
}

extern "C" int mt_bound_triangles(const float* xy, const float* depths, int64_t V,
                                  const int32_t* faces, int64_t F, int32_t B, int32_t tile_h,
                                  int32_t tile_w, int32_t row_min, int32_t col_min,
                                  int32_t row_max, int32_t col_max, float near_plane,
                                  int32_t* tile_min, int32_t* tile_max, int32_t* counts,
                                  const int64_t* num_faces_x3, mt_stream_t stream) {
  MT_REQUIRE(F >= 0 && B >= 0 && tile_h > 0 && tile_w > 0, "mt_bound_triangles: bad sizes");
  if (F == 0 || B == 0) return MT_OK;
  mt::bound_kernel<<<mt::blocks_for(F, mt::kBlock), mt::kBlock, 0, static_cast<hipStream_t>(stream)>>>(
      xy, depths, V, faces, F, B, tile_h, tile_w, row_min, col_min, row_max, col_max, near_plane,
      tile_min, tile_max, counts, num_faces_x3);
  return mt::check_launch("mt_bound_triangles");
}

extern "C" int mt_generate_keys(const float* depths, int64_t V, const int32_t* faces, int64_t F,
                                const int32_t* shell_face_boundaries, int32_t S,
                                const int32_t* tile_min, const int32_t* tile_max,
                                const int64_t* offsets, int32_t B, int32_t grid_cols,
                                mt_key_layout layout, int64_t* keys, int32_t* values,
                                int64_t key_capacity, int32_t* overflow, mt_stream_t stream) {
  MT_REQUIRE(F >= 0 && B >= 0 && B <= 65535 && key_capacity >= 0, "mt_generate_keys: bad sizes");
  if (F == 0 || B == 0) return MT_OK;
  dim3 grid(mt::blocks_for(F, mt::kBlock), static_cast<unsigned>(B));
  mt::keys_kernel<<<grid, mt::kBlock, 0, static_cast<hipStream_t>(stream)>>>(
      depths, V, faces, F, shell_face_boundaries, S, tile_min, tile_max, offsets, grid_cols,
      mt::to_fields(layout), keys, values, key_capacity, overflow);
  return mt::check_launch("mt_generate_keys");
}

extern "C" int mt_key_block_counts(const float* xy, const float* depths, int64_t V,
                                   const int32_t* faces, int64_t F, const int64_t* num_faces_x3,
                                   int32_t B, int32_t tile_h, int32_t tile_w, int32_t row_min,
                                   int32_t col_min, int32_t row_max, int32_t col_max,
                                   float near_plane, const int32_t* cam_dims, int32_t cull,
                                   int32_t* block_counts, mt_stream_t stream) {
  MT_REQUIRE(F >= 0 && B >= 0 && tile_h > 0 && tile_w > 0, "mt_key_block_counts: bad sizes");
  if (F == 0 || B == 0) return MT_OK;
  mt::key_block_counts_kernel<<<mt::blocks_for(F, mt::kBlock), mt::kBlock, 0,
                                static_cast<hipStream_t>(stream)>>>(
      xy, depths, V, faces, F, B, tile_h, tile_w, row_min, col_min, row_max, col_max, near_plane,
      cam_dims, cull, num_faces_x3, block_counts);
  return mt::check_launch("mt_key_block_counts");
}

extern "C" int mt_generate_keys_from_geometry(
    const float* xy, const float* depths, int64_t V, const int32_t* faces, int64_t F,
    const int64_t* num_faces_x3, const int32_t* shell_face_boundaries, int32_t S, int32_t B,
    int32_t tile_h, int32_t tile_w, int32_t row_min, int32_t col_min, int32_t row_max,
    int32_t col_max, float near_plane, const int32_t* cam_dims, int32_t cull,
    const int64_t* block_offsets, int32_t grid_cols, mt_key_layout layout, int64_t* keys,
    int32_t* values, int64_t key_capacity, int32_t* overflow, mt_stream_t stream) {
  MT_REQUIRE(F >= 0 && B >= 0 && B <= 65535 && key_capacity >= 0 && tile_h > 0 && tile_w > 0,
             "mt_generate_keys_from_geometry: bad sizes");
  if (F == 0 || B == 0) return MT_OK;
  dim3 grid(mt::blocks_for(F, mt::kBlock), static_cast<unsigned>(B));
  mt::keys_from_geometry_kernel<<<grid, mt::kBlock, 0, static_cast<hipStream_t>(stream)>>>(
      xy, depths, V, faces, F, num_faces_x3, shell_face_boundaries, S, tile_h, tile_w, row_min,
      col_min, row_max, col_max, near_plane, cam_dims, cull, block_offsets, grid_cols,
      mt::to_fields(layout), keys, values, key_capacity, overflow);
  return mt::check_launch("mt_generate_keys_from_geometry");
}

extern "C" size_t mt_generate_keys_single_pass_workspace_bytes(int64_t F, int32_t B) {
  return mt::single_pass_state_bytes(F < 0 ? 0 : F, B < 0 ? 0 : B);
}

extern "C" int mt_generate_keys_single_pass(
    const float* xy, const float* depths, int64_t V, const int32_t* faces, int64_t F,
    const int64_t* num_faces_x3, const int32_t* shell_face_boundaries, int32_t S, int32_t B,
    int32_t tile_h, int32_t tile_w, int32_t row_min, int32_t col_min, int32_t row_max,
    int32_t col_max, float near_plane, const int32_t* cam_dims, int32_t cull, int32_t grid_cols,
    mt_key_layout layout, int64_t* keys, int32_t* values, int64_t key_capacity, int64_t* n_keys,
    int32_t* overflow, int64_t* scratch_keys, int32_t* scratch_values, void* workspace,
    size_t workspace_bytes, mt_stream_t stream) {
  MT_REQUIRE(F >= 0 && B >= 0 && B <= 65535 && key_capacity >= 0 && tile_h > 0 && tile_w > 0,
             "mt_generate_keys_single_pass: bad sizes");
  MT_REQUIRE(n_keys != nullptr, "mt_generate_keys_single_pass: null key count");
  {
    // The keys are written to the scratch arrays first and then moved to keys / values.
    const auto lo = [](const void* p) { return reinterpret_cast<uintptr_t>(p); };
    const uintptr_t n8 = static_cast<uintptr_t>(key_capacity) * 8, n4 = static_cast<uintptr_t>(key_capacity) * 4;
    const auto apart = [&](const void* a, uintptr_t na, const void* b, uintptr_t nb) {
      return lo(a) + na <= lo(b) || lo(b) + nb <= lo(a);
    };
    MT_REQUIRE(key_capacity == 0 ||
                   (keys && values && scratch_keys && scratch_values &&
                    apart(keys, n8, scratch_keys, n8) && apart(keys, n8, scratch_values, n4) &&
                    apart(values, n4, scratch_keys, n8) && apart(values, n4, scratch_values, n4)),
               "mt_generate_keys_single_pass: keys, values and the two scratch arrays must be "
               "separate device arrays of key_capacity entries");
  }
  MT_REQUIRE(S >= 0 && S <= mt::kMaxPassShells, "mt_generate_keys_single_pass: %d shells (at most %d)",
             S, mt::kMaxPassShells);
  MT_REQUIRE(row_min >= 0 && col_min >= 0 && row_max <= 65535 && col_max <= 65535 &&
                 row_min <= row_max && col_min <= col_max,
             "mt_generate_keys_single_pass: the tile grid must lie in [0, 65535]^2");
  // A block's key count (kPassFaces pairs, each at most the whole tile grid) is an int32;
  // with per-camera grids (cam_dims) row_max and col_max bound every camera's grid.
  MT_REQUIRE(int64_t(mt::kPassFaces) * (int64_t(row_max) - row_min) *
                     (int64_t(col_max) - col_min) < (int64_t(1) << 31),
             "mt_generate_keys_single_pass: %d x %d tiles per camera overflow a block's key count",
             row_max - row_min, col_max - col_min);
  const size_t need = mt::single_pass_state_bytes(F, B);
  if (workspace == nullptr || workspace_bytes < need) {
    mt::set_error("mt_generate_keys_single_pass: workspace of %zu bytes is smaller than the %zu required",
                  workspace_bytes, need);
    return MT_ERR_WORKSPACE_TOO_SMALL;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  MT_REQUIRE(B <= mt::kPassFaces, "mt_generate_keys_single_pass: more than %d cameras", mt::kPassFaces);
  MT_REQUIRE(B <= 1 || (layout.camera_bits < 31 && (int64_t(1) << layout.camera_bits) >= B),
             "mt_generate_keys_single_pass: the key layout's camera field cannot tell %d cameras apart", B);
  const int64_t units = mt::single_pass_blocks(F, B);
  const int lfpb = mt::single_pass_lfpb(B);
  const mt::SinglePassState state = mt::single_pass_state(workspace, F, B);
  if (units == 0) {
    mt::fill_i64_kernel<<<1, 64, 0, st>>>(n_keys, 1, 0);
    return mt::check_launch("mt_generate_keys_single_pass");
  }
  MT_REQUIRE(units < (int64_t(1) << 31) - 8, "mt_generate_keys_single_pass: too many face blocks");
  // The key counter starts at zero on every call (a fill kernel, so the call can be captured).
  mt::fill_i64_kernel<<<1, 64, 0, st>>>(reinterpret_cast<int64_t*>(state.counter), 1, 0);
  auto kernel = mt::pow2_tiles(tile_h, tile_w) ? mt::keys_single_pass_kernel<true>
                                                : mt::keys_single_pass_kernel<false>;
  const int64_t grid = (units + 7) / 8 * 8;  // every XCD the same number of blocks
  kernel<<<static_cast<unsigned>(grid), mt::kPassThreads, 0, st>>>(
      xy, depths, V, faces, F, num_faces_x3, shell_face_boundaries, S, B, lfpb,
      mt::single_pass_slices(B, lfpb), tile_h, tile_w, row_min, col_min, row_max, col_max, near_plane, cam_dims, cull, grid_cols,
      mt::to_fields(layout), scratch_keys, scratch_values, key_capacity, n_keys, overflow,
      state.counter, state.unit_count, state.unit_at);
  int rc = mt::check_launch("mt_generate_keys_single_pass");
  if (rc != MT_OK) return rc;
  rc = mt_inclusive_scan_i32_to_i64(state.unit_count, state.incl, units, state.scan,
                                    state.scan_bytes, stream);
  if (rc != MT_OK) return rc;
  mt::reorder_keys_kernel<<<static_cast<unsigned>((units + mt::kReorderWaves - 1) / mt::kReorderWaves),
                            mt::kReorderWaves * mt::kWave, 0, st>>>(
      scratch_keys, scratch_values, state.unit_count, state.unit_at, state.incl, units,
      key_capacity, keys, values, n_keys);
  return mt::check_launch("mt_generate_keys_single_pass");
}

extern "C" int mt_delineate_tiles(const int64_t* sorted_keys, int64_t n, mt_key_layout layout,
                                  int32_t separate_shells, int32_t B, int32_t S_out, int32_t T,
                                  int64_t* tile_ranges, const int64_t* n_device,
                                  const int64_t* tile_base, mt_stream_t stream) {
  MT_REQUIRE(n >= 0, "mt_delineate_tiles: negative length");
  MT_REQUIRE(tile_base == nullptr || (separate_shells == 0 && S_out == 1),
             "mt_delineate_tiles: per-camera tile grids need one range per (camera, tile)");
  if (n == 0) return MT_OK;
  mt::delineate_kernel<<<mt::blocks_for(n, mt::kBlock), mt::kBlock, 0, static_cast<hipStream_t>(stream)>>>(
      sorted_keys, n, mt::to_fields(layout), separate_shells, B, S_out, T, tile_ranges, n_device,
      tile_base);
  return mt::check_launch("mt_delineate_tiles");
}

extern "C" int mt_camera_positions(const float* extrinsics, int32_t B, float* out,
                                   mt_stream_t stream) {
  MT_REQUIRE(B >= 0, "mt_camera_positions: negative size");
  if (B == 0) return MT_OK;
  mt::camera_positions_kernel<<<mt::blocks_for(B, 64), 64, 0, static_cast<hipStream_t>(stream)>>>(
      extrinsics, B, out);
  return mt::check_launch("mt_camera_positions");
}

#define MT_SH_DISPATCH(KERNEL, ...)                                                          \
  switch (sh_degree) {                                                                       \
    case 0: mt::KERNEL<0><<<grid, mt::kBlock, 0, s>>>(__VA_ARGS__); break;                   \
    case 1: mt::KERNEL<1><<<grid, mt::kBlock, 0, s>>>(__VA_ARGS__); break;                   \
    case 2: mt::KERNEL<2><<<grid, mt::kBlock, 0, s>>>(__VA_ARGS__); break;                   \
    default: mt::KERNEL<3><<<grid, mt::kBlock, 0, s>>>(__VA_ARGS__); break;                  \
  }

extern "C" int mt_evaluate_sh_forward(const float* vertices, int64_t V, const float* sh,
                                      int32_t num_sh, int32_t sh_degree,
                                      const float* camera_positions, int32_t B, float* colors,
                                      mt_stream_t stream) {
  MT_REQUIRE(sh_degree >= 0 && sh_degree <= 3, "mt_evaluate_sh_forward: degree must be in [0, 3]");
  MT_REQUIRE(num_sh >= (sh_degree + 1) * (sh_degree + 1),
             "mt_evaluate_sh_forward: %d coefficients < (deg+1)^2", num_sh);
  if (V == 0 || B == 0) return MT_OK;
  const dim3 grid(mt::blocks_for(V, mt::kBlock));
  hipStream_t s = static_cast<hipStream_t>(stream);
  MT_SH_DISPATCH(sh_forward_kernel, vertices, V, sh, camera_positions, B, colors);
  return mt::check_launch("mt_evaluate_sh_forward");
}

extern "C" int mt_evaluate_sh_backward(const float* vertices, int64_t V, const float* sh,
                                       int32_t num_sh, int32_t sh_degree,
                                       const float* camera_positions, int32_t B,
                                       const float* grad_colors, float* grad_sh,
                                       mt_stream_t stream) {
  MT_REQUIRE(sh_degree >= 0 && sh_degree <= 3, "mt_evaluate_sh_backward: degree must be in [0, 3]");
  MT_REQUIRE(num_sh >= (sh_degree + 1) * (sh_degree + 1),
             "mt_evaluate_sh_backward: %d coefficients < (deg+1)^2", num_sh);
  if (V == 0 || B == 0) return MT_OK;
  const dim3 grid(mt::blocks_for(V, mt::kBlock));
  hipStream_t s = static_cast<hipStream_t>(stream);
  MT_SH_DISPATCH(sh_backward_kernel, vertices, V, sh, camera_positions, B, grad_colors, grad_sh);
  return mt::check_launch("mt_evaluate_sh_backward");
}
