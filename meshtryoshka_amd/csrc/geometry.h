// Ray-triangle test terms shared by the compositing kernels, plus the candidate-pixel
// rectangle that lets a workgroup test a staged triangle only where it can hit.
//
// Everything here is plain IEEE single/double arithmetic on scalars, so the same code
// compiles for the GPU and, in the unit test that checks the rectangle against brute
// force, for the host (define MT_HD as `inline` there). Build with -ffp-contract=off.
#pragma once

#ifndef MT_HD
#define MT_HD __host__ __device__ __forceinline__
#endif

namespace mt {

constexpr float kEpsBary = 1e-8f;  // intersect_and_composite.slang:99
constexpr float kSumLo = 0.999f;   // :132
constexpr float kSumHi = 1.001f;

// Per-triangle terms of the barycentric solve (intersect_and_composite.slang:65-110):
// edge vectors from p1, their Gram entries and D = (chosen 2x2 Gram determinant) + eps.
// The determinant is the one whose off-diagonal |dot| is smallest.
struct TriTerms {
  float p1x, p1y, p2x, p2y;
  float e12x, e12y, e13x, e13y;
  float d1313, d1213, d1212, d1223, d1323, D;
};

MT_HD TriTerms tri_terms(float p1x, float p1y, float p2x, float p2y, float p3x, float p3y) {
  TriTerms t;
  const float e12x = p2x - p1x, e12y = p2y - p1y;
  const float e23x = p3x - p2x, e23y = p3y - p2y;
  const float e13x = p3x - p1x, e13y = p3y - p1y;
  const float d_12_12 = e12x * e12x + e12y * e12y;
  const float d_12_23 = e12x * e23x + e12y * e23y;
  const float d_12_13 = e12x * e13x + e12y * e13y;
  const float d_13_13 = e13x * e13x + e13y * e13y;
  const float d_13_23 = e13x * e23x + e13y * e23y;
  const float d_23_23 = e23x * e23x + e23y * e23y;
  const float a = __builtin_fabsf(d_12_13), b = __builtin_fabsf(d_12_23), c = __builtin_fabsf(d_13_23);
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
  t.p1x = p1x;
  t.p1y = p1y;
  t.p2x = p2x;
  t.p2y = p2y;
  t.e12x = e12x;
  t.e12y = e12y;
  t.e13x = e13x;
  t.e13y = e13y;
  t.d1313 = d_13_13;
  t.d1213 = d_12_13;
  t.d1212 = d_12_12;
  t.d1223 = d_12_23;
  t.d1323 = d_13_23;
  t.D = denom + kEpsBary;
  return t;
}

// Numerators of the three barycentric coordinates at pixel centre (px, py)
// (intersect_and_composite.slang:111-125): u belongs to p1, v to p2, w to p3.
struct BaryNum {
  float un, vn, wn;
};

MT_HD BaryNum bary_numerators(const TriTerms& t, float px, float py) {
  const float a1x = px - t.p1x, a1y = py - t.p1y;
  const float a2x = px - t.p2x, a2y = py - t.p2y;
  const float d_1x_12 = a1x * t.e12x + a1y * t.e12y;
  const float d_1x_13 = a1x * t.e13x + a1y * t.e13y;
  const float d_2x_12 = a2x * t.e12x + a2y * t.e12y;
  const float d_2x_13 = a2x * t.e13x + a2y * t.e13y;
  BaryNum n;
  n.vn = t.d1313 * d_1x_12 - t.d1213 * d_1x_13;
  n.wn = t.d1212 * d_1x_13 - t.d1213 * d_1x_12;
  n.un = t.d1223 * d_2x_13 - t.d1323 * d_2x_12;
  return n;
}

// The reference's acceptance rule (intersect_and_composite.slang:126-145).
MT_HD bool bary_accept(float u, float v, float w) {
  const float sum = (u + v) + w;
  const bool miss = __builtin_isnan(u) || __builtin_isnan(v) || __builtin_isnan(w) || u < 0.0f ||
                    u > 1.0f || v < 0.0f || v > 1.0f || w < 0.0f || w > 1.0f || sum < kSumLo ||
                    sum > kSumHi;
  return !miss;
}

// Miss proof without division: with r an approximate reciprocal of D (relative error
// <= 2^-22, e.g. the hardware rcp) the product q * r is within a factor (1 +- 2^-21) of
// q / D. If it is < -2^-100 the exactly rounded q / D is a nonzero negative number; if it
// is > 1 + 2^-19 the rounded quotient exceeds 1. Either way the reference rejects the
// pixel, so the divisions can be skipped. r == 0 disables the shortcut (callers pass 0
// unless |D| is in [2^-60, 2^60]).
MT_HD bool rcp_proves_miss(const BaryNum& n, float r) {
  if (r == 0.0f) return false;
  const float ua = n.un * r, va = n.vn * r, wa = n.wn * r;
  constexpr float kNeg = -0x1p-100f;
  constexpr float kBig = 1.0f + 0x1p-19f;
  return ua < kNeg || va < kNeg || wa < kNeg || ua > kBig || va > kBig || wa > kBig;
}

// Candidate rectangle of a triangle inside one tile.
//
// The tile's pixel centres are (x0 + i + 0.5, y0 + j + 0.5), 0 <= i < tw, 0 <= j < th.
// The rectangle [rx, rx + rw) x [ry, ry + rh) (tile-local pixel indices) contains every
// pixel whose bary test can accept this triangle; pixels outside it provably fail.
//
// Why: each numerator is an affine function of the pixel centre evaluated in floats,
// e.g. vn = d1313 (a1 . e12) - d1213 (a1 . e13) with a1 = p - p1. Every term passes
// through at most five roundings, so |vn_float - vn_exact| <= g5 (A |a1x| + B |a1y|)
// with A = |d1313 e12x| + |d1213 e13x|, B likewise in y, g5 = 5u / (1 - 5u), plus an
// absolute underflow allowance. Acceptance needs every coordinate >= 0, i.e.
// sign(D) * numerator >= 0 (or the quotient underflowing to -0, covered by |D| 2^-125).
// Over the tile |a1x| is at most the tile's reach from p1, so acceptance implies three
// half-planes sign(D) (alpha x + beta y) + delta >= 0 with exact (double) alpha, beta.
// When their normals positively span the plane with a safe angle between them, the
// half-planes intersect in a triangle whose vertices bound the rectangle (plus slack
// for the double rounding). Before that, a triangle whose |numerators| are too small over
// the whole tile for the coordinates to reach the 0.999 sum (degenerate or distant
// triangles) gets an empty rectangle. Otherwise, or for non-finite input, the whole tile
// is returned. Overflow inside the float test yields inf/NaN, which the test rejects.
struct Rect {
  int x, y, w, h;
};

// 1 / x in double for the rectangle's vertices and distance scale: on the GPU the hardware
// reciprocal refined by one Newton step (relative error near 2^-50, far inside the 1e-8
// relative slack), on the host the division. Callers keep |x| within [2^-1000, 2^1000].
MT_HD double rcp_refined(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double r = __builtin_amdgcn_rcp(x);
  return __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
#else
  return 1.0 / x;
#endif
}

MT_HD bool rcp_range(double x) {
  const double a = __builtin_fabs(x);
  return a >= 0x1p-1000 && a <= 0x1p1000;
}

MT_HD Rect candidate_rect(const TriTerms& t, int tile_x0, int tile_y0, int tw, int th) {
#if defined(__clang__)
  // Only double arithmetic below (the float inputs are converted exactly): fused
  // multiply-adds round less than the separate operations the bound allows for.
#pragma clang fp contract(fast)
#endif
  Rect full{0, 0, tw, th};
  const float vals[14] = {t.p1x, t.p1y, t.p2x, t.p2y, t.e12x, t.e12y, t.e13x,
                          t.e13y, t.d1313, t.d1213, t.d1212, t.d1223, t.d1323, t.D};
  bool ok = t.D != 0.0f;
#pragma unroll
  for (int i = 0; i < 14; ++i) ok = ok && __builtin_isfinite(vals[i]);
  // Nonzero edge components >= 2^-90 keep (pixel offset) x (edge) products normal.
  const float es[4] = {t.e12x, t.e12y, t.e13x, t.e13y};
#pragma unroll
  for (int i = 0; i < 4; ++i) ok = ok && (es[i] == 0.0f || __builtin_fabsf(es[i]) >= 0x1p-90f);
  if (!ok) return full;

  const double sg = t.D > 0.0f ? 1.0 : -1.0;
  const double xlo = 0.5, xhi = double(tw) - 0.5, ylo = 0.5, yhi = double(th) - 0.5;
  const double o1x = double(t.p1x) - tile_x0, o1y = double(t.p1y) - tile_y0;
  const double o2x = double(t.p2x) - tile_x0, o2y = double(t.p2y) - tile_y0;
  auto reach = [](double o, double lo, double hi) {
    return __builtin_fmax(__builtin_fabs(lo - o), __builtin_fabs(hi - o));
  };
  const double r1x = reach(o1x, xlo, xhi), r1y = reach(o1y, ylo, yhi);
  const double r2x = reach(o2x, xlo, xhi), r2y = reach(o2y, ylo, yhi);
  const double rmax = __builtin_fmax(__builtin_fmax(r1x, r1y), __builtin_fmax(r2x, r2y));
  const double e12x = t.e12x, e12y = t.e12y, e13x = t.e13x, e13y = t.e13y;
  const double esum = __builtin_fabs(e12x) + __builtin_fabs(e12y) + __builtin_fabs(e13x) +
                      __builtin_fabs(e13y);
  const double eta = 0x1p-120 * (1.0 + __builtin_fabs(double(t.D)) + rmax * esum);
  constexpr double kGamma = 8.0 * 0x1p-24;  // > 5u / (1 - 5u), also absorbs double rounding

  // Half-planes nx * x + ny * y + c >= 0 in tile-local coordinates.
  double nx[3], ny[3], c[3];
  auto plane = [&](int i, double k1, double k2, double ax, double ay, double bx, double by,
                   double ox, double oy, double rx, double ry) {
    // numerator = k1 (a . (ax, ay)) - k2 (a . (bx, by)), a = p - origin
    const double al = k1 * ax - k2 * bx, be = k1 * ay - k2 * by;
    const double A = __builtin_fabs(k1 * ax) + __builtin_fabs(k2 * bx);
    const double B = __builtin_fabs(k1 * ay) + __builtin_fabs(k2 * by);
    const double delta = kGamma * (A * rx + B * ry) + eta;
    nx[i] = sg * al;
    ny[i] = sg * be;
    c[i] = delta - (nx[i] * ox + ny[i] * oy);
  };
  plane(0, t.d1223, t.d1323, e13x, e13y, e12x, e12y, o2x, o2y, r2x, r2y);  // u
  plane(1, t.d1313, t.d1213, e12x, e12y, e13x, e13y, o1x, o1y, r1x, r1y);  // v
  plane(2, t.d1212, t.d1213, e13x, e13y, e12x, e12y, o1x, o1y, r1x, r1y);  // w

  // Whole-tile rejection: |numerator| <= M_i over the tile (affine, so its largest |value|
  // over the pixel-centre box is |value at the centre| + |nx| hx + |ny| hy with hx, hy the
  // half extents, plus the rounding allowance). If the coordinates cannot sum to 0.999
  // anywhere (degenerate or distant triangles), no pixel can hit.
  const double cx = 0.5 * (xlo + xhi), cy = 0.5 * (ylo + yhi);
  const double hx = 0.5 * (xhi - xlo), hy = 0.5 * (yhi - ylo);
  double msum = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double delta = c[i] + (nx[i] * (i == 0 ? o2x : o1x) + ny[i] * (i == 0 ? o2y : o1y));
    const double centre = nx[i] * cx + ny[i] * cy + (c[i] - delta);
    msum += __builtin_fabs(centre) + __builtin_fabs(nx[i]) * hx + __builtin_fabs(ny[i]) * hy +
            2.0 * __builtin_fabs(delta);
  }
  if (msum < 0.998 * __builtin_fabs(double(t.D))) return Rect{0, 0, 0, 0};

  double nn[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) nn[i] = nx[i] * nx[i] + ny[i] * ny[i];
  if (!(nn[0] > 0.0) || !(nn[1] > 0.0) || !(nn[2] > 0.0)) return full;
  const double c01 = nx[0] * ny[1] - ny[0] * nx[1];
  const double c12 = nx[1] * ny[2] - ny[1] * nx[2];
  const double c20 = nx[2] * ny[0] - ny[2] * nx[0];
  constexpr double kSin2 = 0x1p-40;  // sin^2 of the smallest accepted angle (2^-20 rad)
  const bool spread = c01 * c01 > kSin2 * nn[0] * nn[1] && c12 * c12 > kSin2 * nn[1] * nn[2] &&
                      c20 * c20 > kSin2 * nn[2] * nn[0];
  const bool pos = c01 > 0.0 && c12 > 0.0 && c20 > 0.0;
  const bool neg = c01 < 0.0 && c12 < 0.0 && c20 < 0.0;
  if (!spread || !(pos || neg)) return full;
  if (!rcp_range(c01) || !rcp_range(c12) || !rcp_range(c20)) return full;

  // Vertices: line i meets line j at ((c_j ny_i - c_i ny_j), (c_i nx_j - c_j nx_i)) / cross.
  // One reciprocal per pair of coordinates (the extra rounding is far inside the slack).
  const double i01 = rcp_refined(c01), i12 = rcp_refined(c12), i20 = rcp_refined(c20);
  const double q0x = (c[1] * ny[0] - c[0] * ny[1]) * i01, q0y = (c[0] * nx[1] - c[1] * nx[0]) * i01;
  const double q1x = (c[2] * ny[1] - c[1] * ny[2]) * i12, q1y = (c[1] * nx[2] - c[2] * nx[1]) * i12;
  const double q2x = (c[0] * ny[2] - c[2] * ny[0]) * i20, q2y = (c[2] * nx[0] - c[0] * nx[2]) * i20;
  // A NaN anywhere makes the sum NaN (fmin / fmax below would drop it).
  const double qsum = ((q0x + q0y) + (q1x + q1y)) + (q2x + q2y);
  if (!(qsum == qsum)) return full;
  const double xmin = __builtin_fmin(__builtin_fmin(q0x, q1x), q2x);
  const double xmax = __builtin_fmax(__builtin_fmax(q0x, q1x), q2x);
  const double ymin = __builtin_fmin(__builtin_fmin(q0y, q1y), q2y);
  const double ymax = __builtin_fmax(__builtin_fmax(q0y, q1y), q2y);
  // |c_i| / |n_i| <= |c_i| / max(|nx_i|, |ny_i|): distance scale of each line (over-
  // estimated by 1.001 against the reciprocal's rounding).
  double dist = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double m = __builtin_fmax(__builtin_fabs(nx[i]), __builtin_fabs(ny[i]));
    if (!rcp_range(m)) return full;
    dist += 1.001 * __builtin_fabs(c[i]) * rcp_refined(m);
  }
  const double qmag = __builtin_fabs(xmin) + __builtin_fabs(xmax) + __builtin_fabs(ymin) +
                      __builtin_fabs(ymax);
  const double slack = 1e-3 + 1e-8 * (dist + qmag);
  // Pixel i has its centre at i + 0.5.
  const double ix0 = __builtin_fmax(__builtin_ceil(xmin - slack - 0.5), 0.0);
  const double ix1 = __builtin_fmin(__builtin_floor(xmax + slack - 0.5), double(tw - 1));
  const double iy0 = __builtin_fmax(__builtin_ceil(ymin - slack - 0.5), 0.0);
  const double iy1 = __builtin_fmin(__builtin_floor(ymax + slack - 0.5), double(th - 1));
  Rect r;
  if (ix1 < ix0 || iy1 < iy0) {
    r.x = 0;
    r.y = 0;
    r.w = 0;
    r.h = 0;
    return r;
  }
  r.x = int(ix0);
  r.y = int(iy0);
  r.w = int(ix1) - r.x + 1;
  r.h = int(iy1) - r.y + 1;
  return r;
}

// Pixel centres that a small, well-shaped triangle can hit, from its bounding box alone.
//
// Claim: if the certificate below holds, every pixel centre P farther than m / 2 outside the
// triangle's bounding box (along x or y) fails the reference's barycentric test.
// Why: let lambda be P's exact barycentric coordinates for the float corners p1, p2, p3,
// W the box's larger side and d > 0 P's distance outside the box along one axis, say
// P.x < xlo - d. Then sum_i lambda_i (p_i.x - xlo) = P.x - xlo < -d with every
// p_i.x - xlo in [0, W], so some lambda_i <= -d / (2 W) (at most two are negative). The
// float numerator of that coordinate is cr^2 lambda_i (cr the exact cross product of the
// edges: each of the three Gram determinants equals cr^2) plus at most 24 u L^3 |P - p_k|
// (u = 2^-24, L the longest edge, p_k the numerator's base corner; half of it from the
// rounding of the edges and Gram entries, half from the five operations that evaluate the
// numerator), and |P - p_k| <= 2 (d + W). So it is negative with room to spare, and its
// quotient by D > 0 a negative number that does not round to -0, once
// cr^2 > 192 u L^3 W (1 + W / d). The computed determinant then also has D > 0. With
// L^3 <= sqrt(2) L^2 W and d >= m / 2 this is implied by the certificate
//     cr^2 (1 - 2^-6) >= 2^-15 L^2 W^2 (1 + 2 W / m),
// 2^-15 exceeding 192 sqrt(2) u by a factor 1.9 for the second-order terms; the float
// evaluation of both sides is off by far less than the 2^-6 and 2^-10 factors. Corners
// lie within (-2^12, 2^12), so the box ends widened by m in float are off by at most
// 2^-10 < m / 2, and cr >= 2^-20 keeps every product normal.
//
// Result (tile-box-local pixel indices, like Rect): kCentresNone when no pixel centre of
// the box [tile_x0, tile_x0 + tw) x [tile_y0, tile_y0 + th) is within the widened bounding
// box; kCentresFew with their rectangle when there are at most max_pixels; otherwise, or
// when the certificate fails, kCentresUnknown (the caller falls back to candidate_rect).
constexpr float kCentreMargin = 0x1p-5f;
enum CentreState { kCentresNone = 0, kCentresFew = 1, kCentresUnknown = 2 };
struct CentreRange {
  int x, y, w, h;
  int state;
};

// The certificate of centre_range (corners inside (-2^12, 2^12), cr >= 2^-20 and the shape
// inequality above).
MT_HD bool centre_certificate(float ax, float ay, float bx, float by, float cx, float cy) {
  constexpr float kLim = 0x1p12f;
  // NaN fails every comparison.
  const bool inside = __builtin_fabsf(ax) < kLim && __builtin_fabsf(ay) < kLim &&
                      __builtin_fabsf(bx) < kLim && __builtin_fabsf(by) < kLim &&
                      __builtin_fabsf(cx) < kLim && __builtin_fabsf(cy) < kLim;
  const float e12x = bx - ax, e12y = by - ay, e13x = cx - ax, e13y = cy - ay;
  const float e23x = cx - bx, e23y = cy - by;
  const float cr = __builtin_fabsf(e12x * e13y - e12y * e13x);
  const float l12 = e12x * e12x + e12y * e12y, l13 = e13x * e13x + e13y * e13y;
  const float l23 = e23x * e23x + e23y * e23y;
  const float lsq = __builtin_fmaxf(__builtin_fmaxf(l12, l13), l23) * (1.0f + 0x1p-10f);
  const float xlo = __builtin_fminf(__builtin_fminf(ax, bx), cx);
  const float xhi = __builtin_fmaxf(__builtin_fmaxf(ax, bx), cx);
  const float ylo = __builtin_fminf(__builtin_fminf(ay, by), cy);
  const float yhi = __builtin_fmaxf(__builtin_fmaxf(ay, by), cy);
  const float wb = __builtin_fmaxf(xhi - xlo, yhi - ylo) * (1.0f + 0x1p-10f);
  const float rhs = 0x1p-15f * lsq * wb * wb * (1.0f + (2.0f / kCentreMargin) * wb);
  return inside && cr >= 0x1p-20f && cr * cr * (1.0f - 0x1p-6f) >= rhs * (1.0f + 0x1p-10f);
}

// True when no pixel centre lies within the triangle's bounding box widened by
// kCentreMargin (no clamp to a tile box). With centre_certificate it proves that no pixel
// accepts the triangle; centre_range then reports kCentresNone for every tile box.
MT_HD bool centre_box_empty(float ax, float ay, float bx, float by, float cx, float cy) {
  const float xlo = __builtin_fminf(__builtin_fminf(ax, bx), cx);
  const float xhi = __builtin_fmaxf(__builtin_fmaxf(ax, bx), cx);
  const float ylo = __builtin_fminf(__builtin_fminf(ay, by), cy);
  const float yhi = __builtin_fmaxf(__builtin_fmaxf(ay, by), cy);
  return __builtin_ceilf(xlo - kCentreMargin - 0.5f) > __builtin_floorf(xhi + kCentreMargin - 0.5f) ||
         __builtin_ceilf(ylo - kCentreMargin - 0.5f) > __builtin_floorf(yhi + kCentreMargin - 0.5f);
}

MT_HD CentreRange centre_range(float ax, float ay, float bx, float by, float cx, float cy,
                               int tile_x0, int tile_y0, int tw, int th, int max_pixels) {
  CentreRange out{0, 0, 0, 0, kCentresUnknown};
  if (!centre_certificate(ax, ay, bx, by, cx, cy)) return out;
  const float xlo = __builtin_fminf(__builtin_fminf(ax, bx), cx);
  const float xhi = __builtin_fmaxf(__builtin_fmaxf(ax, bx), cx);
  const float ylo = __builtin_fminf(__builtin_fminf(ay, by), cy);
  const float yhi = __builtin_fmaxf(__builtin_fmaxf(ay, by), cy);
  // Pixel X has its centre at X + 0.5.
  int x0 = int(__builtin_ceilf(xlo - kCentreMargin - 0.5f));
  int x1 = int(__builtin_floorf(xhi + kCentreMargin - 0.5f));
  int y0 = int(__builtin_ceilf(ylo - kCentreMargin - 0.5f));
  int y1 = int(__builtin_floorf(yhi + kCentreMargin - 0.5f));
  x0 = x0 > tile_x0 ? x0 : tile_x0;
  y0 = y0 > tile_y0 ? y0 : tile_y0;
  x1 = x1 < tile_x0 + tw - 1 ? x1 : tile_x0 + tw - 1;
  y1 = y1 < tile_y0 + th - 1 ? y1 : tile_y0 + th - 1;
  if (x1 < x0 || y1 < y0) {
    out.state = kCentresNone;
    return out;
  }
  if ((x1 - x0 + 1) * (y1 - y0 + 1) > max_pixels) return out;
  out.x = x0 - tile_x0;
  out.y = y0 - tile_y0;
  out.w = x1 - x0 + 1;
  out.h = y1 - y0 + 1;
  out.state = kCentresFew;
  return out;
}

}  // namespace mt
