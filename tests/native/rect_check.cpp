// Brute-force check of the candidate rectangle (meshtryoshka_amd/csrc/geometry.h): for
// seeded random triangles of every kind (sub-pixel, large, slivers, collinear, repeated
// corners, far-away or huge coordinates) and random tiles, every pixel whose barycentric
// test accepts the triangle must lie inside the rectangle. Host IEEE float arithmetic
// with -ffp-contract=off evaluates the same expressions as the GPU kernels.
//
// The same sweep checks the bounding-box route (centre_range): when it reports no
// candidate pixel centre, no pixel of the region may accept the triangle; when it reports a
// few, every accepted pixel must lie among them. Likewise, when centre_certificate holds and
// centre_box_empty reports an empty widened box, no pixel may accept the triangle.
//
// Prints "pairs=<tested> hits=<accepted> outside=<violations> area=<sum of rect areas>
// full=<whole-tile fallbacks> pixels=<sum of tile / region pixel counts> centre_none=<...>
// centre_few=<...> centre_unknown=<...> centre_outside=<violations> box_none=<...>
// box_outside=<violations>" and exits 1 on any violation.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#define MT_HD inline
#include "../../meshtryoshka_amd/csrc/geometry.h"

namespace {

uint64_t g_state = 0x9E3779B97F4A7C15ull;
uint64_t next_u64() {
  g_state ^= g_state << 13;
  g_state ^= g_state >> 7;
  g_state ^= g_state << 17;
  return g_state;
}
double uniform() { return double(next_u64() >> 11) * 0x1p-53; }
double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }

struct Tri {
  float x[3], y[3];
};

Tri random_triangle(int kind, double cx, double cy) {
  Tri t;
  switch (kind) {
    case 0: {  // sub-pixel to a few pixels
      const double s = std::pow(10.0, uniform(-2.5, 0.7));
      for (int i = 0; i < 3; ++i) {
        t.x[i] = float(cx + uniform(-s, s));
        t.y[i] = float(cy + uniform(-s, s));
      }
      break;
    }
    case 1: {  // tile-sized to large
      const double s = std::pow(10.0, uniform(0.7, 3.5));
      for (int i = 0; i < 3; ++i) {
        t.x[i] = float(cx + uniform(-s, s));
        t.y[i] = float(cy + uniform(-s, s));
      }
      break;
    }
    case 2: {  // sliver: third corner near the segment of the first two
      const double s = std::pow(10.0, uniform(-1.0, 2.5));
      const double ax = cx + uniform(-s, s), ay = cy + uniform(-s, s);
      const double bx = cx + uniform(-s, s), by = cy + uniform(-s, s);
      const double f = uniform(-0.5, 1.5), eps = std::pow(10.0, uniform(-9.0, -1.0));
      t.x[0] = float(ax);
      t.y[0] = float(ay);
      t.x[1] = float(bx);
      t.y[1] = float(by);
      t.x[2] = float(ax + f * (bx - ax) + uniform(-eps, eps));
      t.y[2] = float(ay + f * (by - ay) + uniform(-eps, eps));
      break;
    }
    case 3: {  // repeated corners / a point
      const double s = std::pow(10.0, uniform(-2.0, 1.5));
      const float px = float(cx + uniform(-s, s)), py = float(cy + uniform(-s, s));
      for (int i = 0; i < 3; ++i) {
        t.x[i] = px;
        t.y[i] = py;
      }
      if (next_u64() & 1) {
        t.x[2] = float(cx + uniform(-s, s));
        t.y[2] = float(cy + uniform(-s, s));
      }
      break;
    }
    case 4: {  // corners snapped to pixel centres / edges
      for (int i = 0; i < 3; ++i) {
        t.x[i] = float(std::floor(cx + uniform(-3.0, 3.0)) + 0.5 * double(next_u64() % 3));
        t.y[i] = float(std::floor(cy + uniform(-3.0, 3.0)) + 0.5 * double(next_u64() % 3));
      }
      break;
    }
    case 5: {  // small, marching-cubes-like triangles (0.01 .. 3 px), any orientation
      const double s = std::pow(10.0, uniform(-2.0, 0.5));
      const double a = uniform(0.0, 6.283185307179586);
      const double b = a + uniform(0.05, 3.09), r1 = s * uniform(0.2, 1.0), r2 = s * uniform(0.2, 1.0);
      t.x[0] = float(cx);
      t.y[0] = float(cy);
      t.x[1] = float(cx + r1 * std::cos(a));
      t.y[1] = float(cy + r1 * std::sin(a));
      t.x[2] = float(cx + r2 * std::cos(b));
      t.y[2] = float(cy + r2 * std::sin(b));
      break;
    }
    case 6: {  // one small angle (2..6 degrees) at the first corner
      const double s = std::pow(10.0, uniform(-2.0, 2.0));
      const double a = uniform(0.0, 6.283185307179586), d = uniform(0.035, 0.105);
      const double r1 = s * uniform(0.3, 1.0), r2 = s * uniform(0.3, 1.0);
      t.x[0] = float(cx);
      t.y[0] = float(cy);
      t.x[1] = float(cx + r1 * std::cos(a));
      t.y[1] = float(cy + r1 * std::sin(a));
      t.x[2] = float(cx + r2 * std::cos(a + d));
      t.y[2] = float(cy + r2 * std::sin(a + d));
      break;
    }
    default: {  // far away or huge coordinates
      const double s = std::pow(10.0, uniform(3.0, 9.0));
      for (int i = 0; i < 3; ++i) {
        t.x[i] = float(uniform(-s, s));
        t.y[i] = float(uniform(-s, s));
      }
      if (next_u64() & 1) {
        t.x[0] = float(cx);
        t.y[0] = float(cy);
      }
      break;
    }
  }
  return t;
}

}  // namespace

int main(int argc, char** argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 200000;
  if (argc > 2) g_state ^= (std::strtoull(argv[2], nullptr, 10) + 1) * 0xD1B54A32D192ED03ull;
  // (h, w): the autotuner's tile shapes, and multi-tile regions (the key kernels bound a
  // face's whole tile-aligned bounding box at once).
  static const int shapes[8][2] = {{8, 8},   {16, 16}, {32, 8}, {16, 32},
                                   {24, 40}, {40, 72}, {8, 96}, {64, 16}};
  long long pairs = 0, hits = 0, outside = 0, area = 0, full = 0, pixels = 0;
  long long cnone = 0, cfew = 0, cunknown = 0, coutside = 0, bnone = 0, boutside = 0;
  for (long it = 0; it < n; ++it) {
    const int* sh = shapes[next_u64() % 8];
    const int th = sh[0], tw = sh[1];
    const int x0 = int(next_u64() % 160) * tw, y0 = int(next_u64() % 100) * th;
    const double cx = x0 + uniform(-4.0, tw + 4.0), cy = y0 + uniform(-4.0, th + 4.0);
    const Tri t = random_triangle(int(next_u64() % 8), cx, cy);
    const mt::TriTerms terms = mt::tri_terms(t.x[0], t.y[0], t.x[1], t.y[1], t.x[2], t.y[2]);
    const mt::Rect r = mt::candidate_rect(terms, x0, y0, tw, th);
    area += (long long)r.w * r.h;
    pixels += (long long)tw * th;
    if (r.w == tw && r.h == th) ++full;
    if (r.x < 0 || r.y < 0 || r.w < 0 || r.h < 0 || r.x + r.w > tw || r.y + r.h > th) {
      std::printf("bad rectangle %d %d %d %d\n", r.x, r.y, r.w, r.h);
      return 1;
    }
    const mt::CentreRange cen =
        mt::centre_range(t.x[0], t.y[0], t.x[1], t.y[1], t.x[2], t.y[2], x0, y0, tw, th, 16);
    const bool box_none = mt::centre_certificate(t.x[0], t.y[0], t.x[1], t.y[1], t.x[2], t.y[2]) &&
                          mt::centre_box_empty(t.x[0], t.y[0], t.x[1], t.y[1], t.x[2], t.y[2]);
    bnone += box_none;
    cnone += cen.state == mt::kCentresNone;
    cfew += cen.state == mt::kCentresFew;
    cunknown += cen.state == mt::kCentresUnknown;
    if (cen.state == mt::kCentresFew &&
        (cen.x < 0 || cen.y < 0 || cen.w <= 0 || cen.h <= 0 || cen.x + cen.w > tw ||
         cen.y + cen.h > th || cen.w * cen.h > 16)) {
      std::printf("bad centre range %d %d %d %d\n", cen.x, cen.y, cen.w, cen.h);
      return 1;
    }
    for (int j = 0; j < th; ++j) {
      for (int i = 0; i < tw; ++i) {
        const float px = float(x0 + i) + 0.5f, py = float(y0 + j) + 0.5f;
        const mt::BaryNum q = mt::bary_numerators(terms, px, py);
        const float v = q.vn / terms.D, w = q.wn / terms.D, u = q.un / terms.D;
        ++pairs;
        if (!mt::bary_accept(u, v, w)) continue;
        ++hits;
        const bool in_centres = cen.state == mt::kCentresUnknown ||
                                (cen.state == mt::kCentresFew && i >= cen.x && i < cen.x + cen.w &&
                                 j >= cen.y && j < cen.y + cen.h);
        if (!in_centres) {
          ++coutside;
          if (coutside <= 10)
            std::printf("centre violation: tri (%a,%a) (%a,%a) (%a,%a) tile (%d,%d,%dx%d) pixel "
                        "(%d,%d) state %d range (%d,%d,%d,%d)\n",
                        t.x[0], t.y[0], t.x[1], t.y[1], t.x[2], t.y[2], x0, y0, th, tw, i, j,
                        cen.state, cen.x, cen.y, cen.w, cen.h);
        }
        if (box_none && ++boutside <= 10)
          std::printf("box violation: tri (%a,%a) (%a,%a) (%a,%a) pixel (%d,%d)\n", t.x[0], t.y[0],
                      t.x[1], t.y[1], t.x[2], t.y[2], x0 + i, y0 + j);
        const bool in = i >= r.x && i < r.x + r.w && j >= r.y && j < r.y + r.h;
        if (!in) {
          ++outside;
          if (outside <= 10)
            std::printf("violation: tri (%a,%a) (%a,%a) (%a,%a) tile (%d,%d,%dx%d) pixel (%d,%d) "
                        "rect (%d,%d,%d,%d)\n",
                        t.x[0], t.y[0], t.x[1], t.y[1], t.x[2], t.y[2], x0, y0, th, tw, i, j, r.x,
                        r.y, r.w, r.h);
        }
      }
    }
  }
  std::printf("pairs=%lld hits=%lld outside=%lld area=%lld full=%lld pixels=%lld centre_none=%lld "
              "centre_few=%lld centre_unknown=%lld centre_outside=%lld box_none=%lld box_outside=%lld\n",
              pairs, hits, outside, area, full, pixels, cnone, cfew, cunknown, coutside, bnone, boutside);
  return outside == 0 && coutside == 0 && boutside == 0 ? 0 : 1;
}
