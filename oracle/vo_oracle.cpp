// vo_oracle.cpp -- CPU oracle for the dexel morphology path.
//
// TEST INFRASTRUCTURE ONLY.  This library restates, on the CPU and in FP64,
// the algorithm of voroffset's `ours` method (VoronoiMorphoVorPower) and of its
// `brute_force` method, so that the HIP implementation in voroffset_amd/ can be
// checked against it.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load it; the product never does.
//
// Reference files restated (geometryprocessing/voroffset, src/vor3d/):
//   Common.h:8-10,91-200         EPS and the segment/point orderings/predicates
//   MorphologyOperators.cpp      1-D merges (appendSegment, unionSegs,
//   MorphologyOperators.hpp      addSegmentAtTheEnd) and complements
//                                (negate_ray, negate_ray_range)
//   Voronoi2D.cpp:67-744         pass 1: half-space Voronoi diagram of segments
//   SeparatePower2D.cpp:71-541   pass 2: power diagram of endpoint disks
//   HalfDilationOperator.cpp:6-29, .hpp:5-15   per-line sweep and unionMap
//   VoronoiVorPower.cpp:24-96    orchestration of the two passes
//   Voronoi.cpp:8-89             erosion = crop o complement o dilation o complement
//   VoronoiBruteForce.cpp:16-100 closed-form dilation
//   app/cli3d/offset3d.cpp:124-133  closing / opening composition
//   Dexelize.cpp:56-225          mesh -> dexels (ray casting, SOS predicates)
//   Voronoi.cpp:91-111, MorphologyOperators.cpp:319-374  calculateXor
//
// How each part is pinned (the reference itself cannot be compiled here: its
// core includes <Eigen/Dense>, which this image does not have):
//   * brute_force: closed forms (single dexel -> lattice ball, single segment
//     -> capsule), brute force over the lattice on tiny inputs, invariants.
//   * ours (both passes): equality with brute_force on random volumes and on
//     ball unions, for all four operations (the reference's own two methods
//     agree to 1.4e-14 on such inputs).
//   * the pass-1 intermediate: the envelope identity (max over triples of
//     rho^2 - dist_z^2 equals max over input segments within |dy| <= R of
//     R^2 - dy^2 - dist_z^2), and its mean count per column against the counts
//     measured on the compiled reference for the same input distribution.
//     The exact triples themselves are "parity unpinned".
//   * 1-D merges and complements: textbook interval union / complement on
//     hand-made cases and random property checks.
//   * calculateXor: set identities (xor(A, A) empty, xor(A, empty) = A,
//     measure = |A| + |B| - 2|A n B|) and pointwise membership on samples.
//   * rasterizer: closed forms (axis-aligned boxes, dexel centres exactly on
//     shared edges and vertices, where SOS must count a point exactly once),
//     and brute force over all triangles (the tile pruning only skips
//     triangles whose bounding box misses the dexel centre).

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr double kEps = 1e-8;  // Common.h:8-10

[[noreturn]] void fail(const char* what) { throw std::runtime_error(what); }

using Column = std::vector<double>;

// ---------------------------------------------------------------------------
// 1-D column primitives (MorphologyOperators.{cpp,hpp})
// ---------------------------------------------------------------------------

// Sorted append with backward merge; accepts a start that precedes earlier
// starts (MorphologyOperators.cpp:15-34).
void append_interval(Column& col, double a, double b) {
  while (!col.empty() && col[col.size() - 2] >= a) {
    b = std::max(b, col.back());
    col.resize(col.size() - 2);
  }
  if (col.empty() || col.back() < a) {
    col.push_back(a);
    col.push_back(b);
  } else if (col.back() < b) {
    col.back() = b;
  }
}

// Append at the tail, merging touching or overlapping intervals
// (MorphologyOperators.hpp:7-15).
void push_merge(Column& dst, double a, double b) {
  if (dst.empty() || dst.back() < a) {
    dst.push_back(a);
    dst.push_back(b);
  } else if (b > dst.back()) {
    dst.back() = b;
  }
}

// Union of two sorted disjoint interval lists; on equal starts the second list
// goes first (MorphologyOperators.cpp:117-158).
Column union_columns(const Column& a, const Column& b) {
  if (a.empty()) return b;
  if (b.empty()) return a;
  Column out;
  out.reserve(a.size() + b.size());
  size_t ia = 0, ib = 0;
  while (ia + 1 < a.size() || ib + 1 < b.size()) {
    bool take_b = ib + 1 < b.size() && (ia + 1 >= a.size() || b[ib] <= a[ia]);
    if (take_b) {
      push_merge(out, b[ib], b[ib + 1]);
      ib += 2;
    } else {
      push_merge(out, a[ia], a[ia + 1]);
      ia += 2;
    }
  }
  return out;
}

// Complement inside [lo, hi] by toggling the two ends; exact compares on the
// bounds, so an endpoint outside [lo, hi] yields an inverted interval
// (MorphologyOperators.cpp:230-259).
void negate_ray(Column& c, double lo, double hi) {
  if (c.empty()) {
    c = {lo, hi};
    return;
  }
  if (c.front() == lo)
    c.erase(c.begin());
  else
    c.insert(c.begin(), lo);
  if (c.back() == hi)
    c.pop_back();
  else
    c.push_back(hi);
}

// Complement clipped to [lo, hi] with parity counting of the dropped ends
// (MorphologyOperators.cpp:282-315).
void negate_ray_range(Column& c, double lo, double hi) {
  if (c.empty()) {
    c = {lo, hi};
    return;
  }
  size_t head = 0;
  while (head < c.size() && c[head] <= lo) ++head;
  c.erase(c.begin(), c.begin() + head);
  if (head % 2 == 0) c.insert(c.begin(), lo);
  size_t tail = 0;
  while (!c.empty() && c.back() >= hi) {
    c.pop_back();
    ++tail;
  }
  if (tail % 2 == 0) c.push_back(hi);
}

// ---------------------------------------------------------------------------
// Radius-tagged segments (pass-1 output) and their merges
// ---------------------------------------------------------------------------

struct RSeg {  // SegmentWithRadius, Common.h:119-155
  double y1, y2, r;
  bool point() const { return std::abs(y1 - y2) < kEps; }
  bool splits(const RSeg& o) const { return r < o.r && !point(); }
  bool hides(const RSeg& o) const {
    return y1 < o.y1 + kEps && o.y2 < y2 + kEps && r > o.r - kEps;
  }
};
using RColumn = std::vector<RSeg>;

// Append one tagged segment to a z-sorted disjoint list, larger radius wins on
// overlap (MorphologyOperators.cpp:37-95).
void append_rseg(RColumn& col, RSeg seg) {
  RColumn later;  // pieces that end up above `seg`, in reverse order
  while (!col.empty() && col.back().y1 >= seg.y1) {
    RSeg& t = col.back();
    if (t.y1 > seg.y2) {
      later.push_back(t);
    } else if (t.y2 >= seg.y2) {
      if (t.r > seg.r) {
        seg.y2 = t.y1;
        later.push_back(t);
      } else if (t.r == seg.r) {
        seg.y2 = t.y2;
      } else {
        t.y1 = seg.y2;
        later.push_back(t);
      }
    } else if (t.r > seg.r) {
      later.push_back(RSeg{t.y2, seg.y2, seg.r});
      later.push_back(t);
      seg.y2 = t.y1;
    }
    col.pop_back();
  }
  if (col.empty() || col.back().y2 < seg.y1) {
    col.push_back(seg);
  } else if (col.back().y2 < seg.y2) {
    RSeg& t = col.back();
    if (t.r == seg.r) {
      t.y2 = seg.y2;
    } else if (t.r > seg.r) {
      seg.y1 = t.y2;
      later.push_back(seg);
    } else {
      t.y2 = seg.y1;
      later.push_back(seg);
    }
  }
  col.insert(col.end(), later.rbegin(), later.rend());
}

// Tail append with priority by radius (MorphologyOperators.hpp:21-71).
void push_priority(RColumn& dst, RSeg s) {
  if (dst.empty()) {
    dst.push_back(s);
    return;
  }
  RSeg& t = dst.back();
  if (t.hides(s)) return;
  if (s.hides(t)) {
    dst.back() = s;
    return;
  }
  if (t.y2 < s.y1) {
    dst.push_back(s);
    return;
  }
  if (s.y2 > t.y2) {
    if (t.splits(s))
      t.y2 = s.y1;
    else
      s.y1 = t.y2;
    dst.push_back(s);
    return;
  }
  if (t.splits(s)) {
    RSeg upper = t;
    upper.y1 = s.y2;
    t.y2 = s.y1;
    dst.push_back(s);
    dst.push_back(upper);
  }
}

// Merge of two radius-tagged lists; on equal starts the second list goes
// first (MorphologyOperators.hpp:75-127).
RColumn union_rcolumns(const RColumn& a, const RColumn& b) {
  if (a.empty()) return b;
  if (b.empty()) return a;
  RColumn out;
  size_t ia = 0, ib = 0;
  while (ia < a.size() || ib < b.size()) {
    bool take_b = ib < b.size() && (ia >= a.size() || b[ib].y1 <= a[ia].y1);
    push_priority(out, take_b ? b[ib++] : a[ia++]);
  }
  return out;
}

// ---------------------------------------------------------------------------
// Pass 1: half-space Voronoi diagram of vertical segments (Voronoi2D.cpp)
// ---------------------------------------------------------------------------

struct VPt {
  double x, y;
};

struct VSeg {  // Segment, Common.h:79-117
  double x, y1, y2;
  bool point() const { return std::abs(y1 - y2) < kEps; }
  VPt lo() const { return {x, y1}; }
  VPt hi() const { return {x, y2}; }
  bool covers(const VSeg& o) const { return y1 <= o.y1 && o.y2 <= y2; }
};

// Ordered by midpoint, then by sweep coordinate (Common.h:91-94).
struct VSegLess {
  bool operator()(const VSeg& a, const VSeg& b) const {
    const double sa = a.y1 + a.y2, sb = b.y1 + b.y2;
    return sa < sb || (sa == sb && a.x < b.x);
  }
};

// Line a*X + b*Y = c.
struct Line {
  double a, b, c;
};

// Perpendicular bisector of two points (Voronoi2D.cpp:76-82).
Line bisector(VPt p, VPt q) {
  return {2 * (p.x - q.x), 2 * (p.y - q.y),
          p.x * p.x - q.x * q.x + p.y * p.y - q.y * q.y};
}

inline double det2(double a, double b, double c, double d) { return a * d - b * c; }
inline double sgn(double a) { return a > 0 ? 1.0 : -1.0; }
inline bool below_sqrt(double a, double x) { return a <= 0 || a * a <= x; }  // a <= sqrt(x)
inline bool above_sqrt(double b, double x) { return b >= 0 && x <= b * b; }  // sqrt(x) <= b

// Circumcentre abscissa of p, q, r, accepted when its ordinate lies in
// [ylo, yhi] (with EPS slack) and it is strictly beyond q
// (ray_intersect_{range,after,before}, Voronoi2D.cpp:109-204).
bool circum_x_in(VPt p, VPt q, VPt r, double ylo, double yhi, double& xout) {
  if (std::abs(p.x - q.x) < kEps && std::abs(q.x - r.x) < kEps) return false;
  const Line l1 = bisector(p, q), l2 = bisector(q, r);
  const double den = det2(l1.a, l2.a, l1.b, l2.b);
  if (std::abs(den) < kEps) return false;
  const double y = det2(l1.a, l2.a, l1.c, l2.c) / den;
  if (!(ylo < y + kEps && y < yhi + kEps)) return false;
  const double x = det2(l1.c, l2.c, l1.b, l2.b) / den;
  if (x > q.x) {
    xout = x;
    return true;
  }
  return false;
}

constexpr double kInf = std::numeric_limits<double>::infinity();

// Point equidistant from the interior of vertical segment s and from the
// points p, q, with ordinate inside s and abscissa >= s.x
// (voronoi_vertex, Voronoi2D.cpp:210-283).
bool seg_vertex_x(const VSeg& s, VPt p, VPt q, double& xout) {
  const Line r = bisector(p, q);
  const double u = 2.0 * (p.x - s.x);
  const double w = s.x * s.x - p.x * p.x;
  const double A = r.a;
  const double B = r.b * u - 2.0 * r.a * p.y;
  const double C = r.a * p.y * p.y - w * r.a - r.c * u;
  const double disc = B * B - 4.0 * A * C;
  const double lo = 2.0 * std::abs(A) * s.y1 + sgn(A) * B;
  const double hi = 2.0 * std::abs(A) * s.y2 + sgn(A) * B;
  const bool up = A > 0;
  double xm;
  if (std::abs(A) < kEps) {
    if (!(2.0 * s.y1 <= p.y + q.y && p.y + q.y <= 2.0 * s.y2)) return false;
    xm = ((p.y - q.y) * (p.y - q.y) - 4.0 * w) / (4.0 * u);
  } else if (disc < 0) {
    return false;
  } else if (std::abs(disc) < kEps && lo < kEps && -kEps < hi) {
    xm = (2.0 * A * r.c + r.b * B) / (2.0 * r.a * A);
  } else if (up ? (below_sqrt(lo, disc) && above_sqrt(hi, disc))
                : (above_sqrt(-lo, disc) && below_sqrt(-hi, disc))) {
    xm = (2.0 * A * r.c + r.b * B - r.b * std::sqrt(disc)) / (2.0 * r.a * A);
  } else if (!up ? (below_sqrt(lo, disc) && above_sqrt(hi, disc))
                 : (above_sqrt(-lo, disc) && below_sqrt(-hi, disc))) {
    xm = (2.0 * A * r.c + r.b * B + r.b * std::sqrt(disc)) / (2.0 * r.a * A);
  } else {
    return false;
  }
  if (xm >= s.x) {
    xout = xm;
    return true;
  }
  return false;
}

class HalfSpaceVoronoi {
 public:
  HalfSpaceVoronoi(int lines, double radius) : n_(lines), R_(radius), events_(lines) {}

  void reset() { front_.clear(); }

  // Remove seeds whose cell closed or that left the radius (Voronoi2D.cpp:674-692).
  void retire(int i) {
    auto& ev = events_[i];
    std::sort(ev.begin(), ev.end(), VSegLess());
    while (!ev.empty()) {
      auto it = front_.find(ev.back());
      ev.pop_back();
      if (it == front_.end()) continue;
      it = front_.erase(it);
      if (it != front_.begin() && it != front_.end()) {
        walk_down(std::prev(it), *it, i);
        walk_up(it, *std::prev(it), i);
      }
    }
    ev.clear();
  }

  // Insert seed segment (i, j1)-(i, j2) (Voronoi2D.cpp:591-669).
  void insert(int i, double j1, double j2) {
    const VSeg s{double(i), j1, j2};
    std::vector<VSeg> pieces;
    auto it = front_.lower_bound(s);
    while (it != front_.end() && s.covers(*it)) it = front_.erase(it);
    while (it != front_.begin() && s.covers(*std::prev(it))) front_.erase(std::prev(it));
    if (it != front_.end() && it->y1 <= s.y2) {  // overlapping neighbour above
      if (it->y1 < s.y1) pieces.push_back({it->x, it->y1, s.y1});
      pieces.push_back({it->x, s.y2, it->y2});
      it = front_.erase(it);
    }
    if (it != front_.begin()) {  // overlapping neighbour below
      auto jt = std::prev(it);
      if (jt->y2 >= s.y1) {
        if (jt->y2 > s.y2) pieces.push_back({jt->x, s.y2, jt->y2});
        pieces.push_back({jt->x, jt->y1, s.y1});
        front_.erase(jt);
      }
    }
    for (const VSeg& p : pieces) {
      front_.insert(it, p);
      expire_at(p);
    }
    auto jt = front_.insert(it, s);
    it = std::next(jt);
    expire_at(s);
    if (it != front_.end()) walk_up(it, s, i);
    if (jt != front_.begin()) walk_down(std::prev(jt), s, i);
  }

  // Emit the front at line i, tagged with rho = sqrt(R^2 - (i - x)^2)
  // (Voronoi2D.cpp:699-723).
  template <class Emit>
  void emit(int i, Emit&& out) const {
    for (const VSeg& s : front_) {
      const double d = (i - s.x < R_) ? i - s.x : R_;
      const double rho = (i == s.x) ? R_ : std::sqrt(R_ * R_ - d * d);
      out(s.y1, s.y2, rho);
    }
  }

  size_t size() const { return front_.size(); }

 private:
  using Iter = std::set<VSeg, VSegLess>::const_iterator;

  void expire_at(const VSeg& s) {
    const int x = int(std::floor(s.x + R_)) + 1;
    if (x < n_) events_[x].push_back(s);
  }

  // Abscissa where the circumcentre of three points lies beyond q, or n_
  // (VoronoiMorpho2D::rayIntersect, Voronoi2D.cpp:295-324).
  double circum_or_end(VPt p, VPt q, VPt r) const {
    if (std::abs(p.x - q.x) < kEps && std::abs(q.x - r.x) < kEps) return n_;
    const Line l1 = bisector(p, q), l2 = bisector(q, r);
    const double den = det2(l1.a, l2.a, l1.b, l2.b);
    if (std::abs(den) < kEps) return n_;
    const double x = det2(l1.c, l2.c, l1.b, l2.b) / den;
    return (x < q.x + kEps) ? double(n_) : x;
  }

  // Sweep abscissa at which the cell of `b` (between `a` below and `c` above)
  // closes, trying the candidate feature triples in a fixed order
  // (treatSegments, Voronoi2D.cpp:329-534).
  double close_x(const VSeg& a, const VSeg& b, const VSeg& c) const {
    if (b.x + kEps > a.x && b.x + kEps > c.x) return n_;
    double x = n_;
    auto below = [&](VPt p, VPt q, VPt r, double y) { return circum_x_in(p, q, r, -kInf, y, x); };
    auto above = [&](VPt p, VPt q, VPt r, double y) { return circum_x_in(p, q, r, y, kInf, x); };
    auto between = [&](VPt p, VPt q, VPt r, double y1, double y2) {
      return circum_x_in(p, q, r, y1, y2, x);
    };
    auto inside = [&](const VSeg& s, VPt p, VPt q) { return seg_vertex_x(s, p, q, x); };
    const bool pa = a.point(), pb = b.point(), pc = c.point();
    if (pb) {
      const VPt m = b.lo();
      if (pa && pc) return circum_or_end(a.lo(), m, c.lo());
      if (pa) {
        below(a.lo(), m, c.lo(), c.y1) || above(a.lo(), m, c.hi(), c.y2) || inside(c, a.lo(), m);
      } else if (pc) {
        below(a.lo(), m, c.lo(), a.y1) || above(a.hi(), m, c.lo(), a.y2) || inside(a, m, c.lo());
      } else {
        below(a.lo(), m, c.lo(), a.y1) || above(a.hi(), m, c.hi(), c.y2) ||
            between(a.hi(), m, c.lo(), a.y2, c.y1) || inside(a, m, c.lo()) ||
            inside(c, a.hi(), m);
      }
      return x;
    }
    if (pa && pc) {
      below(a.lo(), b.lo(), c.lo(), b.y1) || above(a.lo(), b.hi(), c.lo(), b.y2) ||
          inside(b, a.lo(), c.lo());
    } else if (pa) {
      below(a.lo(), b.lo(), c.lo(), b.y1) || above(a.lo(), b.hi(), c.hi(), c.y2) ||
          between(a.lo(), b.hi(), c.lo(), b.y2, c.y1) || inside(b, a.lo(), c.lo()) ||
          inside(c, a.lo(), b.hi());
    } else if (pc) {
      below(a.lo(), b.lo(), c.lo(), a.y1) || above(a.hi(), b.hi(), c.lo(), b.y2) ||
          between(a.hi(), b.lo(), c.lo(), a.y2, b.y1) || inside(b, a.hi(), c.lo()) ||
          inside(a, b.lo(), c.lo());
    } else {
      below(a.lo(), b.lo(), c.lo(), a.y1) || above(a.hi(), b.hi(), c.hi(), c.y2) ||
          between(a.hi(), b.lo(), c.lo(), a.y2, b.y1) ||
          between(a.hi(), b.hi(), c.lo(), b.y2, c.y1) || inside(b, a.hi(), c.lo()) ||
          inside(a, b.lo(), c.lo()) || inside(c, a.hi(), b.hi());
    }
    return x;
  }

  // exploreLeft (Voronoi2D.cpp:539-561): `it` and the seeds below it.
  void walk_down(Iter it, const VSeg& top, int i) {
    while (it != front_.begin()) {
      const double x = close_x(*std::prev(it), *it, top);
      if (x < i + kEps) {
        it = std::prev(front_.erase(it));
        continue;
      }
      schedule(x, *it);
      return;
    }
  }

  // exploreRight (Voronoi2D.cpp:564-586): `it` and the seeds above it.
  void walk_up(Iter it, const VSeg& bottom, int i) {
    while (std::next(it) != front_.end()) {
      const double x = close_x(bottom, *it, *std::next(it));
      if (x < i + kEps) {
        it = front_.erase(it);
        continue;
      }
      schedule(x, *it);
      return;
    }
  }

  void schedule(double x, const VSeg& s) {
    const int xs = int(std::floor(x)) + 1;
    if (x < n_ && xs < n_) events_[xs].push_back(s);
  }

  int n_;
  double R_;
  std::set<VSeg, VSegLess> front_;
  std::vector<std::vector<VSeg>> events_;
};

// ---------------------------------------------------------------------------
// Pass 2: power diagram of endpoint disks + priority list of segment
// interiors (SeparatePower2D.cpp)
// ---------------------------------------------------------------------------

struct PDisk {  // PointWithRadiusX, Common.h:57-74
  double x, y, r;
  bool swallows(const PDisk& o) const {
    return r - o.r > std::sqrt((x - o.x) * (x - o.x) + (y - o.y) * (y - o.y)) - kEps;
  }
};
struct PDiskLess {
  bool operator()(const PDisk& a, const PDisk& b) const {
    return a.y < b.y || (a.y == b.y && a.x < b.x) || (a.y == b.y && a.x == b.x && a.r < b.r);
  }
};

struct XSeg {  // SegmentWithRadiusX, Common.h:158-200
  double x, y1, y2, r;
  bool point() const { return std::abs(y1 - y2) < kEps; }
  bool splits(const XSeg& o) const { return (r + x < o.r + o.x) && !point(); }
  bool hides(const XSeg& o) const {
    return y1 < o.y1 + kEps && o.y2 < y2 + kEps && r + x > o.r + o.x - kEps;
  }
};

class PowerSweep {
 public:
  explicit PowerSweep(int lines) : n_(lines), events_(lines) {}

  void reset() {
    segs_.clear();
    fresh_segs_.clear();
    disks_.clear();
    fresh_disks_.clear();
  }

  // removeInactiveSegments (SeparatePower2D.cpp:258-273).
  void retire_segments(int i) {
    size_t j = 0;
    for (size_t k = 0; k < segs_.size(); ++k)
      if (int(std::floor(segs_[k].x + segs_[k].r)) + 1 > i) segs_[j++] = segs_[k];
    segs_.resize(j);
  }

  // removeInactivePoints (SeparatePower2D.cpp:275-293).
  void retire_disks(int i) {
    auto& ev = events_[i];
    std::sort(ev.begin(), ev.end(), PDiskLess());
    while (!ev.empty()) {
      auto it = disks_.find(ev.back());
      ev.pop_back();
      if (it == disks_.end()) continue;
      it = disks_.erase(it);
      if (it != disks_.begin() && it != disks_.end()) {
        walk_down(std::prev(it), *it, i);
        walk_up(it, *std::prev(it), i);
      }
    }
    ev.clear();
  }

  // insertSegment buffers the line's seeds (SeparatePower2D.cpp:215-226).
  void stage(int i, double j1, double j2, double r) {
    if (std::abs(j1 - j2) < kEps) {
      fresh_disks_.push_back({double(i), j1, r});
      return;
    }
    fresh_segs_.push_back({double(i), j1, j2, r});
    fresh_disks_.push_back({double(i), j1, r});
    fresh_disks_.push_back({double(i), j2, r});
  }

  // flushLine (SeparatePower2D.cpp:456-505): add the staged disks that no
  // current segment interior dominates, then merge the staged interiors.
  void flush() {
    for (const PDisk& p : fresh_disks_) {
      const int k = locate(p);
      if (k >= 0) {
        const XSeg& s = segs_[k];
        const double room = std::min(std::min(s.r - p.x + s.x, p.y - s.y1), s.y2 - p.y);
        if (p.r <= room) continue;
      }
      insert_disk(p);
    }
    merge_segments();
    fresh_segs_.clear();
    fresh_disks_.clear();
  }

  // getLine (SeparatePower2D.cpp:295-341): the slice of every live interior
  // and disk at line i, united.
  Column slice(int i) const {
    Column rs, rp;
    for (const XSeg& s : segs_) {
      if ((i - s.x) > s.r + kEps) fail("pass-2 segment outlived its radius");
      if (i - s.x > s.r) {
        append_interval(rs, s.y1, s.y2);
      } else {
        const double h = std::sqrt(s.r * s.r - (i - s.x) * (i - s.x));
        append_interval(rs, s.y1 - h, s.y2 + h);
      }
    }
    for (const PDisk& p : disks_) {
      if ((i - p.x) > p.r + kEps) fail("pass-2 disk outlived its radius");
      if (i - p.x > p.r) continue;
      const double h = std::sqrt(p.r * p.r - (i - p.x) * (i - p.x));
      append_interval(rp, p.y - h, p.y + h);
    }
    return union_columns(rp, rs);
  }

  size_t disk_count() const { return disks_.size(); }
  size_t seg_count() const { return segs_.size(); }

 private:
  using Iter = std::set<PDisk, PDiskLess>::const_iterator;

  // Abscissa at which the power cell of q (between p below and r above)
  // closes, or n_ when it does not close within q's reach
  // (rayIntersect, SeparatePower2D.cpp:118-158).
  double close_x(const PDisk& p, const PDisk& q, const PDisk& r) const {
    if (std::abs(p.x - q.x) < kEps && std::abs(q.x - r.x) < kEps) return n_;
    Line l1 = power_bisector(p, q), l2 = power_bisector(q, r);
    if (l1.b * l2.b < 0) {
      l1.a = -l1.a;
      l1.b = -l1.b;
      l1.c = -l1.c;
    }
    const double den = det2(l1.a, l2.a, l1.b, l2.b);
    if (std::abs(den) < kEps || den > 0) return n_;
    const double x = det2(l1.c, l2.c, l1.b, l2.b) / den;
    if (x < q.x + kEps || x > q.x + q.r) return n_;
    return x;
  }

  static Line power_bisector(const PDisk& p, const PDisk& q) {  // SeparatePower2D.cpp:82-88
    return {2 * (p.x - q.x), 2 * (p.y - q.y),
            p.x * p.x - q.x * q.x + p.y * p.y - q.y * q.y + q.r * q.r - p.r * p.r};
  }

  void walk_down(Iter it, const PDisk& top, int i) {  // exploreLeft, :163-185
    while (it != disks_.begin()) {
      const double x = close_x(*std::prev(it), *it, top);
      if (x < i + kEps) {
        it = std::prev(disks_.erase(it));
        continue;
      }
      schedule(x, *it);
      return;
    }
  }

  void walk_up(Iter it, const PDisk& bottom, int i) {  // exploreRight, :188-210
    while (std::next(it) != disks_.end()) {
      const double x = close_x(bottom, *it, *std::next(it));
      if (x < i + kEps) {
        it = disks_.erase(it);
        continue;
      }
      schedule(x, *it);
      return;
    }
  }

  void schedule(double x, const PDisk& d) {
    const int xs = int(std::floor(x)) + 1;
    if (x < n_ && xs < n_) events_[xs].push_back(d);
  }

  // insertPoint (SeparatePower2D.cpp:228-253).
  void insert_disk(const PDisk& p) {
    auto it = disks_.lower_bound(p);
    while (it != disks_.end() && p.swallows(*it)) it = disks_.erase(it);
    while (it != disks_.begin() && p.swallows(*std::prev(it))) disks_.erase(std::prev(it));
    auto jt = disks_.insert(it, p);
    const int xs = int(std::floor(p.x + p.r)) + 1;
    if (xs < n_) events_[xs].push_back(p);
    it = std::next(jt);
    if (it != disks_.end()) walk_up(it, p, int(p.x));
    if (jt != disks_.begin()) walk_down(std::prev(jt), p, int(p.x));
  }

  // Index of the interior whose start is the last one <= p.y, if p.y lies
  // strictly below its end; -1 otherwise (locatePoint, :506-541).
  int locate(const PDisk& p) const {
    int lo = 0, hi = int(segs_.size()) - 1;
    if (hi < 0) return -1;
    if (p.y > segs_[hi].y2 || p.y < segs_[lo].y1) return -1;
    while (lo < hi - 1) {
      const int mid = lo + ((hi - lo) >> 1);
      if (segs_[mid].y1 > p.y) {
        hi = mid;
      } else if (segs_[mid].y1 < p.y) {
        lo = mid;
      } else {
        lo = mid;
        break;
      }
    }
    return segs_[lo].y2 > p.y ? lo : -1;
  }

  // unionSegs + addSegment (SeparatePower2D.cpp:351-454): interiors kept
  // z-disjoint, the one reaching further along the sweep wins an overlap.
  void merge_segments() {
    if (segs_.empty()) {
      segs_ = fresh_segs_;
      return;
    }
    if (fresh_segs_.empty()) return;
    std::vector<XSeg> out;
    size_t ia = 0, ib = 0;
    while (ia < segs_.size() || ib < fresh_segs_.size()) {
      const bool take_a =
          ia < segs_.size() && (ib >= fresh_segs_.size() || segs_[ia].y1 <= fresh_segs_[ib].y1);
      push_seg(out, take_a ? segs_[ia++] : fresh_segs_[ib++]);
    }
    segs_.swap(out);
  }

  static void push_seg(std::vector<XSeg>& dst, XSeg s) {
    if (s.point()) return;
    if (dst.empty()) {
      dst.push_back(s);
      return;
    }
    XSeg& t = dst.back();
    if (t.hides(s)) return;
    if (t.y2 < s.y1) {
      dst.push_back(s);
      return;
    }
    if (s.y2 > t.y2) {
      if (t.splits(s)) {
        if (std::abs(t.y1 - s.y1) < kEps)
          dst.pop_back();
        else
          t.y2 = s.y1;
      } else {
        s.y1 = t.y2;
      }
      dst.push_back(s);
      return;
    }
    if (t.splits(s)) {
      XSeg upper = t;
      upper.y1 = s.y2;
      t.y2 = s.y1;
      if (t.point()) dst.pop_back();
      dst.push_back(s);
      if (!upper.point()) dst.push_back(upper);
    }
  }

  int n_;
  std::vector<XSeg> segs_, fresh_segs_;
  std::set<PDisk, PDiskLess> disks_;
  std::vector<PDisk> fresh_disks_;
  std::vector<std::vector<PDisk>> events_;
};

// ---------------------------------------------------------------------------
// Volumes and methods
// ---------------------------------------------------------------------------

struct Volume {
  int nx = 0, ny = 0;
  std::vector<Column> cols;  // index x + nx*y
  Column& at(int x, int y) { return cols[size_t(x) + size_t(nx) * y]; }
  const Column& at(int x, int y) const { return cols[size_t(x) + size_t(nx) * y]; }
};

struct RVolume {
  int nx = 0, ny = 0;
  std::vector<RColumn> cols;
  RColumn& at(int x, int y) { return cols[size_t(x) + size_t(nx) * y]; }
  const RColumn& at(int x, int y) const { return cols[size_t(x) + size_t(nx) * y]; }
};

// Runs fn(line) for line in [0, n) over `threads` workers, in blocks of
// `grain` lines (the dormant TBB structure of VoronoiVorPower.cpp:41-63).
template <class Fn>
void for_lines(int n, int threads, Fn&& fn) {
  if (threads <= 1) {
    for (int l = 0; l < n; ++l) fn(l, 0);
    return;
  }
  constexpr int grain = 10;
  std::atomic<int> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) {
    pool.emplace_back([&, t] {
      for (;;) {
        const int b = next.fetch_add(grain);
        if (b >= n) break;
        for (int l = b; l < std::min(n, b + grain); ++l) fn(l, t);
      }
    });
  }
  for (auto& th : pool) th.join();
}

// Pass 1 of `ours`: both half sweeps along y for every x-line, then their
// union (VoronoiVorPower.cpp:39-66, HalfDilationOperator.cpp:6-29).
RVolume ours_pass1(const Volume& in, double R, int threads) {
  RVolume fwd{in.nx, in.ny, std::vector<RColumn>(in.cols.size())};
  RVolume bwd = fwd;
  std::vector<HalfSpaceVoronoi> ops(std::max(threads, 1), HalfSpaceVoronoi(in.ny, R));
  for_lines(in.nx, threads, [&](int x, int t) {
    HalfSpaceVoronoi& op = ops[t];
    for (int dir = 0; dir < 2; ++dir) {
      RVolume& out = dir == 0 ? fwd : bwd;
      for (int k = 0; k < in.ny; ++k) {
        const int y = dir == 0 ? k : in.ny - 1 - k;
        op.retire(k);
        const Column& c = in.at(x, y);
        for (size_t e = 0; e + 1 < c.size(); e += 2) op.insert(k, c[e], c[e + 1]);
        op.emit(k, [&](double a, double b, double r) { append_rseg(out.at(x, y), RSeg{a, b, r}); });
      }
      op.reset();
    }
  });
  RVolume mid{in.nx, in.ny, std::vector<RColumn>(in.cols.size())};
  for (size_t c = 0; c < mid.cols.size(); ++c) mid.cols[c] = union_rcolumns(fwd.cols[c], bwd.cols[c]);
  return mid;
}

// Pass 2 of `ours` (VoronoiVorPower.cpp:68-95).
Volume ours_pass2(const RVolume& mid, int threads) {
  Volume fwd{mid.nx, mid.ny, std::vector<Column>(mid.cols.size())};
  Volume bwd = fwd;
  std::vector<PowerSweep> ops(std::max(threads, 1), PowerSweep(mid.nx));
  for_lines(mid.ny, threads, [&](int y, int t) {
    PowerSweep& op = ops[t];
    for (int dir = 0; dir < 2; ++dir) {
      Volume& out = dir == 0 ? fwd : bwd;
      for (int k = 0; k < mid.nx; ++k) {
        const int x = dir == 0 ? k : mid.nx - 1 - k;
        op.retire_segments(k);
        op.retire_disks(k);
        for (const RSeg& s : mid.at(x, y)) op.stage(k, s.y1, s.y2, s.r);
        op.flush();
        const Column sl = op.slice(k);
        Column& o = out.at(x, y);
        for (size_t e = 0; e + 1 < sl.size(); e += 2) append_interval(o, sl[e], sl[e + 1]);
      }
      op.reset();
    }
  });
  Volume res{mid.nx, mid.ny, std::vector<Column>(mid.cols.size())};
  for (size_t c = 0; c < res.cols.size(); ++c) res.cols[c] = union_columns(fwd.cols[c], bwd.cols[c]);
  return res;
}

Volume ours_dilation(const Volume& in, double R, int threads, double* t1, double* t2) {
  auto t0 = std::chrono::steady_clock::now();
  RVolume mid = ours_pass1(in, R, threads);
  auto t_1 = std::chrono::steady_clock::now();
  Volume out = ours_pass2(mid, threads);
  auto t_2 = std::chrono::steady_clock::now();
  if (t1) *t1 = std::chrono::duration<double, std::milli>(t_1 - t0).count();
  if (t2) *t2 = std::chrono::duration<double, std::milli>(t_2 - t_1).count();
  return out;
}

// brute_force dilation (VoronoiBruteForce.cpp:16-68,71-100): every lattice
// neighbour within R contributes [a - dz, b + dz].  Events at equal z are
// ordered starts first, so touching intervals merge as in `ours`.
Volume brute_dilation(const Volume& in, double R, int threads) {
  Volume out{in.nx, in.ny, std::vector<Column>(in.cols.size())};
  const double R2 = R * R;
  for_lines(in.ny, threads, [&](int y, int) {
    std::vector<std::pair<double, int>> ev;
    for (int x = 0; x < in.nx; ++x) {
      ev.clear();
      const int y0 = std::max(0, int(std::ceil(y - R))), y1 = std::min(in.ny - 1, int(std::floor(y + R)));
      const int x0 = std::max(0, int(std::ceil(x - R))), x1 = std::min(in.nx - 1, int(std::floor(x + R)));
      for (int yy = y0; yy <= y1; ++yy) {
        for (int xx = x0; xx <= x1; ++xx) {
          const double dx = xx - x, dy = yy - y;
          if (dx * dx + dy * dy > R2) continue;
          const double dz = std::sqrt(R2 - dx * dx - dy * dy);
          const Column& c = in.at(xx, yy);
          for (size_t e = 0; e + 1 < c.size(); e += 2) {
            ev.push_back({c[e] - dz, -1});
            ev.push_back({c[e + 1] + dz, +1});
          }
        }
      }
      std::sort(ev.begin(), ev.end());
      Column& o = out.at(x, y);
      int depth = 0;
      for (const auto& e : ev) {
        if (e.second < 0) {
          if (depth++ == 0) o.push_back(e.first);
        } else if (--depth == 0) {
          o.push_back(e.first);
        }
      }
    }
  });
  return out;
}

enum Method { kOurs = 0, kBrute = 1 };
enum Op { kNoop = 0, kDilation = 1, kErosion = 2, kClosing = 3, kOpening = 4 };

Volume dilate(int method, const Volume& in, double R, int threads, double* t1, double* t2) {
  if (method == kOurs) return ours_dilation(in, R, threads, t1, t2);
  auto t0 = std::chrono::steady_clock::now();
  Volume out = brute_dilation(in, R, threads);
  if (t1) *t1 = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (t2) *t2 = 0;
  return out;
}

// Complement on a grid grown by a ring of full columns (negate, Voronoi.cpp:18-55).
Volume complement_with_ring(const Volume& in, double lo, double hi) {
  Volume g{in.nx + 2, in.ny + 2, std::vector<Column>(size_t(in.nx + 2) * (in.ny + 2))};
  for (int y = 0; y < in.ny; ++y)
    for (int x = 0; x < in.nx; ++x) g.at(x + 1, y + 1) = in.at(x, y);
  for (auto& c : g.cols) negate_ray(c, lo, hi);
  return g;
}

// Crop the ring and complement within [lo, hi] (negateInv, Voronoi.cpp:57-89).
Volume crop_complement(const Volume& g, double lo, double hi) {
  Volume out{g.nx - 2, g.ny - 2, std::vector<Column>(size_t(g.nx - 2) * (g.ny - 2))};
  for (int y = 0; y < out.ny; ++y)
    for (int x = 0; x < out.nx; ++x) {
      out.at(x, y) = g.at(x + 1, y + 1);
      negate_ray_range(out.at(x, y), lo, hi);
    }
  return out;
}

// erosion (Voronoi.cpp:8-17); zlo/zhi are origin_z/s and origin_z/s + 2p + extent_z/s.
Volume erode(int method, const Volume& in, double R, double zlo, double zhi, int threads, double* t1,
             double* t2) {
  const double lo = zlo - 1, hi = zhi + 1;
  Volume g = complement_with_ring(in, lo, hi);
  Volume d = dilate(method, g, R, threads, t1, t2);
  return crop_complement(d, lo + 1, hi - 1);
}

Volume run_op(int method, int op, const Volume& in, double R, double zlo, double zhi, int threads,
              double* t1, double* t2) {
  switch (op) {
    case kNoop:
      return in;
    case kDilation:
      return dilate(method, in, R, threads, t1, t2);
    case kErosion:
      return erode(method, in, R, zlo, zhi, threads, t1, t2);
    case kClosing: {  // offset3d.cpp:124-128
      Volume tmp = dilate(method, in, R, threads, t1, t2);
      return erode(method, tmp, R, zlo, zhi, threads, t1, t2);
    }
    case kOpening: {  // offset3d.cpp:129-133
      Volume tmp = erode(method, in, R, zlo, zhi, threads, t1, t2);
      return dilate(method, tmp, R, threads, t1, t2);
    }
  }
  fail("unknown operation");
}

Volume from_csr(int nx, int ny, const int64_t* off, const double* z) {
  Volume v{nx, ny, std::vector<Column>(size_t(nx) * ny)};
  for (size_t c = 0; c < v.cols.size(); ++c) v.cols[c].assign(z + off[c], z + off[c + 1]);
  return v;
}

thread_local std::string g_error;

struct Result {
  int nx = 0, ny = 0;
  std::vector<int64_t> off;
  std::vector<double> a, b, r;  // endpoints (a only) or triples (a, b, r)
};

Result* to_result(const Volume& v) {
  auto* res = new Result;
  res->nx = v.nx;
  res->ny = v.ny;
  res->off.assign(v.cols.size() + 1, 0);
  for (size_t c = 0; c < v.cols.size(); ++c) res->off[c + 1] = res->off[c] + int64_t(v.cols[c].size());
  res->a.reserve(res->off.back());
  for (const auto& c : v.cols) res->a.insert(res->a.end(), c.begin(), c.end());
  return res;
}

}  // namespace

// ---------------------------------------------------------------------------
// Parity helpers (Voronoi.cpp:91-111, MorphologyOperators.cpp:319-374,
// CompressedVolume.cpp:61-83).

constexpr double kMinSegSize = 1e-10;  // MorphologyOperators.cpp:4-5

// calculate_ray_xor: (A \ B) u (B \ A) through complements and unions,
// then segments shorter than kMinSegSize dropped (removepoint).
Column ray_xor(const Column& a, const Column& b, double lo, double hi) {
  Column na = a, nb = b;
  negate_ray(na, lo, hi);
  negate_ray(nb, lo, hi);
  Column t1 = union_columns(na, b), t2 = union_columns(nb, a);
  negate_ray(t1, lo, hi);
  negate_ray(t2, lo, hi);
  Column u = union_columns(t1, t2), out;
  for (size_t i = 0; i + 1 < u.size(); i += 2) {
    if (u[i + 1] - u[i] < kMinSegSize) continue;
    out.push_back(u[i]);
    out.push_back(u[i + 1]);
  }
  return out;
}

// ---------------------------------------------------------------------------
// Mesh -> dexels (Dexelize.cpp:56-225).  For each dexel centre
// (origin_x + (x + 0.5) * spacing, origin_y + (y + 0.5) * spacing)
// (CompressedVolumeBase.cpp:5-11), the z / spacing of every triangle whose
// xy-projection contains it, sorted ascending (Dexelize.cpp:209).  Containment
// is the SOS test of point_in_triangle_2d (Dexelize.cpp:56-92); a triangle
// whose projection has zero inexact orientation is skipped (:107-121,151-158).
// The reference prunes candidates with geogram's AABB tree (:207); pruning
// never changes the result, so this restatement prunes with 8x8-dexel tiles.
// Arithmetic follows the reference's operation order exactly (no FMA).

namespace raster {

int orientation(double x1, double y1, double x2, double y2, double& twice_area) {  // Dexelize.cpp:56-67
  twice_area = y1 * x2 - x1 * y2;
  if (twice_area > 0) return 1;
  if (twice_area < 0) return -1;
  if (y2 > y1) return 1;
  if (y2 < y1) return -1;
  if (x1 > x2) return 1;
  if (x1 < x2) return -1;
  return 0;
}

bool point_in_triangle(double x0, double y0, double x1, double y1, double x2, double y2, double x3,
                       double y3, double& a, double& b, double& c) {  // Dexelize.cpp:73-92
  x1 -= x0;
  x2 -= x0;
  x3 -= x0;
  y1 -= y0;
  y2 -= y0;
  y3 -= y0;
  const int sa = orientation(x2, y2, x3, y3, a);
  if (sa == 0) return false;
  if (orientation(x3, y3, x1, y1, b) != sa) return false;
  if (orientation(x1, y1, x2, y2, c) != sa) return false;
  const double sum = a + b + c;
  a /= sum;
  b /= sum;
  c /= sum;
  return true;
}

// intersect_ray_z (Dexelize.cpp:136-162): nonzero when the ray hits
bool ray_z(const double* p1, const double* p2, const double* p3, double qx, double qy, double& z) {
  double u, v, w;
  if (!point_in_triangle(qx, qy, p1[0], p1[1], p2[0], p2[1], p3[0], p3[1], u, v, w)) return false;
  z = u * p1[2] + v * p2[2] + w * p3[2];
  const double a11 = p2[0] - p1[0], a12 = p2[1] - p1[1];  // orient_2d_inexact, :107-121
  const double a21 = p3[0] - p1[0], a22 = p3[1] - p1[1];
  const double delta = a11 * a22 - a12 * a21;
  return delta != 0;
}

}  // namespace raster

Result* rasterize(const double* V, int64_t nv, const int32_t* F, int64_t nf, int nx, int ny, double ox,
                  double oy, double spacing) {
  if (nx < 0 || ny < 0 || !(spacing > 0)) fail("rasterize: bad grid");
  for (int64_t i = 0; i < 3 * nf; ++i)
    if (F[i] < 0 || F[i] >= nv) fail("rasterize: triangle index out of range");
  for (int64_t i = 0; i < 3 * nv; ++i)
    if (!std::isfinite(V[i])) fail("rasterize: non-finite vertex");
  constexpr int kT = 8;
  const int tx = (nx + kT - 1) / kT, ty = (ny + kT - 1) / kT;
  std::vector<std::vector<int64_t>> bins(size_t(tx) * ty);
  for (int64_t f = 0; f < nf; ++f) {
    double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
    for (int k = 0; k < 3; ++k)
      for (int d = 0; d < 2; ++d) {
        lo[d] = std::min(lo[d], V[3 * F[3 * f + k] + d]);
        hi[d] = std::max(hi[d], V[3 * F[3 * f + k] + d]);
      }
    // dexel index range whose centres may lie in the closed bbox, widened by one
    const double o[2] = {ox, oy};
    int ilo[2], ihi[2];
    const int n[2] = {nx, ny};
    for (int d = 0; d < 2; ++d) {
      const double a = std::floor((lo[d] - o[d]) / spacing - 0.5) - 1.0;
      const double b = std::ceil((hi[d] - o[d]) / spacing - 0.5) + 1.0;
      ilo[d] = int(std::min(std::max(a, 0.0), double(n[d])));
      ihi[d] = int(std::max(std::min(b, double(n[d] - 1)), -1.0));
    }
    if (ilo[0] > ihi[0] || ilo[1] > ihi[1]) continue;
    for (int b = ilo[1] / kT; b <= ihi[1] / kT; ++b)
      for (int a = ilo[0] / kT; a <= ihi[0] / kT; ++a) bins[size_t(b) * tx + a].push_back(f);
  }
  auto* res = new Result;
  res->nx = nx;
  res->ny = ny;
  std::vector<std::vector<double>> cols(size_t(nx) * ny);
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      const double qx = (x + 0.5) * spacing + ox, qy = (y + 0.5) * spacing + oy;
      auto& col = cols[size_t(x) + size_t(nx) * y];
      for (int64_t f : bins[size_t(y / kT) * tx + x / kT]) {
        double z;
        if (raster::ray_z(V + 3 * F[3 * f], V + 3 * F[3 * f + 1], V + 3 * F[3 * f + 2], qx, qy, z))
          col.push_back(z / spacing);
      }
      std::sort(col.begin(), col.end());
    }
  res->off.assign(cols.size() + 1, 0);
  for (size_t c = 0; c < cols.size(); ++c) {
    res->off[c + 1] = res->off[c] + int64_t(cols[c].size());
    res->a.insert(res->a.end(), cols[c].begin(), cols[c].end());
  }
  return res;
}

extern "C" {

const char* vo_oracle_last_error() { return g_error.c_str(); }

// Morphology op on a CSR volume (endpoints in absolute dexel z units).
// method: 0 = ours, 1 = brute_force; op: 0 noop, 1 dilation, 2 erosion,
// 3 closing, 4 opening.  Returns a result handle, or null on error.
void* vo_oracle_morph(int method, int op, int nx, int ny, const int64_t* off, const double* z,
                      double radius, double zlo, double zhi, int threads, double* t1, double* t2) {
  try {
    Volume in = from_csr(nx, ny, off, z);
    return to_result(run_op(method, op, in, radius, zlo, zhi, threads, t1, t2));
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}

// Pass-1 intermediate of `ours` (the union of both half sweeps): triples
// (y1, y2, rho) per column.
void* vo_oracle_pass1(int nx, int ny, const int64_t* off, const double* z, double radius,
                      int threads) {
  try {
    RVolume mid = ours_pass1(from_csr(nx, ny, off, z), radius, threads);
    auto* res = new Result;
    res->nx = nx;
    res->ny = ny;
    res->off.assign(mid.cols.size() + 1, 0);
    for (size_t c = 0; c < mid.cols.size(); ++c) {
      res->off[c + 1] = res->off[c] + int64_t(mid.cols[c].size());
      for (const RSeg& s : mid.cols[c]) {
        res->a.push_back(s.y1);
        res->b.push_back(s.y2);
        res->r.push_back(s.r);
      }
    }
    return res;
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}

// calculateXor per column (the result volume; its measure is the caller's).
void* vo_oracle_xor(int nx, int ny, const int64_t* off_a, const double* z_a, const int64_t* off_b,
                    const double* z_b, double lo, double hi) {
  try {
    Volume a = from_csr(nx, ny, off_a, z_a), b = from_csr(nx, ny, off_b, z_b), r;
    r.nx = nx;
    r.ny = ny;
    r.cols.resize(a.cols.size());
    for (size_t c = 0; c < a.cols.size(); ++c) r.cols[c] = ray_xor(a.cols[c], b.cols[c], lo, hi);
    return to_result(r);
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}

// Mesh -> dexels: V is nv x 3 (x, y, z), F is nf x 3 vertex indices; the
// result holds, per column x + nx*y, the sorted z / spacing of the hits.
void* vo_oracle_rasterize(const double* V, int64_t nv, const int32_t* F, int64_t nf, int nx, int ny,
                          double ox, double oy, double spacing) {
  try {
    return rasterize(V, nv, F, nf, nx, ny, ox, oy, spacing);
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}

int64_t vo_oracle_count(void* h) { return static_cast<Result*>(h)->off.back(); }

void vo_oracle_copy(void* h, int64_t* off, double* a, double* b, double* r) {
  auto* res = static_cast<Result*>(h);
  std::memcpy(off, res->off.data(), res->off.size() * sizeof(int64_t));
  if (a) std::memcpy(a, res->a.data(), res->a.size() * sizeof(double));
  if (b && !res->b.empty()) std::memcpy(b, res->b.data(), res->b.size() * sizeof(double));
  if (r && !res->r.empty()) std::memcpy(r, res->r.data(), res->r.size() * sizeof(double));
}

void vo_oracle_free(void* h) { delete static_cast<Result*>(h); }

// 1-D primitives on one column, in place; `n` is the current length and
// `cap` the buffer capacity.  Returns the new length, or -1 if it exceeds cap.
int vo_oracle_negate_ray(double* c, int n, int cap, double lo, double hi) {
  Column v(c, c + n);
  negate_ray(v, lo, hi);
  if (int(v.size()) > cap) return -1;
  std::copy(v.begin(), v.end(), c);
  return int(v.size());
}

int vo_oracle_negate_ray_range(double* c, int n, int cap, double lo, double hi) {
  Column v(c, c + n);
  negate_ray_range(v, lo, hi);
  if (int(v.size()) > cap) return -1;
  std::copy(v.begin(), v.end(), c);
  return int(v.size());
}

int vo_oracle_union(const double* a, int na, const double* b, int nb, double* out, int cap) {
  Column u = union_columns(Column(a, a + na), Column(b, b + nb));
  if (int(u.size()) > cap) return -1;
  std::copy(u.begin(), u.end(), out);
  return int(u.size());
}

}  // extern "C"
