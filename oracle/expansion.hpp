// ORACLE — test infrastructure only. Nothing in the product path may include,
// link or call this file; only tests/, the smoke check and the CPU-baseline leg
// of bench.py use the oracle library built from it.
//
// Exact arithmetic on binary64 inputs with Shewchuk's floating-point expansions
// (J. R. Shewchuk, "Adaptive Precision Floating-Point Arithmetic and Fast Robust
// Geometric Predicates", Discrete & Computational Geometry 18:305-363, 1997).
// An expansion is a sum of non-overlapping doubles ordered by increasing
// magnitude; sums and products of expansions are exact as long as no
// intermediate overflows or underflows (true for the scenes used here: scene
// coordinates are far from 1e±150).
//
// The reference gets exactness from CGAL's Exact_rational (src/geometry.h:14-19,
// interval filter then exact sign in src/bvh_peel.cpp:60-79). This header is the
// same idea, written from the published expansion algorithms instead of a
// rational type: every predicate the oracle needs is a polynomial in the
// binary64 inputs, so expansions evaluate its sign exactly without division.
#pragma once
#include <cmath>
#include <vector>

namespace oracle {

struct Exp {
  std::vector<double> c;  // non-overlapping components, increasing magnitude
  Exp() {}
  explicit Exp(double x) {
    if (x != 0.0) c.push_back(x);
  }
  int sign() const {
    if (c.empty()) return 0;
    return c.back() > 0 ? 1 : (c.back() < 0 ? -1 : 0);
  }
  double approx() const {
    double s = 0;
    for (double x : c) s += x;
    return s;
  }
};

// Knuth's TwoSum: a + b = s + e exactly.
inline void two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  double bv = s - a;
  double av = s - bv;
  e = (a - av) + (b - bv);
}

// a * b = p + e exactly (the FMA computes the rounding error of the product).
inline void two_prod(double a, double b, double &p, double &e) {
  p = a * b;
  e = std::fma(a, b, -p);
}

// Shewchuk's GROW-EXPANSION with zero elimination: e + b.
inline Exp grow(const Exp &e, double b) {
  Exp h;
  h.c.reserve(e.c.size() + 1);
  double q = b;
  for (double x : e.c) {
    double s, err;
    two_sum(q, x, s, err);
    q = s;
    if (err != 0.0) h.c.push_back(err);
  }
  if (q != 0.0 || h.c.empty()) {
    if (q != 0.0) h.c.push_back(q);
  }
  return h;
}

// EXPANSION-SUM: add every component of f to e (O(m n), simple and exact).
inline Exp add(const Exp &e, const Exp &f) {
  if (e.c.size() < f.c.size()) return add(f, e);
  Exp h = e;
  for (double x : f.c) h = grow(h, x);
  return h;
}

inline Exp neg(const Exp &e) {
  Exp h = e;
  for (double &x : h.c) x = -x;
  return h;
}

inline Exp sub(const Exp &e, const Exp &f) { return add(e, neg(f)); }

// SCALE-EXPANSION with zero elimination: e * b.
inline Exp scale(const Exp &e, double b) {
  Exp h;
  if (e.c.empty() || b == 0.0) return h;
  h.c.reserve(2 * e.c.size());
  double q, hh;
  two_prod(e.c[0], b, q, hh);
  if (hh != 0.0) h.c.push_back(hh);
  for (size_t i = 1; i < e.c.size(); ++i) {
    double t1, t0;
    two_prod(e.c[i], b, t1, t0);
    double s, err;
    two_sum(q, t0, s, err);
    if (err != 0.0) h.c.push_back(err);
    double s2, err2;
    // FAST-TWO-SUM is valid here because |t1| >= |s| (Shewchuk, Thm. 19).
    s2 = t1 + s;
    err2 = s - (s2 - t1);
    if (err2 != 0.0) h.c.push_back(err2);
    q = s2;
  }
  if (q != 0.0) h.c.push_back(q);
  return h;
}

// Product of two expansions: sum of e scaled by each component of f.
inline Exp mul(const Exp &e, const Exp &f) {
  if (e.c.size() < f.c.size()) return mul(f, e);
  Exp h;
  for (double x : f.c) h = add(h, scale(e, x));
  return h;
}

inline Exp two(double a, double b) { return add(Exp(a), Exp(b)); }

}  // namespace oracle
