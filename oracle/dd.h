// Double-double arithmetic for the oracle (test infrastructure only).
//
// Elemental promotes double to QD's dd_real (include/El/core/Element/decl.hpp:116-153,
// include/El/core/imports/qd.hpp:24) for its high-precision residuals. QD is a
// third-party library that is not vendored with the reference, so this header restates
// the published error-free transformations instead:
//   * TwoSum   -- Knuth, TAOCP vol. 2 (s + e == a + b exactly);
//   * TwoProd  -- Dekker 1971, with the product error taken from one FMA;
//   * the "sloppy" double-double add/sub of Hida, Li, Bailey (QD 2.3, 2001), which QD
//     uses by default, and which is accurate to about 2^-104 relative for same-sign data.
// Accuracy of this header is pinned in tests against Python's exact `fractions`.
#pragma once
#include <cmath>

namespace orc {

struct dd {
    double hi, lo;
};

static inline dd two_sum(double a, double b) {
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}

static inline dd quick_two_sum(double a, double b) {
    double s = a + b;
    double e = b - (s - a);
    return {s, e};
}

static inline dd two_prod(double a, double b) {
    double p = a * b;
    double e = std::fma(a, b, -p);
    return {p, e};
}

// Accurate (IEEE-style) double-double addition: both components added with TwoSum.
static inline dd dd_add(dd a, dd b) {
    dd s = two_sum(a.hi, b.hi);
    dd t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return quick_two_sum(s.hi, s.lo);
}

static inline dd dd_neg(dd a) { return {-a.hi, -a.lo}; }

static inline dd dd_sub(dd a, dd b) { return dd_add(a, dd_neg(b)); }

static inline dd dd_from(double a) { return {a, 0.0}; }

}  // namespace orc
