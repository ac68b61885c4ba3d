// ddpow.hpp - pow(x, y) for finite x > 0, evaluated in double-double
// arithmetic (about 100 significant bits) and rounded once, so the result is
// the correctly rounded power except within ~1e-30 relative of a rounding
// midpoint. The time-step controller needs the power bit for bit as the
// reference's libm computes it (f64::powf in analysis.rs:530-536 and :173);
// a correctly rounded power equals that libm's result wherever the libm is
// correctly rounded, and it is the same on the host and on the device, which
// the device's own pow is not.
//
// Method: log by the atanh series log m = 2 atanh((m-1)/(m+1)) after
// reducing x to 2^e m with m in [1/sqrt 2, sqrt 2); exp by k ln 2 + r, r
// scaled by 2^-10, a Taylor series for expm1 and ten expm1 doublings
// (e^2r - 1 = 2u + u^2). Double-double operations follow Dekker's and
// Knuth's error-free transformations (two-sum; two-product with fma).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace rsgpu {

struct DD {
    double hi, lo;
};

__host__ __device__ inline DD dd_two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
__host__ __device__ inline DD dd_quick(double a, double b) {
    const double s = a + b;
    return {s, b - (s - a)};
}
__host__ __device__ inline DD dd_two_prod(double a, double b) {
    const double p = a * b;
    return {p, fma(a, b, -p)};
}
__host__ __device__ inline DD dd_add(DD x, DD y) {
    DD s = dd_two_sum(x.hi, y.hi);
    const DD t = dd_two_sum(x.lo, y.lo);
    s.lo += t.hi;
    s = dd_quick(s.hi, s.lo);
    s.lo += t.lo;
    return dd_quick(s.hi, s.lo);
}
__host__ __device__ inline DD dd_neg(DD x) { return {-x.hi, -x.lo}; }
__host__ __device__ inline DD dd_mul(DD x, DD y) {
    DD p = dd_two_prod(x.hi, y.hi);
    p.lo += x.hi * y.lo + x.lo * y.hi;
    return dd_quick(p.hi, p.lo);
}
__host__ __device__ inline DD dd_mul_d(DD x, double d) {
    DD p = dd_two_prod(x.hi, d);
    p.lo += x.lo * d;
    return dd_quick(p.hi, p.lo);
}
__host__ __device__ inline DD dd_div(DD x, DD y) {
    const double q1 = x.hi / y.hi;
    DD r = dd_add(x, dd_neg(dd_mul_d(y, q1)));
    const double q2 = r.hi / y.hi;
    r = dd_add(r, dd_neg(dd_mul_d(y, q2)));
    const double q3 = r.hi / y.hi;
    const DD q = dd_quick(q1, q2);
    return dd_add(q, DD{q3, 0.0});
}
__host__ __device__ inline DD dd_div_d(DD x, double d) { return dd_div(x, DD{d, 0.0}); }

// ln 2 to double-double precision
__host__ __device__ inline DD dd_ln2() { return {0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56}; }

// natural log of a finite x > 0
__host__ __device__ inline DD dd_log(double x) {
    int e = 0;
    double m = frexp(x, &e);  // m in [0.5, 1)
    if (m < 0x1.6a09e667f3bcdp-1) {  // 1/sqrt(2)
        m *= 2.0;
        e -= 1;
    }
    const DD f = dd_div(DD{m - 1.0, 0.0}, dd_two_sum(m, 1.0));  // m - 1 is exact here
    const DD f2 = dd_mul(f, f);
    DD term = f, sum = f;
    for (int k = 1; k <= 22; ++k) {
        term = dd_mul(term, f2);
        sum = dd_add(sum, dd_div_d(term, (double)(2 * k + 1)));
    }
    const DD lm = dd_mul_d(sum, 2.0);
    return dd_add(dd_mul_d(dd_ln2(), (double)e), lm);
}

// e^z for |z| well inside the double range; returns the double-double of
// e^(z - k ln 2) and k
__host__ __device__ inline DD dd_exp_red(DD z, int* kout) {
    const double k = rint(z.hi / 0x1.62e42fefa39efp-1);
    DD r = dd_add(z, dd_neg(dd_mul_d(dd_ln2(), k)));
    r.hi *= 0x1p-10;
    r.lo *= 0x1p-10;
    // expm1(r) by Horner's rule: r (1 + r/2 (1 + r/3 (... (1 + r/10))))
    DD u = DD{1.0, 0.0};
    for (int j = 10; j >= 2; --j) u = dd_add(DD{1.0, 0.0}, dd_div_d(dd_mul(r, u), (double)j));
    u = dd_mul(r, u);
    for (int s = 0; s < 10; ++s) u = dd_add(dd_mul_d(u, 2.0), dd_mul(u, u));
    *kout = (int)k;
    return dd_add(DD{1.0, 0.0}, u);
}

// x^y for finite x > 0 and finite y, rounded once. *near (optional) is set
// when the exact power lies within `zone` ulp of a rounding midpoint: there a
// libm whose error bound exceeds half an ulp by less than `zone` may round the
// other way, so its result is not fixed by the exact value.
__host__ __device__ inline double pow_dd(double x, double y, bool* near = nullptr, double zone = 0.05) {
    if (near) *near = false;
    if (x == 1.0 || y == 0.0) return 1.0;
    const DD z = dd_mul_d(dd_log(x), y);
    int k = 0;
    const DD e = dd_exp_red(z, &k);
    const double r = e.hi + e.lo;  // e is normalized: r == e.hi
    if (near && e.lo != 0.0) {
        int ex = 0;
        (void)frexp(r, &ex);
        const double ulp = ldexp(1.0, ex - 53);
        if (0.5 * ulp - fabs(e.lo) < zone * ulp) *near = true;
    }
    return ldexp(r, k);
}

}  // namespace rsgpu
