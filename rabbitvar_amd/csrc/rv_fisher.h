// htslib's kt_fisher_exact (kfunc.c: Lanczos log-gamma, hypergeometric accumulator, two-sided
// tail rule), restated for the device scoring kernels and the host's somatic printer.
//
// The method, its constants and its tail rule come from htslib's kfunc.c, which carries this
// notice:
//
//   The MIT License
//
//   Copyright (C) 2010, 2013-2019 Genome Research Ltd.
//   Copyright (C) 2011 Attractive Chaos <attractor@live.co.uk>
//
//   Permission is hereby granted, free of charge, to any person obtaining a copy of this software
//   and associated documentation files (the "Software"), to deal in the Software without
//   restriction, including without limitation the rights to use, copy, modify, merge, publish,
//   distribute, sublicense, and/or sell copies of the Software, and to permit persons to whom the
//   Software is furnished to do so, subject to the following conditions:
//
//   The above copyright notice and this permission notice shall be included in all copies or
//   substantial portions of the Software.
//
//   THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS OR IMPLIED, INCLUDING
//   BUT NOT LIMITED TO THE WARRANTIES OF MERCHANTABILITY, FITNESS FOR A PARTICULAR PURPOSE AND
//   NONINFRINGEMENT. IN NO EVENT SHALL THE AUTHORS OR COPYRIGHT HOLDERS BE LIABLE FOR ANY CLAIM,
//   DAMAGES OR OTHER LIABILITY, WHETHER IN AN ACTION OF CONTRACT, TORT OR OTHERWISE, ARISING FROM,
//   OUT OF OR IN CONNECTION WITH THE SOFTWARE OR THE USE OR OTHER DEALINGS IN THE SOFTWARE.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "rv_common.h"

namespace rv {

RV_HD double kf_lgamma_d(double z) {
    double x = 0;
    x += 0.1659470187408462e-06 / (z + 7);
    x += 0.9934937113930748e-05 / (z + 6);
    x -= 0.1385710331296526 / (z + 5);
    x += 12.50734324009056 / (z + 4);
    x -= 176.6150291498386 / (z + 3);
    x += 771.3234287757674 / (z + 2);
    x -= 1259.139216722289 / (z + 1);
    x += 676.5203681218835 / z;
    x += 0.9999999999995183;
    return log(x) - 5.58106146679532777 - z + (z - 0.5) * log(z + 6.5);
}
// Log-gamma of the integers 1..n-1 as the host's C library computes kf_lgamma_d (table[m] =
// kf_lgamma_d(m), table[0] unused).  The device's log differs from the host's by an ulp now and
// then; kf_lgamma multiplies log(z + 6.5) by z - 0.5, so at a table total of a few thousand such
// an ulp moves a p-value by ~1e-11 relative.  Scoring kernels look the values up in this
// host-built table instead, so their p-values agree with the host's to the last few ulps (only
// exp and IEEE arithmetic remain, and exp's error is not amplified).
constexpr int FISHER_LGT_N = 1 << 17;
inline void fisher_lgamma_table(double* table, int n) {
    if (n > 0) table[0] = 0;
    for (int m = 1; m < n; m++) table[m] = kf_lgamma_d((double)m);
}
struct LGTable {
    const double* t;  // nullptr: compute
    int n;
};
RV_HD double lgamma_int(int m, LGTable lg) { return (lg.t && m < lg.n) ? lg.t[m] : kf_lgamma_d(m); }
RV_HD double lbinom_d(int n, int k, LGTable lg = LGTable{nullptr, 0}) {
    if (k == 0 || n == k) return 0;
    return lgamma_int(n + 1, lg) - lgamma_int(k + 1, lg) - lgamma_int(n - k + 1, lg);
}
struct HG {
    int n11, n1_, n_1, n;
    double p;
    LGTable lg;
};
RV_HD double hg_acc(int n11, int n1_, int n_1, int n, HG& a) {
    if (n1_ || n_1 || n) {
        a.n11 = n11;
        a.n1_ = n1_;
        a.n_1 = n_1;
        a.n = n;
    } else {
        if (n11 % 11 && n11 + a.n - a.n1_ - a.n_1) {
            if (n11 == a.n11 + 1) {
                a.p *= (double)(a.n1_ - a.n11) / n11 * (a.n_1 - a.n11) / (n11 + a.n - a.n1_ - a.n_1);
                a.n11 = n11;
                return a.p;
            }
            if (n11 == a.n11 - 1) {
                a.p *= (double)a.n11 / (a.n1_ - n11) * (a.n11 + a.n - a.n1_ - a.n_1) / (a.n_1 - n11);
                a.n11 = n11;
                return a.p;
            }
        }
        a.n11 = n11;
    }
    a.p = exp(lbinom_d(a.n1_, a.n11, a.lg) + lbinom_d(a.n - a.n1_, a.n_1 - a.n11, a.lg) - lbinom_d(a.n, a.n_1, a.lg));
    return a.p;
}
// two-sided p of the 2x2 table (n11 n12 / n21 n22); left/right tails optional
RV_HD double fisher_two_sided(int n11, int n12, int n21, int n22, double* pl, double* pr,
                              LGTable lg = LGTable{nullptr, 0}) {
    int i, j, max, min;
    double p, q, left, right, two;
    HG aux{0, 0, 0, 0, 0, lg};
    int n1_ = n11 + n12, n_1 = n11 + n21, n = n11 + n12 + n21 + n22;
    max = (n_1 < n1_) ? n_1 : n1_;
    min = n1_ + n_1 - n;
    if (min < 0) min = 0;
    if (pl) *pl = 1.;
    if (pr) *pr = 1.;
    if (min == max) return 1.;
    q = hg_acc(n11, n1_, n_1, n, aux);
    p = hg_acc(min, 0, 0, 0, aux);
    for (left = 0., i = min + 1; p < 0.99999999 * q && i <= max; ++i) left += p, p = hg_acc(i, 0, 0, 0, aux);
    --i;
    if (p < 1.00000001 * q)
        left += p;
    else
        --i;
    p = hg_acc(max, 0, 0, 0, aux);
    for (right = 0., j = max - 1; p < 0.99999999 * q && j >= 0; --j) right += p, p = hg_acc(j, 0, 0, 0, aux);
    ++j;
    if (p < 1.00000001 * q)
        right += p;
    else
        ++j;
    two = left + right;
    if (two > 1.) two = 1.;
    int di = i - n11 < 0 ? n11 - i : i - n11;
    int dj = j - n11 < 0 ? n11 - j : j - n11;
    if (di < dj)
        right = 1. - left + q;
    else
        left = 1.0 - right + q;
    if (pl) *pl = left;
    if (pr) *pr = right;
    return two;
}

}  // namespace rv
