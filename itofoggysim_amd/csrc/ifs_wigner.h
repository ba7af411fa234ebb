// Wigner d^n_{km}(beta) by the three-term recurrence in n (device functions),
// shared by the table generators and the far-field kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace ifs {

// d^{n0}_{km}(beta) at n0 = max(|k|,|m|) from the closed forms
//   d^j_{j,m}  = sqrt((2j)!/((j+m)!(j-m)!)) c^{j+m} (-s)^{j-m}
//   d^j_{-j,m} = sqrt((2j)!/((j+m)!(j-m)!)) c^{j-m} s^{j+m}
//   d^j_{k,m}  = (-1)^{k-m} d^j_{m,k}
// with c = cos(beta/2), s = sin(beta/2), evaluated in logs.
__device__ inline double wigner_seed(int k, int m, double c, double s) {
    double sign = 1.0;
    if ((k < 0 ? -k : k) < (m < 0 ? -m : m)) {
        int t = k;
        k = m;
        m = t;
        if ((k - m) & 1) sign = -sign;
    }
    int j = k < 0 ? -k : k;
    int ec, es;
    if (k >= 0) {
        ec = j + m;
        es = j - m;
        if (es & 1) sign = -sign;
    } else {
        ec = j - m;
        es = j + m;
    }
    double lv = 0.5 * (lgamma(2.0 * j + 1.0) - lgamma((double)(j + m) + 1.0) - lgamma((double)(j - m) + 1.0));
    if (ec > 0) {
        if (c <= 0.0) return 0.0;
        lv += ec * log(c);
    }
    if (es > 0) {
        if (s <= 0.0) return 0.0;
        lv += es * log(s);
    }
    return sign * exp(lv);
}

// One step of the three-term recurrence in n for d^n_{km}, from d^{n-1} (d1) and d^{n-2} (d2).
__device__ inline double wigner_step(int n, int k, int m, double cb, double d1, double d2) {
    double nn = (double)n;
    double k2 = (double)k * k, m2 = (double)m * m;
    double num = (2.0 * nn - 1.0) * (nn * (nn - 1.0) * cb - (double)k * (double)m) * d1 -
                 nn * sqrt(((nn - 1.0) * (nn - 1.0) - k2) * ((nn - 1.0) * (nn - 1.0) - m2)) * d2;
    double den = (nn - 1.0) * sqrt((nn * nn - k2) * (nn * nn - m2));
    return num / den;
}

}  // namespace ifs
