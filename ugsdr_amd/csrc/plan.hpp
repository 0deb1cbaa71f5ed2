// Host-side FFT planning shared with the kernels (plain structs passed by value).
#pragma once

#include <cstdlib>

#include "dev_fft.hpp"

namespace ugsdr_hip {

using ugsdr_dev::FourStep;
using ugsdr_dev::RadixPlan;

struct RowStat {
    float level;   // max of the row's surface (scaled by 1/N)
    int index;     // first index of the max
    float mean;
    float stddev;  // sample sigma (n-1)
    float second;  // runner-up value of the row
};

// Factor n into the radices the device stages implement.  Largest radix first
// (the first stage needs no twiddles).  Returns false if a prime factor is not
// supported or more than 8 stages would be needed.
inline bool make_plan(int n, RadixPlan& p) {
    if (n < 1 || n > ugsdr_dev::kLdsCap) return false;
    p.n = n;
    p.nst = 0;
    int m = n;
    const int odd[] = {31, 19, 13, 11, 7, 5, 3};
    int tmp[16];
    int cnt = 0;
    for (int r : odd)
        while (m % r == 0) {
            if (cnt >= 16) return false;
            tmp[cnt++] = r;
            m /= r;
        }
    int twos = 0;
    while (m % 2 == 0) { ++twos; m /= 2; }
    if (m != 1) return false;
    while (twos >= 4) { tmp[cnt++] = 16; twos -= 4; }
    if (twos == 3) tmp[cnt++] = 8;
    else if (twos == 2) tmp[cnt++] = 4;
    else if (twos == 1) tmp[cnt++] = 2;
    if (cnt > 8) return false;
    // descending order
    for (int i = 0; i < cnt; ++i)
        for (int j = i + 1; j < cnt; ++j)
            if (tmp[j] > tmp[i]) { int t = tmp[i]; tmp[i] = tmp[j]; tmp[j] = t; }
    for (int i = 0; i < cnt; ++i) p.radix[i] = tmp[i];
    p.nst = cnt;
    if (n == 1) { p.nst = 0; }
    return true;
}

// n = n1 * n2 with both factors plannable, as balanced as possible
inline bool make_fourstep(int n, FourStep& f) {
    int best = -1;
    for (int a = 2; (long)a * a <= n; ++a) {
        if (n % a) continue;
        const int b = n / a;
        if (b > ugsdr_dev::kLdsCap) continue;
        RadixPlan pa, pb;
        if (!make_plan(a, pa) || !make_plan(b, pb)) continue;
        best = a;  // keep the largest a <= sqrt(n)
    }
    if (best < 0) return false;
    f.n = n;
    f.n1 = best;
    f.n2 = n / best;
    make_plan(f.n1, f.p1);
    make_plan(f.n2, f.p2);
    return true;
}

}  // namespace ugsdr_hip
