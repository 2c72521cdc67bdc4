// sdca_format.hpp -- number formatting of the coupling writers (libsdca.so).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>

namespace sdca {

// Appends "%.8f" of v (v = double(float) * 0.5 is exact times 1e8 in FP64, so the
// rounded integer equals what printf's exact conversion produces).
inline char *put_fixed8(char *p, double v)
{
    if (!(v >= 0.0) || v >= 9.0e9) return p + sprintf(p, "%.8f", v);
    const double scaled = std::nearbyint(v * 1e8);
    uint64_t u = (uint64_t)scaled;
    const uint64_t ip = u / 100000000ull;
    uint64_t fp = u % 100000000ull;
    char tmp[24];
    int n = 0;
    uint64_t t = ip;
    do {
        tmp[n++] = (char)('0' + t % 10);
        t /= 10;
    } while (t);
    while (n) *p++ = tmp[--n];
    *p++ = '.';
    for (int k = 7; k >= 0; --k) {
        p[k] = (char)('0' + fp % 10);
        fp /= 10;
    }
    return p + 8;
}

inline char *put_uint(char *p, uint64_t v)
{
    char tmp[24];
    int n = 0;
    do {
        tmp[n++] = (char)('0' + v % 10);
        v /= 10;
    } while (v);
    while (n) *p++ = tmp[--n];
    return p;
}

}  // namespace sdca
