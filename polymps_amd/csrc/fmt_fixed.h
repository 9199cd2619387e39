// Text of printf's "%f" and "%d" conversions without printf.
//
// fmtFixed6(x) writes what printf("%f", x) writes with the C library of this
// system: the exact binary value of x rounded to six decimals, ties to even,
// a '-' whenever the sign bit is set (also for -0.0 and for negative values
// that round to zero).  Values of magnitude 1e12 and above, infinities and
// NaNs are handed to snprintf.  fmtInt(v) writes printf("%d", v).
//
// The VTU writer formats hundreds of millions of numbers per file; snprintf
// spends most of its time parsing the format and in its general big-number
// path, which six fixed decimals of values below 1e12 never need.
#ifndef MPS_FMT_FIXED_H_
#define MPS_FMT_FIXED_H_

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

namespace mpsgpu {

inline int fmtUint(char* b, uint64_t u) {
    char t[24];
    int n = 0;
    do {
        t[n++] = (char)('0' + (int)(u % 10));
        u /= 10;
    } while (u);
    for (int k = 0; k < n; k++) b[k] = t[n - 1 - k];
    return n;
}

inline int fmtInt(char* b, int v) {
    if (v < 0) {
        b[0] = '-';
        return 1 + fmtUint(b + 1, (uint64_t)(-(int64_t)v));
    }
    return fmtUint(b, (uint64_t)v);
}

// b must hold at least 330 characters (the longest snprintf("%f") result).
inline int fmtFixed6(char* b, double x) {
    uint64_t bits;
    memcpy(&bits, &x, sizeof bits);
    const int bexp = (int)((bits >> 52) & 0x7ff);
    if (bexp == 0x7ff || std::fabs(x) >= 1e12) return snprintf(b, 330, "%f", x);
    uint64_t m = bits & ((1ull << 52) - 1);
    int s;   // |x| = m / 2^s, s > 0 since |x| < 1e12 < 2^52
    if (bexp == 0) {
        s = 1074;
    } else {
        m |= 1ull << 52;
        s = 1075 - bexp;
    }
    // q = round_half_even(|x| * 10^6) from the exact product m * 10^6 < 2^73
    uint64_t q = 0;
    if (s < 100) {
        const unsigned __int128 t = (unsigned __int128)m * 1000000u;
        const unsigned __int128 one = 1;
        const unsigned __int128 qq = t >> s;
        const unsigned __int128 rem = t & ((one << s) - 1);
        const unsigned __int128 half = one << (s - 1);
        q = (uint64_t)qq;
        if (rem > half || (rem == half && (q & 1))) q++;
    }   // else |x| * 10^6 < 2^73 / 2^100 < 1/2: q = 0
    int k = 0;
    if (bits >> 63) b[k++] = '-';
    k += fmtUint(b + k, q / 1000000u);
    b[k++] = '.';
    uint32_t f = (uint32_t)(q % 1000000u);
    for (int d = 5; d >= 0; d--) {
        b[k + d] = (char)('0' + f % 10);
        f /= 10;
    }
    return k + 6;
}

}  // namespace mpsgpu

#endif  // MPS_FMT_FIXED_H_
