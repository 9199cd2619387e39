// Compares mpsgpu::fmtFixed6 / fmtInt with snprintf("%f") / ("%d") on random
// bit patterns of doubles and floats, a uniform range, exact dyadic values
// (which hit the round-half-to-even ties) and special values.  Prints the
// first mismatches and exits with 1 if there is any.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

#include "fmt_fixed.h"

int main(int argc, char** argv) {
    const long samples = argc > 1 ? atol(argv[1]) : 1000000;
    std::mt19937_64 rng(12345);
    char a[400], b[400];
    long bad = 0, n = 0;
    auto chk = [&](double x) {
        const int k = mpsgpu::fmtFixed6(a, x);
        a[k] = 0;
        snprintf(b, sizeof b, "%f", x);
        n++;
        if (strcmp(a, b) != 0 && bad++ < 10) printf("x=%a fmtFixed6=%s printf=%s\n", x, a, b);
    };
    for (long i = 0; i < samples; i++) {
        const uint64_t u = rng();
        double x;
        memcpy(&x, &u, 8);
        chk(x);
    }
    for (long i = 0; i < samples; i++) {
        const uint32_t u = (uint32_t)rng();
        float f;
        memcpy(&f, &u, 4);
        chk((double)f);
    }
    std::uniform_real_distribution<double> U(-2e4, 2e4);
    for (long i = 0; i < samples; i++) chk(U(rng));
    for (int m = 0; m < 30; m++)
        for (long j = -3000; j <= 3000; j++) chk(std::ldexp((double)j, -m));
    const double special[] = {0.0, -0.0, 1e-300, -1e-300, 5e-324, -5e-324, 5e-7, 1.5e-6, 2.5e-6, 999999.9999995,
                              1e12, -1e12, 1e12 - 1e-4, 9.99999999999e11, 1e300, INFINITY, -INFINITY, NAN};
    for (double x : special) chk(x);
    for (long j = -1000000; j <= 1000000; j++) {
        const int k = mpsgpu::fmtInt(a, (int)j);
        a[k] = 0;
        snprintf(b, sizeof b, "%d", (int)j);
        if (strcmp(a, b) != 0 && bad++ < 10) printf("%ld: fmtInt=%s\n", j, a);
    }
    const int ext[] = {2147483647, -2147483647 - 1};
    for (int v : ext) {
        const int k = mpsgpu::fmtInt(a, v);
        a[k] = 0;
        snprintf(b, sizeof b, "%d", v);
        if (strcmp(a, b) != 0 && bad++ < 10) printf("%d: fmtInt=%s\n", v, a);
    }
    printf("checked %ld values, %ld mismatches\n", n, bad);
    return bad != 0;
}
