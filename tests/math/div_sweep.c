/* wam_div_1p1 (wam_math.h) against IEEE division by 1.1f on every float
   (stride 1) or on a strided sweep: prints {"n": swept, "ne": differences,
   "first": first differing argument's bits or -1}.  Built by
   tests/test_math_cr.py:

       gcc -O2 -mfma -ffp-contract=off -fopenmp -Iwam_amd/csrc tests/math/div_sweep.c -o div_sweep && ./div_sweep 1 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "wam_math.h"

int main(int argc, char **argv)
{
    const long long stride = argc > 1 ? atoll(argv[1]) : 1;
    long long n = 0, ne = 0, first = -1;
#pragma omp parallel for reduction(+ : n, ne) schedule(static)
    for (long long b = 0; b <= 0xffffffffLL; b += stride) {
        const uint32_t u = (uint32_t)b;
        float x;
        memcpy(&x, &u, 4);
        const float a = wam_div_1p1(x), q = x / 1.1f;
        uint32_t ua, uq;
        memcpy(&ua, &a, 4);
        memcpy(&uq, &q, 4);
        n++;
        if (ua != uq && !(x != x)) {
            ne++;
#pragma omp critical
            if (first < 0 || (long long)u < first) first = (long long)u;
        }
    }
    printf("{\"n\": %lld, \"ne\": %lld, \"first\": %lld}\n", n, ne, first);
    return 0;
}
