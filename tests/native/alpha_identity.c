/* Exhaustive check of the frame-cache sweep's lerp weight: for every pair of
 * integers 0 < a < b < 2^16 (a frames into a keyframe gap of b frames), with
 * y = RN32(1/b), q = RN32(a y), r = fma(-b, q, a) and w = fma(r, y, q)
 * (Markstein's correction in float arithmetic), w must equal RN32(a / b) bit
 * for bit.  The device forms y from its reciprocal instruction and checks it
 * against RN32(1/b) itself.  Compiled and run by tests/test_cpu_units.py. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

int main(void)
{
    uint64_t bad = 0, checked = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : bad, checked)
    for (int b = 2; b < 65536; ++b) {
        const float fb = (float)b;
        const float y = 1.0f / fb;
        for (int a = 1; a < b; ++a) {
            const float fa = (float)a;
            const float want = fa / fb;
            const float q = fa * y;
            const float r = fmaf(-fb, q, fa);
            const float w = fmaf(r, y, q);
            uint32_t wb, gb;
            memcpy(&wb, &want, 4);
            memcpy(&gb, &w, 4);
            bad += wb != gb;
            ++checked;
        }
    }
    printf("%llu %llu\n", (unsigned long long)checked, (unsigned long long)bad);
    return bad != 0;
}
