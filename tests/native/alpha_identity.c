/* Exhaustive check of the frame-cache sweep's lerp weight: for every pair of
 * integers 0 < a < b < 2^16 (b - a frames into a keyframe gap of b frames),
 * RN32(RN64(a * RN64(1/b))) must equal RN32(a / b) bit for bit.  Compiled and
 * run by tests/test_cpu_units.py. */
#include <stdint.h>
#include <stdio.h>
#include <string.h>

int main(void)
{
    uint64_t bad = 0, checked = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : bad, checked)
    for (int b = 2; b < 65536; ++b) {
        const double r = 1.0 / (double)b;
        for (int a = 1; a < b; ++a) {
            const float q = (float)a / (float)b;
            const float p = (float)((double)a * r);
            uint32_t qb, pb;
            memcpy(&qb, &q, 4);
            memcpy(&pb, &p, 4);
            bad += qb != pb;
            ++checked;
        }
    }
    printf("%llu %llu\n", (unsigned long long)checked, (unsigned long long)bad);
    return bad != 0;
}
