/* The shading and post passes turn a stored depth u (u32) into distance as
 * RN(RN(u) * C), C = 350000 * 2^-32, in place of the reference's
 * (RN(RN(u) / 4294967296.0f)) * 350000 (cl2.cl:5834).  This program checks
 * that the two agree for every binary32 value RN(u) can take: 0 and every
 * float in [1, 2^32].  Prints the mismatch count (0 expected). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

int main(void) {
    const float kMulint = 4294967296.0f, kDepthFar = 350000.0f;
    volatile float C = kDepthFar / kMulint;
    uint64_t n = 0, bad = 0;
    for (uint32_t bits = 0x3F800000u; bits <= 0x4F800000u; bits++) {   /* 1.0 .. 2^32 */
        float f;
        memcpy(&f, &bits, 4);
        volatile float a = f / kMulint;
        const float ref = a * kDepthFar;
        const float got = f * C;
        n++;
        if (memcmp(&ref, &got, 4) != 0) {
            if (bad < 5) printf("RN(u) = %a: reference %a, one product %a\n", f, ref, got);
            bad++;
        }
    }
    const float z = 0.0f;
    if ((z / kMulint) * kDepthFar != z * C) bad++;
    printf("%llu values, %llu mismatches\n", (unsigned long long)(n + 1), (unsigned long long)bad);
    return bad != 0;
}
