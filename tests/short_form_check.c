/* The raster passes map a sample index to its pixel by the reference's walk
 * (kernel1, cl2.cl:5000-5127: fragments of op+1 samples, the cursor reset at
 * every row change) and, for triangles certified by short_form_ok, by the
 * short form y = row_of(j), x = min_x + (j - (y - min_y) * w)
 * (openclrenderer_amd/csrc/raster.hip).  This program restates both in the
 * same binary32 operations and compares them on every sample of every
 * certified (w, min_y, n): a mismatch prints and fails.
 *
 * usage: short_form_check <op> <w_max> <min_y_max> <min_y_far_samples>
 * (widths 1..w_max at min_y 0..min_y_max, widths up to 4096 in steps of 13
 * at min_y 0..40, and min_y_far_samples random min_y below 2^22 each) */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

static int OP = 500;
static float row_of(float q, float iw, float min_y) { return floorf(fmaf(q, iw, min_y)); }
static float run_start_steps(float Y, float q, float iw, float min_y) {
    for (int g = 0; g < 64 && q > 0.f && row_of(q - 1.f, iw, min_y) >= Y; g++) q -= 1.f;
    for (int g = 0; g < 64 && row_of(q, iw, min_y) < Y; g++) q += 1.f;
    return q;
}
static float run_start(float Y, float w, float iw, float min_y) {
    const float n0 = (Y - min_y) * w;
    const int lead = n0 > 0.f && row_of(n0 - 1.f, iw, min_y) >= Y;
    const int lag = row_of(n0, iw, min_y) < Y;
    return (lead || lag) ? run_start_steps(Y, n0, iw, min_y) : n0;
}
static int general(float slf, float min_x, float min_y, float w, float iw, float* x, float* y) {
    float faf = floorf(slf * (1.0f / (float)(OP + 1)));
    float of = fmaf(-faf, (float)(OP + 1), slf);
    faf = of < 0.f ? faf - 1.f : (of >= (float)(OP + 1) ? faf + 1.f : faf);
    of = fmaf(-faf, (float)(OP + 1), slf);
    const float paf = faf * (float)OP;
    const float qf = paf + of;
    *y = row_of(qf, iw, min_y);
    const float rs = run_start(*y, w, iw, min_y);
    const float r = paf > rs ? paf : rs;
    const float m = floorf(r * iw);
    float e = fmaf(-m, w, r);
    e = e < 0.f ? e + w : e;
    e = e >= w ? e - w : e;
    const float xo = e + (qf - r);
    *x = min_x + xo;
    return !(xo >= w);
}
static int short_ok(uint32_t n, float min_y, float w, float iw) {
    if (n == 0 || n > (uint32_t)OP + 1 || !(w > 0.f)) return 0;
    const float rows = row_of((float)(n - 1), iw, min_y) - min_y;
    if (!(rows <= 32.f)) return 0;
    int ok = 1;
    for (float k = 1.f; k <= rows; k += 1.f) {
        const float n0 = k * w;
        ok = ok && row_of(n0 - 1.f, iw, min_y) == min_y + (k - 1.f) && row_of(n0, iw, min_y) == min_y + k;
    }
    return ok;
}
static uint64_t cert, notcert, checked, bad;
static void one(float w, float min_y) {
    const float iw = 1.0f / w;   /* rcp_cr: correctly rounded */
    const float min_x = 17.f;
    /* the triangle with the most samples decides; smaller n are prefixes */
    for (uint32_t n = (uint32_t)OP + 1; n >= 1; n = n > 64 ? n - 61 : n - 1) {
        if (!short_ok(n, min_y, w, iw)) { notcert++; continue; }
        cert++;
        for (uint32_t j = 0; j < n; j++) {
            float gx, gy;
            const int gl = general((float)j, min_x, min_y, w, iw, &gx, &gy);
            const float y = row_of((float)j, iw, min_y);
            const float xo = (float)j - (y - min_y) * w;
            const float x = min_x + xo;
            const int l = !(xo >= w);
            checked++;
            if (gl != l || gy != y || (gl && gx != x)) {
                if (bad < 5)
                    printf("w %g min_y %g n %u j %u: walk (%g, %g, %d) short (%g, %g, %d)\n", w, min_y, n, j, gx, gy,
                           gl, x, y, l);
                bad++;
            }
        }
    }
}
int main(int argc, char** argv) {
    if (argc < 5) return 2;
    OP = atoi(argv[1]);
    const int w_max = atoi(argv[2]), y_max = atoi(argv[3]), far = atoi(argv[4]);
    /* every width up to w_max, then wide triangles (few rows) up to 4096 */
    for (int wi = 1; wi <= 4096; wi += wi < w_max ? 1 : 13) {
        for (int my = 0; my <= (wi <= w_max ? y_max : 40); my++) one((float)wi, (float)my);
        uint64_t s = 0x9E3779B97F4A7C15ull * (uint64_t)wi;
        for (int k = 0; k < far; k++) {   /* rows far down: min_y up to 2^22 */
            s ^= s << 13; s ^= s >> 7; s ^= s << 17;
            one((float)wi, (float)(s % 4194304u));
        }
    }
    printf("op %d: certified %llu, not %llu, samples %llu, %llu mismatches\n", OP, (unsigned long long)cert,
           (unsigned long long)notcert, (unsigned long long)checked, (unsigned long long)bad);
    return bad != 0;
}
