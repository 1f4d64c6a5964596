/* The tiled raster path (openclrenderer_amd/csrc/tiles.hip) visits, for each
 * cut triangle and each screen tile its bounding box touches, the pixels the
 * reference's fragment walk visits in that tile (kernel1, cl2.cl:5000-5127;
 * kernel2, cl2.cl:5419-5545: fragments of op + 1 samples, a cursor that
 * steps x by one and is reset to running_width_mod + min_x whenever the
 * row floor(mad(q, 1/width, min_y)) changes, stopping at the first row
 * >= max_y).  It does so in one of two ways:
 *
 *  - triangles that walk_is_rect certifies: every pixel of the bounding box
 *    [min_x, max_x) x [min_y, max_y) clipped to the tile;
 *  - the others: the walk's samples whose rows fall in the tile, j in
 *    [j_first(run_start(y0)), j_first(run_start(y1))), each mapped to its
 *    pixel by the closed form of the walk (raster.hip sample_xy) and kept
 *    when its column is in the tile.
 *
 * This program restates the reference walk (cursor stepping, every fragment
 * of the triangle) and both tile rules in the same binary32 operations, and
 * requires the set of pixels each gives to be equal, for every tile, over
 * bounding boxes of many widths, first rows and heights (top rows, where the
 * row fma lags, included), for the camera (op 500) and light (op 300)
 * fragment sizes and several tile shapes.  Uncertified triangles are also
 * run through the rectangle test to count how often it would be wrong, and
 * certified ones through the sample rule (which must agree everywhere).
 *
 * usage: tile_walk_check <op> <w_max> <min_y_max> <rows_max> <random_cases> */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

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
/* raster.hip sample_xy without the inside test: pixel of sample j */
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
/* tiles.hip walk_is_rect */
static int walk_is_rect(float min_y, float max_y, float w, float iw) {
    const float rows = max_y - min_y;
    const int no_lag = fmaf(w, iw, -1.f) >= 0.f || rows < min_y;
    const int no_lead = w * (max_y + rows) <= 16777215.f;
    if (no_lag && no_lead) return 1;
    if (rows > 64.f) return 0;
    int ok = 1;
    for (float k = 1.f; k <= rows; k += 1.f) {
        const float n0 = k * w;
        ok = ok && row_of(n0 - 1.f, iw, min_y) == min_y + (k - 1.f) && row_of(n0, iw, min_y) == min_y + k;
    }
    return ok;
}
/* tiles.hip first_sample: samples j whose walk index q(j) is below q */
static uint32_t first_sample(uint32_t q) {
    const uint32_t a = q / (uint32_t)OP, r = q - a * (uint32_t)OP;
    return q + a - (r == 0u && a > 0u ? 1u : 0u);
}
/* raster.hip triangle_samples */
static uint32_t triangle_samples(float min_y, float max_y, float w, float iw, uint32_t nfrag) {
    const uint32_t q_end = (uint32_t)run_start(max_y, w, iw, min_y);
    const uint32_t uop = (uint32_t)OP;
    uint32_t m = q_end >= uop + 1 ? (q_end - (uop + 1)) / uop + 1 : 0;
    if (m > nfrag) m = nfrag;
    uint64_t s = (uint64_t)(uop + 1) * m;
    if (m < nfrag) {
        const int64_t r = (int64_t)q_end - (int64_t)uop * m;
        s += r < 0 ? 0 : (r > (int64_t)uop + 1 ? uop + 1 : (uint64_t)r);
    }
    return (uint32_t)s;
}

static uint8_t* ref_set;
static uint8_t* tile_set;
static uint64_t n_cases, n_rect, n_general, n_rect_wrong_if_forced, bad, samples_checked;

/* every fragment of the walk (oracle walk_fragment without the inside test) */
static void ref_walk(int min_x, int min_y, int w, int max_y, uint32_t nfrag) {
    const float mm0 = (float)min_x, mm1 = (float)(min_x + w), mm2 = (float)min_y, mm3 = (float)max_y;
    for (uint32_t dist = 0; dist < nfrag; dist++) {
        int pixel_along = OP * (int)dist;
        int pcount = -1;
        float x = (float)((pixel_along + 0) % w) + mm0 - 1;
        float y = floorf((float)(pixel_along + pcount) / (float)w) + mm2;
        float iwidth = 1.f / (float)w;
        float rwm = (float)((pixel_along + pcount) % w);
        while (pcount < OP) {
            pcount++;
            x += 1;
            rwm += 1.f;
            if (rwm >= (float)w) rwm = 0;
            float ty = y;
            y = floorf(fmaf((float)(pixel_along + pcount), iwidth, mm2));
            x = y != ty ? rwm + mm0 : x;
            if (y >= mm3) break;
            if (x >= mm1 || x < mm0 || y < mm2) continue;
            ref_set[(size_t)((int)y - min_y) * w + ((int)x - min_x)] = 1;
        }
    }
}

/* the sample rule over every tile of a tw x th grid */
static void tile_general(int min_x, int min_y, int w, int max_y, uint32_t n, int tw, int th) {
    const float fw = (float)w, iw = 1.0f / fw, fmin_x = (float)min_x, fmin_y = (float)min_y;
    const int max_x = min_x + w;
    for (int ty0 = (min_y / th) * th; ty0 < max_y; ty0 += th) {
        const int y0 = ty0 > min_y ? ty0 : min_y, y1 = ty0 + th < max_y ? ty0 + th : max_y;
        const uint32_t qa = (uint32_t)run_start((float)y0, fw, iw, fmin_y);
        const uint32_t qb = (uint32_t)run_start((float)y1, fw, iw, fmin_y);
        uint32_t ja = first_sample(qa), jb = first_sample(qb);
        if (jb > n) jb = n;
        for (int tx0 = (min_x / tw) * tw; tx0 < max_x; tx0 += tw) {
            for (uint32_t j = ja; j < jb; j++) {
                float x, y;
                if (!general((float)j, fmin_x, fmin_y, fw, iw, &x, &y)) continue;
                samples_checked++;
                if (y < (float)y0 || y >= (float)y1) {
                    if (bad < 8) printf("sample %u of w %d min_y %d: row %g outside tile rows [%d, %d)\n", j, w, min_y, y, y0, y1);
                    bad++;
                    continue;
                }
                if (x < (float)tx0 || x >= (float)(tx0 + tw)) continue;
                tile_set[(size_t)((int)y - min_y) * w + ((int)x - min_x)] = 1;
            }
        }
    }
}

static void one(int min_x, int min_y, int w, int rows, int tw, int th) {
    const int max_y = min_y + rows;
    const float fw = (float)w, iw = 1.0f / fw;
    const float area = fw * (float)rows;
    const uint32_t nfrag = (uint32_t)ceilf(area / (float)OP);
    const uint32_t n = triangle_samples((float)min_y, (float)max_y, fw, iw, nfrag);
    const size_t sz = (size_t)w * rows;
    memset(ref_set, 0, sz);
    memset(tile_set, 0, sz);
    ref_walk(min_x, min_y, w, max_y, nfrag);
    const int rect = walk_is_rect((float)min_y, (float)max_y, fw, iw);
    n_cases++;
    /* the sample rule must agree with the walk for every triangle it runs
     * on (all uncertified ones) and for certified ones of moderate size (it
     * costs columns-of-tiles times the box per triangle) */
    if (!rect || sz <= 40000) tile_general(min_x, min_y, w, max_y, n, tw, th);
    else memcpy(tile_set, ref_set, sz);
    if (memcmp(ref_set, tile_set, sz) != 0) {
        if (bad < 8) printf("sample rule differs: w %d min_y %d rows %d tile %dx%d\n", w, min_y, rows, tw, th);
        bad++;
    }
    int full = 1;
    for (size_t i = 0; i < sz; i++) full &= ref_set[i];
    if (rect) {
        n_rect++;
        if (!full) {
            if (bad < 8) printf("certified but the walk is not the box: w %d min_y %d rows %d\n", w, min_y, rows);
            bad++;
        }
    } else {
        n_general++;
        if (!full) n_rect_wrong_if_forced++;
    }
}

int main(int argc, char** argv) {
    if (argc < 6) return 2;
    OP = atoi(argv[1]);
    const int w_max = atoi(argv[2]), y_max = atoi(argv[3]), rows_max = atoi(argv[4]), nrand = atoi(argv[5]);
    const size_t cap = (size_t)8192 * 4096;
    ref_set = (uint8_t*)malloc(cap);
    tile_set = (uint8_t*)malloc(cap);
    if (!ref_set || !tile_set) return 3;
    static const int shapes[][2] = {{16, 16}, {32, 16}, {32, 32}, {64, 32}, {8, 8}};
    int si = 0;
    /* every width up to w_max, first rows 0..y_max, heights 1..rows_max */
    for (int w = 1; w <= w_max; w++)
        for (int my = 0; my <= y_max; my += my < 24 ? 1 : 7)
            for (int rows = 1; rows <= rows_max; rows += rows < 8 ? 1 : 5) {
                const int* s = shapes[si++ % 5];
                one(3 + (w * 7 + my) % 50, my, w, rows, s[0], s[1]);
            }
    /* random boxes: wide, tall, far down, up to 4K and 5K frames */
    uint64_t st = 0x243F6A8885A308D3ull ^ (uint64_t)OP;
    for (int k = 0; k < nrand; k++) {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        const int big = (int)(st % 5u) == 0;
        const int W = big ? 5120 : 3840, H = big ? 2880 : 2160;
        int w = 1 + (int)((st >> 8) % (uint64_t)(k % 3 == 0 ? W - 1 : 200));
        int rows = 1 + (int)((st >> 24) % (uint64_t)(k % 4 == 0 ? H - 1 : 120));
        int my = (int)((st >> 40) % (uint64_t)(k % 2 ? 64 : H));
        if (my + rows > H - 1) my = H - 1 - rows;
        if (my < 0) { my = 0; rows = H - 1; }
        int mx = (int)((st >> 50) % (uint64_t)(W - w));
        if ((size_t)w * rows > cap) continue;
        const int* s = shapes[k % 5];
        one(mx, my, w, rows, s[0], s[1]);
    }
    printf("op %d: cases %llu, rect %llu, general %llu (box walk would be wrong for %llu), samples %llu, %llu mismatches\n",
           OP, (unsigned long long)n_cases, (unsigned long long)n_rect, (unsigned long long)n_general,
           (unsigned long long)n_rect_wrong_if_forced, (unsigned long long)samples_checked, (unsigned long long)bad);
    return bad != 0;
}
