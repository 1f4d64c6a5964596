/* Checks of the arithmetic identities the shading pass uses in place of
 * full correctly rounded divisions (div_rcp in csrc/ocr_math.h):
 *
 *   const  B      a = 0 and every finite binary32 a >= 2^-100:
 *                 fma(fma(-q, B, a), y, q) with y = RN(1/B), q = RN(a*y)
 *                 equals the IEEE quotient a / B (below 2^-100 the quotient
 *                 or the residual can be subnormal and the identity fails
 *                 for some divisors; the kernels never divide such values)
 *   random N      N random pairs a, b with exponents in [-60, 60]: the same
 *                 with y = RN(1/b)
 *   tan LIB       with the oracle library LIB: tan_f(a) >= 1 for every
 *                 binary32 a from float(pi/4) rounded up to
 *                 rational_acos(0.05), and tan_f just below it is < 1
 *
 * Prints the number of mismatches; exit status 0 when there are none. */
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float fb(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t bf(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

static float div_rcp(float a, float b, float y) {
  const float q = a * y;
  const float r = fmaf(-q, b, a);
  return fmaf(r, y, q);
}

static int check_const(float b) {
  volatile float vb = b;
  const float y = 1.0f / vb;
  uint64_t bad = 0;
  for (uint64_t u = 0x0D800000ull; u < 0x7F800001ull; u = (u == 0x7F800000ull ? 0x7F800001ull : u + 1)) {
    const float a = fb(u == 0x7F800000ull ? 0u : (uint32_t)u);   /* 2^-100 .. max finite, then +0 */
    volatile float va = a;
    if (bf(div_rcp(a, b, y)) != bf(va / vb)) bad++;
  }
  printf("const %g: %llu mismatches\n", b, (unsigned long long)bad);
  return bad != 0;
}

static int check_random(uint64_t n) {
  uint64_t s = 0x9E3779B97F4A7C15ull, bad = 0;
  for (uint64_t i = 0; i < n; i++) {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    const uint32_t ea = 127 - 60 + (uint32_t)(s % 121), eb = 127 - 60 + (uint32_t)((s >> 8) % 121);
    uint32_t ma = (uint32_t)(s >> 16) & 0x7FFFFF, mb = (uint32_t)(s >> 40) & 0x7FFFFF;
    if ((s >> 61) == 7) mb &= 0x7F0000;   /* divisors with few significand bits */
    const float a = fb(ea << 23 | ma), b = fb(eb << 23 | mb);
    volatile float va = a, vb = b;
    const float y = 1.0f / vb;
    if (bf(div_rcp(a, b, y)) != bf(va / vb)) bad++;
  }
  printf("random %llu: %llu mismatches\n", (unsigned long long)n, (unsigned long long)bad);
  return bad != 0;
}

static int check_tan(const char* lib) {
  void* h = dlopen(lib, RTLD_NOW);
  if (!h) { printf("dlopen: %s\n", dlerror()); return 2; }
  float (*tan_f)(float) = (float (*)(float))dlsym(h, "oracle_tan");
  float (*acos_f)(float) = (float (*)(float))dlsym(h, "oracle_rational_acos");
  if (!tan_f || !acos_f) { printf("missing symbols\n"); return 2; }
  const uint32_t from = 0x3f490fdbu;   /* 0x1.921fb6p-1 */
  const float hi = acos_f(0.05f);
  uint64_t bad = 0, n = 0;
  for (uint32_t u = from; fb(u) <= hi; u++, n++)
    if (!(tan_f(fb(u)) >= 1.0f)) bad++;
  const int below = tan_f(fb(from - 1)) < 1.0f;
  printf("tan: %llu values, %llu below 1, tan just below threshold < 1: %d\n", (unsigned long long)n,
         (unsigned long long)bad, below);
  return bad != 0 || !below || n < 1000000;
}

int main(int argc, char** argv) {
  if (argc == 3 && !strcmp(argv[1], "const")) return check_const((float)atof(argv[2]));
  if (argc == 3 && !strcmp(argv[1], "random")) return check_random(strtoull(argv[2], 0, 10));
  if (argc == 3 && !strcmp(argv[1], "tan")) return check_tan(argv[2]);
  fprintf(stderr, "usage: const B | random N | tan LIB\n");
  return 2;
}
