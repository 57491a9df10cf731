/* Host check of the device's division-free x / 1e6 (div_1e6 in vulcan-proteomics_amd/csrc/vh_common.h, the ppm
 * term of every tolerance bound): the same f32 operations (one multiply, two fmas, the same fallback test) are
 * compared bit for bit with the IEEE division x / 1e6f for every one of the 2^32 f32 bit patterns.
 * Prints "<mismatches> <checked>"; built and run by tests/test_div1e6.py. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static float div_1e6(float x) {
  const float ax = fabsf(x);
  if (!(ax >= 0x1p-90f && ax <= 0x1.fffffep127f)) return x / 1000000.0f;
  const float r = 1e-6f;
  const float q0 = x * r;
  return fmaf(fmaf(-q0, 1000000.0f, x), r, q0);
}

int main(void) {
  long long bad = 0, checked = 0;
#pragma omp parallel for reduction(+ : bad, checked) schedule(static, 1 << 20)
  for (long long i = 0; i < (1LL << 32); ++i) {
    const uint32_t u = (uint32_t)i;
    float x;
    memcpy(&x, &u, 4);
    const float a = x / 1000000.0f, b = div_1e6(x);
    uint32_t ua, ub;
    memcpy(&ua, &a, 4);
    memcpy(&ub, &b, 4);
    bad += (ua != ub) && !(isnan(a) && isnan(b));
    checked += 1;
  }
  printf("%lld %lld\n", bad, checked);
  return 0;
}
