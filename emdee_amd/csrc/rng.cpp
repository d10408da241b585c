// rng.cpp -- momentum-refresh generator: Marsaglia's KISS (1999) 32-bit
// integers feeding the Marsaglia & Tsang (2000) ziggurat for normal deviates,
// with the seeding and table construction of the reference (math.f90:42-177)
// so that EmDee_random_momenta reproduces the reference stream draw for draw.
// Built without FMA contraction so that the tables are bit-identical.
#include <cmath>
#include <cstdint>

#include "models.h"

namespace emdee {

static int32_t xorshift_mix(int32_t k, int n) {
  uint32_t u = (uint32_t)k;
  uint32_t s = n > 0 ? (u << n) : (u >> (-n));
  return (int32_t)(u ^ s);
}

int32_t Kiss::i32() {
  uint32_t ux = (uint32_t)x, uy = (uint32_t)y, uz = (uint32_t)z, uw = (uint32_t)w;
  ux = 69069u * ux + 1327217885u;                 // congruential
  uy ^= uy << 13; uy ^= uy >> 17; uy ^= uy << 5;  // 3-shift register
  uz = 18000u * (uz & 65535u) + (uz >> 16);       // multiply-with-carry
  uw = 30903u * (uw & 65535u) + (uw >> 16);
  x = (int32_t)ux; y = (int32_t)uy; z = (int32_t)uz; w = (int32_t)uw;
  return (int32_t)(ux + uy + (uz << 16) + uw);
}

void Kiss::setup(int32_t seed) {
  // seeding: math.f90:150-161
  x = xorshift_mix(xorshift_mix(xorshift_mix(seed, 13), -17), 5);
  y = xorshift_mix(xorshift_mix(xorshift_mix(x, 13), -17), 5);
  z = xorshift_mix(xorshift_mix(xorshift_mix(y, 13), -17), 5);
  w = xorshift_mix(xorshift_mix(xorshift_mix(z, 13), -17), 5);
  seeding_required = false;
  // ziggurat tables with 128 strips: math.f90:80-104
  const double m1 = 2147483648.0;
  double dn = 3.442619855899, tn = dn, vn = 0.00991256303526217;
  double q = vn * std::exp(0.5 * dn * dn);
  kn[0] = (int32_t)((dn / q) * m1);
  kn[1] = 0;
  wn[0] = q / m1;
  wn[127] = dn / m1;
  fn[0] = 1.0;
  fn[127] = std::exp(-0.5 * dn * dn);
  for (int i = 126; i >= 1; i--) {
    dn = std::sqrt(-2.0 * std::log(vn / dn + std::exp(-0.5 * dn * dn)));
    kn[i + 1] = (int32_t)((dn / tn) * m1);
    tn = dn;
    fn[i] = std::exp(-0.5 * dn * dn);
    wn[i] = dn / m1;
  }
}

static inline int32_t wrap_abs(int32_t h) { return h == INT32_MIN ? INT32_MIN : (h < 0 ? -h : h); }

// math.f90:108-146
double Kiss::normal() {
  const double r = 3.442620, s = 0.2328306e-9;
  int32_t hz = i32(), iz = hz & 127;
  if (wrap_abs(hz) < kn[iz]) return (double)hz * wn[iz];
  for (;;) {
    if (iz == 0) {
      double xx, yy;
      do {
        xx = -0.2904764 * std::log(s * (double)i32() + 0.5);
        yy = -std::log(s * (double)i32() + 0.5);
      } while (!(yy + yy >= xx * xx));
      double v = r + xx;
      return hz <= 0 ? -v : v;
    }
    double xx = (double)hz * wn[iz];
    if (fn[iz] + (s * (double)i32() + 0.5) * (fn[iz - 1] - fn[iz]) < std::exp(-0.5 * xx * xx)) return xx;
    hz = i32();
    iz = hz & 127;
    if (wrap_abs(hz) < kn[iz]) return (double)hz * wn[iz];
  }
}

}  // namespace emdee
