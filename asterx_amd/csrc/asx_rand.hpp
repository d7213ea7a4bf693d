// Counter-based random numbers of the zone samplers: splitmix64 (Steele,
// Lea & Flood, OOPSLA 2014) of (seed, zone, draw), so a zone's draws do not
// depend on the launch shape.
#pragma once

#include "asx_device.hpp"

namespace asx {

ASX_DEV uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
// uniform in (0, 1)
ASX_DEV double urand(uint64_t seed, int64_t z, int draw) {
  const uint64_t h = mix64(seed ^ mix64((uint64_t)z * 64u + (uint64_t)draw));
  return ((double)(h >> 11) + 0.5) * 0x1.0p-53;
}
ASX_DEV double loguni(uint64_t seed, int64_t z, int d, double a, double b) {
  return exp(log(a) + (log(b) - log(a)) * urand(seed, z, d));
}
// three standard normals (Box-Muller), draws d .. d+3
ASX_DEV void normal3(uint64_t seed, int64_t z, int d, double n[3]) {
  const double u1 = urand(seed, z, d), u2 = urand(seed, z, d + 1);
  const double u3 = urand(seed, z, d + 2), u4 = urand(seed, z, d + 3);
  const double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
  const double tp = 6.283185307179586;
  n[0] = r1 * cos(tp * u2);
  n[1] = r1 * sin(tp * u2);
  n[2] = r2 * cos(tp * u4);
}
ASX_DEV void unit_in(const Sym3 &g, double v[3]) {
  double l[3];
  lower(g, v, l);
  const double n = sqrt(dot3(v, l));
  v[0] /= n;
  v[1] /= n;
  v[2] /= n;
}

} // namespace asx
