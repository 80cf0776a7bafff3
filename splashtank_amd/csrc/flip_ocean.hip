// flip_ocean.hip - the ocean's wave list (host code; touches no GPU).
//
// The ocean boundary of the FLIP substep is a sum of linear (Airy) waves whose
// parameters the host computes once per setting: st_ocean_waves restates
// oceanWaves (src/shared.js:769-791) in f64, the arithmetic of the reference's
// JavaScript, with the same sequence of pseudo-random draws, and rounds each
// wave's numbers to f32 as its Float32Array does.
#include <algorithm>
#include <cmath>
#include <cstdint>

#include "splashtank.h"

namespace {

// mulberry32 (src/shared.js:765): a 32-bit state advanced by 0x6D2B79F5, mixed
// by two multiply-xorshift rounds, returned as a fraction of 2^32
struct Mulberry32 {
  uint32_t a;
  double next() {
    a = a + 0x6D2B79F5u;
    uint32_t t = (a ^ (a >> 15)) * (1u | a);
    t = (t + ((t ^ (t >> 7)) * (61u | t))) ^ t;
    return double(t ^ (t >> 14)) / 4294967296.0;
  }
};

// wavenumber of angular frequency om in water of depth D: the fixed point of
// k = om^2 / (g tanh(k D)), 40 iterations from the deep-water value
double wavenumber(double om, double G, double D) {
  double k = om * om / G;
  for (int i = 0; i < 40; i++) k = om * om / (G * std::tanh(std::max(k * D, 1e-6)));
  return k;
}

}  // namespace

extern "C" int st_ocean_waves(const st_ocean* oc, double g, double depth, float* w, uint32_t cap, uint32_t* n) {
  if (!oc || !n) return ST_ERR_ARG;
  const int NW = 24;
  *n = uint32_t(NW);
  // seed: (seed | 0) || 1, times 7919 plus 13, wrapped to 32 bits by mulberry32's a |= 0
  int64_t s = oc->seed != 0 ? int64_t(oc->seed) : 1;
  Mulberry32 rnd{uint32_t(uint64_t(s * 7919 + 13))};
  const double D = std::max(depth, 1e-3), G = std::max(g, 1e-4), lp = std::max(oc->length, 1e-3);
  const double kp = 2.0 * M_PI / lp;
  const double wp = std::sqrt(G * kp * std::tanh(kp * D));
  const double th0 = oc->dir * M_PI / 180.0;
  const double spr = std::max(0.0, oc->spread) * M_PI / 180.0;
  double k[NW], om[NW], S[NW], th[NW], ph[NW], sum = 0.0;
  for (int i = 0; i < NW; i++) {
    const double r = 0.6 + 1.6 * (double(i) + rnd.next()) / double(NW);
    om[i] = r * wp;
    k[i] = wavenumber(om[i], G, D);
    // Pierson-Moskowitz shape in r = omega / omega_peak
    S[i] = std::pow(r, -5.0) * std::exp(-1.25 * std::pow(r, -4.0));
    const double u1 = rnd.next(), u2 = rnd.next(), u3 = rnd.next();
    th[i] = th0 + spr * (u1 + u2 + u3 - 1.5) / 1.5;
    ph[i] = rnd.next() * 2.0 * M_PI;
    sum += S[i];
  }
  // significant height = 4 standard deviations of the surface: m0 = (Hs / 4)^2
  const double m0 = std::pow(std::max(oc->height, 0.0) / 4.0, 2.0);
  for (int i = 0; i < NW && uint32_t(i) < cap && w; i++) {
    const double a = std::min(std::sqrt(2.0 * m0 * S[i] / sum), 0.3 / k[i]);
    float* o = w + size_t(i) * 8;
    o[0] = float(k[i] * std::cos(th[i]));
    o[1] = float(k[i] * std::sin(th[i]));
    o[2] = float(a);
    o[3] = float(om[i]);
    o[4] = float(ph[i]);
    o[5] = float(k[i]);
    o[6] = 0.0f;
    o[7] = 0.0f;
  }
  return ST_OK;
}
