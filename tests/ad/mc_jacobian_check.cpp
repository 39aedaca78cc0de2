// Host check of the forward-mode Mohr-Coulomb update (mpm_amd/csrc/mpmx_ad.h):
// mc_update<double> against the oracle's return mapping bit for bit, and the
// Dual<1> tangents against central differences of mc_update<double>.
// Prints "cases plastic mismatches maxrel" on one line.
#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>

#include "mpmx_ad.h"

using namespace mpmx::ad;
typedef int (*kat_t)(const double*, const double*, const double*, double, double*, int, double*);

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  void* h = dlopen(argv[1], RTLD_NOW);
  if (!h) return 3;
  kat_t kat = (kat_t)dlsym(h, "orc_kat_stress");
  if (!kat) return 4;
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(-1, 1);
  int nbad = 0, nplastic = 0;
  const int n = 4000;
  double maxrel = 0;
  for (int t = 0; t < n; ++t) {
    const double E = 1e6 * (1 + 9 * (U(rng) + 1) / 2), nu = 0.3;
    const double phi_deg = 25 + 15 * (U(rng) + 1) / 2;
    const double coh = (t % 3 == 0) ? 0. : 2000. * (U(rng) + 1);
    const double psi_deg = (t % 2) ? 0. : 5.;
    const double tcut = (t % 5 == 0) ? 100. : 0.;
    const double params[16] = {1, 1800, E, nu, phi_deg, psi_deg, coh, phi_deg,
                               psi_deg, coh, 0, 0, tcut, 0, 0, 0};
    double st[6], de[6];
    const double p0 = -5e3 * (U(rng) + 1.2);
    for (int k = 0; k < 3; ++k) st[k] = p0 + 3e3 * U(rng);
    st[3] = 2e3 * U(rng);
    st[4] = (t % 4 == 0) ? 0 : 1e3 * U(rng);
    st[5] = (t % 4 == 0) ? 0 : 1e3 * U(rng);
    for (int k = 0; k < 6; ++k) de[k] = 2e-3 * U(rng);
    if (t % 4 == 0) de[2] = de[4] = de[5] = 0;
    double sv[9], out_o[6];
    kat(params, st, de, 1800., sv, 1, out_o);
    const double bulk = E / (3.0 * (1. - 2. * nu)), G = E / (2.0 * (1 + nu));
    const double a1 = bulk + (4.0 / 3.0) * G, a2 = bulk - (2.0 / 3.0) * G;
    const double phi = phi_deg * M_PI / 180., psi = psi_deg * M_PI / 180.;
    const McParams<double> m{a1, a2, G, phi, psi, coh, sv[4], coh};
    double out[6];
    const int yt = mc_update(m, st, de, out);
    if (yt) nplastic++;
    for (int k = 0; k < 6; ++k)
      if (out[k] != out_o[k]) {
        nbad++;
        break;
      }
    if (!yt || t % 10) continue;
    for (int q = 0; q < 14; ++q) {
      Dual<1> S[6], DE[6];
      for (int k = 0; k < 6; ++k) {
        S[k] = cst<Dual<1>>(st[k]);
        DE[k] = cst<Dual<1>>(de[k]);
      }
      McParams<Dual<1>> md{cst<Dual<1>>(a1), cst<Dual<1>>(a2), cst<Dual<1>>(G),
                           cst<Dual<1>>(phi), psi, coh, cst<Dual<1>>(sv[4]), coh};
      double hq;
      if (q < 6) {
        S[q].d[0] = 1;
        hq = 1e-3;
      } else if (q < 12) {
        DE[q - 6].d[0] = 1;
        hq = 1e-9;
      } else if (q == 12) {
        md.a1.d[0] = a1 / E;
        md.a2.d[0] = a2 / E;
        md.G.d[0] = G / E;
        hq = E * 1e-7;
      } else {
        md.phi.d[0] = 1;
        hq = 1e-7;
      }
      Dual<1> O[6];
      mc_update(md, S, DE, O);
      auto run = [&](double sgn, double* o) {
        double s2[6], d2[6];
        for (int k = 0; k < 6; ++k) {
          s2[k] = st[k];
          d2[k] = de[k];
        }
        McParams<double> m2 = m;
        if (q < 6)
          s2[q] += sgn * hq;
        else if (q < 12)
          d2[q - 6] += sgn * hq;
        else if (q == 12) {
          const double E2 = E + sgn * hq;
          const double b2 = E2 / (3.0 * (1. - 2. * nu)), G2 = E2 / (2.0 * (1 + nu));
          m2.a1 = b2 + (4.0 / 3.0) * G2;
          m2.a2 = b2 - (2.0 / 3.0) * G2;
          m2.G = G2;
        } else {
          m2.phi += sgn * hq;
        }
        return mc_update(m2, s2, d2, o);
      };
      double op[6], om[6];
      // skip differences that cross a kink (yield type change)
      if (run(1, op) != yt || run(-1, om) != yt) continue;
      double scale = 0;
      for (int k = 0; k < 6; ++k) scale = std::max(scale, std::fabs(O[k].d[0]));
      if (scale == 0) continue;
      for (int k = 0; k < 6; ++k) {
        const double fd = (op[k] - om[k]) / (2 * hq);
        maxrel = std::max(maxrel, std::fabs(fd - O[k].d[0]) / scale);
      }
    }
  }
  printf("%d %d %d %.6e\n", n, nplastic, nbad, maxrel);
  return 0;
}
