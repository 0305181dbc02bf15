// Seeded synthetic porous media for the benchmark and test lattices.
//
// Random numbers: xoshiro256** (Blackman & Vigna, 2018) seeded through
// splitmix64; a uniform double is the top 53 bits times 2^-53.  The streams
// are fixed by the seed, so a mask is reproducible bit for bit (its SHA-256
// is recorded by the tests).
#include <stdint.h>
#include <string.h>

#include <cmath>
#include <vector>

#include "lbm_mi355x.h"
#include "lbm_internal.h"

namespace {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int q = 0; q < 4; q++) {
      uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      s[q] = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
};

inline double min_image(double d, double L) {
  if (d > 0.5 * L) d -= L;
  if (d < -0.5 * L) d += L;
  return d;
}

}  // namespace

extern "C" int lbm_geom_discs2d(int ni, int nj, uint64_t seed, double rmin, double rmax,
                                double phi, int walls, uint8_t *mask, double *phi_out) {
  if (ni <= 0 || nj <= 2 || !mask || rmin <= 0 || rmax < rmin)
    return lbm_fail("lbm_geom_discs2d: bad arguments");
  long N = (long)ni * nj, fluid = N;
  memset(mask, 0, (size_t)N);
  if (walls) {
    for (int i = 0; i < ni; i++) {
      mask[i] = 1;
      mask[(long)(nj - 1) * ni + i] = 1;
    }
    fluid -= 2L * ni;
  }
  Rng rng(seed);
  struct Disc { double x, y, r; };
  std::vector<Disc> discs;
  long tries = 0;
  const long max_tries = 20000000;
  while ((double)fluid / (double)N > phi) {
    if (++tries > max_tries) return lbm_fail("lbm_geom_discs2d: porosity target not reached");
    double r = rng.uniform(rmin, rmax);
    double x = rng.uniform(0.0, (double)ni);
    double ylo = walls ? 1.0 + r : 0.0, yhi = walls ? (double)(nj - 2) - r : (double)nj;
    if (yhi <= ylo) return lbm_fail("lbm_geom_discs2d: radius too large for the channel");
    double y = rng.uniform(ylo, yhi);
    bool ok = true;
    for (const Disc &d : discs) {
      double dx = min_image(x - d.x, ni), dy = walls ? y - d.y : min_image(y - d.y, nj);
      if (dx * dx + dy * dy < (r + d.r) * (r + d.r)) {
        ok = false;
        break;
      }
    }
    if (!ok) continue;
    discs.push_back({x, y, r});
    int j0 = (int)std::floor(y - r), j1 = (int)std::ceil(y + r);
    int i0 = (int)std::floor(x - r), i1 = (int)std::ceil(x + r);
    for (int jj = j0; jj <= j1; jj++) {
      int j = walls ? jj : ((jj % nj) + nj) % nj;
      if (j < 0 || j >= nj) continue;
      for (int ii = i0; ii <= i1; ii++) {
        int i = ((ii % ni) + ni) % ni;
        double dx = (double)ii - x, dy = (double)jj - y;
        if (dx * dx + dy * dy <= r * r) {
          uint8_t &m = mask[(long)j * ni + i];
          if (!m) {
            m = 1;
            fluid--;
          }
        }
      }
    }
  }
  if (phi_out) *phi_out = (double)fluid / (double)N;
  return 0;
}

extern "C" int lbm_geom_spheres3d(int ni, int nj, int nk, uint64_t seed, double rmin, double rmax,
                                  double phi, int zbuf, uint8_t *mask, double *phi_out) {
  if (ni <= 0 || nj <= 0 || nk <= 2 * zbuf || !mask || rmin <= 0 || rmax < rmin)
    return lbm_fail("lbm_geom_spheres3d: bad arguments");
  long E = (long)ni * nj, N = E * nk;
  long sample = E * (nk - 2L * zbuf), fluid = sample;
  memset(mask, 0, (size_t)N);
  Rng rng(seed);
  long tries = 0;
  while ((double)fluid / (double)sample > phi) {
    if (++tries > 100000000L) return lbm_fail("lbm_geom_spheres3d: porosity target not reached");
    double r = rng.uniform(rmin, rmax);
    double x = rng.uniform(0.0, (double)ni);
    double y = rng.uniform(0.0, (double)nj);
    double z = rng.uniform((double)zbuf, (double)(nk - zbuf));
    int k0 = (int)std::floor(z - r), k1 = (int)std::ceil(z + r);
    int j0 = (int)std::floor(y - r), j1 = (int)std::ceil(y + r);
    int i0 = (int)std::floor(x - r), i1 = (int)std::ceil(x + r);
    if (k0 < zbuf) k0 = zbuf;
    if (k1 > nk - zbuf - 1) k1 = nk - zbuf - 1;
    for (int k = k0; k <= k1; k++) {
      double dz = (double)k - z;
      for (int jj = j0; jj <= j1; jj++) {
        int j = ((jj % nj) + nj) % nj;
        double dy = (double)jj - y;
        for (int ii = i0; ii <= i1; ii++) {
          double dx = (double)ii - x;
          if (dx * dx + dy * dy + dz * dz <= r * r) {
            int i = ((ii % ni) + ni) % ni;
            uint8_t &m = mask[(long)k * E + (long)j * ni + i];
            if (!m) {
              m = 1;
              fluid--;
            }
          }
        }
      }
    }
  }
  if (phi_out) *phi_out = (double)fluid / (double)sample;
  return 0;
}
