// generators.cpp -- host-side input generators of the C-ABI: psf_gauss
// (src/psf-gauss.jl:24-55) and the seeded synthetic phantom, mu-map and
// Poisson counts used by the tests and the benchmark.  Being in the library,
// every host (Julia, Python, C++) gets identical bits for a seed.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/spectrecon_amd.h"

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

// 1-D Gaussian of src/psf-gauss.jl:47-55: exp(-0.5 (r/σ)^2)/sum, σ = fwhm/sqrt(log 256);
// fwhm == 0 is the Kronecker impulse
void gauss1(std::vector<double>& v, int p, double fwhm) {
  const int h = (p - 1) / 2;
  v.assign(p, 0.0);
  if (fwhm == 0) {
    v[h] = 1.0;
    return;
  }
  const double sigma = fwhm / std::sqrt(std::log(256.0));
  double s = 0;
  for (int k = 0; k < p; ++k) {
    const double r = (double)(k - h) / sigma;
    v[k] = std::exp(-0.5 * (r * r));
    s += v[k];
  }
  for (int k = 0; k < p; ++k) v[k] /= s;
}

struct Ellipsoid {
  double cx, cy, cz, ax, ay, az, phi, val;
};

void add_ellipsoid(float* img, int nx, int ny, int nz, const Ellipsoid& e, bool replace) {
  const double c = std::cos(e.phi), s = std::sin(e.phi);
  for (int iz = 1; iz <= nz; ++iz) {
    const double dz = (iz - e.cz) / e.az;
    if (dz * dz > 1) continue;
    for (int iy = 1; iy <= ny; ++iy)
      for (int ix = 1; ix <= nx; ++ix) {
        const double dx = ix - e.cx, dy = iy - e.cy;
        const double u = (dx * c + dy * s) / e.ax, w = (-dx * s + dy * c) / e.ay;
        if (u * u + w * w + dz * dz <= 1) {
          float& dst = img[(ix - 1) + (long)nx * (iy - 1) + (long)nx * ny * (iz - 1)];
          dst = replace ? (float)e.val : (float)(dst + e.val);
        }
      }
  }
}

// Poisson sample: Knuth's product method for small means, Hörmann's PTRS
// (transformed rejection with squeeze, 1993) otherwise.
double poisson(SplitMix64& rng, double lam) {
  if (!(lam > 0)) return 0;
  if (lam < 30) {
    const double L = std::exp(-lam);
    double p = 1;
    long k = 0;
    do {
      ++k;
      p *= rng.uniform();
    } while (p > L);
    return (double)(k - 1);
  }
  const double slam = std::sqrt(lam), loglam = std::log(lam);
  const double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2);
  for (;;) {
    const double U = rng.uniform() - 0.5, V = rng.uniform();
    const double us = 0.5 - std::fabs(U);
    const double k = std::floor((2 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return k;
    if (k < 0 || (us < 0.013 && V > us)) continue;
    if (std::log(V) + std::log(invalpha) - std::log(a / (us * us) + b) <= -lam + k * loglam - std::lgamma(k + 1))
      return k;
  }
}

}  // namespace

extern "C" {

// src/psf-gauss.jl:24-45: psf[:,:,y] = g(rx, fwhm_x[y]) * g(rz, fwhm_z[y])' in
// Float64, stored as Float32
int sr_psf_gauss(float* psf, int32_t ny, int32_t px, int32_t pz, const double* fwhm_x, const double* fwhm_z) {
  if (!psf || !fwhm_x || !fwhm_z || ny <= 0 || px <= 0 || pz <= 0) return SR_ERR_ARG;
  std::vector<double> gx, gz;
  for (int y = 0; y < ny; ++y) {
    gauss1(gx, px, fwhm_x[y]);
    gauss1(gz, pz, fwhm_z[y]);
    for (int j = 0; j < pz; ++j)
      for (int i = 0; i < px; ++i) psf[i + (long)px * j + (long)px * pz * y] = (float)(gx[i] * gz[j]);
  }
  return SR_OK;
}

// 6 ellipsoids: centre U(0.3,0.7)*(nx,ny,nz), semi-axes U(0.06,0.2)*nx in plane
// and U(0.1,0.3)*nz axially, in-plane angle U(0,pi), intensity U(0.5,4), summed
// on a zero background, evaluated at 1-based voxel centres.
int sr_phantom(float* image, int32_t nx, int32_t ny, int32_t nz, uint64_t seed) {
  if (!image || nx <= 0 || ny <= 0 || nz <= 0) return SR_ERR_ARG;
  std::memset(image, 0, sizeof(float) * (size_t)nx * ny * nz);
  SplitMix64 rng(seed);
  for (int k = 0; k < 6; ++k) {
    Ellipsoid e;
    e.cx = rng.uniform(0.3, 0.7) * nx;
    e.cy = rng.uniform(0.3, 0.7) * ny;
    e.cz = rng.uniform(0.3, 0.7) * nz;
    e.ax = rng.uniform(0.06, 0.2) * nx;
    e.ay = rng.uniform(0.06, 0.2) * nx;
    e.az = rng.uniform(0.1, 0.3) * nz;
    e.phi = rng.uniform(0, 3.14159265358979323846);
    e.val = rng.uniform(0.5, 4.0);
    add_ellipsoid(image, nx, ny, nz, e, false);
  }
  return SR_OK;
}

// uniform 0.15 outer ellipsoid (semi-axes 0.42nx, 0.36ny, 0.48nz about the
// volume centre) with two 0.05 inner ellipsoids at +-0.18nx (semi-axes 0.10nx,
// 0.14ny, 0.30nz)
int sr_mumap(float* mumap, int32_t nx, int32_t ny, int32_t nz) {
  if (!mumap || nx <= 0 || ny <= 0 || nz <= 0) return SR_ERR_ARG;
  std::memset(mumap, 0, sizeof(float) * (size_t)nx * ny * nz);
  const double cx = (nx + 1) / 2.0, cy = (ny + 1) / 2.0, cz = (nz + 1) / 2.0;
  add_ellipsoid(mumap, nx, ny, nz, Ellipsoid{cx, cy, cz, 0.42 * nx, 0.36 * ny, 0.48 * nz, 0, 0.15}, true);
  for (int sgn = -1; sgn <= 1; sgn += 2)
    add_ellipsoid(mumap, nx, ny, nz,
                  Ellipsoid{cx + sgn * 0.18 * nx, cy, cz, 0.10 * nx, 0.14 * ny, 0.30 * nz, 0, 0.05}, true);
  return SR_OK;
}

int sr_poisson(float* out, const float* mean, size_t n, uint64_t seed) {
  if (!out || !mean) return SR_ERR_ARG;
  SplitMix64 rng(seed);
  for (size_t i = 0; i < n; ++i) out[i] = (float)poisson(rng, (double)mean[i]);
  return SR_OK;
}

}  // extern "C"
