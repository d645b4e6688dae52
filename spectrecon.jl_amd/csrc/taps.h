// taps.h -- host-side PSF tap tables shared by the plan runtime and the unit
// operators: the rank-1 test/factorisation of a PSF plane, the tail/head sums
// of the replicate-pad fold, and the tap-table lengths the kernels are
// instantiated for.  Plain pointers in and out: no allocation.
#pragma once
#include <cmath>

namespace sr {

// tap-table length for a PSF side (the VALU kernels are instantiated per
// multiple of 8 up to 48, then 64)
inline int round_taps(int p) {
  return p <= 8 ? 8 : p <= 16 ? 16 : p <= 24 ? 24 : p <= 32 ? 32 : p <= 40 ? 40 : p <= 48 ? 48 : 64;
}

// Rank-1 factorisation h[a,b] ~ u[a] w[b] of a px x pz plane (u: px values,
// w: pz values), accepted when every entry agrees to 1e-6 of max|h|
// (psf_gauss planes are outer products rounded to Float32, src/psf-gauss.jl:40).
inline bool factor_rank1(const float* h, int px, int pz, double* u, double* w) {
  int ia = 0, ib = 0;
  double hm = 0;
  for (int b = 0; b < pz; ++b)
    for (int a = 0; a < px; ++a)
      if (std::fabs(h[a + px * b]) > hm) {
        hm = std::fabs(h[a + px * b]);
        ia = a;
        ib = b;
      }
  for (int a = 0; a < px; ++a) u[a] = 0.0;
  for (int b = 0; b < pz; ++b) w[b] = 0.0;
  if (hm == 0) return true;
  const double piv = h[ia + px * ib];
  for (int a = 0; a < px; ++a) u[a] = h[a + px * ib];
  for (int b = 0; b < pz; ++b) w[b] = h[ia + px * b] / piv;
  for (int b = 0; b < pz; ++b)
    for (int a = 0; a < px; ++a)
      if (std::fabs(u[a] * w[b] - (double)h[a + px * b]) > 1e-6 * hm) return false;
  return true;
}

// Per plane tail/head sums of the separable taps hu, hw ([plane][nt] each),
// for the adjoint's replicate-pad fold: out [plane][u tail, u head, w tail,
// w head][nt], summed in double.
inline void cum_sums(const float* hu, const float* hw, int nplane, int nt, float* out) {
  for (int pl = 0; pl < nplane; ++pl)
    for (int which = 0; which < 2; ++which) {
      const float* t = (which == 0 ? hu : hw) + (long)pl * nt;
      float* tail = out + ((long)pl * 4 + 2 * which) * nt;
      float* head = tail + nt;
      double acc = 0;
      for (int k = nt - 1; k >= 0; --k) tail[k] = (float)(acc += t[k]);
      acc = 0;
      for (int k = 0; k < nt; ++k) head[k] = (float)(acc += t[k]);
    }
}

}  // namespace sr
