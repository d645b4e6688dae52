// fft_psf.hip -- the depth-dependent PSF convolution of src/fft_convolve.jl
// (fft_conv!, fft_conv_adj!, imfilterz!, :16-130; plan sizes src/plan-psf.jl,
// pads src/helper.jl:8-12) as an LDS-resident FFT per padded plane, for the
// x-fastest work layout ([b][z][y][x] volumes, [x][z] views).
//
// A padded plane is L1 x L2 (L = Power2(n + p - 1) per axis; up / left =
// ceil((L - n) / 2) pads before x / z).  Its rows are real, so rows z = 2j and
// 2j + 1 travel as the real and imaginary parts of one complex row ("row
// pair" j): the x transform of a pair holds both rows' spectra, separated on
// the fly as A = (Z[f] + conj Z[-f]) / 2, B = -i (Z[f] - conj Z[-f]) / 2.  The z
// transform then runs over the half spectrum f = 0..L1/2 only, as L1/2
// complex columns: f = 0 and f = L1/2 are real per row, so they travel as the
// real and imaginary parts of one packed column 0 (its spectral product is
// taken apart with the column's mirror frequencies, below).  Each 1-D
// transform is two radix-R passes in LDS (L = R1 R2, R <= 16, in-register
// DFTs): the forward passes are decimation in frequency, the inverse passes
// decimation in time.  The x passes leave the x spectrum in natural order
// (the z passes read neighbouring frequencies on distinct banks); the z
// passes leave the z spectrum in digit-reversed order (position k1 + R1 k2
// holds frequency k2 + R2 k1) and the inverse z passes take that order back
// to natural order, so no reordering pass exists.  The kernel's spectrum is
// stored in the z passes' order.
//
// The PSFs are point-symmetric (src/spectplan.jl:83), so the spectrum of a
// plane's kernel, centred at index 0 (what pad2sizezero! followed by
// fftshift2! amounts to, src/fft_convolve.jl:17-23), is real: H[f][g] =
// sum_{a,b} h[a,b] cos(2 pi (f (a - cx) / L1 + g (b - cz) / L2)), computed in
// double at plan time and stored as float with the 1 / (L1 L2) of the
// inverse folded in: [plane][L2 positions][L1/2 + 1] floats, where column 0
// holds (H[0] + H[L1/2]) / 2 and column L1/2 holds (H[0] - H[L1/2]) / 2.
//
// The packed column: with P = C0 + i CN (C0, CN the real rows' f = 0 and
// f = L1/2 columns) and its z transform P^, the wanted product is
// C0^ H[0] + i CN^ H[L1/2] = P^[g] Hp[g] + conj(P^[-g]) Hm[g], Hp and Hm the
// stored half sum and half difference (H[0][g] and H[L1/2][g] are even in g).
// The forward sums P^ Hp in registers and W = P^ Hm in LDS over the chunk and
// adds conj(W[-g]) once at its end; the adjoint keeps the view's P^ in LDS and
// adds conj(P^[-g]) Hm per plane.
//
// Forward (project): per block one view and a chunk of planes.  Each plane's
// rows are read with clamped x (the replicate pad, padrepl!) and its z
// transform reads clamped rows, so no padded copy exists; the plane's
// spectrum times H_y is summed in registers over the chunk (the view sum is
// linear), and one inverse transform per chunk writes the chunk's partial
// view (cropped).  Adjoint (back-project): per block one view and a chunk of
// planes; the zero-padded view is transformed once, then per plane its
// spectrum times H_y is transformed back, the pad rows and columns are
// folded onto the edge rows and columns (src/fft_convolve.jl:99-122), and the
// cropped plane times the attenuation factors is written to the work volume.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "kernels.h"

namespace sr {
namespace {

constexpr double kPiD = 3.14159265358979323846;
constexpr int kReadAll = 2;    // items per thread of the pass that reads all, then writes
// threads per block: 512 (8 waves, one block per CU, up to 256 VGPRs: the
// forward's chunk sums and the adjoint's view spectrum live in registers)
constexpr int kNTF = 512, kNTA = 512;

template <int N>
struct Split;
template <>
struct Split<16> { static constexpr int R1 = 4, R2 = 4; };
template <>
struct Split<32> { static constexpr int R1 = 4, R2 = 8; };
template <>
struct Split<64> { static constexpr int R1 = 8, R2 = 8; };
template <>
struct Split<128> { static constexpr int R1 = 8, R2 = 16; };
template <>
struct Split<256> { static constexpr int R1 = 16, R2 = 16; };

int split_r1(int n) { return n <= 32 ? 4 : n <= 128 ? 8 : 16; }

__host__ __device__ constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }
__host__ __device__ constexpr int brev(int i, int bits) {
  int r = 0;
  for (int b = 0; b < bits; ++b) r |= ((i >> b) & 1) << (bits - 1 - b);
  return r;
}
// cos / sin (2 pi m / 16)
__host__ __device__ constexpr float c16(int m) {
  constexpr float c[16] = {1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
                           0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f,
                           -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508977f,
                           0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f};
  return c[m & 15];
}
__host__ __device__ constexpr float s16(int m) { return c16(m + 12); }  // sin x = cos(x - pi/2)

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -a.x * b.y));
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// The thread's index for the passes' item loops.  OPQ (the adjoint): read
// through an empty asm, so that the item indices and LDS offsets derived from
// it are recomputed in every pass of the plane loop instead of being hoisted
// out of it and held in registers for every plane.  At 256 x 128 that hoisting
// spilled 67 VGPRs to scratch and 184 SGPRs; with OPQ in every adjoint pass
// the adjoint holds 184 VGPRs and spills nothing, and a 64-view launch takes
// 0.85 ms instead of 1.41.  The forward is as fast or faster without it.
template <bool OPQ>
__device__ __forceinline__ int item_tid() {
  int t = threadIdx.x;
  if constexpr (OPQ) asm volatile("" : "+v"(t));
  return t;
}

// In-register DFT of R points, natural order in and out (radix-2 stages):
// forward a[k] = sum_n a[n] e^{-2 pi i n k / R}; INV: e^{+...}, unnormalised.
template <int R, bool INV>
__device__ __forceinline__ void dft(float2 (&a)[R]) {
  constexpr int B = ilog2c(R);
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int j = brev(i, B);
    if (j > i) {
      const float2 t = a[i];
      a[i] = a[j];
      a[j] = t;
    }
  }
#pragma unroll
  for (int len = 2; len <= R; len <<= 1) {
#pragma unroll
    for (int i = 0; i < R; i += len) {
#pragma unroll
      for (int k = 0; k < len / 2; ++k) {
        float2 t = a[i + k + len / 2];
        const int m = k * (16 / len);
        if (k != 0) {
          if (m == 4) t = INV ? make_float2(-t.y, t.x) : make_float2(t.y, -t.x);
          else t = cmul(t, make_float2(c16(m), INV ? s16(m) : -s16(m)));
        }
        const float2 u = a[i + k];
        a[i + k] = cadd(u, t);
        a[i + k + len / 2] = csub(u, t);
      }
    }
  }
}

// LDS regions (float2): the twiddles W_L1^m, W_L2^m; region W, the packed
// column's mirror terms [g] (L2); region A, the x-buffer [pos][pair] (stride
// NPS) and in turn the rows' half spectra I [r][f] (L1 / 2 columns at the
// odd stride NIS = L1 / 2 + 1: x pass 2' reads rows 2j of neighbouring pairs
// on different banks);
// region B, the z-buffer of one half of the columns [pos][fl] (stride
// NFH = L1 / 4); in the adjoint, region P, the z pads' partial sums
// [2][R1z][NFH].  The z transforms run over the two column halves in
// turn, so a z pass never overwrites what another thread has yet to read.

// z pass 1 of the forward transform of column half h: item (fl, n1) reads
// rows q = n1 + R1z n2 of column f = h NFH + fl, the real row r = q - left
// separated from its row pair in A (ZERO_PAD 0: r clamped, the replicate pad;
// 1: zero outside the image), and writes Y to B at n1 + R1z k2.
template <int NT, int L1, int L2, int ZERO_PAD, bool OPQ = false>
__device__ __forceinline__ void z_pass1_fwd(const float2* A, float2* B, const float2* tw2, int NPS, int nz, int left,
                                            int h) {
  constexpr int R1z = Split<L2>::R1, R2z = Split<L2>::R2, NFH = L1 / 4;
  for (int it = item_tid<OPQ>(); it < NFH * R1z; it += NT) {
    const int fl = it % NFH, n1 = it / NFH, f = h * NFH + fl;
    const int pf = f, pm = f ? L1 - f : L1 / 2;  // (column 0: f = 0 and f = L1/2 packed)
    float2 a[R2z];
#pragma unroll
    for (int n2 = 0; n2 < R2z; ++n2) {
      int r = n1 + R1z * n2 - left;
      float2 val = make_float2(0.f, 0.f);
      if (!ZERO_PAD || (r >= 0 && r < nz)) {
        r = min(max(r, 0), nz - 1);
        const float2 zf = A[pf * NPS + (r >> 1)], zm = A[pm * NPS + (r >> 1)];
        val = f ? ((r & 1) ? make_float2(0.5f * (zf.y + zm.y), -0.5f * (zf.x - zm.x))
                           : make_float2(0.5f * (zf.x + zm.x), 0.5f * (zf.y - zm.y)))
                : ((r & 1) ? make_float2(zf.y, zm.y) : make_float2(zf.x, zm.x));
      }
      a[n2] = val;
    }
    dft<R2z, false>(a);
#pragma unroll
    for (int k2 = 0; k2 < R2z; ++k2) B[(n1 + R1z * k2) * NFH + fl] = k2 ? cmul(a[k2], tw2[n1 * k2]) : a[k2];
  }
}

// The spectrum items of z pass 2 and 2': slot s of a thread is item
// it = tid + s NT, in half h = it / (NFH R2z), column fl, k2 (the item's
// R1z values are the frequencies k2 + R2z k1 at B positions k1 + R1z k2).
template <int NT, int L1, int L2, bool OPQ = false>
struct SpecItems {
  static constexpr int R2z = Split<L2>::R2, NFH = L1 / 4;
  static constexpr int PER_HALF = NFH * R2z;
  static constexpr int SLOTS = (2 * PER_HALF + NT - 1) / NT;
  // false when slot s is not an item of half h (decided at compile time
  // when a half is whole slots)
  __device__ static __forceinline__ bool at(int s, int h, int* fl, int* k2) {
    int rem;
    if constexpr (PER_HALF % NT == 0) {
      if (s / (PER_HALF / NT) != h) return false;
      rem = item_tid<OPQ>() + (s - h * (PER_HALF / NT)) * NT;
    } else {
      const int it = item_tid<OPQ>() + s * NT;
      if (it / PER_HALF != h) return false;
      rem = it - h * PER_HALF;
    }
    *fl = rem % NFH;
    *k2 = rem / NFH;
    return true;
  }
};

// x passes of the forward transform of a plane's row pairs into A: pass 1
// gathers row pair j at padded positions n1 + R1x n2 straight from global
// memory (LOAD(p, j) gives the pair's two values at padded position p), pass 2
// completes the transform.  Items of pass 1 are (n1, j) with n1 fastest, so a
// wave reads consecutive x.
template <int NT, int L1, bool OPQ = false, class Load>
__device__ __forceinline__ void x_passes_fwd(float2* A, const float2* tw1, int NPS, int NP, Load load) {
  constexpr int R1x = Split<L1>::R1, R2x = Split<L1>::R2;
  for (int it = item_tid<OPQ>(); it < NP * R1x; it += NT) {
    const int n1 = it % R1x, j = it / R1x;
    float2 a[R2x];
#pragma unroll
    for (int n2 = 0; n2 < R2x; ++n2) a[n2] = load(n1 + R1x * n2, j);
    dft<R2x, false>(a);
#pragma unroll
    for (int k2 = 0; k2 < R2x; ++k2) A[(n1 + R1x * k2) * NPS + j] = k2 ? cmul(a[k2], tw1[n1 * k2]) : a[k2];
  }
  __syncthreads();
  // pass 2 leaves frequency f at position f (natural order, read all items
  // first): the z passes then read neighbouring frequencies from
  // neighbouring rows of A, NPS (odd) float2 apart, on distinct LDS banks
  constexpr int RA = kReadAll * kNTF / NT;
  float2 xr[RA][R1x];
#pragma unroll
  for (int s = 0; s < RA; ++s) {
    const int it = item_tid<OPQ>() + s * NT;
    if (it < NP * R2x) {
      const int j = it % NP, k2 = it / NP;
#pragma unroll
      for (int n1 = 0; n1 < R1x; ++n1) xr[s][n1] = A[(n1 + R1x * k2) * NPS + j];
      dft<R1x, false>(xr[s]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < RA; ++s) {
    const int it = item_tid<OPQ>() + s * NT;
    if (it < NP * R2x) {
      const int j = it % NP, k2 = it / NP;
#pragma unroll
      for (int k1 = 0; k1 < R1x; ++k1) A[(k2 + R2x * k1) * NPS + j] = xr[s][k1];
    }
  }
  __syncthreads();
}

// z pass 1' of column half h: item (fl, n1) transforms B back along z; rows
// r = q - left inside the image go to I [r][f] in A; with PADS the rows above
// and below the image are summed per item into P [lo / hi][n1][fl].
template <int NT, int L1, int L2, bool PADS, bool OPQ = false>
__device__ __forceinline__ void z_pass1_inv(float2* A, const float2* B, float2* P, int nz, int left, int h) {
  constexpr int R1z = Split<L2>::R1, R2z = Split<L2>::R2, NIS = L1 / 2 + 1, NFH = L1 / 4;
  for (int it = item_tid<OPQ>(); it < NFH * R1z; it += NT) {
    const int fl = it % NFH, n1 = it / NFH, f = h * NFH + fl;
    float2 a[R2z];
#pragma unroll
    for (int k2 = 0; k2 < R2z; ++k2) a[k2] = B[(n1 + R1z * k2) * NFH + fl];
    dft<R2z, true>(a);
    float2 lo = make_float2(0.f, 0.f), hi = make_float2(0.f, 0.f);
#pragma unroll
    for (int n2 = 0; n2 < R2z; ++n2) {
      const int r = n1 + R1z * n2 - left;
      if (r < 0) lo = cadd(lo, a[n2]);
      else if (r >= nz) hi = cadd(hi, a[n2]);
      else A[r * NIS + f] = a[n2];
    }
    if (PADS) {
      P[n1 * NFH + fl] = lo;
      P[(R1z + n1) * NFH + fl] = hi;
    }
  }
}

// the z pads folded onto rows 0 and nz - 1 (src/fft_convolve.jl:100-110),
// for the columns of half h, in a fixed order
template <int NT, int L1, int L2, bool OPQ = false>
__device__ __forceinline__ void z_fold(float2* A, const float2* P, int nz, int h) {
  constexpr int R1z = Split<L2>::R1, NIS = L1 / 2 + 1, NFH = L1 / 4;
  for (int fl = item_tid<OPQ>(); fl < NFH; fl += NT) {
    const int f = h * NFH + fl;
    float2 lo = make_float2(0.f, 0.f), hi = make_float2(0.f, 0.f);
#pragma unroll
    for (int m = 0; m < R1z; ++m) {
      lo = cadd(lo, P[m * NFH + fl]);
      hi = cadd(hi, P[(R1z + m) * NFH + fl]);
    }
    A[f] = cadd(A[f], lo);
    A[(nz - 1) * NIS + f] = cadd(A[(nz - 1) * NIS + f], hi);
  }
}

// x pass 2' of the inverse transform from the half spectra I [r][f] in A:
// row pair j at frequency f is X_2j[f] + i X_2j+1[f] (X_r[f] = conj X_r[L1 - f]
// above L1/2; X_r[0] and X_r[L1/2] the real and imaginary parts of I[r][0]).
// Reads all its items (at most kReadAll per thread), then after a barrier
// writes the x-buffer over I.
template <int NT, int L1, bool OPQ = false>
__device__ __forceinline__ void x_pass2_inv(float2* A, const float2* tw1, int NPS, int NP, int nz) {
  constexpr int R1x = Split<L1>::R1, R2x = Split<L1>::R2, NIS = L1 / 2 + 1;
  constexpr int RA = kReadAll * kNTF / NT;  // (the same item bound for every block size)
  float2 xr[RA][R1x];
#pragma unroll
  for (int s = 0; s < RA; ++s) {
    const int it = item_tid<OPQ>() + s * NT;
    if (it < NP * R2x) {
      const int j = it % NP, k2 = it / NP;
#pragma unroll
      for (int k1 = 0; k1 < R1x; ++k1) {
        const int f = k2 + R2x * k1;
        const bool upper = f > L1 / 2, nyq = f == L1 / 2;
        const int fh = upper ? L1 - f : nyq ? 0 : f;
        float2 x0 = A[(2 * j) * NIS + fh];
        float2 x1 = 2 * j + 1 < nz ? A[(2 * j + 1) * NIS + fh] : make_float2(0.f, 0.f);
        if (upper) {
          x0 = cconj(x0);
          x1 = cconj(x1);
        } else if (f == 0) {
          x0.y = 0.f;
          x1.y = 0.f;
        } else if (nyq) {
          x0 = make_float2(x0.y, 0.f);
          x1 = make_float2(x1.y, 0.f);
        }
        xr[s][k1] = make_float2(x0.x - x1.y, x0.y + x1.x);
      }
      dft<R1x, true>(xr[s]);
#pragma unroll
      for (int n1 = 1; n1 < R1x; ++n1) xr[s][n1] = cmulc(xr[s][n1], tw1[n1 * k2]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < RA; ++s) {
    const int it = item_tid<OPQ>() + s * NT;
    if (it < NP * R2x) {
      const int j = it % NP, k2 = it / NP;
#pragma unroll
      for (int n1 = 0; n1 < R1x; ++n1) A[(n1 + R1x * k2) * NPS + j] = xr[s][n1];
    }
  }
  __syncthreads();
}

// --- forward -----------------------------------------------------------------
// grid (nchunk, nb), block NT; LDS: twiddles, W, A, B (boff float2 after A).
template <int L1, int L2>
__global__ void __launch_bounds__(kNTF) k_psf_fwd_fft(const float* __restrict__ work, const float* __restrict__ spec,
                                                     int vstride, const int* __restrict__ vids, int Yc,
                                                     float* __restrict__ part, Geo g, int up, int left, int NPS,
                                                     int boff, const float2* __restrict__ twg) {
  constexpr int R1x = Split<L1>::R1, R2x = Split<L1>::R2, R1z = Split<L2>::R1, R2z = Split<L2>::R2;
  constexpr int NF = L1 / 2 + 1, NFH = L1 / 4, NT = kNTF;
  // the opaque item index (as in the adjoint) for the largest planes, where the
  // hoisted per-slot offsets spill (69 VGPRs at 256 x 128: 0.81 ms per 64-view
  // launch, 0.45 ms without); smaller planes do not spill and are faster
  // without it (128 x 64: 92 against 102 us)
  constexpr bool OPQ = L1 * L2 >= 32768;
  using SI = SpecItems<NT, L1, L2, OPQ>;
  extern __shared__ float2 smf[];
  float2* tw1 = smf;
  float2* tw2 = smf + L1;
  float2* W = smf + L1 + L2;
  float2* A = W + L2;
  float2* B = A + boff;
  const int tid = threadIdx.x;
  const int nx = g.nx, ny = g.ny, nz = g.nz, NP = (nz + 1) / 2;
  const int b = blockIdx.y, c = blockIdx.x, v = vids[b];
  const int y0 = c * Yc, y1 = min(ny, y0 + Yc);
  const int S = nx * ny;  // (offsets within one volume fit an int)
  const float* wv = work + (long)S * nz * b;
  for (int m = tid; m < L1 + L2; m += NT) smf[m] = twg[m];
  for (int m = tid; m < L2; m += NT) W[m] = make_float2(0.f, 0.f);
  float2 acc[SI::SLOTS][R1z];
#pragma unroll
  for (int s = 0; s < SI::SLOTS; ++s)
#pragma unroll
    for (int k = 0; k < R1z; ++k) acc[s][k] = make_float2(0.f, 0.f);
  __syncthreads();
  for (int y = y0; y < y1; ++y) {
    const int oy = nx * y;  // element (x, z) of the plane at wv[oy + x + S z]
    x_passes_fwd<NT, L1>(A, tw1, NPS, NP, [&](int p, int j) {
      const int x = min(max(p - up, 0), nx - 1), z0 = 2 * j;
      const int o = oy + x + S * z0;
      return make_float2(wv[o], z0 + 1 < nz ? wv[o + S] : 0.f);
    });
    const float* hy = spec + (long)(v * vstride + y) * L2 * NF;
    for (int h = 0; h < 2; ++h) {
      z_pass1_fwd<NT, L1, L2, 0, OPQ>(A, B, tw2, NPS, nz, left, h);
      __syncthreads();
      // z pass 2, times the plane's kernel spectrum, summed over the chunk
#pragma unroll
      for (int s = 0; s < SI::SLOTS; ++s) {
        int fl, k2;
        if (!SI::at(s, h, &fl, &k2)) continue;
        float2 a[R1z];
#pragma unroll
        for (int n1 = 0; n1 < R1z; ++n1) a[n1] = B[(n1 + R1z * k2) * NFH + fl];
        dft<R1z, false>(a);
        const int f = h * NFH + fl;
        // (Hm loaded by every item, outside the packed column's branch: the
        // items of a wave share k2, so the loads share one address)
        float hm[R1z];
#pragma unroll
        for (int k1 = 0; k1 < R1z; ++k1) {
          const float hv = hy[(k1 + R1z * k2) * NF + f];
          hm[k1] = hy[(k1 + R1z * k2) * NF + L1 / 2];
          acc[s][k1] = make_float2(fmaf(a[k1].x, hv, acc[s][k1].x), fmaf(a[k1].y, hv, acc[s][k1].y));
        }
        if (f == 0) {  // the packed column: W[g] += P^[g] Hm[g]
#pragma unroll
          for (int k1 = 0; k1 < R1z; ++k1) {
            float2& w = W[k2 + R2z * k1];
            w = make_float2(fmaf(a[k1].x, hm[k1], w.x), fmaf(a[k1].y, hm[k1], w.y));
          }
        }
      }
      __syncthreads();  // (B is rewritten next)
    }
  }
  // the packed column's sums: + conj(W[-g]) (W complete after the last barrier)
#pragma unroll
  for (int s = 0; s < SI::SLOTS; ++s) {
    int fl, k2;
    if (!SI::at(s, 0, &fl, &k2) || fl != 0) continue;
#pragma unroll
    for (int k1 = 0; k1 < R1z; ++k1) acc[s][k1] = cadd(acc[s][k1], cconj(W[(L2 - k2 - R2z * k1) & (L2 - 1)]));
  }
  // inverse, column half by column half: z pass 2' from the sums, then z
  // pass 1' into the rows' half spectra
  for (int h = 0; h < 2; ++h) {
#pragma unroll
    for (int s = 0; s < SI::SLOTS; ++s) {
      int fl, k2;
      if (!SI::at(s, h, &fl, &k2)) continue;
      dft<R1z, true>(acc[s]);
#pragma unroll
      for (int n1 = 0; n1 < R1z; ++n1)
        B[(n1 + R1z * k2) * NFH + fl] = n1 ? cmulc(acc[s][n1], tw2[n1 * k2]) : acc[s][n1];
    }
    __syncthreads();
    z_pass1_inv<NT, L1, L2, false>(A, B, nullptr, nz, left, h);
    __syncthreads();
  }
  x_pass2_inv<NT, L1>(A, tw1, NPS, NP, nz);
  // x pass 1': the chunk's partial view [z][x], cropped
  float* pv = part + ((long)b * gridDim.x + c) * ((long)nx * nz);
  for (int it = tid; it < NP * R1x; it += NT) {
    const int n1 = it % R1x, j = it / R1x;
    float2 a[R2x];
#pragma unroll
    for (int k2 = 0; k2 < R2x; ++k2) a[k2] = A[(n1 + R1x * k2) * NPS + j];
    dft<R2x, true>(a);
#pragma unroll
    for (int n2 = 0; n2 < R2x; ++n2) {
      const int x = n1 + R1x * n2 - up;
      if (x >= 0 && x < nx) {
        pv[x + nx * (2 * j)] = a[n2].x;
        if (2 * j + 1 < nz) pv[x + nx * (2 * j + 1)] = a[n2].y;
      }
    }
  }
}

// --- adjoint -----------------------------------------------------------------
// grid (nchunk, nb), block NT; LDS: twiddles, W, A, B, P (boff, poff after A).
template <int L1, int L2>
__global__ void __launch_bounds__(kNTA) k_psf_adj_fft(const float* __restrict__ views, const float* __restrict__ spec,
                                                     int vstride, const int* __restrict__ vids, int Yc,
                                                     float* __restrict__ work, const float* __restrict__ atten,
                                                     long atten_vstride, Geo g, int up, int left, int NPS, int boff,
                                                     int poff, const float2* __restrict__ twg) {
  constexpr int R1x = Split<L1>::R1, R2x = Split<L1>::R2, R1z = Split<L2>::R1, R2z = Split<L2>::R2;
  constexpr int NF = L1 / 2 + 1, NFH = L1 / 4, NT = kNTA;
  using SI = SpecItems<NT, L1, L2, true>;
  extern __shared__ float2 smf[];
  float2* tw1 = smf;
  float2* tw2 = smf + L1;
  float2* W = smf + L1 + L2;
  float2* A = W + L2;
  float2* B = A + boff;
  float2* P = A + poff;
  const int tid = threadIdx.x;
  const int nx = g.nx, ny = g.ny, nz = g.nz, NP = (nz + 1) / 2;
  const int b = blockIdx.y, c = blockIdx.x, v = vids[b];
  const int y0 = c * Yc, y1 = min(ny, y0 + Yc);
  const int S = nx * ny;  // (offsets within one volume fit an int)
  const float* vin = views + (long)b * nx * nz;
  float* wout = work + (long)S * nz * b;
  const float* av = atten ? atten + atten_vstride * v : nullptr;
  for (int m = tid; m < L1 + L2; m += NT) smf[m] = twg[m];
  __syncthreads();
  // the zero-padded view's spectrum, kept in registers for the chunk
  x_passes_fwd<NT, L1, true>(A, tw1, NPS, NP, [&](int p, int j) {
    const int x = p - up, z0 = 2 * j;
    if (x < 0 || x >= nx) return make_float2(0.f, 0.f);
    return make_float2(vin[x + nx * z0], z0 + 1 < nz ? vin[x + nx * (z0 + 1)] : 0.f);
  });
  float2 vs[SI::SLOTS][R1z];
  for (int h = 0; h < 2; ++h) {
    z_pass1_fwd<NT, L1, L2, 1, true>(A, B, tw2, NPS, nz, left, h);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < SI::SLOTS; ++s) {
      int fl, k2;
      if (!SI::at(s, h, &fl, &k2)) continue;
#pragma unroll
      for (int n1 = 0; n1 < R1z; ++n1) vs[s][n1] = B[(n1 + R1z * k2) * NFH + fl];
      dft<R1z, false>(vs[s]);
      if (h == 0 && fl == 0)  // the packed column's P^ for its mirror terms
#pragma unroll
        for (int k1 = 0; k1 < R1z; ++k1) W[k2 + R2z * k1] = vs[s][k1];
    }
    __syncthreads();
  }
  const int nx1 = (NP * R1x + NT - 1) / NT;  // x pass 1' rounds (uniform over the block)
  for (int y = y0; y < y1; ++y) {
    const float* hy = spec + (long)(v * vstride + y) * L2 * NF;
    for (int h = 0; h < 2; ++h) {
      // z pass 2' of the view spectrum times the plane's kernel spectrum
#pragma unroll
      for (int s = 0; s < SI::SLOTS; ++s) {
        int fl, k2;
        if (!SI::at(s, h, &fl, &k2)) continue;
        const int f = h * NFH + fl;
        float2 a[R1z];
        float hm[R1z];  // (loaded by every item, as in the forward)
#pragma unroll
        for (int k1 = 0; k1 < R1z; ++k1) {
          const float hv = hy[(k1 + R1z * k2) * NF + f];
          hm[k1] = hy[(k1 + R1z * k2) * NF + L1 / 2];
          a[k1] = make_float2(vs[s][k1].x * hv, vs[s][k1].y * hv);
        }
        if (f == 0)  // the packed column: + conj(P^[-g]) Hm[g]
#pragma unroll
          for (int k1 = 0; k1 < R1z; ++k1) {
            const float2 m = W[(L2 - k2 - R2z * k1) & (L2 - 1)];
            a[k1] = make_float2(fmaf(m.x, hm[k1], a[k1].x), fmaf(-m.y, hm[k1], a[k1].y));
          }
        dft<R1z, true>(a);
#pragma unroll
        for (int n1 = 0; n1 < R1z; ++n1) B[(n1 + R1z * k2) * NFH + fl] = n1 ? cmulc(a[n1], tw2[n1 * k2]) : a[n1];
      }
      __syncthreads();
      z_pass1_inv<NT, L1, L2, true, true>(A, B, P, nz, left, h);
      __syncthreads();
      z_fold<NT, L1, L2, true>(A, P, nz, h);
      __syncthreads();
    }
    x_pass2_inv<NT, L1, true>(A, tw1, NPS, NP, nz);
    // x pass 1', the x pads folded onto x = 0 and nx - 1 (sums over the R1x
    // lanes of a row pair, src/fft_convolve.jl:112-122), times the
    // attenuation, into the work volume (the item index opaque as in the
    // passes: otherwise the per-position edge and crop masks of both rounds
    // are hoisted out of the plane loop, about 130 SGPRs spilled)
    for (int rd = 0; rd < nx1; ++rd) {
      const int it = item_tid<true>() + rd * NT;
      const bool ok = it < NP * R1x;
      const int n1 = it % R1x, j = ok ? it / R1x : 0;
      float2 a[R2x];
#pragma unroll
      for (int k2 = 0; k2 < R2x; ++k2) a[k2] = ok ? A[(n1 + R1x * k2) * NPS + j] : make_float2(0.f, 0.f);
      dft<R2x, true>(a);
      float2 lo = make_float2(0.f, 0.f), hi = make_float2(0.f, 0.f);
#pragma unroll
      for (int n2 = 0; n2 < R2x; ++n2) {
        const int p = n1 + R1x * n2;
        if (p < up) lo = cadd(lo, a[n2]);
        else if (p >= up + nx) hi = cadd(hi, a[n2]);
      }
#pragma unroll
      for (int m = 1; m < R1x; m <<= 1) {
        lo.x += __shfl_xor(lo.x, m, 64);
        lo.y += __shfl_xor(lo.y, m, 64);
        hi.x += __shfl_xor(hi.x, m, 64);
        hi.y += __shfl_xor(hi.y, m, 64);
      }
      if (ok) {
        const int z0 = 2 * j;
        const bool two = z0 + 1 < nz;
        const int ob = nx * y + S * z0 + n1 - up;  // + R1x n2: position n1 + R1x n2
        // the folds and the attenuation as selects and unconditional loads
        // (pad positions read a clamped in-volume element and are not
        // stored; a single last row reads its own row twice), so that all
        // the loads issue together: 0.84 -> 0.70 ms per 64-view launch at
        // 256 x 128 against per-element branches
#pragma unroll
        for (int n2 = 0; n2 < R2x; ++n2) {
          const int x = n1 + R1x * n2 - up;
          a[n2] = cadd(a[n2], x == 0 ? lo : x == nx - 1 ? hi : make_float2(0.f, 0.f));
        }
        if (av) {
          const int ol = nx * y + S * z0, oh = two ? ol + S : ol;
#pragma unroll
          for (int n2 = 0; n2 < R2x; ++n2) {
            const int x = min(max(n1 + R1x * n2 - up, 0), nx - 1);
            a[n2].x *= av[ol + x];
            a[n2].y *= av[oh + x];
          }
        }
#pragma unroll
        for (int n2 = 0; n2 < R2x; ++n2) {
          const int x = n1 + R1x * n2 - up;
          if (x < 0 || x >= nx) continue;
          wout[ob + R1x * n2] = a[n2].x;
          if (two) wout[ob + R1x * n2 + S] = a[n2].y;
        }
      }
    }
    __syncthreads();  // (the next plane's z pass 2' rewrites B and A)
  }
}

// --- kernel spectra ------------------------------------------------------------
// spec[plane][pos][f] = H[f][g(pos)] / (L1 L2), g(pos) = pos / R1z + R2z (pos % R1z),
// except columns 0 and L1/2: (H[0] + H[L1/2]) / 2 and (H[0] - H[L1/2]) / 2,
// for psfs [plane][pz][px] (each plane [px, pz] column-major), in double:
// C[f][b] = sum_a h[a,b] cos(2 pi f (a - cx) / L1), S likewise with sin, then
// H = sum_b C cos(2 pi g (b - cz) / L2) - S sin(...).  grid (planes), block 256.
__global__ void __launch_bounds__(256) k_psf_spectrum(const float* __restrict__ psfs, int px, int pz, int L1,
                                                      int L2, int R1z, float* __restrict__ spec) {
  extern __shared__ double sdb[];
  const int NF = L1 / 2 + 1, R2z = L2 / R1z;
  double* cx = sdb;
  double* sx = cx + L1;
  double* cz = sx + L1;
  double* sz = cz + L2;
  double* C = sz + L2;
  double* Sn = C + (long)NF * pz;
  const float* h = psfs + (long)blockIdx.x * px * pz;
  for (int m = threadIdx.x; m < L1; m += 256) sincospi(2.0 * m / L1, &sx[m], &cx[m]);
  for (int m = threadIdx.x; m < L2; m += 256) sincospi(2.0 * m / L2, &sz[m], &cz[m]);
  __syncthreads();
  const int hx = (px - 1) / 2, hz = (pz - 1) / 2;
  for (int k = threadIdx.x; k < NF * pz; k += 256) {
    const int f = k % NF, bb = k / NF;
    double c = 0, s = 0;
    for (int a = 0; a < px; ++a) {
      const int m = ((f * (a - hx)) % L1 + L1) % L1;
      const double hv = h[a + px * bb];
      c += hv * cx[m];
      s += hv * sx[m];
    }
    C[k] = c;
    Sn[k] = s;
  }
  __syncthreads();
  const double scale = 1.0 / ((double)L1 * L2);
  float* out = spec + (long)blockIdx.x * L2 * NF;
  for (int k = threadIdx.x; k < L2 * NF; k += 256) {
    const int f = k % NF, pos = k / NF;
    if (f == L1 / 2) continue;  // (written with column 0)
    const int gz = pos / R1z + R2z * (pos % R1z);
    double acc = 0, accn = 0;
    for (int bb = 0; bb < pz; ++bb) {
      const int m = ((gz * (bb - hz)) % L2 + L2) % L2;
      acc += C[bb * NF + f] * cz[m] - Sn[bb * NF + f] * sz[m];
      if (f == 0) accn += C[bb * NF + L1 / 2] * cz[m] - Sn[bb * NF + L1 / 2] * sz[m];
    }
    if (f == 0) {
      out[k] = (float)(0.5 * (acc + accn) * scale);
      out[k + L1 / 2] = (float)(0.5 * (acc - accn) * scale);
    } else {
      out[k] = (float)(acc * scale);
    }
  }
}

std::mutex g_fft_lds_mu;
std::unordered_map<const void*, size_t> g_fft_lds;
template <class K>
void fft_allow_lds(K kernel, size_t bytes) {
  if (bytes <= 65536) return;
  const void* f = reinterpret_cast<const void*>(kernel);
  std::lock_guard<std::mutex> lk(g_fft_lds_mu);
  size_t& done = g_fft_lds[f];
  if (bytes <= done) return;
  if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess) done = bytes;
  else (void)hipGetLastError();
}

int pow2_up(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

#define SR_FFT_SWITCH(L1V, L2V, ...)                                                      \
  switch ((L1V) * 1000 + (L2V)) {                                                         \
    case 16016: { constexpr int L1 = 16, L2 = 16; __VA_ARGS__; } break;                 \
    case 16032: { constexpr int L1 = 16, L2 = 32; __VA_ARGS__; } break;                 \
    case 16064: { constexpr int L1 = 16, L2 = 64; __VA_ARGS__; } break;                 \
    case 32016: { constexpr int L1 = 32, L2 = 16; __VA_ARGS__; } break;                 \
    case 32032: { constexpr int L1 = 32, L2 = 32; __VA_ARGS__; } break;                 \
    case 32064: { constexpr int L1 = 32, L2 = 64; __VA_ARGS__; } break;                 \
    case 32128: { constexpr int L1 = 32, L2 = 128; __VA_ARGS__; } break;                \
    case 64016: { constexpr int L1 = 64, L2 = 16; __VA_ARGS__; } break;                 \
    case 64032: { constexpr int L1 = 64, L2 = 32; __VA_ARGS__; } break;                 \
    case 64064: { constexpr int L1 = 64, L2 = 64; __VA_ARGS__; } break;                 \
    case 64128: { constexpr int L1 = 64, L2 = 128; __VA_ARGS__; } break;                \
    case 64256: { constexpr int L1 = 64, L2 = 256; __VA_ARGS__; } break;                \
    case 128016: { constexpr int L1 = 128, L2 = 16; __VA_ARGS__; } break;               \
    case 128032: { constexpr int L1 = 128, L2 = 32; __VA_ARGS__; } break;               \
    case 128064: { constexpr int L1 = 128, L2 = 64; __VA_ARGS__; } break;               \
    case 128128: { constexpr int L1 = 128, L2 = 128; __VA_ARGS__; } break;              \
    case 128256: { constexpr int L1 = 128, L2 = 256; __VA_ARGS__; } break;              \
    case 256016: { constexpr int L1 = 256, L2 = 16; __VA_ARGS__; } break;               \
    case 256032: { constexpr int L1 = 256, L2 = 32; __VA_ARGS__; } break;               \
    case 256064: { constexpr int L1 = 256, L2 = 64; __VA_ARGS__; } break;               \
    case 256128: { constexpr int L1 = 256, L2 = 128; __VA_ARGS__; } break;              \
    default: return hipErrorInvalidValue;                                                 \
  }

}  // namespace

bool psf_fft_plan(const Geo& g, FftPlan* fp) {
  FftPlan f{};
  f.L1 = pow2_up(g.nx + g.px - 1);
  f.L2 = pow2_up(g.nz + g.pz - 1);
  if (f.L1 < 16 || f.L2 < 16 || f.L1 > 256 || f.L2 > 256 || (f.L1 > 128 && f.L2 > 128)) return false;
  if (f.L1 < 64 && f.L2 > 64 + 64 * (f.L1 == 32)) return false;
  f.up = (f.L1 - g.nx + 1) / 2;  // ceil((L - n) / 2), src/helper.jl:9-12
  f.left = (f.L2 - g.nz + 1) / 2;
  const int NP = (g.nz + 1) / 2;
  f.NPS = NP | 1;  // (odd: rows of A start on different LDS banks)
  const int NIS = f.L1 / 2 + 1, NFH = f.L1 / 4;
  const int R1z = split_r1(f.L2), R2x = f.L1 / split_r1(f.L1);
  const long a = std::max((long)f.L1 * f.NPS, (long)g.nz * NIS);  // x-buffer / rows' half spectra
  const long bz = (long)f.L2 * NFH;                              // z-buffer of a column half
  f.boff = (int)a;
  f.poff = (int)(a + bz);
  f.lds_fwd = sizeof(float2) * (f.L1 + 2L * f.L2 + a + bz);  // twiddles, W, A, B
  f.lds_adj = f.lds_fwd + sizeof(float2) * 2 * R1z * NFH;
  // the x pass 2' reads at most kReadAll items per thread before writing
  if ((long)NP * R2x > (long)kReadAll * kNTF) return false;
  if (f.lds_adj > kMaxLds) return false;
  if (fp) *fp = f;
  return true;
}

void psf_fft_twiddles(const FftPlan& f, std::vector<float2>& tw) {
  tw.resize(f.L1 + f.L2);
  for (int m = 0; m < f.L1; ++m) {
    const double a = -2.0 * kPiD * m / f.L1;
    tw[m] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  for (int m = 0; m < f.L2; ++m) {
    const double a = -2.0 * kPiD * m / f.L2;
    tw[f.L1 + m] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
}

hipError_t launch_psf_spectrum(const float* psfs, int nplanes, const Geo& g, const FftPlan& f, float* spec,
                               hipStream_t s) {
  if (nplanes <= 0) return hipSuccess;
  const int NF = f.L1 / 2 + 1;
  const size_t lds = sizeof(double) * (2 * f.L1 + 2 * f.L2 + 2L * NF * g.pz);
  if (lds > kMaxLds) return hipErrorInvalidConfiguration;
  fft_allow_lds(k_psf_spectrum, lds);
  k_psf_spectrum<<<nplanes, 256, lds, s>>>(psfs, g.px, g.pz, f.L1, f.L2, split_r1(f.L2), spec);
  return hipGetLastError();
}

hipError_t launch_psf_fwd_fft(const float* work, const float* spec, int vstride, const int* vids, int nb, float* part,
                              int nchunk, int* nchunk_used, const Geo& g, const FftPlan& f, const float2* tw,
                              hipStream_t s) {
  const int Yc = (g.ny + nchunk - 1) / nchunk;
  nchunk = (g.ny + Yc - 1) / Yc;
  if (nchunk_used) *nchunk_used = nchunk;
  if (nb <= 0) return hipSuccess;
  SR_FFT_SWITCH(f.L1, f.L2, {
    fft_allow_lds(k_psf_fwd_fft<L1, L2>, f.lds_fwd);
    k_psf_fwd_fft<L1, L2><<<dim3(nchunk, nb), kNTF, f.lds_fwd, s>>>(work, spec, vstride, vids, Yc, part, g, f.up,
                                                                   f.left, f.NPS, f.boff, tw);
  });
  return hipGetLastError();
}

hipError_t launch_psf_adj_fft(const float* views, const float* spec, int vstride, const int* vids, int nb,
                              float* work, const float* atten, long atten_vstride, int nchunk, const Geo& g,
                              const FftPlan& f, const float2* tw, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  const int Yc = (g.ny + nchunk - 1) / nchunk;
  nchunk = (g.ny + Yc - 1) / Yc;
  SR_FFT_SWITCH(f.L1, f.L2, {
    fft_allow_lds(k_psf_adj_fft<L1, L2>, f.lds_adj);
    k_psf_adj_fft<L1, L2><<<dim3(nchunk, nb), kNTA, f.lds_adj, s>>>(views, spec, vstride, vids, Yc, work, atten,
                                                                   atten_vstride, g, f.up, f.left, f.NPS, f.boff,
                                                                   f.poff, tw);
  });
  return hipGetLastError();
}

}  // namespace sr
