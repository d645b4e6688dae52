// kernels.hip -- CDNA4 (gfx950) kernels of the SPECT projector path.
//
// Compiled with -ffp-contract=off: every a*b+c below rounds twice, exactly as
// the reference's Float32 code does, unless fmaf() is written out (only the
// PSF convolutions, whose reference arithmetic is an FFT, use fmaf).
#include "kernels.h"

namespace sr {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// ===========================================================================
// Rotation, single-pass bilinear (:two)
// ===========================================================================

// value of the zero-padded slice at image index (ix, iy)
__device__ __forceinline__ float slice_tap(const float* s, int N, int ix, int iy) {
  return (ix >= 0 && ix < N && iy >= 0 && iy < N) ? s[ix + N * iy] : 0.f;
}

// Sample coordinates of output pixel (x, y) (0-based, crop region) under
// R = c + rotate(2pi - theta, I - c), in padded 1-based coordinates, applied
// in Float32 as ((xx*i1 + xy*i2) + x0) (src/rotatez.jl:348-357).
struct BiCoef {
  int j1, j2;          // 0-based image index of the lower taps
  float w11, w12, w21, w22;
};
__device__ __forceinline__ BiCoef bilinear_coef(const RotView& r, int pad, int x, int y) {
  const float f1 = (float)(x + pad + 1), f2 = (float)(y + pad + 1);
  const float x1 = (r.xx * f1 + r.xy * f2) + r.x0;
  const float x2 = (r.yx * f1 + r.yy * f2) + r.y0;
  const float fl1 = floorf(x1), fl2 = floorf(x2);
  const float t1 = x1 - fl1, t2 = x2 - fl2;
  BiCoef c;
  c.j1 = (int)fl1 - pad - 1;
  c.j2 = (int)fl2 - pad - 1;
  c.w11 = 1.f - t1;
  c.w12 = t1;
  c.w21 = 1.f - t2;
  c.w22 = t2;
  return c;
}

// Taps that fall outside the image lie in the zero pad (the Flat boundary of
// the reference's LinearSpline clamps into that pad too, so both give 0).
__device__ __forceinline__ float bilinear_two(const float* s, int N, int pad, const RotView& r, int x, int y) {
  const BiCoef c = bilinear_coef(r, pad, x, y);
  const float a = slice_tap(s, N, c.j1, c.j2) * c.w11 + slice_tap(s, N, c.j1 + 1, c.j2) * c.w12;
  const float b = slice_tap(s, N, c.j1, c.j2 + 1) * c.w11 + slice_tap(s, N, c.j1 + 1, c.j2 + 1) * c.w12;
  return a * c.w21 + b * c.w22;
}

// grid (nz, ceil(nb/vpg)), block 256 = 64 (x) x 4 (y).  The image slice z is
// staged once in LDS and rotated for vpg views.
template <bool kLds>
__global__ void __launch_bounds__(256) k_rot2_fwd(const float* __restrict__ img, float* __restrict__ out,
                                                  const float* __restrict__ atten, long atten_vstride,
                                                  const RotView* __restrict__ rv, const int* __restrict__ vids,
                                                  int nb, int vpg, Geo g) {
  extern __shared__ float s_img[];
  const int N = g.nx, z = blockIdx.x;
  const long S = (long)N * N, V = S * g.nz;
  const float* gsrc = img + S * z;
  if (kLds) {
    for (long k = threadIdx.x; k < S; k += blockDim.x) s_img[k] = gsrc[k];
    __syncthreads();
  }
  const float* src = kLds ? s_img : gsrc;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int b0 = blockIdx.y * vpg, b1 = min(nb, b0 + vpg);
  for (int b = b0; b < b1; ++b) {
    const int v = vids[b];
    const RotView r = rv[v];
    float* o = out + V * b + S * z;
    const float* a = atten ? atten + atten_vstride * v + S * z : nullptr;
    for (int y = ty; y < N; y += 4)
      for (int x = tx; x < N; x += 64) {
        const long k = x + (long)N * y;
        const float val = (r.kind == 0) ? src[k] : bilinear_two(src, N, g.pad, r, x, y);
        o[k] = a ? val * a[k] : val;
      }
  }
}

// Adjoint: scatter of each work pixel into an LDS accumulator of the image
// slice (ds_add_f32), over the vpg views of the group.
__global__ void __launch_bounds__(256) k_rot2_adj(const float* __restrict__ work, float* __restrict__ part,
                                                  const RotView* __restrict__ rv, const int* __restrict__ vids,
                                                  int nb, int vpg, Geo g) {
  extern __shared__ float acc[];
  const int N = g.nx, z = blockIdx.x, grp = blockIdx.y;
  const long S = (long)N * N, V = S * g.nz;
  for (long k = threadIdx.x; k < S; k += blockDim.x) acc[k] = 0.f;
  __syncthreads();
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int b0 = grp * vpg, b1 = min(nb, b0 + vpg);
  for (int b = b0; b < b1; ++b) {
    const RotView r = rv[vids[b]];
    const float* src = work + V * b + S * z;
    for (int y = ty; y < N; y += 4)
      for (int x = tx; x < N; x += 64) {
        const float p = src[x + (long)N * y];
        const BiCoef c = bilinear_coef(r, g.pad, x, y);
        const bool in1a = c.j1 >= 0 && c.j1 < N, in1b = c.j1 + 1 >= 0 && c.j1 + 1 < N;
        const bool in2a = c.j2 >= 0 && c.j2 < N, in2b = c.j2 + 1 >= 0 && c.j2 + 1 < N;
        if (in1a && in2a) atomicAdd(&acc[c.j1 + N * c.j2], (c.w11 * c.w21) * p);
        if (in1b && in2a) atomicAdd(&acc[c.j1 + 1 + N * c.j2], (c.w12 * c.w21) * p);
        if (in1a && in2b) atomicAdd(&acc[c.j1 + N * (c.j2 + 1)], (c.w11 * c.w22) * p);
        if (in1b && in2b) atomicAdd(&acc[c.j1 + 1 + N * (c.j2 + 1)], (c.w12 * c.w22) * p);
      }
  }
  __syncthreads();
  float* o = part + V * grp + S * z;
  for (long k = threadIdx.x; k < S; k += blockDim.x) o[k] = acc[k];
}

// ===========================================================================
// Rotation, 3-pass 1-D linear interpolation (:one), one slice per block in LDS
// ===========================================================================
constexpr int R1_THREADS = 1024;  // 64 (x) x 16 (y)
constexpr int R1_MX = 4;          // 64*4 >= Np
constexpr int R1_MY = 13;         // 16*13 >= Np (Np <= kMaxNpOne)

// Source position of rot_f90 (adj=false) / rot_f90_adj (adj=true) as a gather,
// B(x,y) = A(sx,sy), src/rotatez.jl:155-230 (0-based):
//   rotl90: B(x,y) = A(y, Np-1-x);  rotr90: B(x,y) = A(Np-1-y, x);
//   rot180: B(x,y) = A(Np-1-x, Np-1-y); the adjoint swaps rotl90 and rotr90.
__device__ __forceinline__ void rotf90_src(int m, bool adj, int Np, int x, int y, int& sx, int& sy) {
  const int mm = adj ? ((4 - m) & 3) : m;
  if (mm == 0) { sx = x; sy = y; }
  else if (mm == 1) { sx = y; sy = Np - 1 - x; }
  else if (mm == 2) { sx = Np - 1 - x; sy = Np - 1 - y; }
  else { sx = Np - 1 - y; sy = x; }
}

// One shear pass over the Np x Np LDS matrix L (column-major, L[x + Np*y]),
// in place.  Row table: tb[i] = base b, tw[2i], tw[2i+1] = (C_odd, C_even) of
// linearinterp! (src/rotatez.jl:21-51): output k (0-based) reads
//   J1 = clamp(b+k, 0, Np-1) with weight w1 and J2 = clamp(b+k+1, 0, Np-1)
// with weight w2, summed as (0 + w1*x[J1]) + w2*x[J2].  kAlongX: interpolate
// along dim 1 with the table indexed by column y (rotate_x!); else along
// dim 2 indexed by row x (rotate_y!).  kAdj: the transpose, evaluated as a
// gather in the reference's scatter order (rows ascending, J1 before J2).
template <bool kAlongX>
__device__ __forceinline__ float shear_fwd_at(const float* L, int Np, const int* tb, const float* tw, int x, int y) {
  const int i = kAlongX ? y : x;  // table row
  const int k = kAlongX ? x : y;  // position along the interpolated axis
  const int b = tb[i];
  const float w1 = tw[2 * i], w2 = tw[2 * i + 1];
  const int t1 = clampi(b + k, 0, Np - 1), t2 = clampi(b + k + 1, 0, Np - 1);
  const float v1 = kAlongX ? L[t1 + Np * y] : L[x + Np * t1];
  const float v2 = kAlongX ? L[t2 + Np * y] : L[x + Np * t2];
  return w1 * v1 + w2 * v2;
}

template <bool kAlongX>
__device__ __forceinline__ float shear_adj_at(const float* L, int Np, const int* tb, const float* tw, int x, int y) {
  const int i = kAlongX ? y : x;
  const int j = kAlongX ? x : y;
  const int b = tb[i];
  const float w1 = tw[2 * i], w2 = tw[2 * i + 1];
  auto ld = [&](int k) { return kAlongX ? L[k + Np * y] : L[x + Np * k]; };
  float s = 0.f;
  if (j > 0 && j < Np - 1) {
    const int k2 = j - b - 1, k1 = j - b;  // J2(k2) == j precedes J1(k1) == j
    if (k2 >= 0 && k2 < Np) s = s + w2 * ld(k2);
    if (k1 >= 0 && k1 < Np) s = s + w1 * ld(k1);
  } else if (j == 0) {
    const int kend = min(Np - 1, -b);
    for (int k = 0; k <= kend; ++k) {
      if (b + k <= 0) s = s + w1 * ld(k);
      if (b + k + 1 <= 0) s = s + w2 * ld(k);
    }
  } else {  // j == Np-1
    for (int k = max(0, Np - 2 - b); k < Np; ++k) {
      if (b + k >= Np - 1) s = s + w1 * ld(k);
      if (b + k + 1 >= Np - 1) s = s + w2 * ld(k);
    }
  }
  return s;
}

template <bool kAlongX, bool kAdj>
__device__ __forceinline__ void shear_pass(float* L, int Np, const int* __restrict__ tb, const float* __restrict__ tw) {
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  float r[R1_MY][R1_MX];
#pragma unroll
  for (int ys = 0; ys < R1_MY; ++ys)
#pragma unroll
    for (int xs = 0; xs < R1_MX; ++xs) {
      const int x = tx + 64 * xs, y = ty + 16 * ys;
      if (x < Np && y < Np)
        r[ys][xs] = kAdj ? shear_adj_at<kAlongX>(L, Np, tb, tw, x, y) : shear_fwd_at<kAlongX>(L, Np, tb, tw, x, y);
    }
  __syncthreads();
#pragma unroll
  for (int ys = 0; ys < R1_MY; ++ys)
#pragma unroll
    for (int xs = 0; xs < R1_MX; ++xs) {
      const int x = tx + 64 * xs, y = ty + 16 * ys;
      if (x < Np && y < Np) L[x + Np * y] = r[ys][xs];
    }
  __syncthreads();
}

// grid (nz, nb): imrotate! :one of slice z for view vids[b] (src/rotatez.jl:240-278)
__global__ void __launch_bounds__(R1_THREADS) k_rot1_fwd(const float* __restrict__ img, float* __restrict__ out,
                                                         const float* __restrict__ atten, long atten_vstride,
                                                         const RotView* __restrict__ rv,
                                                         const int* __restrict__ vids, const int* __restrict__ tab_b,
                                                         const float* __restrict__ tab_w, Geo g) {
  extern __shared__ float L[];
  const int N = g.nx, Np = g.Np, pad = g.pad, z = blockIdx.x, b = blockIdx.y;
  const int v = vids[b];
  const RotView r = rv[v];
  const long S = (long)N * N, V = S * g.nz;
  const float* src = img + S * z;
  float* o = out + V * b + S * z;
  const float* a = atten ? atten + atten_vstride * v + S * z : nullptr;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  if (r.kind == 0) {
    for (int y = ty; y < N; y += 16)
      for (int x = tx; x < N; x += 64) {
        const long k = x + (long)N * y;
        o[k] = a ? src[k] * a[k] : src[k];
      }
    return;
  }
  for (int y = ty; y < Np; y += 16)
    for (int x = tx; x < Np; x += 64) {
      int sx, sy;
      rotf90_src(r.m, false, Np, x, y, sx, sy);
      L[x + Np * y] = slice_tap(src, N, sx - pad, sy - pad);
    }
  __syncthreads();
  const int* tbx = tab_b + (long)v * 2 * Np;
  const int* tby = tbx + Np;
  const float* twx = tab_w + (long)v * 4 * Np;
  const float* twy = twx + 2 * Np;
  if (r.kind == 2) {
    shear_pass<true, false>(L, Np, tbx, twx);
    shear_pass<false, false>(L, Np, tby, twy);
  }
  for (int y = ty; y < N; y += 16)
    for (int x = tx; x < N; x += 64) {
      const float val = (r.kind == 2) ? shear_fwd_at<true>(L, Np, tbx, twx, x + pad, y + pad)
                                      : L[(x + pad) + Np * (y + pad)];
      const long k = x + (long)N * y;
      o[k] = a ? val * a[k] : val;
    }
}

// grid (nz, nb): imrotate_adj! :one (src/rotatez.jl:286-322) of work[b]
// slice z into part[b] (one view per block; the views are summed by
// k_img_reduce).
__global__ void __launch_bounds__(R1_THREADS) k_rot1_adj(const float* __restrict__ work, float* __restrict__ part,
                                                         const RotView* __restrict__ rv,
                                                         const int* __restrict__ vids, const int* __restrict__ tab_b,
                                                         const float* __restrict__ tab_w, Geo g) {
  extern __shared__ float L[];
  const int N = g.nx, Np = g.Np, pad = g.pad, z = blockIdx.x, b = blockIdx.y;
  const long S = (long)N * N, V = S * g.nz;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int v = vids[b];
  const RotView r = rv[v];
  const float* src = work + V * b + S * z;
  float* o = part + V * b + S * z;
  if (r.kind == 0) {
    for (int y = ty; y < N; y += 16)
      for (int x = tx; x < N; x += 64) o[x + (long)N * y] = src[x + (long)N * y];
    return;
  }
  for (int y = ty; y < Np; y += 16)
    for (int x = tx; x < Np; x += 64) L[x + Np * y] = slice_tap(src, N, x - pad, y - pad);
  __syncthreads();
  if (r.kind == 2) {
    const int* tbx = tab_b + (long)v * 2 * Np;
    const int* tby = tbx + Np;
    const float* twx = tab_w + (long)v * 4 * Np;
    const float* twy = twx + 2 * Np;
    shear_pass<true, true>(L, Np, tbx, twx);
    shear_pass<false, true>(L, Np, tby, twy);
    shear_pass<true, true>(L, Np, tbx, twx);
  }
  for (int y = ty; y < N; y += 16)
    for (int x = tx; x < N; x += 64) {
      int sx, sy;
      rotf90_src(r.m, true, Np, x + pad, y + pad, sx, sy);
      o[x + (long)N * y] = L[sx + Np * sy];
    }
}

// ===========================================================================
// Attenuation factors in the reference's summation order
// ===========================================================================
// grid (nz, nb), block = N threads (one per x column).  In place: mur[b] holds
// the rotated mu volume on entry and the factors on exit.
// src/project.jl:103-113: e = -0.5*mu_r[y]; e += mu_r[j] for j = 1..y; e *= -dy; exp.
__global__ void __launch_bounds__(1024) k_atten_exact(float* __restrict__ mur, Geo g, float dy) {
  extern __shared__ float s[];
  const int N = g.nx, z = blockIdx.x, b = blockIdx.y;
  const long S = (long)N * N, V = S * g.nz;
  float* p = mur + V * b + S * z;
  for (long k = threadIdx.x; k < S; k += blockDim.x) s[k] = p[k];
  __syncthreads();
  const float mdy = -dy;
  for (int x = threadIdx.x; x < N; x += blockDim.x) {
    for (int y = 0; y < N; ++y) {
      float e = -0.5f * s[x + N * y];
      for (int j = 0; j <= y; ++j) e = e + s[x + N * j];
      e = e * mdy;
      p[x + (long)N * y] = expf(e);
    }
  }
}

// ===========================================================================
// PSF convolution with replicate padding (src/fft_convolve.jl:33-130)
//
// Closed form of the reference's FFT path (no circular wrap since
// L >= n+p-1):  out[x,z] = sum_{a,b} h[a,b] img[clamp(x+hx-a), clamp(z+hz-b)],
// and its transpose.  Taps are zero-padded to NT (a multiple of 8); each
// thread computes strips of 8 outputs from a register window of NT+7 inputs.
// ===========================================================================

// o[k] += sum_a t[a] * w[k + NT-1-a]   (forward: convolution)
template <int NT>
__device__ __forceinline__ void conv8_fwd(const float (&w)[NT + 7], const float* __restrict__ t, float (&o)[8]) {
#pragma unroll
  for (int a = 0; a < NT; ++a) {
    const float ta = t[a];
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = fmaf(ta, w[k + NT - 1 - a], o[k]);
  }
}
// o[k] += sum_a t[a] * w[k + a]   (transpose: correlation)
template <int NT>
__device__ __forceinline__ void conv8_adj(const float (&w)[NT + 7], const float* __restrict__ t, float (&o)[8]) {
#pragma unroll
  for (int a = 0; a < NT; ++a) {
    const float ta = t[a];
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = fmaf(ta, w[k + a], o[k]);
  }
}

constexpr int PSF_MAXS = 4;  // strips per thread per phase

// Strip of SL outputs, taps processed in groups of 8 so that only SL+7
// inputs are live at a time.  ld(i) returns input i of the strip's window.
//   forward (convolution): o[k] += sum_a t[a] * in[k + NT-1-a]
//   transpose (correlation): o[k] += sum_a t[a] * in[k + a]
template <int NT, int SL, class Ld>
__device__ __forceinline__ void conv_strip_fwd(Ld ld, const float* __restrict__ t, float (&o)[SL]) {
#pragma unroll
  for (int g = 0; g < NT; g += 8) {
    const int base = NT - 8 - g;
    float w[SL + 7];
#pragma unroll
    for (int j = 0; j < SL + 7; ++j) w[j] = ld(base + j);
#pragma unroll
    for (int aa = 0; aa < 8; ++aa) {
      const float ta = t[g + aa];
#pragma unroll
      for (int k = 0; k < SL; ++k) o[k] = fmaf(ta, w[k + 7 - aa], o[k]);
    }
  }
}
template <int NT, int SL, class Ld>
__device__ __forceinline__ void conv_strip_adj(Ld ld, const float* __restrict__ t, float (&o)[SL]) {
#pragma unroll
  for (int g = 0; g < NT; g += 8) {
    float w[SL + 7];
#pragma unroll
    for (int j = 0; j < SL + 7; ++j) w[j] = ld(g + j);
#pragma unroll
    for (int aa = 0; aa < 8; ++aa) {
      const float ta = t[g + aa];
#pragma unroll
      for (int k = 0; k < SL; ++k) o[k] = fmaf(ta, w[k + aa], o[k]);
    }
  }
}


__host__ __device__ __forceinline__ int odd_up(int v) { return v | 1; }

// --- forward, separable: work -> per-chunk partial views -------------------
// grid (nchunk, nb), block >= max(#x strips, #z strips) (one strip of SL
// outputs per thread per phase).  Each block convolves the planes y of its
// chunk and keeps the running sum over y in registers; partials are reduced
// by k_view_reduce.  LDS: the plane, replicate-padded along x (the x-pass
// result is written back over it), and the plane's taps.
template <int NT, int SL>
__global__ void __launch_bounds__(1024) k_psf_fwd_sep(const float* __restrict__ work, Taps t,
                                                      const int* __restrict__ vids, int Yc,
                                                      float* __restrict__ part, Geo g) {
  extern __shared__ float L[];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int hx = (g.px - 1) / 2, hz = (g.pz - 1) / 2;
  const int nxs = (nx + SL - 1) / SL, nzs = (nz + SL - 1) / SL;
  const int WX = odd_up(SL * nxs + NT - 1);
  const int LP = NT - 1 - hx;
  const int b = blockIdx.y, c = blockIdx.x, nchunk = gridDim.x;
  const int v = vids[b];
  const int y0 = c * Yc, y1 = min(ny, y0 + Yc);
  const long S = (long)nx * ny, V = S * nz;
  const float* wb = work + V * b;
  float* s_tu = L + nz * WX;
  float* s_tw = s_tu + NT;
  const int tid = threadIdx.x, nthr = blockDim.x;
  // x strip (z fastest) and z strip (x fastest) of this thread
  const bool hasx = tid < nz * nxs, hasz = tid < nx * nzs;
  const int xs_z = tid % nz, xs_x0 = (tid / nz) * SL;
  const int zs_x = tid % nx, zs_z0 = (tid / nx) * SL;

  float acc[SL];
#pragma unroll
  for (int k = 0; k < SL; ++k) acc[k] = 0.f;

  for (int y = y0; y < y1; ++y) {
    const int plane = v * t.vstride + y;
    if (tid < NT) {
      s_tu[tid] = t.u[(long)plane * NT + tid];
      s_tw[tid] = t.w[(long)plane * NT + tid];
    }
    for (int k = tid; k < nz * WX; k += nthr) {
      const int z = k / WX, col = k - z * WX;
      L[k] = wb[clampi(col - LP, 0, nx - 1) + (long)nx * y + S * z];
    }
    __syncthreads();
    float q[SL];
#pragma unroll
    for (int k = 0; k < SL; ++k) q[k] = 0.f;
    if (hasx) {
      const float* row = L + xs_z * WX + xs_x0;
      conv_strip_fwd<NT, SL>([&](int j) { return row[j]; }, s_tu, q);
    }
    __syncthreads();
    if (hasx) {
#pragma unroll
      for (int k = 0; k < SL; ++k) L[xs_z * WX + xs_x0 + k] = q[k];
    }
    __syncthreads();
    if (hasz) {
      const int zb = zs_z0 + hz - (NT - 1);
      conv_strip_fwd<NT, SL>([&](int j) { return L[clampi(zb + j, 0, nz - 1) * WX + zs_x]; }, s_tw, acc);
    }
    __syncthreads();
  }
  if (hasz) {
    float* pb = part + ((long)b * nchunk + c) * nx * nz;
#pragma unroll
    for (int k = 0; k < SL; ++k)
      if (zs_z0 + k < nz) pb[zs_x + (long)nx * (zs_z0 + k)] = acc[k];
  }
}

// --- forward, direct 2-D ----------------------------------------------------
template <int NT, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_psf_fwd_2d(const float* __restrict__ work, Taps t,
                                                      const int* __restrict__ vids, int Yc,
                                                      float* __restrict__ part, Geo g) {
  extern __shared__ float L[];
  const int nx = g.nx, ny = g.ny, nz = g.nz, pz = g.pz;
  const int hx = (g.px - 1) / 2, hz = (pz - 1) / 2;
  const int nxs = (nx + 7) / 8;
  const int WX = odd_up(8 * nxs + NT - 1);
  const int LP = NT - 1 - hx;
  const int NR = nz + pz - 1;  // rows z-hz .. z+hz of every output row
  const int b = blockIdx.y, c = blockIdx.x, nchunk = gridDim.x;
  const int v = vids[b];
  const int y0 = c * Yc, y1 = min(ny, y0 + Yc);
  const long S = (long)nx * ny, V = S * nz;
  const float* wb = work + V * b;
  const int nsx = nz * nxs;

  float acc[PSF_MAXS][8];
#pragma unroll
  for (int i = 0; i < PSF_MAXS; ++i)
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[i][k] = 0.f;

  for (int y = y0; y < y1; ++y) {
    const float* th = t.h + (long)(v * t.vstride + y) * pz * NT;
    // L[r][col] = P[clamp(col-LP), clamp(r-hz)]
    for (int k = threadIdx.x; k < NR * WX; k += BLOCK) {
      const int r = k / WX, col = k - r * WX;
      L[k] = wb[clampi(col - LP, 0, nx - 1) + (long)nx * y + S * clampi(r - hz, 0, nz - 1)];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PSF_MAXS; ++i) {
      const int s = threadIdx.x + i * BLOCK;
      if (s < nsx) {
        const int z = s % nz, x0 = (s / nz) * 8;
        for (int bb = 0; bb < pz; ++bb) {
          // output row z reads input row z+hz-bb -> L row z+2hz-bb
          const float* row = L + (z + 2 * hz - bb) * WX + x0;
          float w[NT + 7];
#pragma unroll
          for (int j = 0; j < NT + 7; ++j) w[j] = row[j];
          conv8_fwd<NT>(w, th + bb * NT, acc[i]);
        }
      }
    }
    __syncthreads();
  }
  float* pb = part + ((long)b * nchunk + c) * nx * nz;
#pragma unroll
  for (int i = 0; i < PSF_MAXS; ++i) {
    const int s = threadIdx.x + i * BLOCK;
    if (s < nsx) {
      const int z = s % nz, x0 = (s / nz) * 8;
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (x0 + k < nx) pb[x0 + k + (long)nx * z] = acc[i][k];
    }
  }
}

// grid (ceil(nx*nz/256), nb): views = sum of chunk partials (+ ML-EM ratio)
__global__ void __launch_bounds__(256) k_view_reduce(const float* __restrict__ part, int nchunk, int P,
                                                     const int* __restrict__ vids, float* __restrict__ out,
                                                     const float* __restrict__ yobs, const float* __restrict__ bkg,
                                                     int epi) {
  const int b = blockIdx.y;
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= P) return;
  const float* p = part + (long)b * nchunk * P + k;
  float s = 0.f;
  for (int c = 0; c < nchunk; ++c) s += p[(long)c * P];
  if (epi == EPI_RATIO) {
    const long o = (long)vids[b] * P + k;
    s = yobs[o] / (s + bkg[o]);  // src/ml-os-em.jl:44, :124-125
  }
  out[(long)b * P + k] = s;
}

// --- adjoint, separable -------------------------------------------------------
// Transpose of the clamped 1-D convolution along one axis, output j of n:
//   interior: sum_a t[a] in[j - h + a] (zero outside)
//   j == 0:   sum_{i=0}^{min(n-1,h)} in[i] * Tail(i+h),   Tail(k) = sum_{a>=k} t[a]
//   j == n-1: sum_{i>=max(0,n-1-h)} in[i] * Head(i+h-n+1), Head(k) = sum_{a<=k} t[a]
//   n == 1:   in[0] * Tail(0)
// (the fold of the replicate pad, src/fft_convolve.jl:99-122).  Tail/Head
// are per-plane tables in LDS.
__device__ __forceinline__ float adj_edge_lo(const float* in, int stride, int n, int h, const float* tail) {
  float s = 0.f;
  const int i1 = min(n - 1, h);
  for (int i = 0; i <= i1; ++i) s = fmaf(in[i * stride], tail[i + h], s);
  return s;
}
__device__ __forceinline__ float adj_edge_hi(const float* in, int stride, int n, int h, const float* head) {
  float s = 0.f;
  for (int i = max(0, n - 1 - h); i < n; ++i) s = fmaf(in[i * stride], head[i + h - n + 1], s);
  return s;
}

// grid (nchunk, nb), one strip of SL outputs per thread per phase.  LDS: the
// view V[z][x] (stride nx), Qp = the z-adjoint of V with zero columns around
// it (stride WQ, data at column offset hx), and the plane's taps.
// kUniformRows: nx is a multiple of 64, so every wave's z strips share z0 and
// the row bounds of the z pass are wave-uniform (scalar).
template <int NT, int SL, bool kUniformRows>
__global__ void __launch_bounds__(1024) k_psf_adj_sep(const float* __restrict__ views, Taps t,
                                                      const int* __restrict__ vids, int Yc,
                                                      float* __restrict__ work, const float* __restrict__ atten,
                                                      long atten_vstride, Geo g) {
  extern __shared__ float L[];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int hx = (g.px - 1) / 2, hz = (g.pz - 1) / 2;
  const int nxs = (nx + SL - 1) / SL, nzs = (nz + SL - 1) / SL;
  const int WQ = odd_up(SL * nxs + NT - 1);
  const int b = blockIdx.y, c = blockIdx.x;
  const int v = vids[b];
  const int y0 = c * Yc, y1 = min(ny, y0 + Yc);
  const long S = (long)nx * ny, V = S * nz;
  const int P = nx * nz;
  const int tid = threadIdx.x, nthr = blockDim.x;
  float* Vs = L;
  float* Qp = L + P;
  float* s_tu = Qp + nz * WQ;  // taps u, w and their tail/head sums
  float* s_tw = s_tu + NT;
  float* s_ut = s_tw + NT;
  float* s_uh = s_ut + NT;
  float* s_wt = s_uh + NT;
  float* s_wh = s_wt + NT;
  const float* vin = views + (long)b * P;
  for (int k = tid; k < P; k += nthr) Vs[k] = vin[k];
  for (int k = tid; k < nz * WQ; k += nthr) Qp[k] = 0.f;
  const bool hasx = tid < nz * nxs, hasz = tid < nx * nzs;
  const int xs_z = tid % nz, xs_x0 = (tid / nz) * SL;
  const int zs_x = tid % nx, zs_z0 = (tid / nx) * SL;
  float* wout = work + V * b;
  const float* av = atten ? atten + atten_vstride * v : nullptr;

  for (int y = y0; y < y1; ++y) {
    const int plane = v * t.vstride + y;
    if (tid < NT) {
      const float* gu = t.u + (long)plane * NT;
      const float* gw = t.w + (long)plane * NT;
      s_tu[tid] = gu[tid];
      s_tw[tid] = gw[tid];
      float ut = 0.f, uh = 0.f, wt = 0.f, wh = 0.f;
      for (int a = tid; a < NT; ++a) {
        ut += gu[a];
        wt += gw[a];
      }
      for (int a = 0; a <= tid; ++a) {
        uh += gu[a];
        wh += gw[a];
      }
      s_ut[tid] = ut;
      s_uh[tid] = uh;
      s_wt[tid] = wt;
      s_wh[tid] = wh;
    }
    __syncthreads();
    float q[SL];
    // z adjoint: strip (x, z0..z0+SL-1) -> Qp[z][hx + x]
    if (hasz) {
#pragma unroll
      for (int k = 0; k < SL; ++k) q[k] = 0.f;
      const int zb = kUniformRows ? __builtin_amdgcn_readfirstlane(zs_z0 - hz) : zs_z0 - hz;
      conv_strip_adj<NT, SL>(
          [&](int j) {
            const int zr = zb + j;
            return (zr >= 0 && zr < nz) ? Vs[zr * nx + zs_x] : 0.f;
          },
          s_tw, q);
      // replicate-pad fold at z = 0 and z = nz-1 (one evaluation each)
      const int khi = nz - 1 - zs_z0;
      const bool lo = zs_z0 == 0, hi = khi < SL;
      float elo = 0.f, ehi = 0.f;
      if (nz == 1) {
        elo = Vs[zs_x] * s_wt[0];
      } else {
        if (lo) elo = adj_edge_lo(Vs + zs_x, nx, nz, hz, s_wt);
        if (hi) ehi = adj_edge_hi(Vs + zs_x, nx, nz, hz, s_wh);
      }
#pragma unroll
      for (int k = 0; k < SL; ++k)
        if (hi && k == khi && nz > 1) q[k] = ehi;
      if (lo) q[0] = elo;
#pragma unroll
      for (int k = 0; k < SL; ++k)
        if (zs_z0 + k < nz) Qp[(zs_z0 + k) * WQ + hx + zs_x] = q[k];
    }
    __syncthreads();
    // x adjoint: strip (z, x0..x0+SL-1)
    if (hasx) {
#pragma unroll
      for (int k = 0; k < SL; ++k) q[k] = 0.f;
      const float* row = Qp + xs_z * WQ + xs_x0;
      conv_strip_adj<NT, SL>([&](int j) { return row[j]; }, s_tu, q);
      const float* qrow = Qp + xs_z * WQ + hx;
      const int khi = nx - 1 - xs_x0;
      const bool lo = xs_x0 == 0, hi = khi < SL;
      float elo = 0.f, ehi = 0.f;
      if (nx == 1) {
        elo = qrow[0] * s_ut[0];
      } else {
        if (lo) elo = adj_edge_lo(qrow, 1, nx, hx, s_ut);
        if (hi) ehi = adj_edge_hi(qrow, 1, nx, hx, s_uh);
      }
#pragma unroll
      for (int k = 0; k < SL; ++k)
        if (hi && k == khi && nx > 1) q[k] = ehi;
      if (lo) q[0] = elo;
    }
    __syncthreads();
    // stage the plane through Qp's data columns, then write it coalesced,
    // times the attenuation factor (mul3dj!, src/backproject.jl:316)
    if (hasx) {
#pragma unroll
      for (int k = 0; k < SL; ++k)
        if (xs_x0 + k < nx) Qp[xs_z * WQ + hx + xs_x0 + k] = q[k];
    }
    __syncthreads();
    for (int k = tid; k < P; k += nthr) {
      const int z = k / nx, x = k - z * nx;
      const long o = x + (long)nx * y + S * z;
      const float val = Qp[z * WQ + hx + x];
      wout[o] = av ? val * av[o] : val;
    }
    __syncthreads();
  }
}

// --- adjoint, direct 2-D ------------------------------------------------------
// Extended correlation E over [-hx, nx+hx) x [-hz, nz+hz) (the reference's
// padded work matrix after the inverse FFT), then the replicate-pad fold:
// rows (x) first, then columns (z), src/fft_convolve.jl:99-122, then crop.
template <int NT, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_psf_adj_2d(const float* __restrict__ views, Taps t,
                                                      const int* __restrict__ vids, int Yc,
                                                      float* __restrict__ work, const float* __restrict__ atten,
                                                      long atten_vstride, Geo g) {
  extern __shared__ float L[];
  const int nx = g.nx, ny = g.ny, nz = g.nz, px = g.px, pz = g.pz;
  const int hx = (px - 1) / 2, hz = (pz - 1) / 2;
  const int EX = nx + 2 * hx, EZ = nz + 2 * hz;
  const int WE = odd_up(EX);
  const int b = blockIdx.y, c = blockIdx.x;
  const int v = vids[b];
  const int y0 = c * Yc, y1 = min(ny, y0 + Yc);
  const long S = (long)nx * ny, V = S * nz;
  const int P = nx * nz;
  const float* vin = views + (long)b * P;
  float* wout = work + V * b;
  const float* av = atten ? atten + atten_vstride * v : nullptr;
  for (int y = y0; y < y1; ++y) {
    const float* th = t.h + (long)(v * t.vstride + y) * pz * NT;
    // E[ze][xe] at extended (xe - hx, ze - hz):
    //   sum_{a,b} h[a,b] V[xe - 2hx + a, ze - 2hz + b]
    for (int k = threadIdx.x; k < EX * EZ; k += BLOCK) {
      const int ze = k / EX, xe = k - ze * EX;
      float s = 0.f;
      for (int bb = 0; bb < pz; ++bb) {
        const int zz = ze - 2 * hz + bb;
        if (zz < 0 || zz >= nz) continue;
        for (int a = 0; a < px; ++a) {
          const int xx = xe - 2 * hx + a;
          if (xx >= 0 && xx < nx) s = fmaf(th[bb * NT + a], vin[xx + nx * zz], s);
        }
      }
      L[ze * WE + xe] = s;
    }
    __syncthreads();
    // fold x pads onto x = 0 and x = nx-1, for every extended row
    for (int ze = threadIdx.x; ze < EZ; ze += BLOCK) {
      float* row = L + ze * WE;
      float lo = 0.f, hi = 0.f;
      for (int i = 0; i < hx; ++i) lo += row[i];
      for (int i = hx + nx; i < EX; ++i) hi += row[i];
      row[hx] += lo;
      row[hx + nx - 1] += hi;
    }
    __syncthreads();
    // fold z pads onto z = 0 and z = nz-1, for every extended column
    for (int xe = threadIdx.x; xe < EX; xe += BLOCK) {
      float lo = 0.f, hi = 0.f;
      for (int j = 0; j < hz; ++j) lo += L[j * WE + xe];
      for (int j = hz + nz; j < EZ; ++j) hi += L[j * WE + xe];
      L[hz * WE + xe] += lo;
      L[(hz + nz - 1) * WE + xe] += hi;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < P; k += BLOCK) {
      const int z = k / nx, x = k - z * nx;
      const long o = x + (long)nx * y + S * z;
      const float val = L[(z + hz) * WE + hx + x];
      wout[o] = av ? val * av[o] : val;
    }
    __syncthreads();
  }
}

// ===========================================================================
// Image reduction, elementwise
// ===========================================================================
__global__ void __launch_bounds__(256) k_img_reduce(const float* __restrict__ part, int ng, long n,
                                                    float* __restrict__ out, const float* __restrict__ asum,
                                                    int mode) {
  const long k = (long)blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  float s = part[k];
  for (int g = 1; g < ng; ++g) s += part[(long)g * n + k];
  if (mode == IMG_WRITE) out[k] = s;
  else if (mode == IMG_ACCUM) out[k] = out[k] + s;
  else out[k] = out[k] * (s / asum[k]);  // out .*= back ./ asum, src/ml-os-em.jl:46
}

__global__ void __launch_bounds__(256) k_fill(float* __restrict__ p, float v, long n) {
  const long k = (long)blockIdx.x * 256 + threadIdx.x;
  if (k < n) p[k] = v;
}

__global__ void __launch_bounds__(256) k_zero_to_inf(float* __restrict__ p, long n) {
  const long k = (long)blockIdx.x * 256 + threadIdx.x;
  if (k < n && p[k] == 0.f) p[k] = __builtin_inff();  // asum[asum .== 0] .= Inf
}

// ===========================================================================
// Launchers
// ===========================================================================
static inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

// Strip length of the separable kernels: one strip per thread per phase, at
// most 1024 threads per block.
static int sep_strip_len(const Geo& g) {
  for (int sl : {8, 16, 32}) {
    const long sx = (long)g.nz * ((g.nx + sl - 1) / sl), sz = (long)g.nx * ((g.nz + sl - 1) / sl);
    if (std::max(sx, sz) <= 1024) return sl;
  }
  return 0;
}

// kernels whose dynamic LDS exceeds 64 KiB must opt in
template <typename K>
static void allow_lds(K kernel, size_t bytes) {
  if (bytes > 65536)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)bytes);
}

#define SR_SL_SWITCH(SLV, ...)                              \
  switch (SLV) {                                            \
    case 8: { constexpr int SL = 8; __VA_ARGS__; } break;   \
    case 16: { constexpr int SL = 16; __VA_ARGS__; } break; \
    default: { constexpr int SL = 32; __VA_ARGS__; } break; \
  }


hipError_t launch_rot_fwd(int interp, const float* img, float* out, const float* atten, long atten_vstride,
                          const RotView* rv, const int* vids, int nb, const int* tab_b, const float* tab_w,
                          const Geo& g, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  const long S = (long)g.nx * g.nx;
  if (interp == 1) {
    const size_t lds = sizeof(float) * g.Np * g.Np;
    if (g.Np > kMaxNpOne || lds > kMaxLds) return hipErrorInvalidConfiguration;
    allow_lds(k_rot1_fwd, lds);
    k_rot1_fwd<<<dim3(g.nz, nb), R1_THREADS, lds, s>>>(img, out, atten, atten_vstride, rv, vids, tab_b, tab_w, g);
  } else {
    // views per block: enough blocks to fill the device, reuse of the staged slice otherwise
    int vpg = 1;
    while (vpg < 8 && (long)g.nz * cdiv(nb, vpg * 2) >= 1024) vpg *= 2;
    const size_t lds = sizeof(float) * S;
    if (lds <= kMaxLds) {
      allow_lds(k_rot2_fwd<true>, lds);
      k_rot2_fwd<true><<<dim3(g.nz, cdiv(nb, vpg)), 256, lds, s>>>(img, out, atten, atten_vstride, rv, vids, nb,
                                                                  vpg, g);
    } else
      k_rot2_fwd<false><<<dim3(g.nz, cdiv(nb, vpg)), 256, 0, s>>>(img, out, atten, atten_vstride, rv, vids, nb, vpg,
                                                                 g);
  }
  return hipGetLastError();
}

hipError_t launch_rot_adj(int interp, const float* work, float* part, const RotView* rv, const int* vids, int nb,
                          int vpg, const int* tab_b, const float* tab_w, const Geo& g, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  if (interp == 1) {  // one view per block: part holds nb slices-volumes
    const size_t lds = sizeof(float) * g.Np * g.Np;
    if (g.Np > kMaxNpOne || lds > kMaxLds) return hipErrorInvalidConfiguration;
    allow_lds(k_rot1_adj, lds);
    k_rot1_adj<<<dim3(g.nz, nb), R1_THREADS, lds, s>>>(work, part, rv, vids, tab_b, tab_w, g);
  } else {
    const int ng = (nb + vpg - 1) / vpg;
    const size_t lds = sizeof(float) * g.nx * g.nx;
    if (lds > kMaxLds) return hipErrorInvalidConfiguration;
    allow_lds(k_rot2_adj, lds);
    k_rot2_adj<<<dim3(g.nz, ng), 256, lds, s>>>(work, part, rv, vids, nb, vpg, g);
  }
  return hipGetLastError();
}

hipError_t launch_img_reduce(const float* part, int ngroups, float* out, const float* asum, long n, int mode,
                             hipStream_t s) {
  k_img_reduce<<<cdiv(n, 256), 256, 0, s>>>(part, ngroups, n, out, asum, mode);
  return hipGetLastError();
}

hipError_t launch_atten_exact(float* mur, int nb, const Geo& g, float dy, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  const int threads = std::min(1024, (int)(cdiv(g.nx, 64) * 64));
  const size_t lds = sizeof(float) * g.nx * g.nx;
  if (lds > kMaxLds) return hipErrorInvalidConfiguration;
  allow_lds(k_atten_exact, lds);
  k_atten_exact<<<dim3(g.nz, nb), threads, lds, s>>>(mur, g, dy);
  return hipGetLastError();
}

static int pick_block(long strips) {
  if (strips <= (long)PSF_MAXS * 256) return 256;
  if (strips <= (long)PSF_MAXS * 512) return 512;
  return 1024;
}

#define SR_NT_SWITCH(NTV, ...)                                   \
  switch (NTV) {                                                 \
    case 8: { constexpr int NT = 8; __VA_ARGS__; } break;        \
    case 16: { constexpr int NT = 16; __VA_ARGS__; } break;      \
    case 24: { constexpr int NT = 24; __VA_ARGS__; } break;      \
    case 32: { constexpr int NT = 32; __VA_ARGS__; } break;      \
    case 40: { constexpr int NT = 40; __VA_ARGS__; } break;      \
    case 48: { constexpr int NT = 48; __VA_ARGS__; } break;      \
    case 64: { constexpr int NT = 64; __VA_ARGS__; } break;      \
    default: return hipErrorInvalidValue;                        \
  }
#define SR_BLOCK_SWITCH(BV, ...)                                 \
  switch (BV) {                                                  \
    case 256: { constexpr int BLOCK = 256; __VA_ARGS__; } break; \
    case 512: { constexpr int BLOCK = 512; __VA_ARGS__; } break; \
    default: { constexpr int BLOCK = 1024; __VA_ARGS__; } break; \
  }


hipError_t launch_psf_fwd(int separable, const float* work, const Taps& t, const int* vids, int nb, float* part,
                          int nchunk, int* nchunk_used, const Geo& g, hipStream_t s) {
  const int Yc = (g.ny + nchunk - 1) / nchunk;
  nchunk = (g.ny + Yc - 1) / Yc;
  if (nchunk_used) *nchunk_used = nchunk;
  if (nb <= 0) return hipSuccess;
  if (separable) {
    const int SLv = sep_strip_len(g);
    if (SLv == 0) return hipErrorInvalidValue;
    const int sx = (g.nx + SLv - 1) / SLv, sz = (g.nz + SLv - 1) / SLv;
    const int block = std::max(64, (int)cdiv(std::max(g.nz * sx, g.nx * sz), 64) * 64);
    const size_t lds = sizeof(float) * ((size_t)g.nz * odd_up(SLv * sx + t.ntx - 1) + 2 * t.ntx);
    if (lds > kMaxLds) return hipErrorInvalidConfiguration;
    SR_NT_SWITCH(t.ntx, SR_SL_SWITCH(SLv, {
      allow_lds(k_psf_fwd_sep<NT, SL>, lds);
      k_psf_fwd_sep<NT, SL><<<dim3(nchunk, nb), block, lds, s>>>(work, t, vids, Yc, part, g);
    }));
  } else {
    const int nxs = (g.nx + 7) / 8;
    const int WX = odd_up(8 * nxs + t.ntx - 1);
    const long strips = (long)g.nz * nxs;
    const int block = pick_block(strips);
    const size_t lds = sizeof(float) * (g.nz + g.pz - 1) * WX;
    if (strips > (long)PSF_MAXS * 1024 || lds > kMaxLds) return hipErrorInvalidConfiguration;
    SR_NT_SWITCH(t.ntx, SR_BLOCK_SWITCH(block, {
      allow_lds(k_psf_fwd_2d<NT, BLOCK>, lds);
      k_psf_fwd_2d<NT, BLOCK><<<dim3(nchunk, nb), BLOCK, lds, s>>>(work, t, vids, Yc, part, g);
    }));
  }
  return hipGetLastError();
}

hipError_t launch_view_reduce(const float* part, int nchunk, const int* vids, int nb, float* out, const float* yobs,
                              const float* bkg, int epi, const Geo& g, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  const int P = g.nx * g.nz;
  k_view_reduce<<<dim3(cdiv(P, 256), nb), 256, 0, s>>>(part, nchunk, P, vids, out, yobs, bkg, epi);
  return hipGetLastError();
}

hipError_t launch_psf_adj(int separable, const float* views, const Taps& t, const int* vids, int nb, float* work,
                          const float* atten, long atten_vstride, int nchunk, const Geo& g, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  const int Yc = (g.ny + nchunk - 1) / nchunk;
  nchunk = (g.ny + Yc - 1) / Yc;
  if (separable) {
    const int SLv = sep_strip_len(g);
    if (SLv == 0) return hipErrorInvalidValue;
    const int sx = (g.nx + SLv - 1) / SLv, sz = (g.nz + SLv - 1) / SLv;
    const int block = std::max(64, (int)cdiv(std::max(g.nz * sx, g.nx * sz), 64) * 64);
    const size_t lds =
        sizeof(float) * ((size_t)g.nx * g.nz + (size_t)g.nz * odd_up(SLv * sx + t.ntx - 1) + 6 * t.ntx);
    if (lds > kMaxLds) return hipErrorInvalidConfiguration;
    if (g.nx % 64 == 0) {
      SR_NT_SWITCH(t.ntx, SR_SL_SWITCH(SLv, {
        allow_lds(k_psf_adj_sep<NT, SL, true>, lds);
        k_psf_adj_sep<NT, SL, true>
            <<<dim3(nchunk, nb), block, lds, s>>>(views, t, vids, Yc, work, atten, atten_vstride, g);
      }));
    } else {
      SR_NT_SWITCH(t.ntx, SR_SL_SWITCH(SLv, {
        allow_lds(k_psf_adj_sep<NT, SL, false>, lds);
        k_psf_adj_sep<NT, SL, false>
            <<<dim3(nchunk, nb), block, lds, s>>>(views, t, vids, Yc, work, atten, atten_vstride, g);
      }));
    }
  } else {
    const int hx = (g.px - 1) / 2, hz = (g.pz - 1) / 2;
    const size_t lds = sizeof(float) * (size_t)odd_up(g.nx + 2 * hx) * (g.nz + 2 * hz);
    if (lds > kMaxLds) return hipErrorInvalidConfiguration;
    SR_NT_SWITCH(t.ntx, {
      allow_lds(k_psf_adj_2d<NT, 256>, lds);
      k_psf_adj_2d<NT, 256><<<dim3(nchunk, nb), 256, lds, s>>>(views, t, vids, Yc, work, atten, atten_vstride, g);
    });
  }
  return hipGetLastError();
}

hipError_t launch_fill(float* p, float v, long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_fill<<<cdiv(n, 256), 256, 0, s>>>(p, v, n);
  return hipGetLastError();
}

hipError_t launch_zero_to_inf(float* p, long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_zero_to_inf<<<cdiv(n, 256), 256, 0, s>>>(p, n);
  return hipGetLastError();
}

}  // namespace sr
