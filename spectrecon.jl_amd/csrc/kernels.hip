// kernels.hip -- CDNA4 (gfx950) kernels of the SPECT projector path.
//
// Compiled with -ffp-contract=off: every a*b+c below rounds twice, exactly as
// the reference's Float32 code does, unless fmaf() is written out (only the
// PSF convolutions, whose reference arithmetic is an FFT, use fmaf).
#include "kernels.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace sr {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// A pointer known to address global memory: pointers read from memory (not
// kernel arguments) are otherwise generic, and generic (flat) accesses also
// count against the LDS wait counter, which the LDS-only barriers rely on.
template <class T>
__device__ __forceinline__ T* glob(T* p) {
  using G = __attribute__((address_space(1))) T;
  return (T*)(G*)(uintptr_t)p;
}
__device__ __forceinline__ Taps glob(Taps t) {
  t.u = glob(t.u), t.w = glob(t.w), t.h = glob(t.h), t.cum = glob(t.cum);
  return t;
}

// descriptor of a grouped launch that owns this block: the last i with
// d[i].h.blk0 <= blockIdx.x (sorted ranges; block-uniform, scalar loads)
template <class D>
__device__ __forceinline__ int grp_find(const D* __restrict__ d, int n) {
  int lo = 0, hi = n - 1;
  const int bx = (int)blockIdx.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].h.blk0 <= bx)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Walks the elements k = start, start+step, ... of a row-major (rows x ncol)
// array as (row, col) without a division per element.
struct RowIter {
  int row, col, drow, dcol, ncol;
  __device__ __forceinline__ RowIter(int start, int step, int ncol_) : ncol(ncol_) {
    row = start / ncol;
    col = start - row * ncol;
    drow = step / ncol;
    dcol = step - drow * ncol;
  }
  __device__ __forceinline__ void next() {
    col += dcol;
    row += drow;
    if (col >= ncol) {
      col -= ncol;
      ++row;
    }
  }
};

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
__host__ __device__ __forceinline__ BiCoef bilinear_coef(const RotView& r, int pad, int x, int y) {
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

// Vectorized form for N % 4 == 0: block 1024, the slice staged with float4
// loads, each thread producing float4 groups of 4 consecutive x; the
// attenuation float4 loads of a view are issued before the LDS gathers.
constexpr int R2F_THREADS = 1024;
constexpr int R2F_ITEMS = 4;  // float4 groups per thread and pass
__global__ void __launch_bounds__(R2F_THREADS) k_rot2_fwd4(const float* __restrict__ img, float* __restrict__ out,
                                                          const float* __restrict__ atten, long atten_vstride,
                                                          const RotView* __restrict__ rv,
                                                          const int* __restrict__ vids, int nb, int vpg, Geo g) {
  extern __shared__ float s_img[];
  const int N = g.nx, z = blockIdx.x, tid = threadIdx.x;
  const int NN = N * N, n4 = NN >> 2, q = N >> 2;
  const long S = (long)NN, V = S * g.nz;
  const float4* g4 = reinterpret_cast<const float4*>(img + S * z);
  float4* s4 = reinterpret_cast<float4*>(s_img);
  for (int i0 = 0; i0 < n4; i0 += R2F_THREADS * R2F_ITEMS) {
    float4 t4[R2F_ITEMS];
#pragma unroll
    for (int k = 0; k < R2F_ITEMS; ++k) {
      const int i = i0 + tid + k * R2F_THREADS;
      if (i < n4) t4[k] = g4[i];
    }
#pragma unroll
    for (int k = 0; k < R2F_ITEMS; ++k) {
      const int i = i0 + tid + k * R2F_THREADS;
      if (i < n4) s4[i] = t4[k];
    }
  }
  __syncthreads();
  const int b0 = blockIdx.y * vpg, b1 = min(nb, b0 + vpg);
  for (int b = b0; b < b1; ++b) {
    const int v = vids[b];
    const RotView r = rv[v];
    float4* o4 = reinterpret_cast<float4*>(out + V * b + S * z);
    const float4* a4 = atten ? reinterpret_cast<const float4*>(atten + atten_vstride * v + S * z) : nullptr;
    for (int i0 = 0; i0 < n4; i0 += R2F_THREADS * R2F_ITEMS) {
      float4 av[R2F_ITEMS];
#pragma unroll
      for (int k = 0; k < R2F_ITEMS; ++k) {
        const int i = i0 + tid + k * R2F_THREADS;
        av[k] = (a4 && i < n4) ? a4[i] : make_float4(1.f, 1.f, 1.f, 1.f);
      }
#pragma unroll
      for (int k = 0; k < R2F_ITEMS; ++k) {
        const int i = i0 + tid + k * R2F_THREADS;
        if (i < n4) {
          const int y = i / q, x = 4 * (i - y * q);
          float vv[4];
#pragma unroll
          for (int e = 0; e < 4; ++e)
            vv[e] = (r.kind == 0) ? s_img[x + e + N * y] : bilinear_two(s_img, N, g.pad, r, x + e, y);
          float4 res;
          if (a4) {
            res.x = vv[0] * av[k].x;
            res.y = vv[1] * av[k].y;
            res.z = vv[2] * av[k].z;
            res.w = vv[3] * av[k].w;
          } else {
            res = make_float4(vv[0], vv[1], vv[2], vv[3]);
          }
          o4[i] = res;
        }
      }
    }
  }
}

// Adjoint as a gather: image pixel j collects w(i,j) * p[i] from every work
// pixel i whose forward sample lands within one pixel of j.  Since the map is
// a rotation, those i lie within sqrt(2) of q = R^-1(j), i.e. in a 3x3 window.
// Each candidate's sample x(i) is recomputed exactly as the forward does
// ((xx*I1 + xy*I2) + x0 in Float32), and its weight max(0, 1-|x1-J1|) *
// max(0, 1-|x2-J2|) equals the forward's (1-t or t) bit for bit (x - J is
// exact by Sterbenz), so this is the exact transpose, without atomics and
// deterministic.  grid (nz, ngroups), block 1024; the views of a group are
// summed in registers; part[grp] receives the group's sum.
constexpr int RA_THREADS = 1024;
constexpr int RA_MAXP = 16;  // pixels per thread and pass (one pass covers N <= 128)

// kLds: the view's work slice staged in LDS (row stride N + 1, odd: rows fall
// in different banks); else read where it lies (row stride N), for slices
// larger than LDS (N >= 202); the sums are the same either way.
template <bool kLds>
__global__ void __launch_bounds__(RA_THREADS) k_rot2_adj(const float* __restrict__ work, float* __restrict__ part,
                                                         const RotView* __restrict__ rv,
                                                         const int* __restrict__ vids, int nb, int vpg, Geo g) {
  extern __shared__ float sp[];
  const int N = g.nx, pad = g.pad, z = blockIdx.x, grp = blockIdx.y;
  const int NN = N * N, SP = kLds ? N + 1 : N;
  const long S = (long)NN, V = S * g.nz;
  const int tid = threadIdx.x;
  const int b0 = grp * vpg, b1 = min(nb, b0 + vpg);
  float* o = part + V * grp + S * z;
  for (int p0 = 0; p0 < NN; p0 += RA_THREADS * RA_MAXP) {
    float acc[RA_MAXP];
#pragma unroll
    for (int k = 0; k < RA_MAXP; ++k) acc[k] = 0.f;
    for (int b = b0; b < b1; ++b) {
      const RotView r = rv[vids[b]];
      const float* src = work + V * b + S * z;
      if constexpr (kLds) {
        __syncthreads();  // previous view's reads of sp are done
        for (int k0 = 0; k0 < NN; k0 += RA_THREADS * 8) {
          float tv[8];
          int dst[8];
          RowIter it(k0 + tid, RA_THREADS, N);
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const int k = k0 + tid + e * RA_THREADS;
            tv[e] = k < NN ? src[k] : 0.f;
            dst[e] = k < NN ? it.row * SP + it.col : -1;
            it.next();
          }
#pragma unroll
          for (int e = 0; e < 8; ++e)
            if (dst[e] >= 0) sp[dst[e]] = tv[e];
        }
        __syncthreads();
      }
      const float* sl = kLds ? sp : src;  // (the slice, row stride SP)
      RowIter it(p0 + tid, RA_THREADS, N);
#pragma unroll
      for (int k = 0; k < RA_MAXP; ++k) {
        const int jj = p0 + tid + k * RA_THREADS;
        const int j2 = it.row, j1 = it.col;
        it.next();
        if (jj < NN) {
          const float J1 = (float)(j1 + pad + 1), J2 = (float)(j2 + pad + 1);
          // source position of J (approximate inverse; only picks the window)
          const float u1 = J1 - r.x0, u2 = J2 - r.y0;
          const float q1 = r.xx * u1 + r.yx * u2, q2 = r.xy * u1 + r.yy * u2;
          const int lo1 = (int)floorf(q1 - 1.41422f) + 1, lo2 = (int)floorf(q2 - 1.41422f) + 1;
          float A[3], B[3], C[3], D[3];
          int ci[3], cj[3];
          bool vi[3], vj[3];
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            const float F1 = (float)(lo1 + a), F2 = (float)(lo2 + a);
            A[a] = r.xx * F1;
            C[a] = r.yx * F1;
            B[a] = r.xy * F2;
            D[a] = r.yy * F2;
            const int i1 = lo1 + a - pad - 1, i2 = lo2 + a - pad - 1;
            vi[a] = (unsigned)i1 < (unsigned)N;
            vj[a] = (unsigned)i2 < (unsigned)N;
            ci[a] = clampi(i1, 0, N - 1);
            cj[a] = __umul24(clampi(i2, 0, N - 1), SP);
          }
          float sum = acc[k];
#pragma unroll
          for (int bb = 0; bb < 3; ++bb)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              const float x1 = (A[a] + B[bb]) + r.x0;
              const float x2 = (C[a] + D[bb]) + r.y0;
              const float w1 = fmaxf(0.f, 1.f - fabsf(x1 - J1));
              const float w2 = fmaxf(0.f, 1.f - fabsf(x2 - J2));
              const float pv = (vi[a] && vj[bb]) ? sl[ci[a] + cj[bb]] : 0.f;
              sum = sum + (w1 * w2) * pv;
            }
          acc[k] = sum;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < RA_MAXP; ++k) {
      const int jj = p0 + tid + k * RA_THREADS;
      if (jj < NN) o[jj] = acc[k];
    }
  }
}

// Adjoint as a collision-free scatter (N <= 128).  Work pixel i contributes
// w(i,j) p[i] to the 2x2 image pixels around its forward sample x(i) (the
// forward's own Float32 coordinates and weights: the exact transpose).  The
// pixels are processed in 9 phases by (i1 mod 3, i2 mod 3): two pixels of one
// class are >= 3 apart, their samples (a rotation is an isometry) >= 3/sqrt2
// > 2 apart in one coordinate, so their 2x2 footprints are disjoint and the
// LDS read-modify-writes of a phase never touch one address twice: plain
// (non-atomic) updates are exact and deterministic, in a fixed phase order.
// The accumulator has a one-pixel guard ring ((N+2) rows of stride N+8, image
// pixel (j1,j2) at (j1+1,j2+1)), so a footprint that overlaps the image is
// updated unconditionally with two ds_read2/ds_write2 pairs; the ring is
// dropped at the end.  Each wave takes 8x8 tiles of a class (its lanes hit a
// rotated lattice of spacing 3, which the stride N+8 spreads over the banks),
// and the next phase's pixels are loaded while the current one is scattered.
// grid (nz, ngroups), block 1024.
constexpr int RS_THREADS = 1024;
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_TILES = 3;  // tiles per wave and phase: 16 x 3 = 48 >= ceil(43/8)^2 = 36

struct ScatterPix {
  int i1[RS_TILES], i2[RS_TILES];
  float v[RS_TILES];
  int ntl;  // wave-uniform: tiles of this wave in the phase
};

// pixels of phase c (class (c mod 3, c div 3)); i1 = -1 marks an idle lane
__device__ __forceinline__ void scatter_load(ScatterPix& P, const float* __restrict__ src, int N, int c, int wave,
                                             int lane) {
  const int c1 = c % 3, c2 = c / 3;
  const int na = (N - c1 + 2) / 3, nbq = (N - c2 + 2) / 3;
  const int ta = (na + 7) >> 3, ntile = ta * ((nbq + 7) >> 3);
  P.ntl = ntile > wave ? (ntile - wave + RS_WAVES - 1) / RS_WAVES : 0;
#pragma unroll
  for (int k = 0; k < RS_TILES; ++k) {
    const int t = wave + RS_WAVES * k;  // wave-uniform
    const int tb = t / ta, tt = t - tb * ta;
    const int a = tt * 8 + (lane & 7), bq = tb * 8 + (lane >> 3);
    const bool ok = a < na && bq < nbq;
    P.i1[k] = ok ? c1 + 3 * a : -1;
    P.i2[k] = ok ? c2 + 3 * bq : 0;
  }
#pragma unroll
  for (int k = 0; k < RS_TILES; ++k)
    if (k < P.ntl) P.v[k] = src[max(P.i1[k], 0) + N * P.i2[k]];
}

// one phase: scatter the pixels of P into the accumulator
__device__ __forceinline__ void scatter_phase(float* acc, const ScatterPix& P, const RotView& r, int N, int pad,
                                              int SP) {
#pragma unroll
  for (int k = 0; k < RS_TILES; ++k) {
    if (k >= P.ntl) break;
    const BiCoef cf = bilinear_coef(r, pad, P.i1[k], P.i2[k]);
    if (P.i1[k] >= 0 && (unsigned)(cf.j1 + 1) <= (unsigned)N && (unsigned)(cf.j2 + 1) <= (unsigned)N) {
      const float p1 = P.v[k] * cf.w21, p2 = P.v[k] * cf.w22;
      float* q = acc + (cf.j1 + 1) + SP * (cf.j2 + 1);
      const float q0 = q[0], q1 = q[1], q2 = q[SP], q3 = q[SP + 1];
      q[0] = q0 + cf.w11 * p1;
      q[1] = q1 + cf.w12 * p1;
      q[SP] = q2 + cf.w11 * p2;
      q[SP + 1] = q3 + cf.w12 * p2;
    }
  }
}

__global__ void __launch_bounds__(RS_THREADS) k_rot2_adj_scatter(const float* __restrict__ work,
                                                                 float* __restrict__ part,
                                                                 const RotView* __restrict__ rv,
                                                                 const int* __restrict__ vids, int nb, int vpg,
                                                                 Geo g) {
  extern __shared__ float acc[];
  const int N = g.nx, pad = g.pad, z = blockIdx.x, grp = blockIdx.y, tid = threadIdx.x;
  const int NN = N * N, SP = N + 8;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const long S = (long)NN, V = S * g.nz;
  for (int k = tid; k < (N + 2) * SP; k += RS_THREADS) acc[k] = 0.f;
  const int b0 = grp * vpg, b1 = min(nb, b0 + vpg);
  const int st0 = 9 * b0, st1 = 9 * b1;
  const float* zsrc = work + S * z;
  // (view, phase) steps in ping-pong buffers: the next step's pixels load
  // while this one scatters
  ScatterPix A, B;
  if (st0 < st1) scatter_load(A, zsrc + V * b0, N, 0, wave, lane);
  __syncthreads();  // acc zeroed
#pragma unroll 1
  for (int st = st0; st < st1; st += 2) {
    const int b = st / 9, c = st - 9 * b;
    const bool more = st + 1 < st1;
    const int bn = (st + 1) / 9;
    if (more) scatter_load(B, zsrc + V * bn, N, st + 1 - 9 * bn, wave, lane);
    scatter_phase(acc, A, rv[vids[b]], N, pad, SP);
    __syncthreads();  // the phase's updates land before the next phase reads
    if (!more) break;
    const int bm = (st + 2) / 9;
    if (st + 2 < st1) scatter_load(A, zsrc + V * bm, N, st + 2 - 9 * bm, wave, lane);
    scatter_phase(acc, B, rv[vids[bn]], N, pad, SP);
    __syncthreads();
    (void)c;
  }
  float* o = part + V * grp + S * z;
  RowIter it(tid, RS_THREADS, N);
  for (int k = tid; k < NN; k += RS_THREADS) {
    o[k] = acc[(it.row + 1) * SP + it.col + 1];
    it.next();
  }
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
// dim 2 indexed by row x (rotate_y!).  The adjoint is the transpose,
// evaluated as a gather in the reference's scatter order (rows ascending, J1
// before J2).
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

// adjoint at an interior position 0 < j < Np-1 of row i: only J2(j-b-1) and
// J1(j-b) land there, in that order
template <bool kAlongX>
__device__ __forceinline__ float shear_adj_interior(const float* L, int Np, const int* tb, const float* tw, int x,
                                                    int y) {
  const int i = kAlongX ? y : x;
  const int j = kAlongX ? x : y;
  const int b = tb[i];
  const float w1 = tw[2 * i], w2 = tw[2 * i + 1];
  const int k2 = j - b - 1, k1 = j - b;
  float s = 0.f;
  if (k2 >= 0 && k2 < Np) s = s + w2 * (kAlongX ? L[k2 + Np * y] : L[x + Np * k2]);
  if (k1 >= 0 && k1 < Np) s = s + w1 * (kAlongX ? L[k1 + Np * y] : L[x + Np * k1]);
  return s;
}

// adjoint at the clamped end j = 0 (hi false) or j = Np-1 (hi true) of row i:
// every tap whose index clamps there, in scatter order
template <bool kAlongX>
__device__ __forceinline__ float shear_adj_end(const float* L, int Np, const int* tb, const float* tw, int i,
                                               bool hi) {
  const int b = tb[i];
  const float w1 = tw[2 * i], w2 = tw[2 * i + 1];
  auto ld = [&](int k) { return kAlongX ? L[k + Np * i] : L[i + Np * k]; };
  float s = 0.f;
  if (!hi) {
    const int kend = min(Np - 1, -b);
    for (int k = 0; k <= kend; ++k) {
      if (b + k <= 0) s = s + w1 * ld(k);
      if (b + k + 1 <= 0) s = s + w2 * ld(k);
    }
  } else {
    for (int k = max(0, Np - 2 - b); k < Np; ++k) {
      if (b + k >= Np - 1) s = s + w1 * ld(k);
      if (b + k + 1 >= Np - 1) s = s + w2 * ld(k);
    }
  }
  return s;
}

// The block's elements in registers: thread (tx, ty) holds x = tx + 64*xs,
// y = ty + 16*ys.
#define R1_FOR_EACH(...)                                  \
  _Pragma("unroll") for (int ys = 0; ys < R1_MY; ++ys)    \
  _Pragma("unroll") for (int xs = 0; xs < R1_MX; ++xs) {  \
    const int x = tx + 64 * xs, y = ty + 16 * ys;         \
    __VA_ARGS__                                           \
  }

// A pass computes rows y0 <= y < y1 of L only; the other rows keep their
// values (the caller reads none of them, or they are zero and stay zero).
template <bool kAlongX>
__device__ __forceinline__ void shear_pass_fwd(float* L, int Np, const int* tb, const float* tw, int y0, int y1) {
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  float r[R1_MY][R1_MX];
  R1_FOR_EACH(if (x < Np && y >= y0 && y < y1) r[ys][xs] = shear_fwd_at<kAlongX>(L, Np, tb, tw, x, y);)
  __syncthreads();
  R1_FOR_EACH(if (x < Np && y >= y0 && y < y1) L[x + Np * y] = r[ys][xs];)
  __syncthreads();
}

// The clamped ends hold loops over up to Np taps; they are spread over all
// waves (end q = 2*row + hi on lane q/16 of wave q%16) instead of serialising
// in the waves that own columns 0 and Np-1.
template <bool kAlongX>
__device__ __forceinline__ void shear_pass_adj(float* L, int Np, const int* tb, const float* tw, int y0, int y1) {
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int q = tx * 16 + ty, ei = q >> 1;
  const bool hi = q & 1;
  const int ej = hi ? Np - 1 : 0, ey = kAlongX ? ei : ej;
  const bool end = ei < Np && (!hi || Np > 1) && ey >= y0 && ey < y1;
  float e = 0.f;
  if (end) e = shear_adj_end<kAlongX>(L, Np, tb, tw, ei, hi);
  float r[R1_MY][R1_MX];
  R1_FOR_EACH({
    const int j = kAlongX ? x : y;
    if (x < Np && y >= y0 && y < y1 && j > 0 && j < Np - 1)
      r[ys][xs] = shear_adj_interior<kAlongX>(L, Np, tb, tw, x, y);
  })
  __syncthreads();
  R1_FOR_EACH({
    const int j = kAlongX ? x : y;
    if (x < Np && y >= y0 && y < y1 && j > 0 && j < Np - 1) L[x + Np * y] = r[ys][xs];
  })
  if (end) L[kAlongX ? ej + Np * ei : ei + Np * ej] = e;
  __syncthreads();
}

// Table rows of view v (b: x then y, Np ints each; w: x then y, 2*Np floats
// each), copied behind the slice in LDS when kLds (the caller's next
// __syncthreads publishes them), else read in place.
struct R1Tabs {
  const int* bx;
  const int* by;
  const float* wx;
  const float* wy;
};
template <bool kLds>
__device__ __forceinline__ R1Tabs r1_tabs(float* L, int Np, int v, const int* __restrict__ tab_b,
                                          const float* __restrict__ tab_w) {
  const int* gb = tab_b + (long)v * 2 * Np;
  const float* gw = tab_w + (long)v * 4 * Np;
  if (!kLds) return {gb, gb + Np, gw, gw + 2 * Np};
  int* lb = reinterpret_cast<int*>(L + Np * Np);
  float* lw = L + Np * Np + 2 * Np;
  for (int i = threadIdx.x; i < 2 * Np; i += R1_THREADS) lb[i] = gb[i];
  for (int i = threadIdx.x; i < 4 * Np; i += R1_THREADS) lw[i] = gw[i];
  return {lb, lb + Np, lw, lw + 2 * Np};
}

// grid (nz, nb): imrotate! :one of slice z for view vids[b] (src/rotatez.jl:240-278)
template <bool kLds>
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
  const R1Tabs t = r1_tabs<kLds>(L, Np, v, tab_b, tab_w);
  {  // all loads of the slice in flight before the first LDS store; each
     // thread reads source pixel (x, y) (coalesced for every rotation) and
     // stores it where rot_f90 puts it, the inverse map of rotf90_src
    float q[R1_MY][R1_MX];
    R1_FOR_EACH(if (x < Np && y < Np) q[ys][xs] = slice_tap(src, N, x - pad, y - pad);)
    R1_FOR_EACH(if (x < Np && y < Np) {
      int dx, dy;
      rotf90_src(r.m, true, Np, x, y, dx, dy);
      L[dx + Np * dy] = q[ys][xs];
    })
  }
  __syncthreads();
  if (r.kind == 2) {
    shear_pass_fwd<true>(L, Np, t.bx, t.wx, 0, Np);
    shear_pass_fwd<false>(L, Np, t.by, t.wy, pad, pad + N);  // the last pass reads rows pad..pad+N-1
  }
  R1_FOR_EACH(if (x < N && y < N) {
    const float val = (r.kind == 2) ? shear_fwd_at<true>(L, Np, t.bx, t.wx, x + pad, y + pad)
                                    : L[(x + pad) + Np * (y + pad)];
    const long k = x + (long)N * y;
    o[k] = a ? val * a[k] : val;
  })
}

// grid (nz, nb): imrotate_adj! :one (src/rotatez.jl:286-322) of work[b]
// slice z into part[b] (one view per block; the views are summed by
// k_img_reduce).
template <bool kLds>
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
  const R1Tabs t = r1_tabs<kLds>(L, Np, v, tab_b, tab_w);
  {
    float q[R1_MY][R1_MX];
    R1_FOR_EACH(if (x < Np && y < Np) q[ys][xs] = slice_tap(src, N, x - pad, y - pad);)
    R1_FOR_EACH(if (x < Np && y < Np) L[x + Np * y] = q[ys][xs];)
  }
  __syncthreads();
  if (r.kind == 2) {
    // rows outside pad..pad+N-1 are zero on entry (and stay zero) and are
    // not read by the crop, which maps that square onto itself
    shear_pass_adj<true>(L, Np, t.bx, t.wx, pad, pad + N);
    shear_pass_adj<false>(L, Np, t.by, t.wy, 0, Np);
    shear_pass_adj<true>(L, Np, t.bx, t.wx, pad, pad + N);
  }
  R1_FOR_EACH(if (x < N && y < N) {
    int sx, sy;
    rotf90_src(r.m, true, Np, x + pad, y + pad, sx, sy);
    o[x + (long)N * y] = L[sx + Np * sy];
  })
}
#undef R1_FOR_EACH

// ===========================================================================
// Attenuation factors: running sums along y
// ===========================================================================
// In place over the rotated mu volumes of a batch, per column (x, z):
//   a(y) = exp(-dy (S_y - mu_y / 2)),  S_y = sum_{j<=y} mu_j.
// src/project.jl:103-113 forms each y's sum anew (-mu_y/2 + mu_1 + ... + mu_y,
// O(ny^2) per column); here S_y is carried from y - 1 (O(ny)), which moves the
// result by rounding only (the sum is associated differently).
// SEG lanes of a wave share a column: lane (seg, col) holds y in
// [seg L, (seg + 1) L), L = ceil(ny / SEG), with W = 64 / SEG columns per
// wave and the W lanes of one segment on adjacent columns (coalesced).  With
// SEG > 1 the segment sums are scanned across the column's lanes with wave
// shuffles and each lane then writes its segment from its prefix; SEG = 1 is
// one streaming pass.  zfast: volumes [b][y][x][z] (column c = x nz + z, y
// stride nx nz); else [b][z][y][x] (c = z nx + x, y stride nx).
// grid (ceil(cols / (4 W)), nb), block 256.
template <int SEG>
__global__ void __launch_bounds__(256) k_atten_scan(float* __restrict__ mur, Geo g, float dy, int zfast) {
  constexpr int W = 64 / SEG;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int seg = lane / W;
  const long cols = (long)g.nx * g.nz;
  const long c = ((long)blockIdx.x * 4 + wave) * W + lane % W;
  const int ny = g.ny, L = (ny + SEG - 1) / SEG;
  const int y0 = min(ny, seg * L), y1 = min(ny, y0 + L);
  const bool ok = c < cols;
  const long base = zfast ? c : (c / g.nx) * (long)g.nx * ny + c % g.nx;
  const long ys = zfast ? cols : (long)g.nx;
  float* p = mur + (long)blockIdx.y * cols * ny + (ok ? base : 0);
  const float mdy = -dy;
  float s = 0.f;
  if (SEG > 1) {
    float t = 0.f;
    if (ok)
      for (int y = y0; y < y1; ++y) t = t + p[y * ys];
    // inclusive scan over the column's segments (lanes col, col + W, ...)
#pragma unroll
    for (int d = W; d < 64; d <<= 1) {
      const float v = __shfl_up(t, d, 64);
      if (lane >= d) t = t + v;
    }
    s = __shfl_up(t, W, 64);  // the sum of the earlier segments
    if (seg == 0) s = 0.f;
  }
  if (!ok) return;
  for (int y = y0; y < y1; ++y) {
    const float m = p[y * ys];
    s = s + m;
    p[y * ys] = expf((s + (-0.5f * m)) * mdy);
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
// separable PSF kernels: one strip per thread and pass, at most this many
// threads (3 waves per SIMD: 168 VGPRs for a fully unrolled strip)
constexpr int kSepThreads = 768;

// Strip of SL outputs, taps processed in groups of 8 so that only SL+8
// inputs are live at a time.  row[i] is input i of the strip's window; row,
// the taps t and NT are multiples of 16 bytes / 4 floats, so windows and taps
// load as float4 (the window's last float is loaded but unused).
//   forward (convolution): o[k] += sum_a t[a] * in[k + NT-1-a]
//   transpose (correlation): o[k] += sum_a t[a] * in[k + a]
template <int N>
__device__ __forceinline__ void lds_load4(const float* __restrict__ p, float (&w)[N]) {
  const float4* p4 = reinterpret_cast<const float4*>(__builtin_assume_aligned(p, 16));
#pragma unroll
  for (int j = 0; j < N / 4; ++j) {
    const float4 q = p4[j];
    w[4 * j] = q.x;
    w[4 * j + 1] = q.y;
    w[4 * j + 2] = q.z;
    w[4 * j + 3] = q.w;
  }
}
template <int NT, int SL>
__device__ __forceinline__ void conv_strip_fwd(const float* __restrict__ row, const float* __restrict__ t,
                                               float (&o)[SL]) {
#pragma unroll
  for (int g = 0; g < NT; g += 8) {
    float w[SL + 8], ta[8];
    lds_load4(row + NT - 8 - g, w);
    lds_load4(t + g, ta);
#pragma unroll
    for (int aa = 0; aa < 8; ++aa)
#pragma unroll
      for (int k = 0; k < SL; ++k) o[k] = fmaf(ta[aa], w[k + 7 - aa], o[k]);
  }
}
template <int NT, int SL>
__device__ __forceinline__ void conv_strip_adj(const float* __restrict__ row, const float* __restrict__ t,
                                               float (&o)[SL]) {
#pragma unroll 1
  for (int g = 0; g < NT; g += 8) {
    float w[SL + 8], ta[8];
    lds_load4(row + g, w);
    lds_load4(t + g, ta);
#pragma unroll
    for (int aa = 0; aa < 8; ++aa)
#pragma unroll
      for (int k = 0; k < SL; ++k) o[k] = fmaf(ta[aa], w[k + aa], o[k]);
  }
}


// Row stride of the strip layouts: room for v floats plus the one extra float
// a window's float4 load reads, a multiple of 4 floats (16-byte rows) and an
// odd number of float4s, so lanes that read rows r, r+1, ... hit distinct
// bank quads.
__host__ __device__ __forceinline__ int quad_stride(int v) {
  const int q = (v + 4) / 4;  // float4s holding v + 1 floats
  return 4 * (q | 1);
}

__host__ __device__ __forceinline__ int odd_up(int v) { return v | 1; }

// --- forward, separable: per-chunk partial views -------------------------------
// grid (nchunk, nb), block >= max(#x strips, #z strips) (one strip of SL
// outputs per thread per phase).  Each block convolves the planes y of its
// chunk (read from the batch work volume, rotated and attenuated already) and
// keeps the running sum over y in registers; partials are reduced by
// k_view_reduce.  The next plane is prefetched into registers while the
// current one is convolved.  LDS holds the plane replicate-padded along x (x
// pass), then, in the same space, the x-pass result transposed to
// z-contiguous columns Qt[x][PT + z] with replicate pads along z (z pass), so
// every window read of both passes is an immediate offset; then the taps.
constexpr int PF_MAX = 4;  // prefetched float4 per thread (plane <= 16*nthreads floats)

template <int NT, int SL>
__global__ void __launch_bounds__(kSepThreads)
k_psf_fwd_sep(const float* __restrict__ work, Taps t, const int* __restrict__ vids, int Yc,
              float* __restrict__ part, Geo g) {
  extern __shared__ float L[];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int hx = (g.px - 1) / 2, hz = (g.pz - 1) / 2;
  const int nxs = (nx + SL - 1) / SL, nzs = (nz + SL - 1) / SL;
  const int WX = quad_stride(SL * nxs + NT - 1);  // x-pass row stride
  const int LP = NT - 1 - hx;                 // replicate columns left of x = 0
  const int PT = NT - 1 - hz;                 // replicate rows above z = 0
  const int WZ = quad_stride(SL * nzs + NT - 1);  // z-pass column stride
  const int area = max(nz * WX, SL * nxs * WZ);
  const int b = blockIdx.y, c = blockIdx.x, nchunk = gridDim.x;
  const int v = vids[b];
  const int y0 = c * Yc, y1 = min(ny, y0 + Yc);
  float* s_tu = L + area;
  float* s_tw = s_tu + NT;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const long S = (long)nx * ny;
  const float* wb = work + S * nz * b;
  // x strip (z fastest) and z strip (x fastest) of this thread
  const bool hasx = tid < nz * nxs, hasz = tid < nx * nzs;
  const int xs_z = tid % nz, xs_x0 = (tid / nz) * SL;
  const int zs_x = tid % nx, zs_z0 = (tid / nx) * SL;
  // plane elements of this thread as float4 along x: i = tid + k*nthr over (nz x nx/4)
  const int nq = nx >> 2, nv = nz * nq;

  float4 pf[PF_MAX];
  auto prefetch = [&](int y) {
#pragma unroll
    for (int k = 0; k < PF_MAX; ++k) {
      const int i = tid + k * nthr;
      if (i < nv) {
        const int z = i / nq, x4 = i - z * nq;
        pf[k] = *reinterpret_cast<const float4*>(wb + S * z + (long)nx * y + 4 * x4);
      }
    }
  };
  if (y0 < y1) prefetch(y0);

  float acc[SL];
#pragma unroll
  for (int k = 0; k < SL; ++k) acc[k] = 0.f;

  for (int y = y0; y < y1; ++y) {
    const int plane = v * t.vstride + y;
    if (tid < NT) {
      s_tu[tid] = t.u[(long)plane * NT + tid];
      s_tw[tid] = t.w[(long)plane * NT + tid];
    }
#pragma unroll
    for (int k = 0; k < PF_MAX; ++k) {
      const int i = tid + k * nthr;
      if (i < nv) {
        const int z = i / nq, x4 = i - z * nq;
        float* d = L + z * WX + LP + 4 * x4;
        d[0] = pf[k].x;
        d[1] = pf[k].y;
        d[2] = pf[k].z;
        d[3] = pf[k].w;
      }
    }
    if (y + 1 < y1) prefetch(y + 1);
    __syncthreads();
    // replicate pad columns, all threads
    const int rp = WX - LP - nx;
    for (RowIter it(tid, nthr, LP + rp); it.row < nz; it.next()) {
      float* row = L + it.row * WX;
      row[it.col < LP ? it.col : LP + nx + (it.col - LP)] = row[it.col < LP ? LP : LP + nx - 1];
    }
    __syncthreads();
    float q[SL];
#pragma unroll
    for (int k = 0; k < SL; ++k) q[k] = 0.f;
    if (hasx) {
      const float* row = L + xs_z * WX + xs_x0;
      conv_strip_fwd<NT, SL>(row, s_tu, q);
    }
    __syncthreads();
    if (hasx) {
      float* col = L + xs_x0 * WZ + PT + xs_z;  // Qt[x0 + k][PT + z] = col[k*WZ]
#pragma unroll
      for (int k = 0; k < SL; ++k) col[k * WZ] = q[k];
    }
    __syncthreads();
    // replicate pad rows (z) of Qt, all threads
    const int nlow = PT, nhigh = WZ - PT - nz;
    for (RowIter it(tid, nthr, nlow + nhigh); it.row < nx; it.next()) {
      float* colx = L + it.row * WZ;
      colx[it.col < nlow ? it.col : PT + nz + (it.col - nlow)] = colx[it.col < nlow ? PT : PT + nz - 1];
    }
    __syncthreads();
    if (hasz) {
      // output z reads rows z + hz - a (a < NT), i.e. Qt[x][z0 + j], j < NT+SL-1
      const float* col = L + zs_x * WZ + zs_z0;
      conv_strip_fwd<NT, SL>(col, s_tw, acc);
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

// --- forward, separable, on the matrix cores ---------------------------------
// Each pass of the clamped separable convolution is a banded matrix product:
//   x pass  Q[z][x0+n] = sum_k P[z][x0-hx+k] u[n+2hx-k]      (k < 16+px-1)
//   z pass  R[z0+m][x] = sum_k w[m+2hz-k] Q[z0-hz+k][x]      (k < 16+pz-1)
// on zero-padded rows, plus the replicate boundary as one rank-2 step on the
// edge tiles: clamping contributes in[0]*Tail(x+h+1) + in[n-1]*Head(x+h-n),
// Tail(k) = sum_{a>=k} t[a], Head(k) = sum_{a<=k} t[a] (so the pads are zeros
// written once per block).  16x16 output tiles on v_mfma_f32_16x16x4_f32; the
// tap (Toeplitz) fragments are built per plane in registers, the data
// fragments are read from LDS: Pz (the plane, rows z, column x+hx, row stride
// an odd number of float4s: the 16 rows of a fragment hit distinct banks) and
// Qz (the x-pass result, rows z+hz, stride = 16 mod 64 floats: the 4 rows of
// a fragment hit distinct bank ranges).  The z-pass accumulates the view sum
// over the block's planes in the MFMA accumulators; the next plane and taps
// are prefetched in registers.  Lane l: A (m = l%16, k = l/16), B (k = l/16,
// n = l%16), D register r (m = 4(l/16) + r, n = l%16).
// grid (nchunk, nb), block 512 (8 waves); nx, nz multiples of 16.
typedef float mf4 __attribute__((ext_vector_type(4)));
// Workgroup barrier for LDS data only: waits for this wave's LDS operations,
// not for its global loads and stores in flight (__syncthreads' fence drains
// those too, which would expose every prefetch and store).
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }
// keeps the compiler's scheduler from moving instructions across this point
// (the fragment loads of step s+1 stay ahead of the step-s products)
#define SR_SCHED_FENCE() __builtin_amdgcn_sched_barrier(0)
constexpr int MF_THREADS = 512;
constexpr int MF_WAVES = MF_THREADS / 64;
constexpr int MF_TPW = 8;   // output tiles per wave and pass, at most
constexpr int MF_PF = 8;    // prefetched float4 per thread (plane <= 16384 floats)

__host__ __device__ constexpr int mf_ksteps(int nt) { return (16 + nt - 1 + 3) / 4; }
// column of x = 0 in the plane rows: at least hx, a multiple of 4 (16-byte rows stores)
__host__ __device__ __forceinline__ int mf_xo(int hx) { return (hx + 3) & ~3; }
__host__ __device__ __forceinline__ int mf_wx(int nx, int hx, int nt) {
  return quad_stride(max(mf_xo(hx) - hx + nx - 16 + 4 * mf_ksteps(nt), mf_xo(hx) + nx) - 1);
}
__host__ __device__ __forceinline__ int mf_wq(int nx) { return ((nx + 47) / 64) * 64 + 16; }
__host__ __device__ __forceinline__ int mf_rq(int nz, int nt) { return nz - 16 + 4 * mf_ksteps(nt); }

__device__ __forceinline__ mf4 mfma4(float a, float b, mf4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// Transpose of the 4 x 4 blocks (lane group q = lane / 16, register r) of an
// accumulator: register g of group q receives register q of group g.  Applied
// to an output tile with rows m = 4 q + r, register g becomes the B operand of
// the K-step over rows 4 g .. 4 g + 3 (group q holds row 4 g + q), so one
// pass's result feeds the other pass's products without leaving registers.
// Two lane swaps across halves (v_permlane32_swap), two across 16-lane rows
// (v_permlane16_swap).
__device__ __forceinline__ mf4 xpose_groups(mf4 d) {
  const auto p0 = __builtin_amdgcn_permlane32_swap(__float_as_uint(d[0]), __float_as_uint(d[2]), false, false);
  const auto p1 = __builtin_amdgcn_permlane32_swap(__float_as_uint(d[1]), __float_as_uint(d[3]), false, false);
  const auto s0 = __builtin_amdgcn_permlane16_swap(p0[0], p1[0], false, false);
  const auto s1 = __builtin_amdgcn_permlane16_swap(p0[1], p1[1], false, false);
  return mf4{__uint_as_float(s0[0]), __uint_as_float(s0[1]), __uint_as_float(s1[0]), __uint_as_float(s1[1])};
}
// Row groups of the register-resident z passes: output tile j (rows 16 j ..
// 16 j + 15) takes source row groups G = 4 j + d (rows 4 G .. 4 G + 3) for
// d in [zp_dlo(hz), zp_dhi(hz)], the groups that meet its band of 2 hz + 1
// rows; groups outside [0, nz / 4) hold no rows and are skipped, not
// multiplied as zeros.  ZP_DMAX bounds -zp_dlo for taps up to 64 (hz <= 31).
constexpr int ZP_DMAX = 8;
__host__ __device__ constexpr int zp_dlo(int hz) { return -((hz + 3) >> 2); }
__host__ __device__ constexpr int zp_dhi(int hz) { return (15 + hz) >> 2; }
__host__ __device__ constexpr int floor4(int d) { return d >= 0 ? d / 4 : -((3 - d) / 4); }

// tail sum Tail(k) (k >= 0) and head sum Head(k) of nt-entry tables
__device__ __forceinline__ float tail_at(const float* tail, int nt, int k) { return k < nt ? tail[max(k, 0)] : 0.f; }
__device__ __forceinline__ float head_at(const float* head, int nt, int k) { return k < 0 ? 0.f : head[min(k, nt - 1)]; }
// The A operand of a z-pass tile's replicate-boundary step (output row zm):
// lane q4 == 0 Tail(zm + hz + 1) of the w tail sums swt, q4 == 1
// Head(zm + hz - nz) of the head sums swh, other lanes 0.  One LDS load with
// no branch: an LDS tap table holds zeros at entries -1 and nt (and beyond),
// which stand for the zero cases of tail_at / head_at.
__device__ __forceinline__ float bnd_tap(const float* swt, const float* swh, int nt, int q4, int zm, int hz, int nz) {
  const float* p = q4 == 0 ? swt + min(zm + hz + 1, nt) : q4 == 1 ? swh + min(max(zm + hz - nz, -1), nt - 1) : swt + nt;
  return *p;
}

// tap tables of one plane in LDS: u, w and the tail/head sums of u and w,
// each with OFF zeros before and at least as many after (the Toeplitz
// fragments index u[n + 2h - k] and u[k - n] without bounds checks)
struct MfTaps {
  int nt, off, tp, size;
  __host__ __device__ explicit MfTaps(int nt_) : nt(nt_) {
    off = 4 * mf_ksteps(nt_);
    tp = (off + 4 * mf_ksteps(nt_) + 16 + 3) / 4 * 4;
    tp += (76 - tp % 64) % 64;  // tp = 12 (mod 64): lanes reading different tables of one plane hit distinct banks
    size = 6 * tp;  // u, w, u tail, u head, w tail, w head
  }
};
// LDS of the matrix-core kernels: the two plane buffers (the larger of the
// two layouts of the first) and tbp tap groups
__host__ __device__ inline size_t mf_lds_floats(const Geo& g, int nt, int tbp) {
  const size_t p = std::max((size_t)g.nz * mf_wx(g.nx, (g.px - 1) / 2, nt), (size_t)mf_rq(g.nx, nt) * mf_wq(g.nz));
  return p + (size_t)mf_rq(g.nz, nt) * mf_wq(g.nx) + (size_t)tbp * MfTaps(nt).size;
}
// planes whose tap tables are loaded at once: up to 8, within the LDS
static int mf_tap_planes(const Geo& g, int nt, int Yc) {
  const size_t base = mf_lds_floats(g, nt, 0), per = MfTaps(nt).size;
  const long fit = ((long)(kMaxLds / sizeof(float)) - (long)base) / (long)per;
  return (int)std::max(1L, std::min<long>(std::min(8, Yc), fit));
}

// The wave's TPW tiles form a run t = wave*TPW + i (the last wave's run is
// clamped to the last tile: recomputed, not stored).
// (ntx > 0: tiles row-major, x fastest; ntx = -ntz: column-major, z fastest)
template <int TPW>
__device__ __forceinline__ void mf_tiles(int wave, int ntile, int ntx, int (&z0)[TPW], int (&x0)[TPW],
                                         bool (&own)[TPW]) {
#pragma unroll
  for (int i = 0; i < TPW; ++i) {
    const int t0 = wave * TPW + i, tt = min(t0, ntile - 1);
    own[i] = t0 < ntile;
    if (ntx > 0) {
      const int tz = tt / ntx;
      z0[i] = 16 * tz;
      x0[i] = 16 * (tt - tz * ntx);
    } else {
      const int tx = tt / -ntx;
      x0[i] = 16 * tx;
      z0[i] = 16 * (tt - tx * -ntx);
    }
  }
}

// ZF (z-fastest layout, :two rotation fused): plane y of the rotated,
// attenuated view is not read from a work volume but computed while it is
// staged: element (x, z-quad) = atten[v][y][x][:] * bilinear(img_t at the
// rotated position of pixel (x, y)) -- the expression of k_rot2_fwd_z, with
// taps in the zero pad given weight 0 (their clamped addresses read a real,
// finite pixel).  The element slots of a thread are split in two parts:
// MF_RA slots are gathered during the x pass (the pass with more live
// registers) and stored after it, MF_RB slots during the z pass; img_t is
// [y][x][z] (nz % 4 == 0), atten [v][y][x][z].
constexpr int MF_RA = 2, MF_RB = 3;
struct RotSlot {
  float4 s11, s12, s21, s22, a;
};
// the taps of pixel (x, y) (theta = 0: the forward is a copy of (x, y)) and
// their weights, 0 for taps in the zero pad
struct RotTaps {
  int ja, jb, ra, rb;
  float wa1, wa2, wb1, wb2, w21, w22;
};
__device__ __forceinline__ RotTaps rot_taps(const RotView& r, int N, int pad, int x, int y) {
  const BiCoef c = bilinear_coef(r, pad, x, y);
  const bool k0 = r.kind == 0;
  const int j1 = k0 ? x : c.j1, j2 = k0 ? y : c.j2;
  const bool x0 = (unsigned)j1 < (unsigned)N, x1 = (unsigned)(j1 + 1) < (unsigned)N;
  const bool y0 = (unsigned)j2 < (unsigned)N, y1 = (unsigned)(j2 + 1) < (unsigned)N;
  RotTaps t;
  t.ja = clampi(j1, 0, N - 1);
  t.jb = clampi(j1 + 1, 0, N - 1);
  t.ra = clampi(j2, 0, N - 1) * N;
  t.rb = clampi(j2 + 1, 0, N - 1) * N;
  t.wa1 = k0 ? 1.f : (x0 && y0 ? c.w11 : 0.f);
  t.wa2 = k0 ? 0.f : (x1 && y0 ? c.w12 : 0.f);
  t.wb1 = k0 ? 0.f : (x0 && y1 ? c.w11 : 0.f);
  t.wb2 = k0 ? 0.f : (x1 && y1 ? c.w12 : 0.f);
  t.w21 = k0 ? 1.f : c.w21;
  t.w22 = k0 ? 0.f : c.w22;
  return t;
}
__device__ __forceinline__ void rot_issue(RotSlot& q, const float* __restrict__ img_t, const float* __restrict__ av,
                                          const RotView& r, int N, int nz, int pad, int x, int y, int zo) {
  const RotTaps t = rot_taps(r, N, pad, x, y);
  q.s11 = *reinterpret_cast<const float4*>(img_t + (long)(t.ra + t.ja) * nz + zo);
  q.s12 = *reinterpret_cast<const float4*>(img_t + (long)(t.ra + t.jb) * nz + zo);
  q.s21 = *reinterpret_cast<const float4*>(img_t + (long)(t.rb + t.ja) * nz + zo);
  q.s22 = *reinterpret_cast<const float4*>(img_t + (long)(t.rb + t.jb) * nz + zo);
  q.a = *reinterpret_cast<const float4*>(av + ((long)y * N + x) * nz + zo);
}
// the weights are recomputed here rather than held with the loads in flight
__device__ __forceinline__ float4 rot_value(const RotSlot& q, const RotView& r, int N, int pad, int x, int y) {
  const RotTaps t = rot_taps(r, N, pad, x, y);
  float4 v;
  v.x = ((q.s11.x * t.wa1 + q.s12.x * t.wa2) * t.w21 + (q.s21.x * t.wb1 + q.s22.x * t.wb2) * t.w22) * q.a.x;
  v.y = ((q.s11.y * t.wa1 + q.s12.y * t.wa2) * t.w21 + (q.s21.y * t.wb1 + q.s22.y * t.wb2) * t.w22) * q.a.y;
  v.z = ((q.s11.z * t.wa1 + q.s12.z * t.wa2) * t.w21 + (q.s21.z * t.wb1 + q.s22.z * t.wb2) * t.w22) * q.a.z;
  v.w = ((q.s11.w * t.wa1 + q.s12.w * t.wa2) * t.w21 + (q.s21.w * t.wb1 + q.s22.w * t.wb2) * t.w22) * q.a.w;
  return v;
}

template <int TPW, int NW, bool ZF>
__global__ void __launch_bounds__(NW * 64) k_psf_fwd_mfma(const float* __restrict__ work, Taps t,
                                                            const int* __restrict__ vids, int Yc,
                                                            float* __restrict__ part, Geo g, int TBP,
                                                            const float* __restrict__ atten, long atten_vstride,
                                                            const RotView* __restrict__ rv) {
  extern __shared__ float L[];
  const int NT = t.ntx;
  const MfTaps TT(NT);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int hx = (g.px - 1) / 2, hz = (g.pz - 1) / 2;
  const int nsx = (16 + g.px - 1 + 3) / 4, nsz = (16 + g.pz - 1 + 3) / 4;
  const int ntx = nx >> 4, ntile = ntx * (nz >> 4);
  const int WX = mf_wx(nx, hx, NT), WQ = mf_wq(nx), RQ = mf_rq(nz, NT), XO = mf_xo(hx);
  // ZF: the plane is staged as rows x + hx of stride WZ (= 16 mod 64) with
  // zero rows around it
  const int WZ = mf_wq(nz), RX = mf_rq(nx, NT);
  const int b = blockIdx.y, c = blockIdx.x, nchunk = gridDim.x;
  const int v = vids[b];
  const int y0 = c * Yc, y1 = min(ny, y0 + Yc);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int m16 = lane & 15, q4 = lane >> 4;
  const int PSZ = ZF ? RX * WZ : nz * WX;
  float* Pz = L;
  float* Qz = Pz + PSZ;
  float* s_taps = Qz + RQ * WQ;  // TBP planes of TT.size
  for (int k = tid; k < PSZ + RQ * WQ + TBP * TT.size; k += (NW * 64)) L[k] = 0.f;

  const long S = (long)nx * ny;
  const int nq = nx >> 2, nz4 = nz >> 2;
  const int nv = nz * nq;  // float4 elements of a plane
  // x-fastest: the plane as float4 along x, slot k of this thread = element
  // tid + k*threads (clamped to the last element: the loads and stores stay
  // unconditional, so no wait is forced between them; a clamped slot
  // rewrites that element)
  constexpr int PF = ZF ? 1 : MF_PF * MF_WAVES / NW;
  const float* wb = work + S * nz * b;
  int pfo[PF], pfs[PF];
  float4 pf[PF];
  // ZF: slot k = element tid + k*threads of the [x][z] plane, (x, z quad)
  // packed as x << 8 | zq
  constexpr int RS = ZF ? MF_RA + MF_RB : 1;
  int rxz[RS];
  RotSlot ra[ZF ? MF_RA : 1], rb[ZF ? MF_RB : 1];
  const RotView rot = ZF ? rv[v] : RotView{};
  const float* av = ZF ? atten + atten_vstride * v : nullptr;
  if (ZF) {
#pragma unroll
    for (int k = 0; k < RS; ++k) {
      const int i = min(tid + k * (NW * 64), nv - 1), x = i / nz4;
      rxz[k] = x << 8 | (i - x * nz4);
    }
  } else {
#pragma unroll
    for (int k = 0; k < PF; ++k) {
      const int i = min(tid + k * (NW * 64), nv - 1);
      const int z = i / nq, x4 = i - z * nq;
      pfo[k] = (int)(S * z) + 4 * x4;
      pfs[k] = z * WX + XO + 4 * x4;
    }
  }
  auto prefetch = [&](int y) {
#pragma unroll
    for (int k = 0; k < PF; ++k) pf[k] = *reinterpret_cast<const float4*>(wb + nx * y + pfo[k]);
  };
  // (scheduling fences keep each part's gathers where they are issued, so
  // that only one part is live at a time)
  auto rot_half = [&](auto& h, int k0, int y) {
    SR_SCHED_FENCE();
#pragma unroll
    for (int k = 0; k < (int)(sizeof(h) / sizeof(h[0])); ++k)
      rot_issue(h[k], work, av, rot, nx, nz, g.pad, rxz[k0 + k] >> 8, y, 4 * (rxz[k0 + k] & 255));
    SR_SCHED_FENCE();
  };
  auto rot_store = [&](auto& h, int k0, int y) {
    SR_SCHED_FENCE();
#pragma unroll
    for (int k = 0; k < (int)(sizeof(h) / sizeof(h[0])); ++k) {
      const int x = rxz[k0 + k] >> 8, zq = rxz[k0 + k] & 255;
      const float4 val = rot_value(h[k], rot, nx, g.pad, x, y);
      *reinterpret_cast<float4*>(Pz + (hx + x) * WZ + 4 * zq) = make_float4(val.x, val.y, val.z, val.w);
    }
  };
  // thread tid < 6 NT loads entry tk of table tab (u, w, u tail, u head, w tail, w head)
  // the tap tables of up to TBP (<= 8) planes at a time: all loads of a
  // thread first, then the stores (clamped entries repeat the last one)
  auto load_taps = [&](int yg, int ng) {
    constexpr int ME = (8 * 6 * 64 + NW * 64 - 1) / (NW * 64);
    const int n = ng * 6 * NT;
    float tvv[ME];
    int tdst[ME];
#pragma unroll
    for (int j = 0; j < ME; ++j) {
      const int e = min(tid + j * NW * 64, n - 1);
      const int pl = e / (6 * NT), r = e - pl * 6 * NT, tab = r / NT, tk = r - tab * NT;
      const long plane = (long)v * t.vstride + yg + pl;
      tdst[j] = pl * TT.size + tab * TT.tp + TT.off + tk;
      const float* src = tab == 0 ? t.u + plane * NT : tab == 1 ? t.w + plane * NT : t.cum + (plane * 4 + tab - 2) * NT;
      tvv[j] = src[tk];
    }
#pragma unroll
    for (int j = 0; j < ME; ++j) s_taps[tdst[j]] = tvv[j];
  };
  if (y0 < y1 && !ZF) prefetch(y0);
  int z0s[TPW], x0s[TPW];
  bool own[TPW];
  mf_tiles<TPW>(wave, ntile, ntx, z0s, x0s, own);
  int aoff[TPW], zoff[TPW];
#pragma unroll
  for (int i = 0; i < TPW; ++i) {
    aoff[i] = ZF ? (x0s[i] + q4) * WZ + z0s[i] + m16 : (z0s[i] + m16) * WX + XO - hx + x0s[i] + q4;
    zoff[i] = (z0s[i] + q4) * WQ + x0s[i] + m16;
  }
  mf4 dv[TPW];
#pragma unroll
  for (int i = 0; i < TPW; ++i) dv[i] = mf4{0.f, 0.f, 0.f, 0.f};
  __syncthreads();  // pads zeroed before the first plane lands
  if (ZF && y0 < y1) {  // the first plane: its latency is exposed once per block
    rot_half(ra, 0, y0);
    rot_store(ra, 0, y0);
    rot_half(rb, MF_RA, y0);
    rot_store(rb, MF_RA, y0);
  }

  for (int y = y0; y < y1; ++y) {
    if ((y - y0) % TBP == 0) {  // next tap group (the previous group's last reads are done)
      lds_barrier();
      load_taps(y, min(TBP, y1 - y));
    }
    const float* tb = s_taps + ((y - y0) % TBP) * TT.size;
    const int yn = min(y + 1, y1 - 1);
    if (!ZF) {
#pragma unroll
      for (int k = 0; k < PF; ++k)
        *reinterpret_cast<float4*>(Pz + pfs[k]) = make_float4(pf[k].x, pf[k].y, pf[k].z, pf[k].w);
    }
    lds_barrier();
    if (ZF) rot_half(ra, 0, yn);
    else prefetch(yn);
    // Toeplitz fragments: x-pass B[k][n] = u[n+2hx-k] = ub[-k], z-pass A[m][k] = w[m+2hz-k] = wz[-k]
    const float* ub = tb + TT.off + m16 - q4 + 2 * hx;
    const float* wz = tb + TT.tp + TT.off + m16 - q4 + 2 * hz;
    const float* sut = tb + 2 * TT.tp + TT.off;
    const float* suh = sut + TT.tp;
    const float* swt = suh + TT.tp;
    const float* swh = swt + TT.tp;
    // x pass: K steps in ping-pong fragment registers (step s+1 loads while s
    // multiplies).  The operands enter swapped, so the tile comes out
    // transposed: lane (k, n) register r holds Q[z0 + n][x0 + 4k + r], four
    // consecutive x stored as one float4.  The replicate boundary goes first,
    // on every tile (its fragments vanish away from the edges).
    mf4 dx[TPW];
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
      const int xn = x0s[i] + m16;
      // P[z][0] and P[z][nx-1] of the tile's row z = z0 + m
      const float pe = q4 == 0 ? (ZF ? Pz[hx * WZ + z0s[i] + m16] : Pz[(z0s[i] + m16) * WX + XO])
                       : q4 == 1 ? (ZF ? Pz[(hx + nx - 1) * WZ + z0s[i] + m16] : Pz[(z0s[i] + m16) * WX + XO + nx - 1])
                                 : 0.f;
      const float te = q4 == 0 ? tail_at(sut, NT, xn + hx + 1) : q4 == 1 ? head_at(suh, NT, xn + hx - nx) : 0.f;
      dx[i] = mfma4(te, pe, mf4{0.f, 0.f, 0.f, 0.f});
    }
    {
      float a0[TPW], a1[TPW], b0, b1;
      auto ld = [&](float (&a)[TPW], float& bb, int s) {
        bb = ub[-4 * s];
#pragma unroll
        for (int i = 0; i < TPW; ++i) a[i] = Pz[aoff[i] + 4 * s * (ZF ? WZ : 1)];
      };
      ld(a0, b0, 0);
      int s = 0;
      for (; s + 1 < nsx; s += 2) {
        ld(a1, b1, s + 1);
        SR_SCHED_FENCE();
#pragma unroll
        for (int i = 0; i < TPW; ++i) dx[i] = mfma4(b0, a0[i], dx[i]);
        SR_SCHED_FENCE();
        ld(a0, b0, min(s + 2, nsx - 1));
        SR_SCHED_FENCE();
#pragma unroll
        for (int i = 0; i < TPW; ++i) dx[i] = mfma4(b1, a1[i], dx[i]);
      }
      if (s < nsx) {
#pragma unroll
        for (int i = 0; i < TPW; ++i) dx[i] = mfma4(b0, a0[i], dx[i]);
      }
    }
    // (a clamped duplicate tile stores the same values as its owner)
#pragma unroll
    for (int i = 0; i < TPW; ++i)
      *reinterpret_cast<float4*>(Qz + (hz + z0s[i] + m16) * WQ + x0s[i] + 4 * q4) =
          make_float4(dx[i][0], dx[i][1], dx[i][2], dx[i][3]);
    lds_barrier();
    // ZF: every wave's x pass is done, so Pz takes half A of the next plane
    // (half B follows after the z pass)
    if (ZF) {
      rot_store(ra, 0, yn);
      rot_half(rb, MF_RA, yn);
    }
    // z pass: accumulate the view sum, the replicate boundary first
    {
      // (lanes q4 >= 2: zero weight on a real value; the loads first, then the products)
      float ae[TPW], be[TPW];
#pragma unroll
      for (int i = 0; i < TPW; ++i) {
        ae[i] = bnd_tap(swt, swh, NT, q4, z0s[i] + m16, hz, nz);
        be[i] = Qz[(q4 == 1 ? hz + nz - 1 : hz) * WQ + x0s[i] + m16];
      }
#pragma unroll
      for (int i = 0; i < TPW; ++i) dv[i] = mfma4(ae[i], be[i], dv[i]);
    }
    {
      float bq0[TPW], bq1[TPW], a0, a1;
      auto ld = [&](float (&bq)[TPW], float& aa, int s) {
        aa = wz[-4 * s];
#pragma unroll
        for (int i = 0; i < TPW; ++i) bq[i] = Qz[zoff[i] + 4 * s * WQ];
      };
      ld(bq0, a0, 0);
      int s = 0;
      for (; s + 1 < nsz; s += 2) {
        ld(bq1, a1, s + 1);
        SR_SCHED_FENCE();
#pragma unroll
        for (int i = 0; i < TPW; ++i) dv[i] = mfma4(a0, bq0[i], dv[i]);
        SR_SCHED_FENCE();
        ld(bq0, a0, min(s + 2, nsz - 1));
        SR_SCHED_FENCE();
#pragma unroll
        for (int i = 0; i < TPW; ++i) dv[i] = mfma4(a1, bq1[i], dv[i]);
      }
      if (s < nsz) {
#pragma unroll
        for (int i = 0; i < TPW; ++i) dv[i] = mfma4(a0, bq0[i], dv[i]);
      }
    }
    if (ZF) rot_store(rb, MF_RA, yn);
    // the next plane's stores touch Pz (last read before the barrier above)
    // and the other tap buffer; Qz is rewritten only after the next barrier
  }
  float* pb = part + ((long)b * nchunk + c) * nx * nz;
#pragma unroll
  for (int i = 0; i < TPW; ++i) {
    if (own[i]) {
#pragma unroll
      for (int r = 0; r < 4; ++r) pb[(long)(z0s[i] + 4 * q4 + r) * nx + x0s[i] + m16] = dv[i][r];
    }
  }
}

// --- forward, separable, matrix cores, column-owned, :two rotation fused -----
// z-fastest forward with one block barrier per plane.  Wave w owns the 16-wide
// x column w (nx <= 128) with all nz/16 z tiles: the x pass reads the block's
// staged plane, its result goes to a wave-private buffer, and the z pass reads
// only that buffer.  The plane is double-buffered: plane y+1 is gathered
// (rotation and attenuation fused, as in k_psf_fwd_mfma) and stored during
// plane y's passes.  The plane buffers carry no pad rows: the x pass reads row
// clamp(x0 - hx + k), which is the replicate-padded convolution itself, so
// only the z pass needs its boundary step (zero pad rows in the wave buffer).
//   x pass  Q[z][x0+n] = sum_k P[clamp(x0-hx+k)][z] u[n+2hx-k]
//   z pass  R[z0+m][x] = sum_k w[m+2hz-k] Q[z0-hz+k][x] + in[0]Tail + in[nz-1]Head
// LDS: two planes (rows x, stride WZ = 16 mod 64 floats), per wave Qw (rows
// z + hz, 16 columns in four 4-float chunks, chunk c of row r stored at
// c ^ (((r - hz) / 2) & 3): the x pass's float4 stores then spread over all
// banks), TBP tap planes.  grid (nchunk, nb), block 512.
__host__ __device__ inline size_t mf_fcol_lds_floats(const Geo& g, int nt, int tbp) {
  return 2 * (size_t)g.nx * mf_wq(g.nz) + (size_t)MF_WAVES * mf_rq(g.nz, nt) * 16 + (size_t)tbp * MfTaps(nt).size +
         16 * (size_t)g.nx;
}
// One pixel x of plane y of the :two rotation, for the fused forward's tap
// table: e[0] = element offsets ((row * N + col) * nz) of the four taps,
// clamped into the image; e[1] = the bilinear weights with the x masks in
// (wx1, wx2) and the y masks in (wy1, wy2).  A tap in the zero pad gets a
// zero factor, so ((s11 wx1 + s12 wx2) wy1 + (s21 wx1 + s22 wx2) wy2) equals
// the per-tap masked form of rot_taps / rot_value for finite data.
// (theta = 0: the copy of (x, y), all four offsets on that pixel)
__device__ __forceinline__ void rot_entry(const RotView& r, int N, int nz, int pad, int x, int y, float4* e) {
  int o11, o12, o21, o22;
  float wx1 = 1.f, wx2 = 0.f, wy1 = 1.f, wy2 = 0.f;
  if (r.kind == 0) {
    o11 = o12 = o21 = o22 = (y * N + x) * nz;
  } else {
    const BiCoef c = bilinear_coef(r, pad, x, y);
    const int ja = clampi(c.j1, 0, N - 1), jb = clampi(c.j1 + 1, 0, N - 1);
    const int ra = clampi(c.j2, 0, N - 1) * N, rb = clampi(c.j2 + 1, 0, N - 1) * N;
    o11 = (ra + ja) * nz;
    o12 = (ra + jb) * nz;
    o21 = (rb + ja) * nz;
    o22 = (rb + jb) * nz;
    wx1 = (unsigned)c.j1 < (unsigned)N ? c.w11 : 0.f;
    wx2 = (unsigned)(c.j1 + 1) < (unsigned)N ? c.w12 : 0.f;
    wy1 = (unsigned)c.j2 < (unsigned)N ? c.w21 : 0.f;
    wy2 = (unsigned)(c.j2 + 1) < (unsigned)N ? c.w22 : 0.f;
  }
  e[0] = make_float4(__int_as_float(o11), __int_as_float(o12), __int_as_float(o21), __int_as_float(o22));
  e[1] = make_float4(wx1, wx2, wy1, wy2);
}
template <int TPW>
__global__ void __launch_bounds__(MF_THREADS) k_psf_fwd_col(const float* __restrict__ img_t, Taps t,
                                                            const int* __restrict__ vids, int Yc,
                                                            float* __restrict__ part, Geo g, int TBP,
                                                            const float* __restrict__ atten, long atten_vstride,
                                                            const RotView* __restrict__ rv) {
  extern __shared__ float L[];
  const int NT = t.ntx;
  const MfTaps TT(NT);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int hx = (g.px - 1) / 2, hz = (g.pz - 1) / 2;
  const int nsx = (16 + g.px - 1 + 3) / 4, nsz = (16 + g.pz - 1 + 3) / 4;
  const int ntx = nx >> 4;
  const int WZ = mf_wq(nz), RG = mf_rq(nz, NT), PS = nx * WZ;
  const int b = blockIdx.y, c = blockIdx.x, nchunk = gridDim.x;
  const int v = vids[b];
  const int y0 = c * Yc, y1 = min(ny, y0 + Yc);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int m16 = lane & 15, q4 = lane >> 4;
  float* Pz = L;  // two plane buffers
  float* Qw = L + 2 * PS + wave * RG * 16;
  float* s_taps = L + 2 * PS + MF_WAVES * RG * 16;  // TBP planes of TT.size
  // two tap tables of the rotation (planes of even / odd offset from y0),
  // nx entries of two float4 each (rot_entry)
  float4* s_rt = reinterpret_cast<float4*>(s_taps + TBP * TT.size);
  // zeros where no plane data lands: the wave buffers' pad rows and the tap tables' pads
  for (int k = tid; k < MF_WAVES * RG * 16 + TBP * TT.size; k += MF_THREADS) L[2 * PS + k] = 0.f;

  // element slots: element tid + k*threads of the [x][z] plane, (x, z quad)
  // packed as x << 8 | zq; xo = the element's offset in a [x][z] plane
  const int nz4 = nz >> 2, nv = nx * nz4;
  constexpr int RS = MF_RA + MF_RB;
  int rxz[RS], xo[RS];
  RotSlot ra[MF_RA], rb[MF_RB];
  const RotView rot = rv[v];
  const float* av = atten + atten_vstride * v;
#pragma unroll
  for (int k = 0; k < RS; ++k) {
    const int i = min(tid + k * MF_THREADS, nv - 1), x = i / nz4;
    rxz[k] = x << 8 | (i - x * nz4);
    xo[k] = x * nz + 4 * (i - x * nz4);
  }
  auto rot_table = [&](int y) {  // plane y's entries (one pixel per thread)
    if (tid < nx) rot_entry(rot, nx, nz, g.pad, tid, y, s_rt + 2 * nx * ((y - y0) & 1) + 2 * tid);
  };
  // (scheduling fences keep each part's gathers where they are issued, so
  // that only one part is live at a time)
  auto rot_half = [&](auto& h, int k0, int y) {
    const float4* T = s_rt + 2 * nx * ((y - y0) & 1);
    const float* ay = av + (long)y * nx * nz;
    SR_SCHED_FENCE();
#pragma unroll
    for (int k = 0; k < (int)(sizeof(h) / sizeof(h[0])); ++k) {
      const float4 o = T[2 * (rxz[k0 + k] >> 8)];
      const int zo = 4 * (rxz[k0 + k] & 255);
      h[k].s11 = *reinterpret_cast<const float4*>(img_t + (__float_as_int(o.x) + zo));
      h[k].s12 = *reinterpret_cast<const float4*>(img_t + (__float_as_int(o.y) + zo));
      h[k].s21 = *reinterpret_cast<const float4*>(img_t + (__float_as_int(o.z) + zo));
      h[k].s22 = *reinterpret_cast<const float4*>(img_t + (__float_as_int(o.w) + zo));
      h[k].a = *reinterpret_cast<const float4*>(ay + xo[k0 + k]);
    }
    SR_SCHED_FENCE();
  };
  auto rot_store = [&](auto& h, int k0, int y, float* buf) {
    const float4* T = s_rt + 2 * nx * ((y - y0) & 1);
    SR_SCHED_FENCE();
#pragma unroll
    for (int k = 0; k < (int)(sizeof(h) / sizeof(h[0])); ++k) {
      const int x = rxz[k0 + k] >> 8, zq = rxz[k0 + k] & 255;
      const float4 w = T[2 * x + 1];
      const RotSlot& q = h[k];
      float4 val;
      val.x = ((q.s11.x * w.x + q.s12.x * w.y) * w.z + (q.s21.x * w.x + q.s22.x * w.y) * w.w) * q.a.x;
      val.y = ((q.s11.y * w.x + q.s12.y * w.y) * w.z + (q.s21.y * w.x + q.s22.y * w.y) * w.w) * q.a.y;
      val.z = ((q.s11.z * w.x + q.s12.z * w.y) * w.z + (q.s21.z * w.x + q.s22.z * w.y) * w.w) * q.a.z;
      val.w = ((q.s11.w * w.x + q.s12.w * w.y) * w.z + (q.s21.w * w.x + q.s22.w * w.y) * w.w) * q.a.w;
      *reinterpret_cast<float4*>(buf + x * WZ + 4 * zq) = val;
    }
  };
  auto load_taps = [&](int yg, int ng) {
    constexpr int ME = (8 * 6 * 64 + MF_THREADS - 1) / MF_THREADS;
    const int n = ng * 6 * NT;
    float tvv[ME];
    int tdst[ME];
#pragma unroll
    for (int j = 0; j < ME; ++j) {
      const int e = min(tid + j * MF_THREADS, n - 1);
      const int pl = e / (6 * NT), r = e - pl * 6 * NT, tab = r / NT, tk = r - tab * NT;
      const long plane = (long)v * t.vstride + yg + pl;
      tdst[j] = pl * TT.size + tab * TT.tp + TT.off + tk;
      const float* src = tab == 0 ? t.u + plane * NT : tab == 1 ? t.w + plane * NT : t.cum + (plane * 4 + tab - 2) * NT;
      tvv[j] = src[tk];
    }
#pragma unroll
    for (int j = 0; j < ME; ++j) s_taps[tdst[j]] = tvv[j];
  };
  // waves past the last column recompute it and store nothing
  const bool own = wave < ntx;
  const int x0 = 16 * min(wave, ntx - 1);
  mf4 dv[TPW];
#pragma unroll
  for (int i = 0; i < TPW; ++i) dv[i] = mf4{0.f, 0.f, 0.f, 0.f};
  if (y0 < y1) {  // the first plane and tap group: all their loads in flight together
    rot_table(y0);
    if (y0 + 1 < y1) rot_table(y0 + 1);
    __syncthreads();  // zero fill and the first two rotation tables done
    rot_half(ra, 0, y0);
    rot_half(rb, MF_RA, y0);
    load_taps(y0, min(TBP, y1 - y0));
    rot_store(ra, 0, y0, Pz);
    rot_store(rb, MF_RA, y0, Pz);
  }

  for (int y = y0; y < y1; ++y) {
    const int cur = (y - y0) & 1;
    lds_barrier();  // plane y staged; every wave is done with plane y-1 (its buffer, taps and rotation table)
    if (y > y0 && (y - y0) % TBP == 0) {
      load_taps(y, min(TBP, y1 - y));
      lds_barrier();
    }
    const float* tb = s_taps + ((y - y0) % TBP) * TT.size;
    const float* Pc = Pz + cur * PS;
    float* Pn = Pz + (cur ^ 1) * PS;
    const int yn = min(y + 1, y1 - 1);
    rot_half(ra, 0, yn);
    // plane y + 2's rotation table, into plane y's (read before the barrier
    // above; its next readers stage plane y + 2, after the next barrier)
    if (y + 2 < y1) rot_table(y + 2);
    // x pass (operands swapped: lane (k, n) register r = Q[z0 + n][x0 + 4k + r])
    const float* ub = tb + TT.off + m16 - q4 + 2 * hx;
    mf4 dx[TPW];
#pragma unroll
    for (int i = 0; i < TPW; ++i) dx[i] = mf4{0.f, 0.f, 0.f, 0.f};
    {
      float a0[TPW], a1[TPW], b0, b1;
      auto ld = [&](float (&a)[TPW], float& bb, int s) {
        bb = ub[-4 * s];
        const float* pr = Pc + __umul24(clampi(x0 - hx + 4 * s + q4, 0, nx - 1), WZ) + m16;
#pragma unroll
        for (int i = 0; i < TPW; ++i) a[i] = pr[16 * i];
      };
      ld(a0, b0, 0);
      int s = 0;
      for (; s + 1 < nsx; s += 2) {
        ld(a1, b1, s + 1);
        SR_SCHED_FENCE();
#pragma unroll
        for (int i = 0; i < TPW; ++i) dx[i] = mfma4(b0, a0[i], dx[i]);
        SR_SCHED_FENCE();
        ld(a0, b0, min(s + 2, nsx - 1));
        SR_SCHED_FENCE();
#pragma unroll
        for (int i = 0; i < TPW; ++i) dx[i] = mfma4(b1, a1[i], dx[i]);
      }
      if (s < nsx) {
#pragma unroll
        for (int i = 0; i < TPW; ++i) dx[i] = mfma4(b0, a0[i], dx[i]);
      }
    }
#pragma unroll
    for (int i = 0; i < TPW; ++i)
      *reinterpret_cast<float4*>(Qw + (hz + 16 * i + m16) * 16 + 4 * (q4 ^ ((m16 >> 1) & 3))) =
          make_float4(dx[i][0], dx[i][1], dx[i][2], dx[i][3]);
    // the wave reads back what its own lanes wrote (LDS keeps a wave's order)
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    rot_store(ra, 0, yn, Pn);
    rot_half(rb, MF_RA, yn);
    // z pass: accumulate the view sum, the replicate boundary first
    const float* wz = tb + TT.tp + TT.off + m16 - q4 + 2 * hz;
    const float* swt = tb + 4 * TT.tp + TT.off;
    const float* swh = swt + TT.tp;
    {
      // (lanes q4 >= 2: zero weight on a real value; the loads first, then the products)
      const float be = Qw[q4 == 1 ? (hz + nz - 1) * 16 + (m16 ^ ((((nz - 1) >> 1) & 3) << 2)) : hz * 16 + m16];
      float ae[TPW];
#pragma unroll
      for (int i = 0; i < TPW; ++i) ae[i] = bnd_tap(swt, swh, NT, q4, 16 * i + m16, hz, nz);
#pragma unroll
      for (int i = 0; i < TPW; ++i) dv[i] = mfma4(ae[i], be, dv[i]);
    }
    {
      // (Qw rows are swizzled in 4-float chunks: row r, column x at chunk
      // (x / 4) ^ (((r - hz) / 2) & 3), which spreads the float4 stores of the
      // x pass over all banks; a load's 4 rows x 16 columns stay on distinct banks)
      const int hq = ((q4 - hz + 64) >> 1) & 3;
      float bq0[TPW], bq1[TPW], a0, a1;
      auto ld = [&](float (&bq)[TPW], float& aa, int s) {
        aa = wz[-4 * s];
        const float* qr = Qw + (q4 + 4 * s) * 16 + (m16 ^ (((hq + 2 * s) & 3) << 2));
#pragma unroll
        for (int i = 0; i < TPW; ++i) bq[i] = qr[256 * i];
      };
      ld(bq0, a0, 0);
      int s = 0;
      for (; s + 1 < nsz; s += 2) {
        ld(bq1, a1, s + 1);
        SR_SCHED_FENCE();
#pragma unroll
        for (int i = 0; i < TPW; ++i) dv[i] = mfma4(a0, bq0[i], dv[i]);
        SR_SCHED_FENCE();
        ld(bq0, a0, min(s + 2, nsz - 1));
        SR_SCHED_FENCE();
#pragma unroll
        for (int i = 0; i < TPW; ++i) dv[i] = mfma4(a1, bq1[i], dv[i]);
      }
      if (s < nsz) {
#pragma unroll
        for (int i = 0; i < TPW; ++i) dv[i] = mfma4(a0, bq0[i], dv[i]);
      }
    }
    asm volatile("" ::: "memory");  // the next plane's Qw stores stay after these reads
    rot_store(rb, MF_RA, yn, Pn);
  }
  if (own) {
    float* pb = part + ((long)b * nchunk + c) * nx * nz;
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) pb[(long)(16 * i + 4 * q4 + r) * nx + x0 + m16] = dv[i][r];
    }
  }
}

// --- forward, z-fastest, producer and consumer waves -----------------------
// The passes of k_psf_fwd_col, with the staging moved to waves of their own:
// waves 0..7 (consumers) own an x column each and run plane y's x pass from
// the staged plane and its z pass from registers (the x pass's tiles
// transposed in 4 x 4 blocks of lane groups), while the NPW producer waves
// stage plane y + 1 (the gathers of the :two rotation through its tap table,
// attenuation, bilinear combination), load plane y + 1's PSF taps and build
// plane y + 2's rotation table.  The gathers and their arithmetic then
// overlap the matrix products instead of alternating with them; one barrier
// per plane.  nz = 16 TPW.
// LDS: two planes (rows x, stride WZ), two tap planes (planes of even / odd
// offset from y0), two rotation tables.
// grid (nchunk, nb), block 512 + 64 NPW.
constexpr int WS_RE = 5;  // element slots per producer thread and round
__host__ __device__ inline size_t mf_fws_lds_floats(const Geo& g, int nt) {
  return 2 * (size_t)g.nx * mf_wq(g.nz) + 2 * (size_t)MfTaps(nt).size + 16 * (size_t)g.nx;
}
// GRP: the grid holds the blocks of the ngrp launches that grp describes
// (the arguments before grp are unused).
template <int TPW, int NPW, bool GRP>
__global__ void __launch_bounds__(MF_THREADS + 64 * NPW) k_psf_fwd_ws(const float* __restrict__ img_t, Taps t,
                                                                    const int* __restrict__ vids, int Yc,
                                                                    float* __restrict__ part, Geo g,
                                                                    const float* __restrict__ atten,
                                                                    long atten_vstride,
                                                                    const RotView* __restrict__ rv,
                                                                    const PsfFwdGrp* __restrict__ grp, int ngrp) {
  constexpr int PT = 64 * NPW;                       // producer threads
  constexpr int RR = (10240 / 4 + PT * WS_RE - 1) / (PT * WS_RE);  // rounds (nx nz <= 10240)
  extern __shared__ float L[];
  int b = blockIdx.y, c = blockIdx.x, nchunk = gridDim.x;
  if constexpr (GRP) {
    const PsfFwdGrp& a = grp[grp_find(grp, ngrp)];
    const int l = (int)blockIdx.x - a.h.blk0;
    img_t = glob(a.img_t), t = glob(a.t), vids = glob(a.vids), Yc = a.Yc, part = glob(a.part), g = a.g;
    atten = glob(a.atten), atten_vstride = a.atten_vstride, rv = glob(a.rv);
    b = __builtin_amdgcn_readfirstlane(l / a.h.gx);  // (uniform: kept in scalar registers)
    c = __builtin_amdgcn_readfirstlane(l - b * a.h.gx);
    nchunk = a.h.gx;
  }
  const int NT = t.ntx;
  const MfTaps TT(NT);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int hx = (g.px - 1) / 2, hz = (g.pz - 1) / 2;
  const int nsx = (16 + g.px - 1 + 3) / 4;
  const int ntx = nx >> 4;
  const int WZ = mf_wq(nz), PS = nx * WZ;
  const int v = vids[b];
  const int y0 = c * Yc, y1 = min(ny, y0 + Yc);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const bool producer = wave >= MF_WAVES;
  const int m16 = lane & 15, q4 = lane >> 4;
  float* Pz = L;  // two plane buffers
  float* s_taps = L + 2 * PS;  // two tap planes of TT.size
  float4* s_rt = reinterpret_cast<float4*>(s_taps + 2 * TT.size);
  for (int k = tid; k < 2 * TT.size; k += MF_THREADS + PT) L[2 * PS + k] = 0.f;

  if (producer) {
    const RotView rot = rv[v];
    const float* av = atten + atten_vstride * v;
    const int nz4 = nz >> 2, nv = nx * nz4;
    const int pt = tid - MF_THREADS;  // producer thread
    // slots: element pt + (r * WS_RE + k) * PT of the [x][z] plane, packed as
    // x << 8 | zq (clamped to the last element: a clamped slot rewrites it)
    int sxz[RR][WS_RE];
#pragma unroll
    for (int r = 0; r < RR; ++r)
#pragma unroll
      for (int k = 0; k < WS_RE; ++k) {
        const int i = min(pt + (r * WS_RE + k) * PT, nv - 1), x = i / nz4;
        sxz[r][k] = x << 8 | (i - x * nz4);
      }
    auto rot_table = [&](int y) {  // plane y's rotation table, one pixel per producer thread
      if (pt < nx) rot_entry(rot, nx, nz, g.pad, pt, y, s_rt + 2 * nx * ((y - y0) & 1) + 2 * pt);
    };
    // plane y's six tap tables: fetched into registers one plane ahead (their
    // loads overlap the gathers instead of adding a second memory round trip
    // per plane), then stored into tap plane (y - y0) & 1
    constexpr int TE = (6 * 64 + PT - 1) / PT;  // (NT <= 64)
    float tv[TE];
    auto fetch_taps = [&](int y) {
      const long plane = (long)v * t.vstride + y;
#pragma unroll
      for (int j = 0; j < TE; ++j) {
        const int e = pt + j * PT;
        if (e < 6 * NT) {
          const int tab = e / NT, tk = e - tab * NT;
          const float* src =
              tab == 0 ? t.u + plane * NT : tab == 1 ? t.w + plane * NT : t.cum + (plane * 4 + tab - 2) * NT;
          tv[j] = src[tk];
        }
      }
    };
    auto put_taps = [&](int y) {
      float* dst = s_taps + ((y - y0) & 1) * TT.size + TT.off;
#pragma unroll
      for (int j = 0; j < TE; ++j) {
        const int e = pt + j * PT;
        if (e < 6 * NT) {
          const int tab = e / NT, tk = e - tab * NT;
          dst[tab * TT.tp + tk] = tv[j];
        }
      }
    };
    auto stage = [&](int y, float* buf) {  // plane y into buf, RR rounds of WS_RE elements
      const float4* T = s_rt + 2 * nx * ((y - y0) & 1);
      const float* ay = av + (long)y * nx * nz;
#pragma unroll
      for (int r = 0; r < RR; ++r) {
        if (r * WS_RE * PT >= nv) break;
        // (the slot words are opaque here, so their addresses are derived per
        // round instead of being hoisted out of the plane loop and held live)
#pragma unroll
        for (int k = 0; k < WS_RE; ++k) asm volatile("" : "+v"(sxz[r][k]));
        RotSlot h[WS_RE];
        SR_SCHED_FENCE();
#pragma unroll
        for (int k = 0; k < WS_RE; ++k) {
          const int x = sxz[r][k] >> 8, zq = sxz[r][k] & 255;
          const float4 o = T[2 * x];
          h[k].s11 = *reinterpret_cast<const float4*>(img_t + (__float_as_int(o.x) + 4 * zq));
          h[k].s12 = *reinterpret_cast<const float4*>(img_t + (__float_as_int(o.y) + 4 * zq));
          h[k].s21 = *reinterpret_cast<const float4*>(img_t + (__float_as_int(o.z) + 4 * zq));
          h[k].s22 = *reinterpret_cast<const float4*>(img_t + (__float_as_int(o.w) + 4 * zq));
          h[k].a = *reinterpret_cast<const float4*>(ay + 4 * __umul24(x, nz4) + 4 * zq);
        }
        SR_SCHED_FENCE();
#pragma unroll
        for (int k = 0; k < WS_RE; ++k) {
          const int x = sxz[r][k] >> 8, zq = sxz[r][k] & 255;
          const float4 w = T[2 * x + 1];
          const RotSlot& q = h[k];
          float4 val;
          val.x = ((q.s11.x * w.x + q.s12.x * w.y) * w.z + (q.s21.x * w.x + q.s22.x * w.y) * w.w) * q.a.x;
          val.y = ((q.s11.y * w.x + q.s12.y * w.y) * w.z + (q.s21.y * w.x + q.s22.y * w.y) * w.w) * q.a.y;
          val.z = ((q.s11.z * w.x + q.s12.z * w.y) * w.z + (q.s21.z * w.x + q.s22.z * w.y) * w.w) * q.a.z;
          val.w = ((q.s11.w * w.x + q.s12.w * w.y) * w.z + (q.s21.w * w.x + q.s22.w * w.y) * w.w) * q.a.w;
          *reinterpret_cast<float4*>(buf + x * WZ + 4 * zq) = val;
        }
        SR_SCHED_FENCE();
      }
    };
    if (y0 < y1) {
      rot_table(y0);
      if (y0 + 1 < y1) rot_table(y0 + 1);
    }
    __syncthreads();  // zero fill and the first two rotation tables done
    if (y0 < y1) {
      fetch_taps(y0);
      stage(y0, Pz);
      put_taps(y0);
      if (y0 + 1 < y1) fetch_taps(y0 + 1);
    }
    // (the same barriers as the consumer loop below, one per plane)
    for (int y = y0; y < y1; ++y) {
      lds_barrier();  // plane y and its taps staged; plane y - 1's buffers free
      if (y + 1 < y1) {
        put_taps(y + 1);
        if (y + 2 < y1) fetch_taps(y + 2);
        stage(y + 1, Pz + (((y - y0) & 1) ^ 1) * PS);
      }
      // plane y + 2's rotation table, into plane y's (read before the barrier above)
      if (y + 2 < y1) rot_table(y + 2);
    }
    return;
  }

  __syncthreads();  // zero fill and the first two rotation tables done
  const bool own = wave < ntx;  // (waves past the last column recompute it and store nothing)
  const int x0 = 16 * min(wave, ntx - 1);
  const int dlo = zp_dlo(hz), dhi = zp_dhi(hz);
  mf4 dv[TPW];
#pragma unroll
  for (int i = 0; i < TPW; ++i) dv[i] = mf4{0.f, 0.f, 0.f, 0.f};

  for (int y = y0; y < y1; ++y) {
    const int cur = (y - y0) & 1;
    lds_barrier();  // plane y and its taps staged; plane y - 1's buffers free
    const float* tb = s_taps + cur * TT.size;
    const float* Pc = Pz + cur * PS;
    // x pass: lane (k, n) register r = Q[16 i + 4 k + r][x0 + n]
    const float* ub = tb + TT.off + m16 - q4 + 2 * hx;
    mf4 dx[TPW];
#pragma unroll
    for (int i = 0; i < TPW; ++i) dx[i] = mf4{0.f, 0.f, 0.f, 0.f};
    {
      float a0[TPW], a1[TPW], b0, b1;
      auto ld = [&](float (&a)[TPW], float& bb, int s) {
        bb = ub[-4 * s];
        const float* pr = Pc + __umul24(clampi(x0 - hx + 4 * s + q4, 0, nx - 1), WZ) + m16;
#pragma unroll
        for (int i = 0; i < TPW; ++i) a[i] = pr[16 * i];
      };
      ld(a0, b0, 0);
      int s = 0;
      for (; s + 1 < nsx; s += 2) {
        ld(a1, b1, s + 1);
        SR_SCHED_FENCE();
#pragma unroll
        for (int i = 0; i < TPW; ++i) dx[i] = mfma4(a0[i], b0, dx[i]);
        SR_SCHED_FENCE();
        ld(a0, b0, min(s + 2, nsx - 1));
        SR_SCHED_FENCE();
#pragma unroll
        for (int i = 0; i < TPW; ++i) dx[i] = mfma4(a1[i], b1, dx[i]);
      }
      if (s < nsx) {
#pragma unroll
        for (int i = 0; i < TPW; ++i) dx[i] = mfma4(a0[i], b0, dx[i]);
      }
    }
    // register g of tile i: lane (k, n) = Q[16 i + 4 g + k][x0 + n], the B
    // operand of the z pass's step over source row group 4 i + g
#pragma unroll
    for (int i = 0; i < TPW; ++i) dx[i] = xpose_groups(dx[i]);
    // z pass: accumulate the view sum.  Output tile j, source group G = 4 j + d:
    // A (m, k) = w[m - k + hz - 4 d]; the replicate boundary rides on the
    // groups holding rows 0 and nz - 1 (k = 0 of G = 0 adds Tail(z + hz + 1),
    // k = 3 of the last group Head(z + hz - nz), z = 16 j + m).
    const float* wz = tb + TT.tp + TT.off + m16 - q4 + hz;
    const float* swt = tb + 4 * TT.tp + TT.off;
    const float* swh = swt + TT.tp;
#pragma unroll
    for (int d = -ZP_DMAX; d <= ZP_DMAX + 3; ++d) {
      if (d < dlo || d > dhi) continue;  // (wave-uniform)
      const int e = floor4(d), gg = d - 4 * e;
      const float a = wz[-4 * d];
#pragma unroll
      for (int j = 0; j < TPW; ++j) {
        const int i = j + e;
        if (i < 0 || i >= TPW) continue;
        float aa = a;
        if (i == 0 && gg == 0) aa += q4 == 0 ? swt[min(16 * j + m16 + hz + 1, NT)] : 0.f;
        if (i == TPW - 1 && gg == 3) aa += q4 == 3 ? swh[min(max(16 * j + m16 + hz - nz, -1), NT - 1)] : 0.f;
        dv[j] = mfma4(aa, dx[i][gg], dv[j]);
      }
    }
  }
  if (own) {
    float* pb = part + ((long)b * nchunk + c) * nx * nz;
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) pb[(long)(16 * i + 4 * q4 + r) * nx + x0 + m16] = dv[i][r];
    }
  }
}

// --- forward, z-fastest, producer waves and column-pair consumers ----------
// As k_psf_fwd_ws with 4 producer waves, but 4 consumer waves, each owning
// two x columns A and B.  A wave's x pass and z pass of one column depend on
// each other through its LDS buffer; here they are taken from different
// columns and planes and merged into one K loop, so every step has 2 TPW
// independent products:
//   loop 1: x pass of A (plane y)   with the z pass of B (plane y - 1)
//   loop 2: x pass of B (plane y)   with the z pass of A (plane y)
// and after the last plane the z pass of B alone.  Each column keeps one
// wave buffer (a loop's z pass reads it before the next x pass rewrites it).
// Taps: a ring of three planes (the z pass of B still reads plane y - 1's
// while the producers load plane y + 1's).  grid (nchunk, nb), block 512.
constexpr int WS2_TAPS = 3;
__host__ __device__ inline size_t mf_fws2_lds_floats(const Geo& g, int nt) {
  return 2 * (size_t)g.nx * mf_wq(g.nz) + (size_t)MF_WAVES * mf_rq(g.nz, nt) * 16 +
         WS2_TAPS * (size_t)MfTaps(nt).size + 16 * (size_t)g.nx;
}
template <int TPW>
__global__ void __launch_bounds__(MF_THREADS) k_psf_fwd_ws2(const float* __restrict__ img_t, Taps t,
                                                            const int* __restrict__ vids, int Yc,
                                                            float* __restrict__ part, Geo g,
                                                            const float* __restrict__ atten, long atten_vstride,
                                                            const RotView* __restrict__ rv) {
  constexpr int NCW = 4;               // consumer waves (two columns each)
  constexpr int PT = MF_THREADS - 64 * NCW;  // producer threads
  constexpr int RR = (10240 / 4 + PT * WS_RE - 1) / (PT * WS_RE);
  extern __shared__ float L[];
  const int NT = t.ntx;
  const MfTaps TT(NT);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int hx = (g.px - 1) / 2, hz = (g.pz - 1) / 2;
  const int nsx = (16 + g.px - 1 + 3) / 4, nsz = (16 + g.pz - 1 + 3) / 4;
  const int ntx = nx >> 4;
  const int WZ = mf_wq(nz), RG = mf_rq(nz, NT), PS = nx * WZ;
  const int b = blockIdx.y, c = blockIdx.x, nchunk = gridDim.x;
  const int v = vids[b];
  const int y0 = c * Yc, y1 = min(ny, y0 + Yc);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const bool producer = wave >= NCW;
  const int m16 = lane & 15, q4 = lane >> 4;
  float* Pz = L;  // two plane buffers
  float* s_taps = L + 2 * PS + MF_WAVES * RG * 16;  // WS2_TAPS tap planes of TT.size
  float4* s_rt = reinterpret_cast<float4*>(s_taps + WS2_TAPS * TT.size);
  for (int k = tid; k < MF_WAVES * RG * 16 + WS2_TAPS * TT.size; k += MF_THREADS) L[2 * PS + k] = 0.f;
  auto tap_plane = [&](int y) { return s_taps + ((y - y0) % WS2_TAPS) * TT.size; };

  if (producer) {
    const RotView rot = rv[v];
    const float* av = atten + atten_vstride * v;
    const int nz4 = nz >> 2, nv = nx * nz4;
    const int pt = tid - 64 * NCW;
    int sxz[RR][WS_RE];
#pragma unroll
    for (int r = 0; r < RR; ++r)
#pragma unroll
      for (int k = 0; k < WS_RE; ++k) {
        const int i = min(pt + (r * WS_RE + k) * PT, nv - 1), x = i / nz4;
        sxz[r][k] = x << 8 | (i - x * nz4);
      }
    auto rot_table = [&](int y) {
      if (pt < nx) rot_entry(rot, nx, nz, g.pad, pt, y, s_rt + 2 * nx * ((y - y0) & 1) + 2 * pt);
    };
    auto load_taps = [&](int y) {
      const long plane = (long)v * t.vstride + y;
      float* dst = tap_plane(y) + TT.off;
      for (int e = pt; e < 6 * NT; e += PT) {
        const int tab = e / NT, tk = e - tab * NT;
        const float* src =
            tab == 0 ? t.u + plane * NT : tab == 1 ? t.w + plane * NT : t.cum + (plane * 4 + tab - 2) * NT;
        dst[tab * TT.tp + tk] = src[tk];
      }
    };
    auto stage = [&](int y, float* buf) {
      const float4* T = s_rt + 2 * nx * ((y - y0) & 1);
      const float* ay = av + (long)y * nx * nz;
#pragma unroll
      for (int r = 0; r < RR; ++r) {
        if (r * WS_RE * PT >= nv) break;
#pragma unroll
        for (int k = 0; k < WS_RE; ++k) asm volatile("" : "+v"(sxz[r][k]));
        RotSlot h[WS_RE];
        SR_SCHED_FENCE();
#pragma unroll
        for (int k = 0; k < WS_RE; ++k) {
          const int x = sxz[r][k] >> 8, zq = sxz[r][k] & 255;
          const float4 o = T[2 * x];
          h[k].s11 = *reinterpret_cast<const float4*>(img_t + (__float_as_int(o.x) + 4 * zq));
          h[k].s12 = *reinterpret_cast<const float4*>(img_t + (__float_as_int(o.y) + 4 * zq));
          h[k].s21 = *reinterpret_cast<const float4*>(img_t + (__float_as_int(o.z) + 4 * zq));
          h[k].s22 = *reinterpret_cast<const float4*>(img_t + (__float_as_int(o.w) + 4 * zq));
          h[k].a = *reinterpret_cast<const float4*>(ay + 4 * __umul24(x, nz4) + 4 * zq);
        }
        SR_SCHED_FENCE();
#pragma unroll
        for (int k = 0; k < WS_RE; ++k) {
          const int x = sxz[r][k] >> 8, zq = sxz[r][k] & 255;
          const float4 w = T[2 * x + 1];
          const RotSlot& q = h[k];
          float4 val;
          val.x = ((q.s11.x * w.x + q.s12.x * w.y) * w.z + (q.s21.x * w.x + q.s22.x * w.y) * w.w) * q.a.x;
          val.y = ((q.s11.y * w.x + q.s12.y * w.y) * w.z + (q.s21.y * w.x + q.s22.y * w.y) * w.w) * q.a.y;
          val.z = ((q.s11.z * w.x + q.s12.z * w.y) * w.z + (q.s21.z * w.x + q.s22.z * w.y) * w.w) * q.a.z;
          val.w = ((q.s11.w * w.x + q.s12.w * w.y) * w.z + (q.s21.w * w.x + q.s22.w * w.y) * w.w) * q.a.w;
          *reinterpret_cast<float4*>(buf + x * WZ + 4 * zq) = val;
        }
        SR_SCHED_FENCE();
      }
    };
    if (y0 < y1) {
      rot_table(y0);
      if (y0 + 1 < y1) rot_table(y0 + 1);
    }
    __syncthreads();  // zero fill and the first two rotation tables done
    if (y0 < y1) {
      stage(y0, Pz);
      load_taps(y0);
    }
    for (int y = y0; y < y1; ++y) {
      lds_barrier();  // plane y and its taps staged; plane y - 1's plane buffer free
      if (y + 1 < y1) {
        stage(y + 1, Pz + (((y - y0) & 1) ^ 1) * PS);
        load_taps(y + 1);  // (into plane y - 2's tap plane)
      }
      if (y + 2 < y1) rot_table(y + 2);
    }
    return;
  }

  __syncthreads();  // zero fill and the first two rotation tables done
  // columns A = 2 wave, B = 2 wave + 1 (past the last column: recomputed, not stored)
  const bool ownA = 2 * wave < ntx, ownB = 2 * wave + 1 < ntx;
  const int x0A = 16 * min(2 * wave, ntx - 1), x0B = 16 * min(2 * wave + 1, ntx - 1);
  float* QA = L + 2 * PS + (2 * wave) * RG * 16;
  float* QB = QA + RG * 16;
  mf4 dvA[TPW], dvB[TPW];
#pragma unroll
  for (int i = 0; i < TPW; ++i) dvA[i] = dvB[i] = mf4{0.f, 0.f, 0.f, 0.f};
  const int ns = max(nsx, nsz);

  // one merged loop: the x pass of the column at x0 from plane buffer Pc
  // with plane tb's taps into Qx (skipped when !xon), and the z pass of the
  // column whose x-pass result is in Qz with plane tz's taps into dv (skipped
  // when !zon)
  auto merged = [&](bool xon, const float* Pc, const float* tb, int x0, float* Qx, bool zon, const float* tz,
                    const float* Qz, mf4 (&dv)[TPW]) {
    mf4 dx[TPW];
#pragma unroll
    for (int i = 0; i < TPW; ++i) dx[i] = mf4{0.f, 0.f, 0.f, 0.f};
    const float* ub = tb + TT.off + m16 - q4 + 2 * hx;
    const float* wz = tz + TT.tp + TT.off + m16 - q4 + 2 * hz;
    if (zon) {  // the z pass's replicate boundary first
      const float* swt = tz + 4 * TT.tp + TT.off;
      const float* swh = swt + TT.tp;
      const float be = q4 == 0 ? Qz[hz * 16 + m16] : q4 == 1 ? Qz[(hz + nz - 1) * 16 + m16] : 0.f;
#pragma unroll
      for (int i = 0; i < TPW; ++i) {
        const int zm = 16 * i + m16;
        const float ae = q4 == 0 ? tail_at(swt, NT, zm + hz + 1) : q4 == 1 ? head_at(swh, NT, zm + hz - nz) : 0.f;
        dv[i] = mfma4(ae, be, dv[i]);
      }
    }
    float ax0[TPW], ax1[TPW], bx0 = 0.f, bx1 = 0.f, bz0[TPW], bz1[TPW], az0 = 0.f, az1 = 0.f;
    auto ldx = [&](float (&a)[TPW], float& bb, int s) {
      bb = ub[-4 * s];
      const float* pr = Pc + __umul24(clampi(x0 - hx + 4 * s + q4, 0, nx - 1), WZ) + m16;
#pragma unroll
      for (int i = 0; i < TPW; ++i) a[i] = pr[16 * i];
    };
    auto ldz = [&](float (&bq)[TPW], float& aa, int s) {
      aa = wz[-4 * s];
#pragma unroll
      for (int i = 0; i < TPW; ++i) bq[i] = Qz[(16 * i + q4 + 4 * s) * 16 + m16];
    };
    auto mmx = [&](const float (&a)[TPW], float bb) {
#pragma unroll
      for (int i = 0; i < TPW; ++i) dx[i] = mfma4(bb, a[i], dx[i]);
    };
    auto mmz = [&](const float (&bq)[TPW], float aa) {
#pragma unroll
      for (int i = 0; i < TPW; ++i) dv[i] = mfma4(aa, bq[i], dv[i]);
    };
    const int nx_s = xon ? nsx : 0, nz_s = zon ? nsz : 0;
    if (nx_s > 0) ldx(ax0, bx0, 0);
    if (nz_s > 0) ldz(bz0, az0, 0);
    for (int s = 0; s < ns; s += 2) {
      if (s + 1 < nx_s) ldx(ax1, bx1, s + 1);
      if (s + 1 < nz_s) ldz(bz1, az1, s + 1);
      SR_SCHED_FENCE();
      if (s < nx_s) mmx(ax0, bx0);
      if (s < nz_s) mmz(bz0, az0);
      SR_SCHED_FENCE();
      if (s + 2 < nx_s) ldx(ax0, bx0, s + 2);
      if (s + 2 < nz_s) ldz(bz0, az0, s + 2);
      SR_SCHED_FENCE();
      if (s + 1 < nx_s) mmx(ax1, bx1);
      if (s + 1 < nz_s) mmz(bz1, az1);
    }
    asm volatile("" ::: "memory");  // the Qx stores below stay after the Qz reads above
    if (xon) {
#pragma unroll
      for (int i = 0; i < TPW; ++i)
        *reinterpret_cast<float4*>(Qx + (hz + 16 * i + m16) * 16 + 4 * q4) =
            make_float4(dx[i][0], dx[i][1], dx[i][2], dx[i][3]);
      // the wave reads back what its own lanes wrote (LDS keeps a wave's order)
      asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    }
  };

  for (int y = y0; y < y1; ++y) {
    const int cur = (y - y0) & 1;
    lds_barrier();  // plane y and its taps staged; plane y - 1's plane buffer free
    const float* Pc = Pz + cur * PS;
    const float* ty = tap_plane(y);
    // loop 1: x of A (plane y) with z of B (plane y - 1)
    merged(true, Pc, ty, x0A, QA, y > y0, y > y0 ? tap_plane(y - 1) : ty, QB, dvB);
    // loop 2: x of B (plane y) with z of A (plane y)
    merged(true, Pc, ty, x0B, QB, true, ty, QA, dvA);
  }
  if (y0 < y1) merged(false, Pz, tap_plane(y1 - 1), x0B, QB, true, tap_plane(y1 - 1), QB, dvB);
  float* pb = part + ((long)b * nchunk + c) * nx * nz;
#pragma unroll
  for (int i = 0; i < TPW; ++i) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (ownA) pb[(long)(16 * i + 4 * q4 + r) * nx + x0A + m16] = dvA[i][r];
      if (ownB) pb[(long)(16 * i + 4 * q4 + r) * nx + x0B + m16] = dvB[i][r];
    }
  }
}

// --- forward, direct 2-D ----------------------------------------------------
// general (non-separable) PSF: strips of 8 x-outputs, pz rows of NT taps each
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
  const long S = (long)nx * ny;
  const float* wb = work + S * nz * b;
  const int nsx = nz * nxs;

  float acc[PSF_MAXS][8];
#pragma unroll
  for (int i = 0; i < PSF_MAXS; ++i)
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[i][k] = 0.f;

  for (int y = y0; y < y1; ++y) {
    const float* th = t.h + (long)(v * t.vstride + y) * pz * NT;
    // L[r][col] = P[clamp(col-LP), clamp(r-hz)] (small kernels: direct clamped loads)
    for (RowIter it(threadIdx.x, BLOCK, WX); it.row < NR; it.next())
      L[it.row * WX + it.col] =
          wb[clampi(it.col - LP, 0, nx - 1) + (long)nx * y + S * clampi(it.row - hz, 0, nz - 1)];
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
// GRP: the grid holds the blocks of the ngrp launches that grp describes
// (the arguments before grp are unused).
template <bool GRP>
__global__ void __launch_bounds__(256) k_view_reduce(const float* __restrict__ part, int nchunk, int P,
                                                     const int* __restrict__ vids, float* __restrict__ out,
                                                     const float* __restrict__ yobs, const float* __restrict__ bkg,
                                                     int epi, const ViewReduceGrp* __restrict__ grp, int ngrp) {
  int b = blockIdx.y, bx = blockIdx.x;
  if constexpr (GRP) {
    const ViewReduceGrp& a = grp[grp_find(grp, ngrp)];
    const int l = (int)blockIdx.x - a.h.blk0;
    part = glob(a.part), nchunk = a.nchunk, P = a.P, vids = glob(a.vids), out = glob(a.out);
    yobs = glob(a.yobs), bkg = glob(a.bkg), epi = a.epi;
    b = __builtin_amdgcn_readfirstlane(l / a.h.gx);
    bx = __builtin_amdgcn_readfirstlane(l - b * a.h.gx);
  }
  const int k = bx * 256 + threadIdx.x;
  if (k >= P) return;
  const float* p = part + (long)b * nchunk * P + k;
  float yo = 0.f, bk = 0.f;
  if (epi == EPI_RATIO) {  // issued ahead of the partials
    const long o = (long)vids[b] * P + k;
    yo = yobs[o];
    bk = bkg[o];
  }
  // chunk partials summed in chunk order, 8 loads in flight at a time
  float s = 0.f;
  int c = 0;
  for (; c + 8 <= nchunk; c += 8) {
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = p[(long)(c + i) * P];
#pragma unroll
    for (int i = 0; i < 8; ++i) s += v[i];
  }
  for (; c < nchunk; ++c) s += p[(long)c * P];
  if (epi == EPI_RATIO) s = yo / (s + bk);  // src/ml-os-em.jl:44, :124-125
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
template <int NT>
__device__ __forceinline__ float adj_edge_lo(const float* in, int n, int h, const float* tail) {
  float s = 0.f;
  const int i1 = min(n - 1, h);
#pragma unroll 8
  for (int i = 0; i <= i1; ++i) s = fmaf(in[i], tail[i + h], s);
  return s;
}
template <int NT>
__device__ __forceinline__ float adj_edge_hi(const float* in, int n, int h, const float* head) {
  float s = 0.f;
  const int i0 = max(0, n - 1 - h);
#pragma unroll 8
  for (int i = i0; i < n; ++i) s = fmaf(in[i], head[i + h - n + 1], s);
  return s;
}

// grid (nchunk, nb), one strip of SL outputs per thread per phase.  LDS: Vz =
// the view with zero rows around it (z-pass windows are immediate offsets),
// Qp = the z-adjoint of the view with zero columns around it (stride WQ, data
// at column offset hx), the taps and their tail/head sums.  The x-pass strip
// (z, x0..x0+SL-1) is written straight to the work volume, times the
// attenuation factor (mul3dj!, src/backproject.jl:108).
template <int NT, int SL>
__global__ void __launch_bounds__(kSepThreads)
k_psf_adj_sep(const float* __restrict__ views, Taps t, const int* __restrict__ vids, int Yc, float* __restrict__ work,
              const float* __restrict__ atten, long atten_vstride, Geo g) {
  extern __shared__ float L[];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int hx = (g.px - 1) / 2, hz = (g.pz - 1) / 2;
  const int nxs = (nx + SL - 1) / SL, nzs = (nz + SL - 1) / SL;
  const int WQ = quad_stride(SL * nxs + NT - 1);
  const int b = blockIdx.y, c = blockIdx.x;
  const int v = vids[b];
  const int y0 = c * Yc, y1 = min(ny, y0 + Yc);
  const long S = (long)nx * ny, V = S * nz;
  const int P = nx * nz;
  const int tid = threadIdx.x, nthr = blockDim.x;
  // Vt: the view transposed to z-contiguous columns Vt[x][hz + z] with zero
  // pads (column stride VZ), so the z-pass windows are immediate offsets
  const int VZ = quad_stride(SL * nzs + NT - 1);
  float* Vt = L;
  float* Qp = L + (long)nx * VZ;
  float* s_taps = Qp + nz * WQ;  // 2 buffers of: taps u, w and their tail/head sums
  const float* vin = views + (long)b * P;
  for (int k = tid; k < nx * VZ; k += nthr) Vt[k] = 0.f;
  for (int k = tid; k < nz * WQ; k += nthr) Qp[k] = 0.f;
  __syncthreads();
  for (RowIter it(tid, nthr, nx); it.row < nz; it.next()) Vt[it.col * VZ + hz + it.row] = vin[it.row * nx + it.col];
  const bool hasx = tid < nz * nxs, hasz = tid < nx * nzs;
  const int xs_z = tid % nz, xs_x0 = (tid / nz) * SL;
  const int zs_x = tid % nx, zs_z0 = (tid / nx) * SL;
  float* wout = work + V * b;
  const float* av = atten ? atten + atten_vstride * v : nullptr;
  const bool vec4 = (nx % 4 == 0) && (SL % 4 == 0);

  // thread tid < 6 NT holds entry tid of the (u, w, u tail, u head, w tail,
  // w head) table of the next plane, loaded one plane ahead
  const int tab = tid / NT, tk = tid - tab * NT;
  auto tap_load = [&](int y) {
    const long plane = (long)v * t.vstride + y;
    return tab == 0 ? t.u[plane * NT + tk] : tab == 1 ? t.w[plane * NT + tk] : t.cum[(plane * 4 + tab - 2) * NT + tk];
  };
  float tv = (tid < 6 * NT && y0 < y1) ? tap_load(y0) : 0.f;
  for (int y = y0; y < y1; ++y) {
    float* s_tu = s_taps + (y & 1) * 6 * NT;
    float* s_tw = s_tu + NT;
    float* s_ut = s_tw + NT;
    float* s_uh = s_ut + NT;
    float* s_wt = s_uh + NT;
    float* s_wh = s_wt + NT;
    if (tid < 6 * NT) s_tu[tid] = tv;
    __syncthreads();
    if (tid < 6 * NT && y + 1 < y1) tv = tap_load(y + 1);
    float q[SL];
    // z adjoint: strip (x, z0..z0+SL-1) -> Qp[z][hx + x]
    if (hasz) {
#pragma unroll
      for (int k = 0; k < SL; ++k) q[k] = 0.f;
      // row zs_z0 - hz + j of the view is Vt[x][zs_z0 + j]
      const float* col = Vt + zs_x * VZ + zs_z0;
      conv_strip_adj<NT, SL>(col, s_tw, q);
      // replicate-pad fold at z = 0 and z = nz-1 (one evaluation each)
      const int khi = nz - 1 - zs_z0;
      const bool lo = zs_z0 == 0, hi = khi < SL;
      float elo = 0.f, ehi = 0.f;
      const float* vcol = Vt + zs_x * VZ + hz;  // the view's column x, z contiguous
      if (nz == 1) {
        elo = vcol[0] * s_wt[0];
      } else {
        if (lo) elo = adj_edge_lo<NT>(vcol, nz, hz, s_wt);
        if (hi) ehi = adj_edge_hi<NT>(vcol, nz, hz, s_wh);
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
    // x adjoint: strip (z, x0..x0+SL-1), then out to the work volume; the
    // strip's attenuation factors are loaded before the convolution
    if (hasx) {
      const long o = xs_x0 + (long)nx * y + S * xs_z;
      const bool v4 = vec4 && xs_x0 + SL <= nx;
      float4 a4[SL / 4];
#pragma unroll
      for (int k = 0; k < SL / 4; ++k)
        a4[k] = (av && v4) ? *reinterpret_cast<const float4*>(av + o + 4 * k) : make_float4(1.f, 1.f, 1.f, 1.f);
#pragma unroll
      for (int k = 0; k < SL; ++k) q[k] = 0.f;
      const float* row = Qp + xs_z * WQ + xs_x0;
      conv_strip_adj<NT, SL>(row, s_tu, q);
      const float* qrow = Qp + xs_z * WQ + hx;
      const int khi = nx - 1 - xs_x0;
      const bool lo = xs_x0 == 0, hi = khi < SL;
      float elo = 0.f, ehi = 0.f;
      if (nx == 1) {
        elo = qrow[0] * s_ut[0];
      } else {
        if (lo) elo = adj_edge_lo<NT>(qrow, nx, hx, s_ut);
        if (hi) ehi = adj_edge_hi<NT>(qrow, nx, hx, s_uh);
      }
#pragma unroll
      for (int k = 0; k < SL; ++k)
        if (hi && k == khi && nx > 1) q[k] = ehi;
      if (lo) q[0] = elo;
      if (v4) {
#pragma unroll
        for (int k = 0; k < SL; k += 4) {
          float4 val = make_float4(q[k], q[k + 1], q[k + 2], q[k + 3]);
          if (av) {
            val.x = val.x * a4[k / 4].x;
            val.y = val.y * a4[k / 4].y;
            val.z = val.z * a4[k / 4].z;
            val.w = val.w * a4[k / 4].w;
          }
          *reinterpret_cast<float4*>(wout + o + k) = val;
        }
      } else {
#pragma unroll
        for (int k = 0; k < SL; ++k)
          if (xs_x0 + k < nx) wout[o + k] = av ? q[k] * av[o + k] : q[k];
      }
    }
    __syncthreads();
  }
}

// --- adjoint, separable, on the matrix cores ---------------------------------
// Transpose of the matrix-core forward, per plane y of the view:
//   z adjoint  G[j0+m][x] = sum_k w[k-m] V[j0-hz+k][x]          (correlation)
//   x adjoint  H[z][j0+n] = sum_k G[z][j0-hx+k] u[k-n]
// on zero-padded rows; the transposed replicate boundary makes image row
// (column) 0 sum_i Tail(i+h) in[i] and row n-1 sum_i Head(i+h-n+1) in[i],
// i.e. the same band with the taps of those rows replaced by their tail /
// head sums (per-tile fragment tables).  Then H times the attenuation factors is
// written to the work volume.  LDS: Vz (the view, rows z+hz, stride 16 mod
// 64), Gx (the z adjoint, rows z, column x+hx), tap buffers.  Each wave takes
// a run of consecutive tiles, so both halves of a 128-byte output line are
// stored back to back.  grid (nchunk, nb), block 512.
template <int TPW, int NW, bool ZF>
__global__ void __launch_bounds__(NW * 64) k_psf_adj_mfma(const float* __restrict__ views, Taps t,
                                                            const int* __restrict__ vids, int Yc,
                                                            float* __restrict__ work, const float* __restrict__ atten,
                                                            long atten_vstride, Geo g, int TBP) {
  extern __shared__ float L[];
  const int NT = t.ntx;
  const MfTaps TT(NT);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int hx = (g.px - 1) / 2, hz = (g.pz - 1) / 2;
  const int nsx = (16 + g.px - 1 + 3) / 4, nsz = (16 + g.pz - 1 + 3) / 4;
  const int ntx = nx >> 4, ntile = ntx * (nz >> 4);
  const int WX = mf_wx(nx, hx, NT), WQ = mf_wq(nx), RQ = mf_rq(nz, NT);
  const int b = blockIdx.y, c = blockIdx.x;
  const int v = vids[b];
  const int y0 = c * Yc, y1 = min(ny, y0 + Yc);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int m16 = lane & 15, q4 = lane >> 4;
  float* Vz = L;
  float* Gx = Vz + RQ * WQ;
  float* s_taps = Gx + nz * WX;  // TBP planes of TT.size
  for (int k = tid; k < RQ * WQ + nz * WX + TBP * TT.size; k += (NW * 64)) L[k] = 0.f;
  __syncthreads();
  const int P = nx * nz;
  const float* vin = views + (long)b * P;
  for (RowIter it(tid, (NW * 64), nx); it.row < nz; it.next()) Vz[(hz + it.row) * WQ + it.col] = vin[it.row * nx + it.col];

  const long S = (long)nx * ny, V = S * nz;
  float* wout = work + V * b;
  const float* av = atten ? atten + atten_vstride * v : nullptr;
  // the tap tables of up to TBP (<= 8) planes at a time: all loads of a
  // thread first, then the stores (clamped entries repeat the last one)
  auto load_taps = [&](int yg, int ng) {
    constexpr int ME = (8 * 6 * 64 + NW * 64 - 1) / (NW * 64);
    const int n = ng * 6 * NT;
    float tvv[ME];
    int tdst[ME];
#pragma unroll
    for (int j = 0; j < ME; ++j) {
      const int e = min(tid + j * NW * 64, n - 1);
      const int pl = e / (6 * NT), r = e - pl * 6 * NT, tab = r / NT, tk = r - tab * NT;
      const long plane = (long)v * t.vstride + yg + pl;
      tdst[j] = pl * TT.size + tab * TT.tp + TT.off + tk;
      const float* src = tab == 0 ? t.u + plane * NT : tab == 1 ? t.w + plane * NT : t.cum + (plane * 4 + tab - 2) * NT;
      tvv[j] = src[tk];
    }
#pragma unroll
    for (int j = 0; j < ME; ++j) s_taps[tdst[j]] = tvv[j];
  };
  int z0s[TPW], x0s[TPW];
  bool own[TPW];
  // ZF: runs of tiles go down z, so a wave writes a whole z-column of its x range
  mf_tiles<TPW>(wave, ntile, ZF ? -(nz >> 4) : ntx, z0s, x0s, own);
  int voff[TPW], goff[TPW];
#pragma unroll
  for (int i = 0; i < TPW; ++i) {
    voff[i] = (z0s[i] + q4) * WQ + x0s[i] + m16;  // z adjoint B fragment (rows of Vz)
    goff[i] = (z0s[i] + m16) * WX + x0s[i] + q4;  // x adjoint A fragment (rows of Gx)
  }

  for (int y = y0; y < y1; ++y) {
    if ((y - y0) % TBP == 0) {  // next tap group (the previous group's last reads are done)
      lds_barrier();
      load_taps(y, min(TBP, y1 - y));
    }
    const float* tb = s_taps + ((y - y0) % TBP) * TT.size;
    lds_barrier();  // taps (and, first time, Vz) in place; the previous x adjoint has read Gx
    // Fragment tables per tile and lane: the transposed replicate boundary
    // replaces the taps of image row/column 0 by their tail sums and of row/
    // column n-1 by their head sums (same fragment index 4s + k - m).
    const float* su = tb + TT.off + q4 - m16;
    const float* sw = su + TT.tp;
    const float* sut = sw + TT.tp;
    const float* suh = sut + TT.tp;
    const float* swt = suh + TT.tp;
    const float* swh = swt + TT.tp;
    // z adjoint: Vz -> Gx
    const float* ap[TPW];
#pragma unroll
    for (int i = 0; i < TPW; ++i)
      ap[i] = (z0s[i] == 0 && m16 == 0) ? swt : (z0s[i] + 16 == nz && m16 == 15) ? swh : sw;
    mf4 dg[TPW];
#pragma unroll
    for (int i = 0; i < TPW; ++i) dg[i] = mf4{0.f, 0.f, 0.f, 0.f};
    {
      float bq0[TPW], bq1[TPW], a0[TPW], a1[TPW];
      auto ld = [&](float (&bq)[TPW], float (&aa)[TPW], int s) {
#pragma unroll
        for (int i = 0; i < TPW; ++i) {
          aa[i] = ap[i][4 * s];
          bq[i] = Vz[voff[i] + 4 * s * WQ];
        }
      };
      ld(bq0, a0, 0);
      int s = 0;
      for (; s + 1 < nsz; s += 2) {
        ld(bq1, a1, s + 1);
        SR_SCHED_FENCE();
#pragma unroll
        for (int i = 0; i < TPW; ++i) dg[i] = mfma4(a0[i], bq0[i], dg[i]);
        SR_SCHED_FENCE();
        ld(bq0, a0, min(s + 2, nsz - 1));
        SR_SCHED_FENCE();
#pragma unroll
        for (int i = 0; i < TPW; ++i) dg[i] = mfma4(a1[i], bq1[i], dg[i]);
      }
      if (s < nsz) {
#pragma unroll
        for (int i = 0; i < TPW; ++i) dg[i] = mfma4(a0[i], bq0[i], dg[i]);
      }
    }
    // (a clamped duplicate tile stores the same values as its owner)
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
      float* go = Gx + (z0s[i] + 4 * q4) * WX + hx + x0s[i] + m16;
#pragma unroll
      for (int r = 0; r < 4; ++r) go[r * WX] = dg[i][r];
    }
    lds_barrier();
    // x adjoint: Gx -> H, times the attenuation factors, to the work volume
    float af[TPW][4];
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
      if (ZF) {  // [y][x][z]: the lane's 4 z are contiguous
        const float4 a4 = av ? *reinterpret_cast<const float4*>(av + ((long)y * nx + x0s[i] + m16) * nz + z0s[i] + 4 * q4)
                             : make_float4(1.f, 1.f, 1.f, 1.f);
        af[i][0] = a4.x;
        af[i][1] = a4.y;
        af[i][2] = a4.z;
        af[i][3] = a4.w;
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          af[i][r] = av ? av[S * (z0s[i] + 4 * q4 + r) + (long)nx * y + x0s[i] + m16] : 1.f;
      }
    }
    const float* bp[TPW];
#pragma unroll
    for (int i = 0; i < TPW; ++i)
      bp[i] = (x0s[i] == 0 && m16 == 0) ? sut : (x0s[i] + 16 == nx && m16 == 15) ? suh : su;
    mf4 dh[TPW];
#pragma unroll
    for (int i = 0; i < TPW; ++i) dh[i] = mf4{0.f, 0.f, 0.f, 0.f};
    {
      float a0[TPW], a1[TPW], b0[TPW], b1[TPW];
      auto ld = [&](float (&a)[TPW], float (&bb)[TPW], int s) {
#pragma unroll
        for (int i = 0; i < TPW; ++i) {
          bb[i] = bp[i][4 * s];
          a[i] = Gx[goff[i] + 4 * s];
        }
      };
      ld(a0, b0, 0);
      int s = 0;
      for (; s + 1 < nsx; s += 2) {
        ld(a1, b1, s + 1);
        SR_SCHED_FENCE();
#pragma unroll
        for (int i = 0; i < TPW; ++i) dh[i] = mfma4(a0[i], b0[i], dh[i]);
        SR_SCHED_FENCE();
        ld(a0, b0, min(s + 2, nsx - 1));
        SR_SCHED_FENCE();
#pragma unroll
        for (int i = 0; i < TPW; ++i) dh[i] = mfma4(a1[i], b1[i], dh[i]);
      }
      if (s < nsx) {
#pragma unroll
        for (int i = 0; i < TPW; ++i) dh[i] = mfma4(a0[i], b0[i], dh[i]);
      }
    }
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
      const int z0 = z0s[i], x0 = x0s[i];
      if (ZF) {
        *reinterpret_cast<float4*>(wout + ((long)y * nx + x0 + m16) * nz + z0 + 4 * q4) =
            make_float4(dh[i][0] * af[i][0], dh[i][1] * af[i][1], dh[i][2] * af[i][2], dh[i][3] * af[i][3]);
      } else {
        float* o = wout + (long)nx * y + x0 + m16;
#pragma unroll
        for (int r = 0; r < 4; ++r) o[S * (z0 + 4 * q4 + r)] = dh[i][r] * af[i][r];
      }
    }
  }
}

// --- adjoint, separable, matrix cores, column-owned (z-fastest) --------------
// The adjoint in the other pass order (the two 1-D passes act on different
// axes and commute): the x adjoint first, from the block's view staged once in
// LDS and only read afterwards, then the z adjoint from registers (the x
// adjoint's tiles transposed in 4 x 4 blocks of lane groups, source row groups
// outside the plane skipped).  Each wave takes (plane, 16-wide x column) items
// with all nz/16 z tiles of the column, so no block barrier separates the
// passes or the planes (two per tap group only):
//   x adjoint  H[z][x0+n] = sum_k V[z][x0-hx+k] u'[k-n]
//   z adjoint  R[z0+m][x] = sum_k w'[k-m] H[z0-hz+k][x]
// (u', w': the taps, or their tail / head sums for output column / row 0 and
// n-1, the transposed replicate boundary).  LDS: Vx (rows z, column XO + x,
// stride WX = 4 odd mod 64: a fragment's 16 rows x 4 columns hit distinct
// banks), tap groups.  nz = 16 TPW.
// grid (nchunk, nb), block 64 NWV.
__host__ __device__ inline size_t mf_col_lds_floats(const Geo& g, int nt, int tbp) {
  return (size_t)g.nz * mf_wx(g.nx, (g.px - 1) / 2, nt) + (size_t)tbp * MfTaps(nt).size;
}
// NWV waves per block (16: four waves per SIMD hide more of each other's LDS
// and memory waits); the block's (plane, column) items of a tap group are
// dealt round robin to the waves.
// GRP: the grid holds the blocks of the ngrp launches that grp describes
// (the arguments before grp are unused).
template <int TPW, int NWV, bool GRP>
__global__ void __launch_bounds__(64 * NWV) k_psf_adj_col(const float* __restrict__ views, Taps t,
                                                            const int* __restrict__ vids, int Yc,
                                                            float* __restrict__ work, const float* __restrict__ atten,
                                                            long atten_vstride, Geo g, int TBP,
                                                            const PsfAdjGrp* __restrict__ grp, int ngrp) {
  extern __shared__ float L[];
  int b = blockIdx.y, c = blockIdx.x;
  if constexpr (GRP) {
    const PsfAdjGrp& a = grp[grp_find(grp, ngrp)];
    const int l = (int)blockIdx.x - a.h.blk0;
    views = glob(a.views), t = glob(a.t), vids = glob(a.vids), Yc = a.Yc, work = glob(a.work);
    atten = glob(a.atten), atten_vstride = a.atten_vstride, g = a.g, TBP = a.TBP;
    b = __builtin_amdgcn_readfirstlane(l / a.h.gx);  // (uniform: kept in scalar registers)
    c = __builtin_amdgcn_readfirstlane(l - b * a.h.gx);
  }
  constexpr int NTH = 64 * NWV;
  const int NT = t.ntx;
  const MfTaps TT(NT);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int hx = (g.px - 1) / 2, hz = (g.pz - 1) / 2;
  const int nsx = (16 + g.px - 1 + 3) / 4;
  const int dlo = zp_dlo(hz), dhi = zp_dhi(hz);
  const int ntx = nx >> 4;
  const int WX = mf_wx(nx, hx, NT), XO = mf_xo(hx);
  const int v = vids[b];
  const int y0 = c * Yc, y1 = min(ny, y0 + Yc);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int m16 = lane & 15, q4 = lane >> 4;
  float* Vx = L;
  float* s_taps = Vx + nz * WX;  // TBP planes of TT.size
  // zeros where no data lands: Vx's pad columns, the tap tables
  {
    const int padc = WX - nx;  // columns [0, XO) and [XO + nx, WX) of every row
    for (int k = tid; k < nz * padc; k += NTH) {
      const int z = k / padc, cc = k - z * padc;
      Vx[z * WX + (cc < XO ? cc : cc + nx)] = 0.f;
    }
    for (int k = tid; k < TBP * TT.size; k += NTH) L[nz * WX + k] = 0.f;
  }
  const float* vin = views + (long)b * nx * nz;
  const int nq = nx >> 2;
  for (int k = tid; k < nz * nq; k += NTH) {
    const int z = k / nq, x4 = k - z * nq;
    *reinterpret_cast<float4*>(Vx + z * WX + XO + 4 * x4) = *reinterpret_cast<const float4*>(vin + z * nx + 4 * x4);
  }
  __syncthreads();  // (the tap tables' zeros land before their entries)
  const long V = (long)nx * ny * nz;
  float* wout = work + V * b;
  const float* av = atten ? atten + atten_vstride * v : nullptr;
  // the tap tables of up to TBP (<= 8) planes at a time; the next group's
  // loads are in flight in registers while the current group is convolved
  constexpr int ME = (8 * 6 * 64 + NTH - 1) / NTH;
  float tvv[ME];
  int tdst[ME];
  auto fetch_taps = [&](int yg, int ng) {
    const int n = ng * 6 * NT;
#pragma unroll
    for (int j = 0; j < ME; ++j) {
      const int e = min(tid + j * NTH, n - 1);
      const int pl = e / (6 * NT), r = e - pl * 6 * NT, tab = r / NT, tk = r - tab * NT;
      const long plane = (long)v * t.vstride + yg + pl;
      tdst[j] = pl * TT.size + tab * TT.tp + TT.off + tk;
      const float* src = tab == 0 ? t.u + plane * NT : tab == 1 ? t.w + plane * NT : t.cum + (plane * 4 + tab - 2) * NT;
      tvv[j] = src[tk];
    }
  };
  fetch_taps(y0, min(TBP, y1 - y0));
  for (int yg = y0; yg < y1; yg += TBP) {
    const int ng = min(TBP, y1 - yg);
    lds_barrier();  // the previous tap group is no longer read (the first time: Vx is in place)
#pragma unroll
    for (int j = 0; j < ME; ++j) s_taps[tdst[j]] = tvv[j];
    lds_barrier();
    if (yg + TBP < y1) fetch_taps(yg + TBP, min(TBP, y1 - yg - TBP));
    for (int it = wave; it < ng * ntx; it += NWV) {
      const int pl = it / ntx, col = it - pl * ntx, y = yg + pl;
      const float* tb = s_taps + pl * TT.size;
      const float* su = tb + TT.off + q4 - m16;
      const float* sw = su + TT.tp;
      const float* sut = sw + TT.tp;
      const float* suh = sut + TT.tp;
      const float* swt = suh + TT.tp;
      const float* swh = swt + TT.tp;
      // z taps: the first tile's row 0 takes the tail sums, the last tile's
      // row 15 the head sums (nz = 16 TPW), every other row the plain taps, so
      // a step reads three tap words: plain, first tile, last tile
      const float* apm = sw + hz;
      const float* ap0 = (m16 == 0 ? swt : (TPW == 1 && m16 == 15) ? swh : sw) + hz;
      const float* apl = (m16 == 15 ? swh : sw) + hz;
      {
        const int x0 = 16 * col;
        const long ob = ((long)y * nx + x0 + m16) * nz + 4 * q4;  // + 16 i: lane's 4 z of tile i
        // attenuation factors of the column's tiles, in flight during the passes
        float4 af[TPW];
#pragma unroll
        for (int i = 0; i < TPW; ++i)
          af[i] = av ? *reinterpret_cast<const float4*>(av + ob + 16 * i) : make_float4(1.f, 1.f, 1.f, 1.f);
        // x adjoint: lane (k, n) register r = H[16 i + 4 k + r][x0 + n]
        const float* bp = (x0 == 0 && m16 == 0) ? sut : (x0 + 16 == nx && m16 == 15) ? suh : su;
        const float* va = Vx + m16 * WX + XO - hx + x0 + q4;
        mf4 dh[TPW];
#pragma unroll
        for (int i = 0; i < TPW; ++i) dh[i] = mf4{0.f, 0.f, 0.f, 0.f};
        {
          float a0[TPW], a1[TPW], b0, b1;
          auto ld = [&](float (&a)[TPW], float& bb, int s) {
            bb = bp[4 * s];
#pragma unroll
            for (int i = 0; i < TPW; ++i) a[i] = va[16 * i * WX + 4 * s];
          };
          ld(a0, b0, 0);
          int s = 0;
          for (; s + 1 < nsx; s += 2) {
            ld(a1, b1, s + 1);
            SR_SCHED_FENCE();
#pragma unroll
            for (int i = 0; i < TPW; ++i) dh[i] = mfma4(a0[i], b0, dh[i]);
            SR_SCHED_FENCE();
            ld(a0, b0, min(s + 2, nsx - 1));
            SR_SCHED_FENCE();
#pragma unroll
            for (int i = 0; i < TPW; ++i) dh[i] = mfma4(a1[i], b1, dh[i]);
          }
          if (s < nsx) {
#pragma unroll
            for (int i = 0; i < TPW; ++i) dh[i] = mfma4(a0[i], b0, dh[i]);
          }
        }
        // register g of tile i: lane (k, n) = H[16 i + 4 g + k][x0 + n]
#pragma unroll
        for (int i = 0; i < TPW; ++i) dh[i] = xpose_groups(dh[i]);
        // z adjoint from registers: output tile j, source group G = 4 j + d,
        // A (m, k) = w'[4 d + k - m + hz]; lane register r = R[16 j + 4 k + r][x0 + n]
        mf4 dz[TPW];
#pragma unroll
        for (int i = 0; i < TPW; ++i) dz[i] = mf4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int d = -ZP_DMAX; d <= ZP_DMAX + 3; ++d) {
          if (d < dlo || d > dhi) continue;  // (wave-uniform)
          const int e = floor4(d), gg = d - 4 * e;
          const float am = apm[4 * d], a0 = ap0[4 * d], al = apl[4 * d];
#pragma unroll
          for (int j = 0; j < TPW; ++j) {
            const int i = j + e;
            if (i < 0 || i >= TPW) continue;
            dz[j] = mfma4(j == 0 ? a0 : j == TPW - 1 ? al : am, dh[i][gg], dz[j]);
          }
        }
#pragma unroll
        for (int i = 0; i < TPW; ++i)
          *reinterpret_cast<float4*>(wout + ob + 16 * i) = make_float4(dz[i][0] * af[i].x, dz[i][1] * af[i].y,
                                                                       dz[i][2] * af[i].z, dz[i][3] * af[i].w);
      }
    }
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

// The same extended correlation and fold, with the view (zero columns on both
// sides) staged in LDS once per block, each plane's taps in LDS, and each
// thread computing an 8-wide x strip of E from a register window, as the
// forward kernel does.  Out-of-range view columns read zeros and tap columns
// past px are zero, so every added term is an exact zero: E equals the
// per-term loop of k_psf_adj_2d for finite data.
__host__ __device__ inline int adj2d_wv(const Geo& g, int nt) { return odd_up(g.nx + 2 * ((g.px - 1) / 2) + nt + 16); }
__host__ __device__ inline size_t adj2d_lds_floats(const Geo& g, int nt) {
  const int hx = (g.px - 1) / 2, hz = (g.pz - 1) / 2;
  return (size_t)odd_up(g.nx + 2 * hx) * (g.nz + 2 * hz) + (size_t)g.nz * adj2d_wv(g, nt) + (size_t)g.pz * nt;
}
template <int NT, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_psf_adj_2d_strips(const float* __restrict__ views, Taps t,
                                                             const int* __restrict__ vids, int Yc,
                                                             float* __restrict__ work,
                                                             const float* __restrict__ atten, long atten_vstride,
                                                             Geo g) {
  extern __shared__ float L[];
  const int nx = g.nx, ny = g.ny, nz = g.nz, px = g.px, pz = g.pz;
  const int hx = (px - 1) / 2, hz = (pz - 1) / 2;
  const int EX = nx + 2 * hx, EZ = nz + 2 * hz;
  const int WE = odd_up(EX), WV = adj2d_wv(g, NT), OX = 2 * hx + 8;  // view column x at OX + x
  const int sx = (EX + 7) / 8;                                         // 8-wide strips per E row
  const int b = blockIdx.y, c = blockIdx.x;
  const int v = vids[b];
  const int y0 = c * Yc, y1 = min(ny, y0 + Yc);
  const long S = (long)nx * ny, V = S * nz;
  const int P = nx * nz;
  float* E = L;
  float* Vs = E + (size_t)WE * EZ;
  float* ts = Vs + (size_t)nz * WV;
  const float* vin = views + (long)b * P;
  float* wout = work + V * b;
  const float* av = atten ? atten + atten_vstride * v : nullptr;
  for (int k = threadIdx.x; k < nz * WV; k += BLOCK) {
    const int z = k / WV, x = k - z * WV - OX;
    Vs[k] = (x >= 0 && x < nx) ? vin[x + nx * z] : 0.f;
  }
  for (int y = y0; y < y1; ++y) {
    const float* th = t.h + (long)(v * t.vstride + y) * pz * NT;
    for (int k = threadIdx.x; k < pz * NT; k += BLOCK) ts[k] = th[k];
    __syncthreads();  // (the view, the first time) and this plane's taps in place
    // E[ze][xe] = sum_{a,b} h[a,b] V[xe - 2hx + a, ze - 2hz + b]
    for (int s = threadIdx.x; s < sx * EZ; s += BLOCK) {
      const int ze = s / sx, xe0 = (s - ze * sx) * 8;
      float acc[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] = 0.f;
      for (int bb = 0; bb < pz; ++bb) {
        const int zz = ze - 2 * hz + bb;
        if (zz < 0 || zz >= nz) continue;
        const float* row = Vs + zz * WV + OX - 2 * hx + xe0;
        const float* hb = ts + bb * NT;
        float w[NT + 7];
#pragma unroll
        for (int j = 0; j < NT + 7; ++j) w[j] = row[j];
#pragma unroll
        for (int a = 0; a < NT; ++a) {
          const float ha = hb[a];
#pragma unroll
          for (int k = 0; k < 8; ++k) acc[k] = fmaf(ha, w[k + a], acc[k]);
        }
      }
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (xe0 + k < EX) E[ze * WE + xe0 + k] = acc[k];
    }
    __syncthreads();
    // fold x pads onto x = 0 and x = nx-1, for every extended row
    for (int ze = threadIdx.x; ze < EZ; ze += BLOCK) {
      float* row = E + ze * WE;
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
      for (int j = 0; j < hz; ++j) lo += E[j * WE + xe];
      for (int j = hz + nz; j < EZ; ++j) hi += E[j * WE + xe];
      E[hz * WE + xe] += lo;
      E[(hz + nz - 1) * WE + xe] += hi;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < P; k += BLOCK) {
      const int z = k / nx, x = k - z * nx;
      const long o = x + (long)nx * y + S * z;
      const float val = E[(z + hz) * WE + hx + x];
      wout[o] = av ? val * av[o] : val;
    }
    __syncthreads();
  }
}

// ===========================================================================
// z-fastest work layout: volumes [y][x][z] (a pixel's z-vector contiguous, a
// PSF plane y one contiguous [x][z] block).  The rotations become gathers
// over z-vectors: lanes run along z (float4 each), the pixel's coefficients
// are the same for all of its lanes.  nz % 4 == 0, nz / 4 <= 64.
// ===========================================================================
// img [z][y][x] (the caller's layout) -> out [y][x][z]; grid (ceil(N/32), ny), block 256
__global__ void __launch_bounds__(256) k_to_zfast(const float* __restrict__ img, float* __restrict__ out, Geo g) {
  __shared__ float t[32 * 257];  // [x][z], z < 256, row stride 257
  const int N = g.nx, nz = g.nz, x0 = blockIdx.x * 32, y = blockIdx.y, tid = threadIdx.x;
  const int nxb = min(32, N - x0);
  const long S = (long)N * g.ny;
  for (int k = tid; k < nz * 32; k += 256) {
    const int z = k >> 5, x = k & 31;
    if (x < nxb) t[x * 257 + z] = img[S * z + (long)N * y + x0 + x];
  }
  __syncthreads();
  float* o = out + ((long)y * N + x0) * nz;
  for (int k = tid; k < nxb * nz; k += 256) {
    const int x = k / nz, z = k - x * nz;
    o[k] = t[x * 257 + z];
  }
}

// Work-groups go round-robin over the 8 XCDs, each with its own L2.  A grid
// of 8*C blocks is remapped so that XCD k runs logical blocks [k*C, (k+1)*C):
// neighbouring pixels, which share their gather sources, stay in one L2.
// Logical blocks >= the real count exit at once.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int C) { return (int)(blockIdx.x % kXcds) * C + (int)(blockIdx.x / kXcds); }
static inline int xcd_chunk(long blocks) { return (int)((blocks + kXcds - 1) / kXcds); }

// lane -> (pixel of the wave's group, z quad); PW = 64 / (nz/4) pixels per wave
struct ZLane {
  int sub, z4, pw;
  __device__ ZLane(int lane, int nz4) : pw(64 / nz4) {
    sub = lane / nz4;
    z4 = lane - sub * nz4;
  }
};

__device__ __forceinline__ float4 ld4_or0(const float* p, bool ok) {
  return ok ? *reinterpret_cast<const float4*>(p) : make_float4(0.f, 0.f, 0.f, 0.f);
}

// out[b][y][x][:] = atten[v][y][x][:] * bilinear(img_t, c + R(2pi - theta)(x,y) - c)[:]
// (the :two forward of k_rot2_fwd, same Float32 expression, over z-vectors)
// Logical block (b, pixel block): views outermost, so an XCD takes whole
// views in turn.  grid 8 * xcd_chunk(nb * PB), PB = ceil(N*N / (4*PW)); block 256
__global__ void __launch_bounds__(256) k_rot2_fwd_z(const float* __restrict__ img_t, float* __restrict__ out,
                                                    const float* __restrict__ atten, long atten_vstride,
                                                    const RotView* __restrict__ rv, const int* __restrict__ vids,
                                                    int nb, int PB, int C, Geo g) {
  const int lb = xcd_block(C);
  if (lb >= nb * PB) return;
  const int b = lb / PB, blk = lb - b * PB;
  const int N = g.nx, nz = g.nz, nz4 = nz >> 2;
  const ZLane zl(threadIdx.x & 63, nz4);
  const int pix = (blk * 4 + (threadIdx.x >> 6)) * zl.pw + zl.sub;
  if (zl.sub >= zl.pw || pix >= N * N) return;
  const int v = vids[b];
  const RotView r = rv[v];
  const int x = pix % N, y = pix / N, zo = 4 * zl.z4;
  const long V = (long)N * N * nz;
  float4 val;
  if (r.kind == 0) {
    val = *reinterpret_cast<const float4*>(img_t + (long)pix * nz + zo);
  } else {
    const BiCoef c = bilinear_coef(r, g.pad, x, y);
    // taps outside the image read 0 (the zero pad); their addresses are clamped in range
    const bool x0 = (unsigned)c.j1 < (unsigned)N, x1 = (unsigned)(c.j1 + 1) < (unsigned)N;
    const bool y0 = (unsigned)c.j2 < (unsigned)N, y1 = (unsigned)(c.j2 + 1) < (unsigned)N;
    const int ja = clampi(c.j1, 0, N - 1), jb = clampi(c.j1 + 1, 0, N - 1);
    const long ra = (long)clampi(c.j2, 0, N - 1) * N, rb = (long)clampi(c.j2 + 1, 0, N - 1) * N;
    const float4 s11 = ld4_or0(img_t + (ra + ja) * nz + zo, x0 && y0);
    const float4 s12 = ld4_or0(img_t + (ra + jb) * nz + zo, x1 && y0);
    const float4 s21 = ld4_or0(img_t + (rb + ja) * nz + zo, x0 && y1);
    const float4 s22 = ld4_or0(img_t + (rb + jb) * nz + zo, x1 && y1);
    // a = s11 w11 + s12 w12, b = s21 w11 + s22 w12, out = a w21 + b w22
    val.x = (s11.x * c.w11 + s12.x * c.w12) * c.w21 + (s21.x * c.w11 + s22.x * c.w12) * c.w22;
    val.y = (s11.y * c.w11 + s12.y * c.w12) * c.w21 + (s21.y * c.w11 + s22.y * c.w12) * c.w22;
    val.z = (s11.z * c.w11 + s12.z * c.w12) * c.w21 + (s21.z * c.w11 + s22.z * c.w12) * c.w22;
    val.w = (s11.w * c.w11 + s12.w * c.w12) * c.w21 + (s21.w * c.w11 + s22.w * c.w12) * c.w22;
  }
  const long o = (long)pix * nz + zo;
  if (atten) {
    const float4 a = *reinterpret_cast<const float4*>(atten + atten_vstride * v + o);
    val.x = val.x * a.x;
    val.y = val.y * a.y;
    val.z = val.z * a.z;
    val.w = val.w * a.w;
  }
  *reinterpret_cast<float4*>(out + V * b + o) = val;
}

// Adjoint of k_rot2_fwd_z as a gather: target pixel j collects w * work[b][i][:]
// over the K (source i, weight w) slots of its plan-time list (the forward's
// coefficients, w = the product of its two 1-D weights; unused slots carry
// w = 0 and i = j and load nothing), ent laid out [view][K][NN] so that neighbouring targets
// read neighbouring slots.  A lane takes Q adjacent z quads of one pixel
// (nz/(4Q) lanes per pixel, PW = 64/(nz/(4Q)) pixels per wave).  Wave s of a
// block sums views s, s + VS, ... of the batch for the block's PW pixels (one
// view per step, few registers: many waves in flight hide the slot -> source
// load chain); the VS wave sums are then added in wave order through LDS and
// written to part [y][x][z].
// MODE 0: sources work[b].  MODE 1 (normaliser): the source volume of view v
// is H .* atten[v] (H one volume, hvol, shared by all views, the product
// rounded as the PSF adjoint rounds it), for the back-projection of a
// view-independent PSF adjoint result.  MODE 2: both sums in one pass (the
// slots are loaded once): the back-projection of work and the normaliser;
// the normaliser goes to em_asum (zeros as Inf) and drives the EM update.
// MODE 3: MODE 2 with the products H .* atten[v] read from `atten` (the
// plan's premultiplied volumes, the same roundings), one load per slot.
// em_asum (MODE 0, optional): part holds the z-fastest image x and is updated
// in place, x .*= back ./ asum (src/ml-os-em.jl:46), asum z-fastest too.
// zero_inf (MODE 0/1): the sums are OS-EM normalisers, zeros written as Inf
// (src/ml-os-em.jl:110).
// Subsets (gridDim.y > 1): grid row k sums the k-th of consecutive view
// lists of nb views, the first nb_rem of them one view longer (the strided
// blocks of src/ml-os-em.jl:76 laid out block after block), into volume k of
// part (and of em_asum): every OS-EM normaliser from one launch.
// grid (8 * xcd_chunk(ceil(NN / PW)), subsets), block 64 * VS
// GRP: the grid holds the blocks of the ngrp one-subset launches that grp
// describes (the arguments before grp are unused); every launch's range
// starts at a multiple of 8, so its blocks land on the XCDs as they would
// alone.
template <int K, int VS, int Q, int MODE, bool GRP>
__global__ void __launch_bounds__(64 * VS) k_rot2_adj_ell(const float* __restrict__ work, float* __restrict__ part,
                                                          const int2* __restrict__ ent, const int* __restrict__ vids,
                                                          int nb, int nb_rem, int C, Geo g,
                                                          const float* __restrict__ hvol,
                                                          const float* __restrict__ atten, long atten_vstride,
                                                          float* __restrict__ em_asum, int zero_inf,
                                                          const EllGrp* __restrict__ grp, int ngrp) {
  constexpr int NS = MODE >= 2 ? 2 : 1;  // sums carried
  __shared__ float4 red[VS][64][Q * NS];
  int lb, sk = blockIdx.y;
  if constexpr (GRP) {
    const EllGrp& a = grp[grp_find(grp, ngrp)];
    const int l = (int)blockIdx.x - a.h.blk0;
    work = glob(a.work), part = glob(a.part), ent = glob(a.ent), vids = glob(a.vids), nb = a.nb, nb_rem = 0;
    C = a.C, g = a.g, hvol = glob(a.hvol), atten = glob(a.atten), atten_vstride = a.atten_vstride;
    em_asum = glob(a.em_asum), zero_inf = a.zero_inf;
    lb = __builtin_amdgcn_readfirstlane((l % kXcds) * C + l / kXcds), sk = 0;
  } else {
    lb = xcd_block(C);
  }
  const int N = g.nx, NN = N * N, nz = g.nz, nl = (nz >> 2) / Q, pw = 64 / nl;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int sub = lane / nl, zl = lane - sub * nl;
  if (lb * pw >= NN) return;  // whole block past the end (uniform)
  const int j = lb * pw + sub;
  const bool act = sub < pw && j < NN;
  const int jj = min(j, NN - 1), zo = 4 * Q * (act ? zl : 0);
  const long V = (long)NN * nz;
  {  // subset sk
    vids += sk * nb + min(sk, nb_rem);
    nb += sk < nb_rem ? 1 : 0;
    part += V * sk;
    if (em_asum) em_asum += V * sk;
  }
  float4 acc[Q * NS];
#pragma unroll
  for (int q = 0; q < Q * NS; ++q) acc[q] = make_float4(0.f, 0.f, 0.f, 0.f);
  auto fma4 = [](float4& a, float w, const float4& p) {
    a.x = a.x + w * p.x;
    a.y = a.y + w * p.y;
    a.z = a.z + w * p.z;
    a.w = a.w + w * p.w;
  };
  for (int b = wave; b < nb; b += VS) {
    const int2* ea = ent + (long)vids[b] * K * NN + jj;
    int2 sa[K];
#pragma unroll
    for (int k = 0; k < K; ++k) sa[k] = ea[(long)k * NN];
    if (MODE != 1) {  // back-projection of work[b]
      const float* wa = work + V * b + zo;
      float4 pa[K][Q];
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (__int_as_float(sa[k].y) != 0.f)  // unused slots load nothing
#pragma unroll
          for (int q = 0; q < Q; ++q) pa[k][q] = *reinterpret_cast<const float4*>(wa + (long)sa[k].x * nz + 4 * q);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const float w = __int_as_float(sa[k].y);
        if (w != 0.f)
#pragma unroll
          for (int q = 0; q < Q; ++q) fma4(acc[q], w, pa[k][q]);
      }
    }
    if (MODE == 3) {  // normaliser: the premultiplied H .* atten[v]
      const float* aa = atten + atten_vstride * vids[b] + zo;
      float4 pa[K][Q];
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (__int_as_float(sa[k].y) != 0.f)
#pragma unroll
          for (int q = 0; q < Q; ++q) pa[k][q] = *reinterpret_cast<const float4*>(aa + (long)sa[k].x * nz + 4 * q);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const float w = __int_as_float(sa[k].y);
        if (w != 0.f)
#pragma unroll
          for (int q = 0; q < Q; ++q) fma4(acc[Q + q], w, pa[k][q]);
      }
    }
    if (MODE == 1 || MODE == 2) {  // normaliser: H .* atten[v]
      const float* ha = hvol + zo;
      const float* aa = atten + atten_vstride * vids[b] + zo;
      float4 pa[K][Q];
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (__int_as_float(sa[k].y) != 0.f)
#pragma unroll
          for (int q = 0; q < Q; ++q) {
            const float4 h = *reinterpret_cast<const float4*>(ha + (long)sa[k].x * nz + 4 * q);
            const float4 fa = *reinterpret_cast<const float4*>(aa + (long)sa[k].x * nz + 4 * q);
            pa[k][q] = make_float4(h.x * fa.x, h.y * fa.y, h.z * fa.z, h.w * fa.w);
          }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const float w = __int_as_float(sa[k].y);
        if (w != 0.f)
#pragma unroll
          for (int q = 0; q < Q; ++q) fma4(acc[(NS - 1) * Q + q], w, pa[k][q]);
      }
    }
  }
  if (VS > 1) {
#pragma unroll
    for (int q = 0; q < Q * NS; ++q) red[wave][lane][q] = acc[q];
    __syncthreads();
    if (wave != 0) return;
#pragma unroll
    for (int s = 1; s < VS; ++s)
#pragma unroll
      for (int q = 0; q < Q * NS; ++q) {
        const float4 o = red[s][lane][q];
        acc[q].x = acc[q].x + o.x;
        acc[q].y = acc[q].y + o.y;
        acc[q].z = acc[q].z + o.z;
        acc[q].w = acc[q].w + o.w;
      }
  }
  if (!act) return;
  const float inf = __builtin_inff();
  auto zinf = [&](const float4& a) {
    return make_float4(a.x == 0.f ? inf : a.x, a.y == 0.f ? inf : a.y, a.z == 0.f ? inf : a.z, a.w == 0.f ? inf : a.w);
  };
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const long o = (long)j * nz + zo + 4 * q;
    float4* po = reinterpret_cast<float4*>(part + o);
    if (MODE >= 2) {
      const float4 a = zinf(acc[Q + q]);
      *reinterpret_cast<float4*>(em_asum + o) = a;
      const float4 x = *po;
      *po = make_float4(x.x * (acc[q].x / a.x), x.y * (acc[q].y / a.y), x.z * (acc[q].z / a.z), x.w * (acc[q].w / a.w));
    } else if (em_asum) {
      const float4 a = *reinterpret_cast<const float4*>(em_asum + o);
      const float4 x = *po;
      *po = make_float4(x.x * (acc[q].x / a.x), x.y * (acc[q].y / a.y), x.z * (acc[q].z / a.z), x.w * (acc[q].w / a.w));
    } else {
      *po = zero_inf ? zinf(acc[q]) : acc[q];
    }
  }
}

// Image reduction from z-fastest partials: part[g] [y][x][z] -> out [z][y][x]
// (the caller's layout) with the modes of k_img_reduce, groups summed in the
// same order.  A block transposes a 32 (x) x 32 (z) tile of row y through LDS:
// float4 reads along z, 128-byte row writes along x.  zout (optional): the
// new out values also in [y][x][z] (the next projection's z-fastest image).
// grid (ceil(N/32), ny, ceil(nz/32)), block 256
__global__ void __launch_bounds__(256) k_img_reduce_z(const float* __restrict__ part, int ng, float* __restrict__ out,
                                                      const float* __restrict__ asum, int mode, Geo g,
                                                      float* __restrict__ zout) {
  __shared__ float t[32 * 33];  // [x][z]
  const int N = g.nx, nz = g.nz, x0 = blockIdx.x * 32, y = blockIdx.y, z0 = blockIdx.z * 32, tid = threadIdx.x;
  const long V = (long)N * g.ny * nz, S = (long)N * g.ny;
  {
    const int x = tid >> 3, z = 4 * (tid & 7);
    if (x0 + x < N && z0 + z < nz) {
      const float* pp = part + ((long)y * N + x0 + x) * nz + z0 + z;
      float4 sv = *reinterpret_cast<const float4*>(pp);
      for (int gg = 1; gg < ng; ++gg) {
        const float4 o = *reinterpret_cast<const float4*>(pp + (long)gg * V);
        sv.x += o.x;
        sv.y += o.y;
        sv.z += o.z;
        sv.w += o.w;
      }
      t[x * 33 + z] = sv.x;
      t[x * 33 + z + 1] = sv.y;
      t[x * 33 + z + 2] = sv.z;
      t[x * 33 + z + 3] = sv.w;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int e = tid + 256 * r, z = e >> 5, x = e & 31;
    if (x0 + x >= N || z0 + z >= nz) continue;
    const long o = S * (z0 + z) + (long)N * y + x0 + x;
    const float sv = t[x * 33 + z];
    float nv;
    if (mode == IMG_WRITE) nv = sv;
    else if (mode == IMG_ACCUM) nv = out[o] + sv;
    else nv = out[o] * (sv / asum[o]);  // out .*= back ./ asum, src/ml-os-em.jl:46
    out[o] = nv;
    t[x * 33 + z] = nv;  // (this thread's own element)
  }
  if (!zout) return;
  __syncthreads();
  const int x = tid >> 3, z = 4 * (tid & 7);
  if (x0 + x < N && z0 + z < nz)
    *reinterpret_cast<float4*>(zout + ((long)y * N + x0 + x) * nz + z0 + z) =
        make_float4(t[x * 33 + z], t[x * 33 + z + 1], t[x * 33 + z + 2], t[x * 33 + z + 3]);
}

// ===========================================================================
// Gradient of L = ||A(x, mu) - y||^2 (no reference counterpart)
// ===========================================================================
// r = 2 (p - y) (r may be p); per-block partial sums of (p - y)^2 in double.
// grid ceil(n / 256), block 256
__global__ void __launch_bounds__(256) k_resid_loss(const float* p, const float* __restrict__ y,
                                                    float* r, long n, double* __restrict__ partial) {
  __shared__ double sh[256];
  const long k = (long)blockIdx.x * 256 + threadIdx.x;
  double d2 = 0;
  if (k < n) {
    const float d = p[k] - y[k];
    r[k] = 2.f * d;
    d2 = (double)d * (double)d;
  }
  sh[threadIdx.x] = d2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// out = sum of the partials in a fixed order (one block of 256)
__global__ void __launch_bounds__(256) k_sum_partials(const double* __restrict__ partial, int n,
                                                      double* __restrict__ out) {
  __shared__ double sh[256];
  double s = 0;
  for (int k = threadIdx.x; k < n; k += 256) s += partial[k];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

// dL/dmu_r of each view in the rotated frame, in place over g: with
// E_y = g_y * w_y (g = conv^T(r), w = R(x) .* atten, z-fastest [b][y][x][z]),
// dL/dmu_r[j] = -dy (sum_{y>j} E_y + E_j / 2), a reverse scan along y per
// (x, z) column.  grid (ceil(nx*nz / 256), nb), block 256
// The same scan with the rotated image formed on the fly instead of read:
// w_y = R(x)_y a_y, R(x) the :two rotation of the z-fastest image img_t at
// the view's coefficients (the expression, masks and clamps of k_rot2_fwd_z,
// one z element per thread), a the view's attenuation factors.  GA: g .* a,
// the attenuated PSF adjoint that the image gradient gathers, into ga.  y
// from the far side in runs of U whose loads are issued together; the sums
// in the order of a one-by-one walk.  grid (ceil(nx nz / 256), nb), block 256.
template <bool GA>
__global__ void __launch_bounds__(256) k_mu_grad_scan_rot(float* __restrict__ g, float* __restrict__ ga,
                                                          const float* __restrict__ img_t,
                                                          const float* __restrict__ atten, long atten_vstride,
                                                          const RotView* __restrict__ rv,
                                                          const int* __restrict__ vids, Geo geo, float dy) {
  const int N = geo.nx, nz = geo.nz;
  const long cols = (long)N * nz, c = (long)blockIdx.x * 256 + threadIdx.x;
  if (c >= cols) return;
  const int x = (int)(c / nz), z = (int)(c - (long)x * nz);
  const long V = cols * geo.ny;
  const int v = vids[blockIdx.y];
  const RotView r = rv[v];
  float* gb = g + V * blockIdx.y + c;
  float* gab = GA ? ga + V * blockIdx.y + c : nullptr;
  const float* ab = atten + atten_vstride * v + c;
  const float* im = img_t + z;  // element z of pixel p at im[p nz]
  auto rot = [&](int y) -> float {
    if (r.kind == 0) return im[(long)(y * N + x) * nz];
    const BiCoef cf = bilinear_coef(r, geo.pad, x, y);
    const bool x0 = (unsigned)cf.j1 < (unsigned)N, x1 = (unsigned)(cf.j1 + 1) < (unsigned)N;
    const bool y0 = (unsigned)cf.j2 < (unsigned)N, y1 = (unsigned)(cf.j2 + 1) < (unsigned)N;
    const long ja = clampi(cf.j1, 0, N - 1), jb = clampi(cf.j1 + 1, 0, N - 1);
    const long ra = (long)clampi(cf.j2, 0, N - 1) * N, rb = (long)clampi(cf.j2 + 1, 0, N - 1) * N;
    const float s11 = (x0 && y0) ? im[(ra + ja) * nz] : 0.f;
    const float s12 = (x1 && y0) ? im[(ra + jb) * nz] : 0.f;
    const float s21 = (x0 && y1) ? im[(rb + ja) * nz] : 0.f;
    const float s22 = (x1 && y1) ? im[(rb + jb) * nz] : 0.f;
    return (s11 * cf.w11 + s12 * cf.w12) * cf.w21 + (s21 * cf.w11 + s22 * cf.w12) * cf.w22;
  };
  const float mdy = -dy;
  float acc = 0.f;
  auto step = [&](long o, float gv, float rv_, float av) {
    const float e = gv * (rv_ * av);
    if constexpr (GA) gab[o] = gv * av;
    gb[o] = mdy * (acc + 0.5f * e);
    acc = acc + e;
  };
  constexpr int U = 8;
  int y = geo.ny - 1;
  for (; y >= U - 1; y -= U) {
    float gv[U], wv[U], av[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      gv[u] = gb[(long)(y - u) * cols];
      av[u] = ab[(long)(y - u) * cols];
      wv[u] = rot(y - u);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) step((long)(y - u) * cols, gv[u], wv[u], av[u]);
  }
  for (; y >= 0; --y) step((long)y * cols, gb[(long)y * cols], rot(y), ab[(long)y * cols]);
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

__global__ void __launch_bounds__(256) k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src, long n) {
  for (long k = (long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long)gridDim.x * 256) dst[k] = src[k];
}

// *flag = 1 when some entry is not > 0 (NaN included); the caller clears it
__global__ void __launch_bounds__(256) k_flag_nonpositive(const float* __restrict__ p, long n, int* __restrict__ flag) {
  bool bad = false;
  for (long k = (long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long)gridDim.x * 256) bad |= !(p[k] > 0.f);
  if (bad) *flag = 1;
}

// ===========================================================================
// Launchers
// ===========================================================================
static inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

// compute units of the current device (cached per device)
static int device_cus() {
  static std::atomic<int> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    (void)hipGetLastError();
    return 256;
  }
  int n = cache[dev].load(std::memory_order_relaxed);
  if (n > 0) return n;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    n = 256;
  }
  cache[dev].store(n, std::memory_order_relaxed);
  return n;
}

// Strip length of the separable kernels: one strip per thread per phase, at
// most 1024 threads per block.
static int sep_strip_len(const Geo& g) {
  for (int sl : {8, 16, 32}) {
    const long sx = (long)g.nz * ((g.nx + sl - 1) / sl), sz = (long)g.nx * ((g.nz + sl - 1) / sl);
    if (std::max(sx, sz) <= kSepThreads) return sl;
  }
  return 0;
}

// kernels whose dynamic LDS exceeds 64 KiB must opt in (once per kernel and size)
static std::mutex g_lds_mu;
static std::unordered_map<const void*, size_t> g_lds_done;
template <typename K>
static void allow_lds(K kernel, size_t bytes) {
  if (bytes <= 65536) return;
  const void* f = reinterpret_cast<const void*>(kernel);
  std::lock_guard<std::mutex> lk(g_lds_mu);
  size_t& done = g_lds_done[f];
  if (bytes <= done) return;
  if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess) done = bytes;
  else (void)hipGetLastError();
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
    const size_t lds = sizeof(float) * g.Np * g.Np, lds_t = lds + sizeof(float) * 6 * g.Np;
    if (g.Np > kMaxNpOne || lds > kMaxLds) return hipErrorInvalidConfiguration;
    if (lds_t <= kMaxLds) {  // interpolation tables beside the slice
      allow_lds(k_rot1_fwd<true>, lds_t);
      k_rot1_fwd<true><<<dim3(g.nz, nb), R1_THREADS, lds_t, s>>>(img, out, atten, atten_vstride, rv, vids, tab_b,
                                                                 tab_w, g);
    } else {
      allow_lds(k_rot1_fwd<false>, lds);
      k_rot1_fwd<false><<<dim3(g.nz, nb), R1_THREADS, lds, s>>>(img, out, atten, atten_vstride, rv, vids, tab_b,
                                                                tab_w, g);
    }
  } else {
    // views per block: enough blocks to fill the device, reuse of the staged slice otherwise
    int vpg = 1;
    while (vpg < 8 && (long)g.nz * cdiv(nb, vpg * 2) >= 1024) vpg *= 2;
    const size_t lds = sizeof(float) * S;
    if (g.nx % 4 == 0 && lds <= kMaxLds) {
      // views per block: blocks run in rounds of (CUs x blocks per CU: LDS
      // and 1024-thread blocks allow two), a block taking about half a view's
      // time to stage its slice and one per view, so the time goes as
      // rounds x (2 views + 1); the fewest views per block at that minimum
      // (at 128x128x80, 64 views: 11 views, one round of 480 blocks, against
      // 8 views in 1.25 rounds)
      const long slots = (long)device_cus() * std::max(1L, std::min(2L, (long)(kMaxLds / lds)));
      int vpg4 = nb;
      long best = -1;
      for (int v = 1; v <= nb; ++v) {
        const long cost = (long)cdiv((long)g.nz * cdiv(nb, v), slots) * (2L * v + 1);
        if (best < 0 || cost < best) best = cost, vpg4 = v;
      }
      allow_lds(k_rot2_fwd4, lds);
      k_rot2_fwd4<<<dim3(g.nz, cdiv(nb, vpg4)), R2F_THREADS, lds, s>>>(img, out, atten, atten_vstride, rv, vids, nb,
                                                                       vpg4, g);
    } else if (lds <= kMaxLds) {
      allow_lds(k_rot2_fwd<true>, lds);
      k_rot2_fwd<true><<<dim3(g.nz, cdiv(nb, vpg)), 256, lds, s>>>(img, out, atten, atten_vstride, rv, vids, nb,
                                                                  vpg, g);
    } else
      k_rot2_fwd<false><<<dim3(g.nz, cdiv(nb, vpg)), 256, 0, s>>>(img, out, atten, atten_vstride, rv, vids, nb, vpg,
                                                                 g);
  }
  return hipGetLastError();
}

int rot2_adj_blocks_per_cu(const Geo& g) {
  const size_t lds = g.nx <= 128 ? sizeof(float) * (g.nx + 2) * (g.nx + 8) : sizeof(float) * g.nx * (g.nx + 1);
  return std::max(1, std::min(2, (int)(kMaxLds / lds)));  // 1024-thread blocks: at most 2 per CU
}

hipError_t launch_rot_adj(int interp, const float* work, float* part, const RotView* rv, const int* vids, int nb,
                          int vpg, const int* tab_b, const float* tab_w, const Geo& g, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  if (interp == 1) {  // one view per block: part holds nb slices-volumes
    const size_t lds = sizeof(float) * g.Np * g.Np, lds_t = lds + sizeof(float) * 6 * g.Np;
    if (g.Np > kMaxNpOne || lds > kMaxLds) return hipErrorInvalidConfiguration;
    if (lds_t <= kMaxLds) {
      allow_lds(k_rot1_adj<true>, lds_t);
      k_rot1_adj<true><<<dim3(g.nz, nb), R1_THREADS, lds_t, s>>>(work, part, rv, vids, tab_b, tab_w, g);
    } else {
      allow_lds(k_rot1_adj<false>, lds);
      k_rot1_adj<false><<<dim3(g.nz, nb), R1_THREADS, lds, s>>>(work, part, rv, vids, tab_b, tab_w, g);
    }
  } else {
    const int ng = (nb + vpg - 1) / vpg;
    if (g.nx <= 128) {
      const size_t lds = sizeof(float) * (g.nx + 2) * (g.nx + 8);
      allow_lds(k_rot2_adj_scatter, lds);
      k_rot2_adj_scatter<<<dim3(g.nz, ng), RS_THREADS, lds, s>>>(work, part, rv, vids, nb, vpg, g);
    } else {
      const size_t lds = sizeof(float) * g.nx * (g.nx + 1);
      if (lds <= kMaxLds) {
        allow_lds(k_rot2_adj<true>, lds);
        k_rot2_adj<true><<<dim3(g.nz, ng), RA_THREADS, lds, s>>>(work, part, rv, vids, nb, vpg, g);
      } else {
        k_rot2_adj<false><<<dim3(g.nz, ng), RA_THREADS, 0, s>>>(work, part, rv, vids, nb, vpg, g);
      }
    }
  }
  return hipGetLastError();
}

hipError_t launch_img_reduce(const float* part, int ngroups, float* out, const float* asum, long n, int mode,
                             hipStream_t s) {
  k_img_reduce<<<cdiv(n, 256), 256, 0, s>>>(part, ngroups, n, out, asum, mode);
  return hipGetLastError();
}

hipError_t launch_atten_scan(float* mur, int nb, const Geo& g, float dy, bool zfast, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  const long cols = (long)g.nx * g.nz;
  // lanes per column: one (a streaming pass) while the batch's columns give
  // a launch about 256K lanes, else the power of two that does, within ny
  int seg = 1;
  while (seg < 64 && 2 * seg <= g.ny && cols * nb * seg < (256L << 10)) seg *= 2;
  const int zf = zfast ? 1 : 0;
  switch (seg) {
#define SR_ATTEN_CASE(S) \
  case S:                \
    k_atten_scan<S><<<dim3(cdiv(cols, 4L * (64 / S)), nb), 256, 0, s>>>(mur, g, dy, zf); \
    break;
    SR_ATTEN_CASE(1) SR_ATTEN_CASE(2) SR_ATTEN_CASE(4) SR_ATTEN_CASE(8) SR_ATTEN_CASE(16) SR_ATTEN_CASE(32)
    SR_ATTEN_CASE(64)
#undef SR_ATTEN_CASE
    default:
      return hipErrorInvalidValue;
  }
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

// tiles per wave of the matrix-core kernels (the launch's template argument)
static int mf_tpw(const Geo& g) { return (int)cdiv((long)(g.nx / 16) * (g.nz / 16), MF_WAVES); }
#define SR_TPW_SWITCH(TV, ...)                                  \
  switch (TV) {                                                 \
    case 1: { constexpr int TPW = 1; __VA_ARGS__; } break;      \
    case 2: { constexpr int TPW = 2; __VA_ARGS__; } break;      \
    case 3: { constexpr int TPW = 3; __VA_ARGS__; } break;      \
    case 4: { constexpr int TPW = 4; __VA_ARGS__; } break;      \
    case 5: { constexpr int TPW = 5; __VA_ARGS__; } break;      \
    case 6: { constexpr int TPW = 6; __VA_ARGS__; } break;      \
    case 7: { constexpr int TPW = 7; __VA_ARGS__; } break;      \
    case 8: { constexpr int TPW = 8; __VA_ARGS__; } break;      \
    default: return hipErrorInvalidValue;                       \
  }

// matrix-core separable kernels: 16-multiple planes that the block's
// registers can prefetch, taps within the fragment tables, LDS within a CU
bool psf_mfma_supported(const Geo& g, int nt) {
  if (g.nx % 16 != 0 || g.nz % 16 != 0 || g.nx < 16 || g.nz < 16) return false;
  if ((long)g.nx * g.nz > 4L * MF_PF * MF_THREADS) return false;
  if (mf_tpw(g) > MF_TPW) return false;
  if (6 * nt > MF_THREADS || std::max(g.px, g.pz) > nt) return false;
  return sizeof(float) * mf_lds_floats(g, nt, 1) <= kMaxLds;
}

bool psf_sep_supported(const Geo& g) {
  const int SLv = sep_strip_len(g);
  if (SLv == 0 || g.nx % 4 != 0) return false;
  const int sx = (g.nx + SLv - 1) / SLv, sz = (g.nz + SLv - 1) / SLv;
  const long need = std::max((long)std::max(g.nz * sx, g.nx * sz), (long)cdiv((long)g.nx * g.nz, 4L * PF_MAX));
  return need <= kSepThreads;  // the separable kernels' launch bound
}

bool psf_fwd_col_ok(const Geo& g, int ntx) {
  return psf_mfma_supported(g, ntx) && g.nx / 16 <= MF_WAVES && g.nz / 16 <= MF_TPW &&
         sizeof(float) * mf_fcol_lds_floats(g, ntx, 1) <= kMaxLds;
}

bool psf_fwd_ws_ok(const Geo& g, int ntx, int npw) {
  return psf_fwd_col_ok(g, ntx) && g.nx <= 64 * npw && (long)g.nx * g.nz <= 10240 &&
         sizeof(float) * mf_fws_lds_floats(g, ntx) <= kMaxLds;
}

bool psf_fwd_ws2_ok(const Geo& g, int ntx) {
  return psf_fwd_col_ok(g, ntx) && g.nx <= 256 && (long)g.nx * g.nz <= 10240 &&
         sizeof(float) * mf_fws2_lds_floats(g, ntx) <= kMaxLds;
}

// producer waves of the built-in forward: 8 while a column's tiles leave the
// 1024-thread kernel room (nz <= 80), else 4; 0 when neither fits
int psf_fwd_auto_npw(const Geo& g, int ntx) {
  if (g.nz / 16 <= 5 && psf_fwd_ws_ok(g, ntx, 8)) return 8;
  return psf_fwd_ws_ok(g, ntx, 4) ? 4 : 0;
}

bool psf_adj_col_ok(const Geo& g, int ntx) {
  return psf_mfma_supported(g, ntx) && g.nz / 16 <= MF_TPW && sizeof(float) * mf_col_lds_floats(g, ntx, 1) <= kMaxLds;
}

hipError_t launch_psf_fwd(int separable, const float* work, const Taps& t, const int* vids, int nb, float* part,
                          int nchunk, int* nchunk_used, const Geo& g, hipStream_t s, bool zfast, const float* atten,
                          long atten_vstride, const RotView* rv, int family) {
  const int Yc = (g.ny + nchunk - 1) / nchunk;
  nchunk = (g.ny + Yc - 1) / Yc;
  if (nchunk_used) *nchunk_used = nchunk;
  if (nb <= 0) return hipSuccess;
  if (separable && zfast && family == PSF_WS2 && psf_fwd_ws2_ok(g, t.ntx)) {
    const size_t lds = sizeof(float) * mf_fws2_lds_floats(g, t.ntx);
    SR_TPW_SWITCH(g.nz / 16, {
      allow_lds(k_psf_fwd_ws2<TPW>, lds);
      k_psf_fwd_ws2<TPW><<<dim3(nchunk, nb), MF_THREADS, lds, s>>>(work, t, vids, Yc, part, g, atten, atten_vstride,
                                                                   rv);
    });
    return hipGetLastError();
  }
  const int npw = family == PSF_WS8 ? 8 : family == PSF_WS4 ? 4 : psf_fwd_auto_npw(g, t.ntx);
  if (separable && zfast && (family == PSF_AUTO || family == PSF_WS8 || family == PSF_WS4) && npw > 0 &&
      psf_fwd_ws_ok(g, t.ntx, npw)) {
    const size_t lds = sizeof(float) * mf_fws_lds_floats(g, t.ntx);
    SR_TPW_SWITCH(g.nz / 16, {
      if (npw == 8) {
        allow_lds(k_psf_fwd_ws<TPW, 8, false>, lds);
        k_psf_fwd_ws<TPW, 8, false><<<dim3(nchunk, nb), MF_THREADS + 512, lds, s>>>(work, t, vids, Yc, part, g, atten,
                                                                                    atten_vstride, rv, nullptr, 0);
      } else {
        allow_lds(k_psf_fwd_ws<TPW, 4, false>, lds);
        k_psf_fwd_ws<TPW, 4, false><<<dim3(nchunk, nb), MF_THREADS + 256, lds, s>>>(work, t, vids, Yc, part, g, atten,
                                                                                    atten_vstride, rv, nullptr, 0);
      }
    });
    return hipGetLastError();
  }
  if (separable && zfast && family != PSF_TILE && psf_fwd_col_ok(g, t.ntx)) {
    const size_t base = mf_fcol_lds_floats(g, t.ntx, 0), per = MfTaps(t.ntx).size;
    const int tbp = (int)std::max(1L, std::min<long>(std::min(8, Yc), ((long)(kMaxLds / sizeof(float)) - (long)base) /
                                                                         (long)per));
    const size_t lds = sizeof(float) * mf_fcol_lds_floats(g, t.ntx, tbp);
    SR_TPW_SWITCH(g.nz / 16, {
      allow_lds(k_psf_fwd_col<TPW>, lds);
      k_psf_fwd_col<TPW><<<dim3(nchunk, nb), MF_THREADS, lds, s>>>(work, t, vids, Yc, part, g, tbp, atten,
                                                                   atten_vstride, rv);
    });
    return hipGetLastError();
  }
  if (separable && psf_mfma_supported(g, t.ntx)) {
    const int tbp = mf_tap_planes(g, t.ntx, Yc);
    const size_t lds = sizeof(float) * mf_lds_floats(g, t.ntx, tbp);
    SR_TPW_SWITCH(mf_tpw(g), {
      if (zfast) {
        allow_lds(k_psf_fwd_mfma<TPW, MF_WAVES, true>, lds);
        k_psf_fwd_mfma<TPW, MF_WAVES, true><<<dim3(nchunk, nb), MF_THREADS, lds, s>>>(
            work, t, vids, Yc, part, g, tbp, atten, atten_vstride, rv);
      } else {
        allow_lds(k_psf_fwd_mfma<TPW, MF_WAVES, false>, lds);
        k_psf_fwd_mfma<TPW, MF_WAVES, false><<<dim3(nchunk, nb), MF_THREADS, lds, s>>>(work, t, vids, Yc, part, g,
                                                                                      tbp, nullptr, 0, nullptr);
      }
    });
    return hipGetLastError();
  }
  const int SLv = separable ? sep_strip_len(g) : 0;
  if (separable && SLv > 0 && g.nx % 4 == 0) {
    const int sx = (g.nx + SLv - 1) / SLv, sz = (g.nz + SLv - 1) / SLv;
    // one strip per thread per pass, and enough threads to prefetch the plane
    const long need = std::max((long)std::max(g.nz * sx, g.nx * sz), (long)cdiv((long)g.nx * g.nz, 4L * PF_MAX));
    const int block = std::max(64, (int)cdiv(need, 64) * 64);
    if (block > kSepThreads) return hipErrorInvalidConfiguration;  // (the kernel's launch bound)
    const size_t area = std::max((size_t)g.nz * quad_stride(SLv * sx + t.ntx - 1),
                                 (size_t)SLv * sx * quad_stride(SLv * sz + t.ntx - 1));
    const size_t lds = sizeof(float) * (area + 2 * t.ntx);
    if (lds > kMaxLds) return hipErrorInvalidConfiguration;
    SR_NT_SWITCH(t.ntx, SR_SL_SWITCH(SLv, {
      allow_lds(k_psf_fwd_sep<NT, SL>, lds);
      k_psf_fwd_sep<NT, SL><<<dim3(nchunk, nb), block, lds, s>>>(work, t, vids, Yc, part, g);
    }));
    return hipGetLastError();
  }
  if (separable || zfast) return hipErrorInvalidConfiguration;
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
  return hipGetLastError();
}

hipError_t launch_view_reduce(const float* part, int nchunk, const int* vids, int nb, float* out, const float* yobs,
                              const float* bkg, int epi, const Geo& g, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  const int P = g.nx * g.nz;
  k_view_reduce<false><<<dim3(cdiv(P, 256), nb), 256, 0, s>>>(part, nchunk, P, vids, out, yobs, bkg, epi, nullptr, 0);
  return hipGetLastError();
}

hipError_t launch_psf_adj(int separable, const float* views, const Taps& t, const int* vids, int nb, float* work,
                          const float* atten, long atten_vstride, int nchunk, const Geo& g, hipStream_t s,
                          bool zfast, int family) {
  if (nb <= 0) return hipSuccess;
  const int Yc = (g.ny + nchunk - 1) / nchunk;
  nchunk = (g.ny + Yc - 1) / Yc;
  if (separable && zfast && family != PSF_TILE && psf_adj_col_ok(g, t.ntx)) {
    const long per = MfTaps(t.ntx).size;
    // tap planes that fit beside the view
    const long room = ((long)(kMaxLds / sizeof(float)) - (long)mf_col_lds_floats(g, t.ntx, 0)) / per;
    const int tbp = (int)std::max(1L, std::min<long>(std::min(8, Yc), room));
    const size_t lds = sizeof(float) * mf_col_lds_floats(g, t.ntx, tbp);
    // (16 waves hold at most 128 registers each: the 8-tile columns take 8)
    SR_TPW_SWITCH(g.nz / 16, {
      constexpr int NWV = TPW >= 8 ? 8 : 16;
      allow_lds(k_psf_adj_col<TPW, NWV, false>, lds);
      k_psf_adj_col<TPW, NWV, false><<<dim3(nchunk, nb), 64 * NWV, lds, s>>>(views, t, vids, Yc, work, atten,
                                                                             atten_vstride, g, tbp, nullptr, 0);
    });
    return hipGetLastError();
  }
  if (separable && psf_mfma_supported(g, t.ntx)) {
    const int tbp = mf_tap_planes(g, t.ntx, Yc);
    const size_t lds = sizeof(float) * mf_lds_floats(g, t.ntx, tbp);
    SR_TPW_SWITCH(mf_tpw(g), {
      if (zfast) {
        allow_lds(k_psf_adj_mfma<TPW, MF_WAVES, true>, lds);
        k_psf_adj_mfma<TPW, MF_WAVES, true><<<dim3(nchunk, nb), MF_THREADS, lds, s>>>(views, t, vids, Yc, work, atten,
                                                                                     atten_vstride, g, tbp);
      } else {
        allow_lds(k_psf_adj_mfma<TPW, MF_WAVES, false>, lds);
        k_psf_adj_mfma<TPW, MF_WAVES, false><<<dim3(nchunk, nb), MF_THREADS, lds, s>>>(views, t, vids, Yc, work, atten,
                                                                                      atten_vstride, g, tbp);
      }
    });
    return hipGetLastError();
  }
  if (zfast) return hipErrorInvalidConfiguration;
  if (separable) {
    const int SLv = sep_strip_len(g);
    if (SLv == 0) return hipErrorInvalidValue;
    const int sx = (g.nx + SLv - 1) / SLv, sz = (g.nz + SLv - 1) / SLv;
    // one strip per thread and pass; at least one thread per tap-table entry
    const int block = (int)cdiv(std::max(std::max(g.nz * sx, g.nx * sz), 6 * t.ntx), 64) * 64;
    if (block > kSepThreads) return hipErrorInvalidConfiguration;  // (the kernel's launch bound)
    const size_t lds = sizeof(float) * ((size_t)g.nx * quad_stride(SLv * sz + t.ntx - 1) +
                                        (size_t)g.nz * quad_stride(SLv * sx + t.ntx - 1) + 12 * t.ntx);
    if (lds > kMaxLds) return hipErrorInvalidConfiguration;
    SR_NT_SWITCH(t.ntx, SR_SL_SWITCH(SLv, {
      allow_lds(k_psf_adj_sep<NT, SL>, lds);
      k_psf_adj_sep<NT, SL><<<dim3(nchunk, nb), block, lds, s>>>(views, t, vids, Yc, work, atten, atten_vstride, g);
    }));
  } else if (sizeof(float) * adj2d_lds_floats(g, t.ntx) <= kMaxLds) {
    const size_t lds = sizeof(float) * adj2d_lds_floats(g, t.ntx);
    // threads per block: about 16 waves per CU over the blocks that fit in
    // its LDS (large planes leave room for one), not more than the strips
    const long per_cu = std::max<long>(1, (long)(kMaxLds / lds));
    const long strips = (long)((g.nx + g.px - 1 + 7) / 8) * (g.nz + g.pz - 1);
    int block = per_cu >= 4 ? 256 : per_cu >= 2 ? 512 : 1024;
    while (block > 256 && strips <= block / 2) block /= 2;
    SR_NT_SWITCH(t.ntx, SR_BLOCK_SWITCH(block, {
      allow_lds(k_psf_adj_2d_strips<NT, BLOCK>, lds);
      k_psf_adj_2d_strips<NT, BLOCK><<<dim3(nchunk, nb), BLOCK, lds, s>>>(views, t, vids, Yc, work, atten,
                                                                          atten_vstride, g);
    }));
  } else {  // (planes too large for the staged view: the per-term loop)
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

bool zfast_supported(const Geo& g) {
  // z-vector lanes, the attenuation kernel's LDS columns, the fused forward's
  // element slots (MF_RA + MF_RB float4 per thread)
  return g.nz % 4 == 0 && g.nz / 4 <= 64 && sizeof(float) * g.ny * 256 <= kMaxLds &&
         (long)g.nx * g.nz <= 4L * (MF_RA + MF_RB) * MF_THREADS;
}

// Per-target source lists of the :two forward (the adjoint's gather form),
// padded to K slots: slot k of target pixel q of view v is
// ent[(v * K + k) * NN + q] = (source pixel p, weight bits), weight = w1 * w2
// of the forward's bilinear coefficients (bilinear_coef, evaluated here on
// the host with the same Float32 operations; the map also for a kind-0 view,
// as the scatter adjoint does), slots ordered by p; zero weights and taps in
// the zero pad are left out, unused slots are (q, 0).  Pixel index = y*N + x.
// K = the largest list, at least 4; false when it exceeds kMaxEll.
bool build_rot2_lists(const RotView* rv, int nview, const Geo& g, int* K, std::vector<int2>& ent) {
  const int N = g.nx;
  const long NN = (long)N * N;
  struct Tap {
    int q;
    float w;
  };
  auto taps = [&](const RotView& r, int pix, Tap (&t)[4]) {
    int n = 0;
    const BiCoef c = bilinear_coef(r, g.pad, pix % N, pix / N);
    const int jx[2] = {c.j1, c.j1 + 1}, jy[2] = {c.j2, c.j2 + 1};
    const float wx[2] = {c.w11, c.w12}, wy[2] = {c.w21, c.w22};
    for (int a = 0; a < 2; ++a)
      for (int bb = 0; bb < 2; ++bb) {
        const float w = wx[bb] * wy[a];
        if (jx[bb] < 0 || jx[bb] >= N || jy[a] < 0 || jy[a] >= N || w == 0.f) continue;
        t[n++] = {jy[a] * N + jx[bb], w};
      }
    return n;
  };
  // views are independent: split them over host threads
  const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  auto over_views = [&](auto&& f) {
    std::vector<std::thread> th;
    for (int w = 0; w < nth; ++w)
      th.emplace_back([&, w] {
        for (int v = w; v < nview; v += nth) f(v);
      });
    for (auto& t : th) t.join();
  };
  std::vector<unsigned char> cnt((size_t)nview * NN, 0);
  std::vector<int> kv(nview, 4);
  over_views([&](int v) {
    unsigned char* cv = cnt.data() + (size_t)v * NN;
    Tap t[4];
    for (long pix = 0; pix < NN; ++pix) {
      const int n = taps(rv[v], (int)pix, t);
      for (int k = 0; k < n; ++k) kv[v] = std::max(kv[v], (int)++cv[t[k].q]);
    }
  });
  const int kmax = *std::max_element(kv.begin(), kv.end());
  if (kmax > kMaxEll) return false;
  *K = kmax;
  ent.assign((size_t)nview * kmax * NN, int2{});
  over_views([&](int v) {
    int2* ev = ent.data() + (size_t)v * kmax * NN;
    for (long q = 0; q < NN; ++q)
      for (int k = 0; k < kmax; ++k) ev[k * NN + q] = int2{(int)q, 0};
    std::vector<unsigned char> fill(NN, 0);
    Tap t[4];
    for (long pix = 0; pix < NN; ++pix) {
      const int n = taps(rv[v], (int)pix, t);
      for (int k = 0; k < n; ++k) {
        int2 e;
        e.x = (int)pix;
        std::memcpy(&e.y, &t[k].w, sizeof(float));
        ev[(long)fill[t[k].q]++ * NN + t[k].q] = e;
      }
    }
  });
  return true;
}

// The same lists built on the device, target by target: target q of view v
// takes the sources p whose forward footprint holds q, in increasing p.  The
// forward samples at A f + b (f = (x + pad + 1, y + pad + 1), A the float32
// rotation of RotView), and q = (qx, qy) is in the footprint of a sample in
// [qx + pad, qx + pad + 2) x [qy + pad, qy + pad + 2).  That square mapped
// back through A (a rotation) lies within sqrt(2) of the preimage of its
// centre; the window searched is 2.5 pixels about it, and every candidate is
// tested with bilinear_coef itself, so membership, weights and order are the
// host builder's.  COUNT: the block's largest list (a block-wide max, no
// atomics) to blockmax[v * gridDim.x + blockIdx.x]; else the K slots of each
// target, unused ones (q, 0).  grid (ceil(NN / 256), nview), block 256.
template <bool COUNT>
__global__ void __launch_bounds__(256) k_rot2_lists(const RotView* __restrict__ rv, Geo g, int K,
                                                    int2* __restrict__ ent, int* __restrict__ blockmax) {
  __shared__ int wmax[4];
  const int N = g.nx, NN = N * N, v = blockIdx.y;
  const int q = blockIdx.x * 256 + threadIdx.x;
  int n = 0;
  if (q < NN) {
    const RotView r = rv[v];
    const int qx = q % N, qy = q / N;
    const double sx = qx + g.pad + 1.0 - (double)r.x0, sy = qy + g.pad + 1.0 - (double)r.y0;
    const double det = (double)r.xx * r.yy - (double)r.xy * r.yx;
    const double ox = ((double)r.yy * sx - (double)r.xy * sy) / det - g.pad - 1.0;
    const double oy = (-(double)r.yx * sx + (double)r.xx * sy) / det - g.pad - 1.0;
    const int x0 = max(0, (int)floor(ox - 2.5)), x1 = min(N - 1, (int)ceil(ox + 2.5));
    const int y0 = max(0, (int)floor(oy - 2.5)), y1 = min(N - 1, (int)ceil(oy + 2.5));
    int2* eq = ent + (long)v * K * NN + q;
    for (int y = y0; y <= y1; ++y)
      for (int x = x0; x <= x1; ++x) {
        const BiCoef c = bilinear_coef(r, g.pad, x, y);
        const int dx = qx - c.j1, dy = qy - c.j2;
        if ((unsigned)dx > 1u || (unsigned)dy > 1u) continue;
        const float w = (dx ? c.w12 : c.w11) * (dy ? c.w22 : c.w21);
        if (w == 0.f) continue;
        if (!COUNT && n < K) eq[(long)n * NN] = make_int2(y * N + x, __float_as_int(w));
        ++n;
      }
    if (!COUNT)
      for (int k = n; k < K; ++k) eq[(long)k * NN] = make_int2(q, 0);
  }
  if (COUNT) {
    for (int m = 32; m >= 1; m >>= 1) n = max(n, __shfl_xor(n, m, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0)
      blockmax[v * gridDim.x + blockIdx.x] = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
  }
}

int rot2_list_blocks(const Geo& g) { return (int)cdiv((long)g.nx * g.nx, 256); }

hipError_t launch_rot2_list_count(const RotView* rv, int nview, const Geo& g, int* blockmax, hipStream_t s) {
  k_rot2_lists<true><<<dim3(rot2_list_blocks(g), nview), 256, 0, s>>>(rv, g, 0, nullptr, blockmax);
  return hipGetLastError();
}

hipError_t launch_rot2_list_fill(const RotView* rv, int nview, const Geo& g, int K, int2* ent, hipStream_t s) {
  k_rot2_lists<false><<<dim3(rot2_list_blocks(g), nview), 256, 0, s>>>(rv, g, K, ent, nullptr);
  return hipGetLastError();
}

hipError_t launch_to_zfast(const float* img, float* out, const Geo& g, hipStream_t s) {
  k_to_zfast<<<dim3(cdiv(g.nx, 32), g.ny), 256, 0, s>>>(img, out, g);
  return hipGetLastError();
}

hipError_t launch_rot2_fwd_z(const float* img_t, float* out, const float* atten, long atten_vstride,
                             const RotView* rv, const int* vids, int nb, const Geo& g, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  const int pw = 64 / (g.nz / 4);
  const int PB = (int)cdiv((long)g.nx * g.nx, 4L * pw), C = xcd_chunk((long)nb * PB);
  k_rot2_fwd_z<<<kXcds * C, 256, 0, s>>>(img_t, out, atten, atten_vstride, rv, vids, nb, PB, C, g);
  return hipGetLastError();
}


// nsubsets > 1: consecutive view lists of nb views (the first nb_rem of them
// nb + 1) summed into consecutive volumes of part (em_asum), one grid row each
static hipError_t launch_ell(int mode, const float* work, float* part, const int2* ent, int K, const int* vids, int nb,
                             const Geo& g, hipStream_t s, const float* hvol, const float* atten, long atten_vstride,
                             float* em_asum, bool zero_inf, int vs_req, int nsubsets = 1, int nb_rem = 0) {
  if (nb <= 0 && nb_rem == 0) return hipSuccess;
  if (nb < 0 || nsubsets <= 0 || nb_rem < 0 || nb_rem >= nsubsets) return hipErrorInvalidValue;
  const int q = (g.nz / 4) % 2 == 0 ? 2 : 1;  // z quads per lane
  const int pw = 64 / (g.nz / 4 / q);
  const int C = xcd_chunk(cdiv((long)g.nx * g.nx, pw));
  // view slices (waves) per block: each wave sums every vs-th view.
  // Built-in: the fewest slices (2, 4, 8) that give the launch (all its
  // subset rows) 4096 waves, at most the longest view list rounded down to a
  // power of two (measured on
  // 4- to 64-view batches: 2 wins at 128x128 pixels, 8 at 64x64 from 8 views)
  int vs = 2;
  if (vs_req == 2 || vs_req == 4 || vs_req == 8)
    vs = vs_req;
  else
    while (vs < 8 && 2 * vs <= nb + (nb_rem > 0 ? 1 : 0) && (long)kXcds * C * vs * nsubsets < 4096) vs *= 2;
  const int zi = zero_inf ? 1 : 0;
#define SR_ELL(KK, VV, QQ)                                                                                            \
  if (K == KK && vs == VV && q == QQ) {                                                                             \
    if (mode == 0)                                                                                                  \
      k_rot2_adj_ell<KK, VV, QQ, 0, false><<<dim3(kXcds * C, nsubsets), 64 * VV, 0, s>>>(                       \
          work, part, ent, vids, nb, nb_rem, C, g, hvol, atten, atten_vstride, em_asum, zi, nullptr, 0);          \
    else if (mode == 1)                                                                                             \
      k_rot2_adj_ell<KK, VV, QQ, 1, false><<<dim3(kXcds * C, nsubsets), 64 * VV, 0, s>>>(                       \
          work, part, ent, vids, nb, nb_rem, C, g, hvol, atten, atten_vstride, em_asum, zi, nullptr, 0);          \
    else if (mode == 2)                                                                                             \
      k_rot2_adj_ell<KK, VV, QQ, 2, false><<<dim3(kXcds * C, nsubsets), 64 * VV, 0, s>>>(                       \
          work, part, ent, vids, nb, nb_rem, C, g, hvol, atten, atten_vstride, em_asum, zi, nullptr, 0);          \
    else                                                                                                            \
      k_rot2_adj_ell<KK, VV, QQ, 3, false><<<dim3(kXcds * C, nsubsets), 64 * VV, 0, s>>>(                       \
          work, part, ent, vids, nb, nb_rem, C, g, hvol, atten, atten_vstride, em_asum, zi, nullptr, 0);          \
    return hipGetLastError();                                                                                       \
  }
#define SR_ELL_K(KK) SR_ELL(KK, 2, 1) SR_ELL(KK, 4, 1) SR_ELL(KK, 8, 1) SR_ELL(KK, 2, 2) SR_ELL(KK, 4, 2) SR_ELL(KK, 8, 2)
  SR_ELL_K(4) SR_ELL_K(5) SR_ELL_K(6) SR_ELL_K(7) SR_ELL_K(8)
#undef SR_ELL_K
#undef SR_ELL
  return hipErrorInvalidConfiguration;
}

// ---- grouped launches: descriptors mirror the single launches above ----------
int psf_fwd_grp_desc(const float* work, const Taps& t, const int* vids, int nb, float* part, int nchunk,
                     int* nchunk_used, const Geo& g, const float* atten, long atten_vstride, const RotView* rv,
                     int family, PsfFwdGrp* d, size_t* lds) {
  // launch_psf_fwd on a z-fastest separable plan: its producer/consumer kernel
  const int Yc = (g.ny + nchunk - 1) / nchunk;
  nchunk = (g.ny + Yc - 1) / Yc;
  if (nchunk_used) *nchunk_used = nchunk;
  if (nb <= 0) return 0;
  if (family == PSF_WS2 && psf_fwd_ws2_ok(g, t.ntx)) return 0;
  const int npw = family == PSF_WS8 ? 8 : family == PSF_WS4 ? 4 : psf_fwd_auto_npw(g, t.ntx);
  if (!((family == PSF_AUTO || family == PSF_WS8 || family == PSF_WS4) && (npw == 4 || npw == 8) &&
        psf_fwd_ws_ok(g, t.ntx, npw)))
    return 0;
  const int tpw = g.nz / 16;
  if (g.nz % 16 != 0 || tpw < 1 || tpw > 8) return 0;
  *d = PsfFwdGrp{GrpHdr{0, nchunk * nb, nchunk, 0}, work, t, vids, Yc, part, g, atten, atten_vstride, rv};
  *lds = sizeof(float) * mf_fws_lds_floats(g, t.ntx);
  return 1 << 16 | tpw << 8 | npw;
}

hipError_t launch_psf_fwd_grp(int key, const PsfFwdGrp* d, int n, int blocks, size_t lds, hipStream_t s) {
  if (n <= 0 || blocks <= 0) return hipSuccess;
  if ((key >> 16) != 1) return hipErrorInvalidValue;
  const int npw = key & 255;
  SR_TPW_SWITCH((key >> 8) & 255, {
    if (npw == 8) {
      allow_lds(k_psf_fwd_ws<TPW, 8, true>, lds);
      k_psf_fwd_ws<TPW, 8, true><<<blocks, MF_THREADS + 512, lds, s>>>(nullptr, Taps{}, nullptr, 0, nullptr, Geo{},
                                                                       nullptr, 0, nullptr, d, n);
    } else if (npw == 4) {
      allow_lds(k_psf_fwd_ws<TPW, 4, true>, lds);
      k_psf_fwd_ws<TPW, 4, true><<<blocks, MF_THREADS + 256, lds, s>>>(nullptr, Taps{}, nullptr, 0, nullptr, Geo{},
                                                                       nullptr, 0, nullptr, d, n);
    } else {
      return hipErrorInvalidValue;
    }
  });
  return hipGetLastError();
}

int view_reduce_grp_desc(const float* part, int nchunk, const int* vids, int nb, float* out, const float* yobs,
                         const float* bkg, int epi, const Geo& g, ViewReduceGrp* d) {
  if (nb <= 0) return 0;
  const int P = g.nx * g.nz, gx = (int)cdiv(P, 256);
  *d = ViewReduceGrp{GrpHdr{0, gx * nb, gx, 0}, part, nchunk, P, vids, out, yobs, bkg, epi};
  return 1;
}

hipError_t launch_view_reduce_grp(const ViewReduceGrp* d, int n, int blocks, hipStream_t s) {
  if (n <= 0 || blocks <= 0) return hipSuccess;
  k_view_reduce<true><<<blocks, 256, 0, s>>>(nullptr, 0, 0, nullptr, nullptr, nullptr, nullptr, 0, d, n);
  return hipGetLastError();
}

int psf_adj_grp_desc(const float* views, const Taps& t, const int* vids, int nb, float* work, const float* atten,
                     long atten_vstride, int nchunk, const Geo& g, int family, PsfAdjGrp* d, size_t* lds) {
  // launch_psf_adj on a z-fastest separable plan: its column-owned kernel
  if (nb <= 0) return 0;
  const int Yc = (g.ny + nchunk - 1) / nchunk;
  nchunk = (g.ny + Yc - 1) / Yc;
  if (family == PSF_TILE || !psf_adj_col_ok(g, t.ntx)) return 0;
  const int tpw = g.nz / 16;
  if (g.nz % 16 != 0 || tpw < 1 || tpw > 8) return 0;
  const long per = MfTaps(t.ntx).size;
  const long room = ((long)(kMaxLds / sizeof(float)) - (long)mf_col_lds_floats(g, t.ntx, 0)) / per;
  const int tbp = (int)std::max(1L, std::min<long>(std::min(8, Yc), room));
  *lds = sizeof(float) * mf_col_lds_floats(g, t.ntx, tbp);
  *d = PsfAdjGrp{GrpHdr{0, nchunk * nb, nchunk, 0}, views, t, vids, Yc, work, atten, atten_vstride, g, tbp};
  return 2 << 16 | tpw << 8 | (tpw >= 8 ? 8 : 16);
}

hipError_t launch_psf_adj_grp(int key, const PsfAdjGrp* d, int n, int blocks, size_t lds, hipStream_t s) {
  if (n <= 0 || blocks <= 0) return hipSuccess;
  if ((key >> 16) != 2) return hipErrorInvalidValue;
  SR_TPW_SWITCH((key >> 8) & 255, {
    constexpr int NWV = TPW >= 8 ? 8 : 16;
    if ((key & 255) != NWV) return hipErrorInvalidValue;
    allow_lds(k_psf_adj_col<TPW, NWV, true>, lds);
    k_psf_adj_col<TPW, NWV, true><<<blocks, 64 * NWV, lds, s>>>(nullptr, Taps{}, nullptr, 0, nullptr, nullptr, 0, Geo{},
                                                                0, d, n);
  });
  return hipGetLastError();
}

int ell_grp_desc(int mode, const float* work, float* part, const int2* ent, int K, const int* vids, int nb,
                 const Geo& g, const float* hvol, const float* atten, long atten_vstride, float* em_asum, int vs_req,
                 EllGrp* d) {
  // launch_ell of one subset, as launch_rot2_adj_z (mode 0) and
  // launch_rot2_adj_z_em_norm (mode 2) call it
  if (nb <= 0 || (mode != 0 && mode != 2 && mode != 3) || K < 4 || K > 8) return 0;
  const int q = (g.nz / 4) % 2 == 0 ? 2 : 1;
  const int pw = 64 / (g.nz / 4 / q);
  const int C = xcd_chunk(cdiv((long)g.nx * g.nx, pw));
  int vs = 2;
  if (vs_req == 2 || vs_req == 4 || vs_req == 8)
    vs = vs_req;
  else
    while (vs < 8 && 2 * vs <= nb && (long)kXcds * C * vs < 4096) vs *= 2;
  *d = EllGrp{GrpHdr{0, kXcds * C, 1, 0}, work, part, ent, vids, nb, C, g, hvol, atten, atten_vstride, em_asum,
              mode >= 2 ? 1 : 0};
  return 3 << 16 | K << 12 | vs << 8 | q << 4 | mode;
}

hipError_t launch_ell_grp(int key, const EllGrp* d, int n, int blocks, hipStream_t s) {
  if (n <= 0 || blocks <= 0) return hipSuccess;
  if ((key >> 16) != 3) return hipErrorInvalidValue;
  const int K = (key >> 12) & 15, vs = (key >> 8) & 15, q = (key >> 4) & 15, mode = key & 15;
#define SR_ELLG(KK, VV, QQ)                                                       \
  if (K == KK && vs == VV && q == QQ) {                                           \
    if (mode == 0)                                                                \
      k_rot2_adj_ell<KK, VV, QQ, 0, true><<<blocks, 64 * VV, 0, s>>>(                \
          nullptr, nullptr, nullptr, nullptr, 0, 0, 0, Geo{}, nullptr, nullptr, 0, nullptr, 0, d, n);    \
    else if (mode == 2)                                                           \
      k_rot2_adj_ell<KK, VV, QQ, 2, true><<<blocks, 64 * VV, 0, s>>>(                \
          nullptr, nullptr, nullptr, nullptr, 0, 0, 0, Geo{}, nullptr, nullptr, 0, nullptr, 0, d, n);    \
    else if (mode == 3)                                                           \
      k_rot2_adj_ell<KK, VV, QQ, 3, true><<<blocks, 64 * VV, 0, s>>>(                \
          nullptr, nullptr, nullptr, nullptr, 0, 0, 0, Geo{}, nullptr, nullptr, 0, nullptr, 0, d, n);    \
    else                                                                          \
      return hipErrorInvalidValue;                                                \
    return hipGetLastError();                                                     \
  }
#define SR_ELLG_K(KK) SR_ELLG(KK, 2, 1) SR_ELLG(KK, 4, 1) SR_ELLG(KK, 8, 1) SR_ELLG(KK, 2, 2) SR_ELLG(KK, 4, 2) SR_ELLG(KK, 8, 2)
  SR_ELLG_K(4) SR_ELLG_K(5) SR_ELLG_K(6) SR_ELLG_K(7) SR_ELLG_K(8)
#undef SR_ELLG_K
#undef SR_ELLG
  return hipErrorInvalidConfiguration;
}

hipError_t launch_rot2_adj_z(const float* work, float* part, const int2* ent, int K, const int* vids, int nb,
                             const Geo& g, hipStream_t s, const float* atten, long atten_vstride,
                             const float* em_asum, bool zero_inf, int vs) {
  // (with atten: work is H, the normaliser form)
  return atten ? launch_ell(1, nullptr, part, ent, K, vids, nb, g, s, work, atten, atten_vstride, nullptr, zero_inf,
                            vs)
               : launch_ell(0, work, part, ent, K, vids, nb, g, s, nullptr, nullptr, 0, const_cast<float*>(em_asum),
                            zero_inf, vs);
}

hipError_t launch_rot2_adj_z_norms(const float* hvol, float* asum, const int2* ent, int K, const int* vids, int nb,
                                   int nb_rem, int nsubsets, const Geo& g, hipStream_t s, const float* atten,
                                   long atten_vstride, int vs) {
  return launch_ell(1, nullptr, asum, ent, K, vids, nb, g, s, hvol, atten, atten_vstride, nullptr, true, vs, nsubsets,
                    nb_rem);
}

hipError_t launch_rot2_adj_z_em_norm(const float* work, float* x, const int2* ent, int K, const int* vids, int nb,
                                     const Geo& g, hipStream_t s, const float* hvol, const float* atten,
                                     long atten_vstride, float* asum, int vs) {
  // (hvol null: atten holds the products H .* atten[v])
  return launch_ell(hvol ? 2 : 3, work, x, ent, K, vids, nb, g, s, hvol, atten, atten_vstride, asum, true, vs);
}

// out[v][i] = h[i] * atten[v][i] over nview volumes of V floats (grid-stride)
__global__ void __launch_bounds__(256) k_mul_volumes(const float* __restrict__ h, const float* __restrict__ atten,
                                                     float* __restrict__ out, long V, int nview) {
  const long n4 = V >> 2, total = n4 * nview;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < total; i += (long)gridDim.x * 256) {
    const long v = i / n4, k = i - v * n4;
    const float4 a = reinterpret_cast<const float4*>(atten)[i];
    const float4 b = reinterpret_cast<const float4*>(h)[k];
    reinterpret_cast<float4*>(out)[i] = make_float4(b.x * a.x, b.y * a.y, b.z * a.z, b.w * a.w);
  }
}

hipError_t launch_mul_volumes(const float* h, const float* atten, float* out, long V, int nview, hipStream_t s) {
  if (V <= 0 || nview <= 0) return hipSuccess;
  if (V % 4 != 0) return hipErrorInvalidValue;
  const long total = (V >> 2) * nview;
  k_mul_volumes<<<(unsigned)std::min<long>(cdiv(total, 256), 8192), 256, 0, s>>>(h, atten, out, V, nview);
  return hipGetLastError();
}

hipError_t launch_img_reduce_z(const float* part, int ngroups, float* out, const float* asum, int mode, const Geo& g,
                               hipStream_t s, float* zout) {
  k_img_reduce_z<<<dim3(cdiv(g.nx, 32), g.ny, cdiv(g.nz, 32)), 256, 0, s>>>(part, ngroups, out, asum, mode, g, zout);
  return hipGetLastError();
}

hipError_t launch_resid_loss(const float* p, const float* y, float* r, long n, double* partial, double* out,
                             hipStream_t s) {
  const unsigned nblk = cdiv(n, 256);
  if (n <= 0) return hipSuccess;
  k_resid_loss<<<nblk, 256, 0, s>>>(p, y, r, n, partial);
  k_sum_partials<<<1, 256, 0, s>>>(partial, (int)nblk, out);
  return hipGetLastError();
}

hipError_t launch_mu_grad_scan_rot(float* g, float* ga, const float* img_t, const float* atten, long atten_vstride,
                                   const RotView* rv, const int* vids, int nb, const Geo& geo, float dy,
                                   hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  const dim3 grid(cdiv((long)geo.nx * geo.nz, 256), nb);
  if (ga)
    k_mu_grad_scan_rot<true><<<grid, 256, 0, s>>>(g, ga, img_t, atten, atten_vstride, rv, vids, geo, dy);
  else
    k_mu_grad_scan_rot<false><<<grid, 256, 0, s>>>(g, nullptr, img_t, atten, atten_vstride, rv, vids, geo, dy);
  return hipGetLastError();
}

hipError_t launch_fill(float* p, float v, long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_fill<<<cdiv(n, 256), 256, 0, s>>>(p, v, n);
  return hipGetLastError();
}

// Rows of nx floats permuted between [z][y][x] (mode 0 source) and
// [y][z][x] (mode 0 target); mode 1 the other way.  grid (ny, nz), block 256.
__global__ void __launch_bounds__(256) k_permute_rows(const float* __restrict__ src, float* __restrict__ dst, int nx,
                                                      int ny, int nz, int mode) {
  const int y = blockIdx.x, z = blockIdx.y;
  const long zy = ((long)z * ny + y) * nx, yz = ((long)y * nz + z) * nx;
  const float* s = src + (mode == 0 ? zy : yz);
  float* d = dst + (mode == 0 ? yz : zy);
  for (int x = threadIdx.x; x < nx; x += 256) d[x] = s[x];
}

hipError_t launch_permute_rows(const float* src, float* dst, int nx, int ny, int nz, int mode, hipStream_t s) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return hipSuccess;
  k_permute_rows<<<dim3(ny, nz), 256, 0, s>>>(src, dst, nx, ny, nz, mode);
  return hipGetLastError();
}

hipError_t launch_copy16(void* dst, const void* src, long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_copy16<<<std::min(cdiv(n, 256), 8192u), 256, 0, s>>>(static_cast<uint4*>(dst), static_cast<const uint4*>(src), n);
  return hipGetLastError();
}

hipError_t launch_flag_nonpositive(const float* p, long n, int* flag, hipStream_t s) {
  k_fill<<<1, 256, 0, s>>>(reinterpret_cast<float*>(flag), 0.f, 1);  // bit pattern of int 0
  if (n > 0) k_flag_nonpositive<<<std::min(cdiv(n, 256), 1024u), 256, 0, s>>>(p, n, flag);
  return hipGetLastError();
}

hipError_t launch_zero_to_inf(float* p, long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_zero_to_inf<<<cdiv(n, 256), 256, 0, s>>>(p, n);
  return hipGetLastError();
}

}  // namespace sr
