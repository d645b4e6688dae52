// spect_oracle.cpp -- CPU restatement of SPECTrecon.jl's projector path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
// library: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it.  The product never links, calls or falls back to it.
//
// Every array is column-major (Julia order): image/mumap [x,y,z] with x
// fastest, views [x,z,view], psfs [px,pz,y,view].  Each function names the
// reference file:line it restates.  Two instantiations exist: T=float is the
// reference's own Float32 arithmetic order (the "FP32 oracle"); T=double is
// the same algorithm in FP64 (the "FP64 oracle").
//
// Parity pins (what fixes each function independently of this restatement)
// are listed in tests/test_oracle.py.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;  // Float64(pi)

inline long jmod(long x, long y) {  // Julia mod for Int (result in [0,y))
  long r = x % y;
  return r < 0 ? r + y : r;
}

// src/plan-rotate.jl:56
inline int rot_padsize(int nx) {
  return (int)std::ceil(1.0 + nx * std::sqrt(2.0) / 2.0 - nx / 2.0);
}

// src/helper.jl:8-12
inline long power2(long x) { return 1L << (long)std::ceil(std::log2((double)x)); }
inline int padup(int n, int p) { return (int)std::ceil((power2(n + p - 1) - n) / 2.0); }
inline int paddown(int n, int p) { return (int)std::floor((power2(n + p - 1) - n) / 2.0); }

// ---------------------------------------------------------------------------
// Rotation, 3-pass 1-D linear interpolation (":one")
// ---------------------------------------------------------------------------

// LinearInterpolators.SparseInterpolator with a LinearSpline kernel: 2
// coefficients per output row, C (values) and J (1-based input indices).
template <class T>
struct SparseLin {
  long ncoeff;
  std::vector<T> C;
  std::vector<long> J;
  explicit SparseLin(int n) : ncoeff(2L * n), C(2L * n), J(2L * n) {}
};

// src/rotatez.jl:21-51 (linearinterp!)
template <class T>
void linearinterp(SparseLin<T>& A, double s, double e, long len) {
  const double dec = std::ceil(s) - s;
  const long nc = A.ncoeff;
  for (long i = 1; i <= nc; ++i) A.C[i - 1] = (i % 2 == 1) ? T(dec) : T(1.0 - dec);
  if (e <= (double)len) {
    A.J[nc - 1] = (long)std::ceil(e);
    for (long i = nc - 1; i >= 1; --i)
      A.J[i - 1] = std::max(1L, A.J[nc - 1] - (long)std::ceil((nc - i) / 2.0));
  } else {
    A.J[0] = (long)std::floor(s);
    for (long i = 2; i <= nc; ++i)
      A.J[i - 1] = std::min(len, A.J[0] + (long)std::ceil((i - 1) / 2.0));
  }
}

// mul!(y, A, x) for the sparse interpolator: y[i] = C1*x[J1] + C2*x[J2],
// accumulated from zero in coefficient order.  x, y strided views.
template <class T>
void sparse_apply(const SparseLin<T>& A, T* y, long ys, const T* x, long xs) {
  const long n = A.ncoeff / 2;
  for (long i = 0; i < n; ++i) {
    T sum = T(0);
    sum += A.C[2 * i] * x[(A.J[2 * i] - 1) * xs];
    sum += A.C[2 * i + 1] * x[(A.J[2 * i + 1] - 1) * xs];
    y[i * ys] = sum;
  }
}

// mul!(y, A', x): scatter, rows in ascending order.
template <class T>
void sparse_apply_adj(const SparseLin<T>& A, T* y, long ys, const T* x, long xs) {
  const long n = A.ncoeff / 2;
  for (long j = 0; j < n; ++j) y[j * ys] = T(0);
  for (long i = 0; i < n; ++i) {
    const T a = x[i * xs];
    if (a == T(0)) continue;
    y[(A.J[2 * i] - 1) * ys] += A.C[2 * i] * a;
    y[(A.J[2 * i + 1] - 1) * ys] += A.C[2 * i + 1] * a;
  }
}

// src/rotatez.jl:57-147 (rotate_x!, rotate_x_adj!, rotate_y!, rotate_y_adj!)
// Matrices are Np x Np column-major; "x" interpolates along dim 1 for each
// column i, "y" along dim 2 for each row i.
template <class T>
void rotate_pass(T* out, const T* img, long Np, double coef, SparseLin<T>& A, bool along_x, bool adj) {
  const double c = (Np + 1) / 2.0;
  for (long i = 1; i <= Np; ++i) {
    const double s = (i - c) * coef + 1;
    const double e = (i - c) * coef + Np;
    linearinterp(A, s, e, Np);
    if (along_x) {  // column i: contiguous
      T* o = out + (i - 1) * Np;
      const T* x = img + (i - 1) * Np;
      if (adj) sparse_apply_adj(A, o, 1, x, 1);
      else sparse_apply(A, o, 1, x, 1);
    } else {  // row i: stride Np
      T* o = out + (i - 1);
      const T* x = img + (i - 1);
      if (adj) sparse_apply_adj(A, o, Np, x, Np);
      else sparse_apply(A, o, Np, x, Np);
    }
  }
}

// src/rotatez.jl:155-230 (rotl90!, rotr90!, rot180!, rot_f90!, rot_f90_adj!)
// 0-based: A(i,j) at A[i + Np*j].
template <class T>
void rotl90(T* B, const T* A, long Np) {  // B[Np-1-j, i] = A[i, j]
  for (long j = 0; j < Np; ++j)
    for (long i = 0; i < Np; ++i) B[(Np - 1 - j) + Np * i] = A[i + Np * j];
}
template <class T>
void rotr90(T* B, const T* A, long Np) {  // B[j, Np-1-i] = A[i, j]
  for (long j = 0; j < Np; ++j)
    for (long i = 0; i < Np; ++i) B[j + Np * (Np - 1 - i)] = A[i + Np * j];
}
template <class T>
void rot180(T* B, const T* A, long Np) {
  for (long j = 0; j < Np; ++j)
    for (long i = 0; i < Np; ++i) B[(Np - 1 - i) + Np * (Np - 1 - j)] = A[i + Np * j];
}
// returns false for an invalid quadrant (the reference throws "invalid m!")
template <class T>
bool rot_f90(T* out, const T* img, long Np, long m, bool adj) {
  if (m == 0) std::memcpy(out, img, sizeof(T) * Np * Np);
  else if (m == 2) rot180(out, img, Np);
  else if ((m == 1 && !adj) || (m == 3 && adj)) rotl90(out, img, Np);
  else if ((m == 3 && !adj) || (m == 1 && adj)) rotr90(out, img, Np);
  else return false;
  return true;
}

// src/helper.jl:20-34 (padzero!) for the square rotation pad
template <class T>
void padzero_sq(T* out, const T* img, long N, long pad) {
  const long Np = N + 2 * pad;
  std::fill(out, out + Np * Np, T(0));
  for (long j = 0; j < N; ++j)
    for (long i = 0; i < N; ++i) out[(i + pad) + Np * (j + pad)] = img[i + N * j];
}
template <class T>
void crop_sq(T* out, const T* work, long N, long pad) {
  const long Np = N + 2 * pad;
  for (long j = 0; j < N; ++j)
    for (long i = 0; i < N; ++i) out[i + N * j] = work[(i + pad) + Np * (j + pad)];
}

// isapprox(x::Float32-or-64, y::Float64) with the default rtol of the wider
// of the two eps: src/rotatez.jl:258 compares a Float32 angle to Float64.
inline bool jl_isapprox(double x, double y, double rtol) {
  return x == y || std::fabs(x - y) <= rtol * std::max(std::fabs(x), std::fabs(y));
}

template <class T>
struct RotWork {
  long N, pad, Np;
  std::vector<T> w1, w2;
  SparseLin<T> A;
  explicit RotWork(long n)
      : N(n), pad(rot_padsize((int)n)), Np(n + 2 * rot_padsize((int)n)),
        w1((size_t)Np * Np), w2((size_t)Np * Np), A((int)Np) {}
};

// src/rotatez.jl:240-322 (imrotate!/imrotate_adj! for RotateMode{:one}).
// theta_rtol is sqrt(eps(T)) of the plan type (Float32 -> 3.4527e-4).
template <class T>
void imrotate_one(T* out, const T* img, double theta, RotWork<T>& W, bool adj) {
  const long N = W.N, Np = W.Np, pad = W.pad;
  const double rtol = std::sqrt(std::numeric_limits<T>::epsilon());
  if (std::fmod(theta, 2 * kPi) == 0.0) {  // mod(θ, 2π) ≈ 0 (atol 0)
    std::memcpy(out, img, sizeof(T) * N * N);
    return;
  }
  const long m = jmod((long)std::floor(0.5 + theta / (kPi / 2)), 4);
  if (jl_isapprox(theta, m * (kPi / 2), rtol)) {
    padzero_sq(W.w2.data(), img, N, pad);
    rot_f90(W.w1.data(), W.w2.data(), Np, m, adj);
  } else {
    const double mod_theta = theta - m * (kPi / 2);
    const double tan_mod = std::tan(mod_theta / 2);
    const double sin_mod = -std::sin(mod_theta);
    padzero_sq(W.w1.data(), img, N, pad);
    if (!adj) {
      rot_f90(W.w2.data(), W.w1.data(), Np, m, false);
      rotate_pass(W.w1.data(), W.w2.data(), Np, tan_mod, W.A, true, false);
      rotate_pass(W.w2.data(), W.w1.data(), Np, sin_mod, W.A, false, false);
      rotate_pass(W.w1.data(), W.w2.data(), Np, tan_mod, W.A, true, false);
    } else {
      rotate_pass(W.w2.data(), W.w1.data(), Np, tan_mod, W.A, true, true);
      rotate_pass(W.w1.data(), W.w2.data(), Np, sin_mod, W.A, false, true);
      rotate_pass(W.w2.data(), W.w1.data(), Np, tan_mod, W.A, true, true);
      rot_f90(W.w1.data(), W.w2.data(), Np, m, true);
    }
  }
  crop_sq(out, W.w1.data(), N, pad);
}

// ---------------------------------------------------------------------------
// Rotation, single-pass bilinear (":two")
// ---------------------------------------------------------------------------

// The affine map c + rotate(2π-θ, I - c) of src/rotatez.jl:348-353, with
// coefficients evaluated in Float64 and stored in T (AffineTransform2D{T}),
// then applied in T as ((xx*i1 + xy*i2) + x0).
template <class T>
struct Affine {
  T xx, xy, x0, yx, yy, y0;
};
template <class T>
Affine<T> rot_affine(double theta, long Np) {
  const double c = (1.0 + Np) / 2.0;
  const double phi = 2 * kPi - theta;
  const double cs = std::cos(phi), sn = std::sin(phi);
  const T cT = T(c);  // I - c: translation (-c, -c) stored in T
  Affine<T> A;
  A.xx = T(cs * 1.0 - sn * 0.0);
  A.xy = T(cs * 0.0 - sn * 1.0);
  const T xr = T(cs * (double)(-cT) - sn * (double)(-cT));
  A.yx = T(cs * 0.0 + sn * 1.0);
  A.yy = T(cs * 1.0 + sn * 0.0);
  const T yr = T(cs * (double)(-cT) + sn * (double)(-cT));
  A.x0 = T((double)xr + c);
  A.y0 = T((double)yr + c);
  return A;
}

// LinearSpline coefficients with Flat boundaries: j1=floor(x), j2=j1+1,
// clamped to [1,len]; weights (1-t, t).
template <class T>
inline void lin_coefs(T x, long len, long& j1, long& j2, T& w1, T& w2) {
  const T fl = std::floor(x);
  const T t = x - fl;
  w1 = T(1) - t;
  w2 = t;
  j1 = (long)fl;
  j2 = j1 + 1;
  j1 = std::min(std::max(j1, 1L), len);
  j2 = std::min(std::max(j2, 1L), len);
}

// src/rotatez.jl:331-362 and :390-416 (imrotate!/imrotate_adj! :two).
template <class T>
void imrotate_two(T* out, const T* img, double theta, RotWork<T>& W, bool adj) {
  const long N = W.N, Np = W.Np, pad = W.pad;
  if (!adj && std::fmod(theta, 2 * kPi) == 0.0) {  // forward-only shortcut
    std::memcpy(out, img, sizeof(T) * N * N);
    return;
  }
  const Affine<T> R = rot_affine<T>(theta, Np);
  padzero_sq(W.w1.data(), img, N, pad);
  T* in = W.w1.data();
  T* o = W.w2.data();
  if (adj) std::fill(o, o + Np * Np, T(0));
  for (long i2 = 1; i2 <= Np; ++i2) {
    for (long i1 = 1; i1 <= Np; ++i1) {
      const T f1 = T(i1), f2 = T(i2);
      const T x1 = (R.xx * f1 + R.xy * f2) + R.x0;
      const T x2 = (R.yx * f1 + R.yy * f2) + R.y0;
      long j11, j12, j21, j22;
      T w11, w12, w21, w22;
      lin_coefs(x1, Np, j11, j12, w11, w12);
      lin_coefs(x2, Np, j21, j22, w21, w22);
      const long io = (i1 - 1) + Np * (i2 - 1);
      if (!adj) {
        const T a = in[(j11 - 1) + Np * (j21 - 1)] * w11 + in[(j12 - 1) + Np * (j21 - 1)] * w12;
        const T b = in[(j11 - 1) + Np * (j22 - 1)] * w11 + in[(j12 - 1) + Np * (j22 - 1)] * w12;
        o[io] = a * w21 + b * w22;
      } else {
        const T v = in[io];
        if (v == T(0)) continue;
        o[(j11 - 1) + Np * (j21 - 1)] += (w11 * w21) * v;
        o[(j12 - 1) + Np * (j21 - 1)] += (w12 * w21) * v;
        o[(j11 - 1) + Np * (j22 - 1)] += (w11 * w22) * v;
        o[(j12 - 1) + Np * (j22 - 1)] += (w12 * w22) * v;
      }
    }
  }
  crop_sq(out, o, N, pad);
}

template <class T>
void imrotate(T* out, const T* img, double theta, RotWork<T>& W, int method, bool adj) {
  if (method == 1) imrotate_one(out, img, theta, W, adj);
  else imrotate_two(out, img, theta, W, adj);
}

// ---------------------------------------------------------------------------
// PSF convolution: literal FFT path and closed-form direct path
// ---------------------------------------------------------------------------

// Iterative radix-2 complex FFT (lengths are powers of two, src/helper.jl:8).
template <class T>
struct FFT1 {
  long n = 0;
  int logn = 0;
  std::vector<std::complex<T>> tw;  // e^{-2πik/n}
  std::vector<long> rev;
  void init(long len) {
    n = len;
    logn = 0;
    while ((1L << logn) < n) ++logn;
    tw.resize(n / 2 + 1);
    for (long k = 0; k < n / 2; ++k)
      tw[k] = std::complex<T>((T)std::cos(2 * kPi * k / n), (T)-std::sin(2 * kPi * k / n));
    rev.resize(n);
    for (long i = 0; i < n; ++i) {
      long r = 0;
      for (int b = 0; b < logn; ++b) r |= ((i >> b) & 1) << (logn - 1 - b);
      rev[i] = r;
    }
  }
  // inverse: conjugate twiddles, no scaling here
  void run(std::complex<T>* a, bool inverse) const {
    for (long i = 0; i < n; ++i)
      if (i < rev[i]) std::swap(a[i], a[rev[i]]);
    for (long len = 2; len <= n; len <<= 1) {
      const long step = n / len;
      for (long i = 0; i < n; i += len)
        for (long k = 0; k < len / 2; ++k) {
          std::complex<T> w = tw[k * step];
          if (inverse) w = std::conj(w);
          const std::complex<T> u = a[i + k], v = a[i + k + len / 2] * w;
          a[i + k] = u + v;
          a[i + k + len / 2] = u - v;
        }
    }
  }
};

template <class T>
struct PsfWork {
  int nx, nz, px, pz;
  int pu, pd, pl, pr;  // (padup, paddown, padleft, padright), src/plan-psf.jl:49-53
  long L1, L2;
  std::vector<std::complex<T>> img_c, ker_c, col;
  std::vector<T> workmat, wvx, wvz;
  FFT1<T> f1, f2;
  PsfWork(int nx_, int nz_, int px_, int pz_) : nx(nx_), nz(nz_), px(px_), pz(pz_) {
    pu = padup(nx, px);
    pd = paddown(nx, px);
    pl = padup(nz, pz);
    pr = paddown(nz, pz);
    L1 = nx + pu + pd;
    L2 = nz + pl + pr;
    img_c.resize(L1 * L2);
    ker_c.resize(L1 * L2);
    col.resize(std::max(L1, L2));
    workmat.resize(L1 * L2);
    wvx.resize(L1);
    wvz.resize(L2);
    f1.init(L1);
    f2.init(L2);
  }
  void fft2(std::complex<T>* a, bool inverse) {
    for (long j = 0; j < L2; ++j) f1.run(a + j * L1, inverse);
    if (L2 > 1) {
      for (long i = 0; i < L1; ++i) {
        for (long j = 0; j < L2; ++j) col[j] = a[i + j * L1];
        f2.run(col.data(), inverse);
        for (long j = 0; j < L2; ++j) a[i + j * L1] = col[j];
      }
    }
    if (inverse) {
      const T s = T(1) / T(L1 * L2);
      for (long k = 0; k < L1 * L2; ++k) a[k] *= s;
    }
  }
  // src/fft_convolve.jl:16-25 (imfilterz!), with fftshift2! (src/helper.jl:111-142)
  void imfilterz() {
    fft2(img_c.data(), false);
    fft2(ker_c.data(), false);
    for (long k = 0; k < L1 * L2; ++k) img_c[k] *= ker_c[k];
    fft2(img_c.data(), true);
    const long m = L1 / 2, n = L2 / 2;  // circshift by (L1/2, L2/2)
    for (long j = 0; j < L2; ++j)
      for (long i = 0; i < L1; ++i) {
        const long si = (i + m) % L1, sj = (L2 == 1) ? 0 : (j + n) % L2;
        workmat[i + j * L1] = std::real(img_c[si + sj * L1]);
      }
  }
  // src/helper.jl:66-80 (pad2sizezero!) for the kernel
  void pad_kernel(const T* ker) {
    std::fill(ker_c.begin(), ker_c.end(), std::complex<T>(0));
    const long d1 = (long)std::ceil((L1 - px) / 2.0), d2 = (long)std::ceil((L2 - pz) / 2.0);
    for (long j = 0; j < pz; ++j)
      for (long i = 0; i < px; ++i) ker_c[(i + d1) + (j + d2) * L1] = ker[i + px * j];
  }
  // src/fft_convolve.jl:33-55 (fft_conv!); img/out have column stride ld
  void conv(T* out, long ldo, const T* img, long ldi, const T* ker) {
    for (long j = 0; j < L2; ++j)  // padrepl!, src/helper.jl:42-55
      for (long i = 0; i < L1; ++i) {
        const long si = std::min(std::max(i - pu, 0L), (long)nx - 1);
        const long sj = std::min(std::max(j - pl, 0L), (long)nz - 1);
        img_c[i + j * L1] = std::complex<T>(img[si + sj * ldi], T(0));
      }
    pad_kernel(ker);
    imfilterz();
    for (long j = 0; j < nz; ++j)
      for (long i = 0; i < nx; ++i) out[i + j * ldo] = workmat[(i + pu) + (j + pl) * L1];
  }
  // src/fft_convolve.jl:82-130 (fft_conv_adj!)
  void conv_adj(T* out, long ldo, const T* img, long ldi, const T* ker) {
    std::fill(img_c.begin(), img_c.end(), std::complex<T>(0));
    for (long j = 0; j < nz; ++j)
      for (long i = 0; i < nx; ++i) img_c[(i + pu) + (j + pl) * L1] = std::complex<T>(img[i + j * ldi], T(0));
    pad_kernel(ker);
    imfilterz();
    T* W = workmat.data();
    auto fold_rows = [&](long r0, long r1, long dst) {  // rows r0..r1 (0-based) onto row dst
      std::fill(wvz.begin(), wvz.end(), T(0));
      for (long i = r0; i <= r1; ++i)
        for (long m = 0; m < L2; ++m) wvz[m] += W[i + m * L1];
      for (long m = 0; m < L2; ++m) W[dst + m * L1] += wvz[m];
    };
    auto fold_cols = [&](long c0, long c1, long dst) {
      std::fill(wvx.begin(), wvx.end(), T(0));
      for (long j = c0; j <= c1; ++j)
        for (long n = 0; n < L1; ++n) wvx[n] += W[n + j * L1];
      for (long n = 0; n < L1; ++n) W[n + dst * L1] += wvx[n];
    };
    fold_rows(0, pu - 1, pu);
    fold_rows(pu + nx, L1 - 1, pu + nx - 1);
    fold_cols(0, pl - 1, pl);
    fold_cols(pl + nz, L2 - 1, pl + nz - 1);
    for (long j = 0; j < nz; ++j)
      for (long i = 0; i < nx; ++i) out[i + j * ldo] = W[(i + pu) + (j + pl) * L1];
  }
};

// Closed form of fft_conv! (no wrap because L >= n+p-1):
//   out[x,z] = Σ_{a,b} h[a,b] img[clamp(x+hx-a), clamp(z+hz-b)]
// and its exact transpose.  Accumulation in T, kernel-index order.
template <class T>
void conv_direct(T* out, long ldo, const T* img, long ldi, int nx, int nz, const T* ker, int px, int pz, bool adj) {
  const int hx = (px - 1) / 2, hz = (pz - 1) / 2;
  auto cl = [](long v, long n) { return std::min(std::max(v, 0L), n - 1); };
  if (!adj) {
    for (long z = 0; z < nz; ++z)
      for (long x = 0; x < nx; ++x) {
        T s = T(0);
        for (long b = 0; b < pz; ++b)
          for (long a = 0; a < px; ++a) s += ker[a + px * b] * img[cl(x + hx - a, nx) + ldi * cl(z + hz - b, nz)];
        out[x + ldo * z] = s;
      }
  } else {
    for (long z = 0; z < nz; ++z)
      for (long x = 0; x < nx; ++x) out[x + ldo * z] = T(0);
    for (long z = 0; z < nz; ++z)
      for (long x = 0; x < nx; ++x) {
        const T v = img[x + ldi * z];
        for (long b = 0; b < pz; ++b)
          for (long a = 0; a < px; ++a) out[cl(x + hx - a, nx) + ldo * cl(z + hz - b, nz)] += ker[a + px * b] * v;
      }
  }
}

// ---------------------------------------------------------------------------
// Projector / back-projector
// ---------------------------------------------------------------------------

template <class T>
struct Problem {
  int nx, ny, nz, px, pz, nview, method, conv_kind;  // conv_kind 0: literal FFT, 1: direct
  T dy;
  const T* mumap;
  const T* psfs;
  const double* angles;  // viewangle[k] promoted to double
};

template <class T>
struct ViewWork {
  RotWork<T> rot;
  PsfWork<T> psf;
  std::vector<T> imgr, mumapr, add_img, e;
  ViewWork(const Problem<T>& P)
      : rot(P.nx), psf(P.nx, P.nz, P.px, P.pz), imgr((size_t)P.nx * P.ny * P.nz),
        mumapr((size_t)P.nx * P.ny * P.nz), add_img((size_t)P.nx * P.ny * P.nz), e((size_t)P.nx * P.nz) {}
};

// attenuation factor for plane y (0-based) into e[x,z]:
// src/project.jl:103-113 (scale3dj! by -0.5, plus3dj! for j=1..y, *(-dy), exp)
template <class T>
void atten_plane(T* e, const T* mumapr, int nx, int ny, int nz, int y, T dy) {
  const long sxy = (long)nx * ny;
  for (long z = 0; z < nz; ++z)
    for (long x = 0; x < nx; ++x) e[x + nx * z] = T(-0.5 * (double)mumapr[x + nx * y + sxy * z]);
  for (int j = 0; j <= y; ++j)
    for (long z = 0; z < nz; ++z)
      for (long x = 0; x < nx; ++x) e[x + nx * z] += mumapr[x + nx * j + sxy * z];
  for (long k = 0; k < (long)nx * nz; ++k) e[k] = e[k] * (-dy);
  for (long k = 0; k < (long)nx * nz; ++k) e[k] = std::exp(e[k]);
}

template <class T>
void conv_plane(const Problem<T>& P, ViewWork<T>& W, T* out, long ldo, const T* in, long ldi, const T* ker, bool adj) {
  if (P.conv_kind == 0) {
    if (adj) W.psf.conv_adj(out, ldo, in, ldi, ker);
    else W.psf.conv(out, ldo, in, ldi, ker);
  } else {
    conv_direct(out, ldo, in, ldi, P.nx, P.nz, ker, P.px, P.pz, adj);
  }
}

// src/project.jl:77-135 (project! for one view, buffer_id form)
template <class T>
void project_view(const Problem<T>& P, ViewWork<T>& W, T* view, const T* image, int viewidx) {
  const int nx = P.nx, ny = P.ny, nz = P.nz;
  const long sxy = (long)nx * ny;
  const double th = P.angles[viewidx];
  for (int z = 0; z < nz; ++z) {
    imrotate(W.imgr.data() + sxy * z, image + sxy * z, th, W.rot, P.method, false);
    imrotate(W.mumapr.data() + sxy * z, P.mumap + sxy * z, th, W.rot, P.method, false);
  }
  const T* psf_v = P.psfs + (long)P.px * P.pz * ny * viewidx;
  for (int y = 0; y < ny; ++y) {
    atten_plane(W.e.data(), W.mumapr.data(), nx, ny, nz, y, P.dy);
    for (long z = 0; z < nz; ++z)  // mul3dj!
      for (long x = 0; x < nx; ++x) W.imgr[x + nx * y + sxy * z] *= W.e[x + nx * z];
    // strided plane view [:, y, :]: column stride nx*ny
    conv_plane(P, W, W.add_img.data() + (long)nx * y, sxy, W.imgr.data() + (long)nx * y, sxy,
               psf_v + (long)P.px * P.pz * y, false);
  }
  for (long z = 0; z < nz; ++z)  // copy3dj! + plus3dj!
    for (long x = 0; x < nx; ++x) view[x + nx * z] = W.add_img[x + sxy * z];
  for (int y = 1; y < ny; ++y)
    for (long z = 0; z < nz; ++z)
      for (long x = 0; x < nx; ++x) view[x + nx * z] += W.add_img[x + nx * y + sxy * z];
}

// src/backproject.jl:69-124 (backproject! for one view, accumulating)
template <class T>
void backproject_view(const Problem<T>& P, ViewWork<T>& W, T* image, const T* view, int viewidx) {
  const int nx = P.nx, ny = P.ny, nz = P.nz;
  const long sxy = (long)nx * ny;
  const double th = P.angles[viewidx];
  for (int z = 0; z < nz; ++z)
    imrotate(W.mumapr.data() + sxy * z, P.mumap + sxy * z, th, W.rot, P.method, false);
  const T* psf_v = P.psfs + (long)P.px * P.pz * ny * viewidx;
  for (int y = 0; y < ny; ++y) {
    atten_plane(W.e.data(), W.mumapr.data(), nx, ny, nz, y, P.dy);
    conv_plane(P, W, W.imgr.data() + (long)nx * y, sxy, view, nx, psf_v + (long)P.px * P.pz * y, true);
    for (long z = 0; z < nz; ++z)
      for (long x = 0; x < nx; ++x) W.imgr[x + nx * y + sxy * z] *= W.e[x + nx * z];
  }
  for (int z = 0; z < nz; ++z)
    imrotate(W.imgr.data() + sxy * z, W.imgr.data() + sxy * z, th, W.rot, P.method, true);
  for (long k = 0; k < sxy * nz; ++k) image[k] += W.imgr[k];
}

// Gradient w.r.t. the mu-map of L = sum_v ||A_v(x, mu) - y_v||^2 through one
// view, accumulated into gmu (no reference counterpart: the reference's
// forward, src/project.jl:77-135, differentiated by hand).  resid = dL/dA_v =
// 2 (A_v x - y_v).  With imgr = R_v x, mur = R_v mu and the attenuation
// a_y = exp(e_y), e_y = -dy (sum_{j<y} mur_j + mur_y / 2):
//   G_y = conv_y^T(resid)                    (dL / d(imgr_y .* a_y))
//   E_y = G_y .* imgr_y .* a_y               (dL / de_y)
//   dL/dmur_j = -dy (sum_{y>j} E_y + E_j / 2)
//   dL/dmu += R_v^T(dL/dmur)
template <class T>
void grad_mu_view(const Problem<T>& P, ViewWork<T>& W, T* gmu, const T* image, const T* resid, int viewidx) {
  const int nx = P.nx, ny = P.ny, nz = P.nz;
  const long sxy = (long)nx * ny;
  const double th = P.angles[viewidx];
  for (int z = 0; z < nz; ++z) {
    imrotate(W.imgr.data() + sxy * z, image + sxy * z, th, W.rot, P.method, false);
    imrotate(W.mumapr.data() + sxy * z, P.mumap + sxy * z, th, W.rot, P.method, false);
  }
  const T* psf_v = P.psfs + (long)P.px * P.pz * ny * viewidx;
  // E into add_img
  for (int y = 0; y < ny; ++y) {
    atten_plane(W.e.data(), W.mumapr.data(), nx, ny, nz, y, P.dy);
    conv_plane(P, W, W.add_img.data() + (long)nx * y, sxy, resid, nx, psf_v + (long)P.px * P.pz * y, true);
    for (long z = 0; z < nz; ++z)
      for (long x = 0; x < nx; ++x) {
        const long k = x + nx * y + sxy * z;
        W.add_img[k] = W.add_img[k] * W.imgr[k] * W.e[x + nx * z];
      }
  }
  // reverse scan along y, into imgr
  for (long z = 0; z < nz; ++z)
    for (long x = 0; x < nx; ++x) {
      T acc = T(0);
      for (int y = ny - 1; y >= 0; --y) {
        const long k = x + nx * y + sxy * z;
        const T ev = W.add_img[k];
        W.imgr[k] = -P.dy * (acc + T(0.5) * ev);
        acc += ev;
      }
    }
  for (int z = 0; z < nz; ++z)
    imrotate(W.imgr.data() + sxy * z, W.imgr.data() + sxy * z, th, W.rot, P.method, true);
  for (long k = 0; k < sxy * nz; ++k) gmu[k] += W.imgr[k];
}

// spawner (src/spawn.jl:10-17): ≤ nthread contiguous chunks of size
// ceil(ntask/nthread); chunk c uses buffer c.
template <class F>
void spawner(int nthread, int ntask, F&& f) {
  if (ntask <= 0) return;
  const int chunk = (ntask + nthread - 1) / nthread;
  const int nchunk = (ntask + chunk - 1) / chunk;
  if (nchunk == 1) {
    for (int t = 0; t < ntask; ++t) f(0, t);
    return;
  }
  std::vector<std::thread> th;
  for (int c = 0; c < nchunk; ++c)
    th.emplace_back([&, c] {
      for (int t = c * chunk; t < std::min(ntask, (c + 1) * chunk); ++t) f(c, t);
    });
  for (auto& t : th) t.join();
}

// src/project.jl:143-163 (:fast): views[:,:,ii] for index[ii]
template <class T>
void project(const Problem<T>& P, T* views, const T* image, const int* index, int nindex, int nthread) {
  std::vector<ViewWork<T>*> W;
  const int nb = std::max(1, std::min(nthread, nindex));
  for (int b = 0; b < nb; ++b) W.push_back(new ViewWork<T>(P));
  spawner(nb, nindex, [&](int b, int ii) {
    project_view(P, *W[b], views + (long)P.nx * P.nz * ii, image, index[ii]);
  });
  for (auto* w : W) delete w;
}

// src/backproject.jl:132-162 (:fast): zero image, per-buffer accumulation,
// then image += add_img[b] in buffer order.
template <class T>
void backproject(const Problem<T>& P, T* image, const T* views, const int* index, int nindex, int nthread) {
  const long V = (long)P.nx * P.ny * P.nz;
  std::fill(image, image + V, T(0));
  const int nb = std::max(1, std::min(nthread, nindex));
  std::vector<ViewWork<T>*> W;
  std::vector<std::vector<T>> acc(nb, std::vector<T>(V, T(0)));
  for (int b = 0; b < nb; ++b) W.push_back(new ViewWork<T>(P));
  spawner(nb, nindex, [&](int b, int ii) {
    backproject_view(P, *W[b], acc[b].data(), views + (long)P.nx * P.nz * ii, index[ii]);
  });
  for (int b = 0; b < nb; ++b)
    for (long k = 0; k < V; ++k) image[k] += acc[b][k];
  for (auto* w : W) delete w;
}

// L = sum over the indexed views of ||A_v x - y_v||^2 (returned), its
// gradients w.r.t. x (the back-projection of 2 (A x - y)) and w.r.t. the
// plan's mu-map (grad_mu_view summed over the views in index order)
template <class T>
double grad(const Problem<T>& P, T* gx, T* gmu, const T* image, const T* yobs, const int* index, int nindex,
            int nthread) {
  const long S = (long)P.nx * P.nz, V = (long)P.nx * P.ny * P.nz;
  std::vector<T> views(S * nindex);
  project(P, views.data(), image, index, nindex, nthread);
  double loss = 0;
  for (long k = 0; k < S * nindex; ++k) {
    const T d = views[k] - yobs[k];
    loss += (double)d * (double)d;
    views[k] = T(2) * d;
  }
  backproject(P, gx, views.data(), index, nindex, nthread);
  std::fill(gmu, gmu + V, T(0));
  ViewWork<T> W(P);
  for (int ii = 0; ii < nindex; ++ii) grad_mu_view(P, W, gmu, image, views.data() + S * ii, index[ii]);
  return loss;
}

// src/ml-os-em.jl:21-49 (mlem!)
template <class T>
void mlem(const Problem<T>& P, T* out, const T* x0, const T* y, const T* b, int niter, int nthread) {
  const long V = (long)P.nx * P.ny * P.nz, M = (long)P.nx * P.nz * P.nview;
  std::vector<int> all(P.nview);
  for (int k = 0; k < P.nview; ++k) all[k] = k;
  std::vector<T> ones(M, T(1)), asum(V), ybar(M), yratio(M), back(V);
  backproject(P, asum.data(), ones.data(), all.data(), P.nview, nthread);
  for (auto& a : asum)
    if (a == T(0)) a = std::numeric_limits<T>::infinity();
  std::copy(x0, x0 + V, out);
  for (int it = 0; it < niter; ++it) {
    project(P, ybar.data(), out, all.data(), P.nview, nthread);
    for (long k = 0; k < M; ++k) yratio[k] = y[k] / (ybar[k] + b[k]);
    backproject(P, back.data(), yratio.data(), all.data(), P.nview, nthread);
    for (long k = 0; k < V; ++k) out[k] *= back[k] / asum[k];
  }
}

// src/ml-os-em.jl:70-131 (Ablock + osem!): block nb uses views nb:nblocks:nview
// (with nblocks not dividing nview, as the batched use allows, the blocks'
// sizes differ by at most one)
template <class T>
void osem(const Problem<T>& P, T* out, const T* x0, const T* y, const T* b, int nblocks, int niter, int nthread) {
  const long V = (long)P.nx * P.ny * P.nz, S = (long)P.nx * P.nz;
  std::vector<std::vector<int>> idx(nblocks);
  for (int nb = 0; nb < nblocks; ++nb)
    for (int v = nb; v < P.nview; v += nblocks) idx[nb].push_back(v);
  const int nmax = (int)idx[0].size();
  std::vector<T> ones(S * nmax, T(1)), ybar(S * nmax), yratio(S * nmax), back(V);
  std::vector<std::vector<T>> asum(nblocks, std::vector<T>(V));
  for (int nb = 0; nb < nblocks; ++nb) {
    backproject(P, asum[nb].data(), ones.data(), idx[nb].data(), (int)idx[nb].size(), nthread);
    for (auto& a : asum[nb])
      if (a == T(0)) a = std::numeric_limits<T>::infinity();
  }
  std::copy(x0, x0 + V, out);
  for (int it = 0; it < niter; ++it)
    for (int nb = 0; nb < nblocks; ++nb) {
      const int nsub = (int)idx[nb].size();
      project(P, ybar.data(), out, idx[nb].data(), nsub, nthread);
      for (int ii = 0; ii < nsub; ++ii) {
        const long v = idx[nb][ii];
        for (long k = 0; k < S; ++k)
          yratio[ii * S + k] = y[v * S + k] / (ybar[ii * S + k] + b[v * S + k]);
      }
      backproject(P, back.data(), yratio.data(), idx[nb].data(), nsub, nthread);
      for (long k = 0; k < V; ++k) out[k] *= back[k] / asum[nb][k];
    }
}

// src/psf-gauss.jl:24-55: Float64 1-D Gaussians (or Kronecker impulse for
// fwhm == 0), outer product, stored as T.
template <class T>
void psf_gauss(T* psf, int ny, int px, int pz, const double* fwhm_x, const double* fwhm_z) {
  auto g1 = [](std::vector<double>& v, int p, double fwhm) {
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
  };
  std::vector<double> gx, gz;
  for (int y = 0; y < ny; ++y) {
    g1(gx, px, fwhm_x[y]);
    g1(gz, pz, fwhm_z[y]);
    for (int j = 0; j < pz; ++j)
      for (int i = 0; i < px; ++i) psf[i + px * j + (long)px * pz * y] = T(gx[i] * gz[j]);
  }
}

template <class T>
Problem<T> make_problem(int nx, int ny, int nz, int px, int pz, int nview, T dy, const T* mumap, const T* psfs,
                        const double* angles, int method, int conv_kind) {
  Problem<T> P;
  P.nx = nx; P.ny = ny; P.nz = nz; P.px = px; P.pz = pz; P.nview = nview;
  P.method = method; P.conv_kind = conv_kind; P.dy = dy;
  P.mumap = mumap; P.psfs = psfs; P.angles = angles;
  return P;
}

}  // namespace

// ---------------------------------------------------------------------------
// C entry points (ctypes).  Suffix _f32 = reference Float32 order, _f64 =
// the same algorithm in Float64.  Angles are always passed as double.
// ---------------------------------------------------------------------------
extern "C" {

int or_rot_padsize(int nx) { return rot_padsize(nx); }
void or_psf_pads(int n, int p, int* up, int* down) {
  *up = padup(n, p);
  *down = paddown(n, p);
}

// linearinterp! + SparseInterpolator mul! for one row: used by the tests to
// pin the J/C tables against linear interpolation.
void or_linearinterp(double s, double e, int len, float* C, long* J) {
  SparseLin<float> A(len);
  linearinterp(A, s, e, len);
  for (long i = 0; i < A.ncoeff; ++i) {
    C[i] = A.C[i];
    J[i] = A.J[i];
  }
}

// quadrant rotation m (0..3) of an Np x Np matrix; returns 0, or -1 for the
// reference's "invalid m!" throw
int or_rot_f90_f32(float* out, const float* img, int Np, int m, int adj) {
  return rot_f90(out, img, Np, m, adj != 0) ? 0 : -1;
}

#define OR_ROT(SUF, T)                                                                    \
  void or_imrotate_##SUF(T* out, const T* img, int N, int nz, double theta, int method, int adj) { \
    RotWork<T> W(N);                                                                      \
    for (int z = 0; z < nz; ++z)                                                          \
      imrotate(out + (long)N * N * z, img + (long)N * N * z, theta, W, method, adj != 0); \
  }
OR_ROT(f32, float)
OR_ROT(f64, double)

#define OR_CONV(SUF, T)                                                                               \
  void or_fft_conv_##SUF(T* out, const T* img, int nx, int nz, const T* ker, int px, int pz, int adj, \
                         int kind) {                                                                  \
    if (kind == 0) {                                                                                  \
      PsfWork<T> W(nx, nz, px, pz);                                                                   \
      if (adj) W.conv_adj(out, nx, img, nx, ker);                                                     \
      else W.conv(out, nx, img, nx, ker);                                                             \
    } else {                                                                                          \
      conv_direct(out, nx, img, nx, nx, nz, ker, px, pz, adj != 0);                                   \
    }                                                                                                 \
  }
OR_CONV(f32, float)
OR_CONV(f64, double)

#define OR_PROJ(SUF, T)                                                                                     \
  void or_project_##SUF(T* views, const T* image, const T* mumap, const T* psfs, int nx, int ny, int nz,    \
                        int px, int pz, int nview, T dy, const double* angles, int method, const int* index, \
                        int nindex, int nthread, int conv_kind) {                                           \
    auto P = make_problem<T>(nx, ny, nz, px, pz, nview, dy, mumap, psfs, angles, method, conv_kind);        \
    project(P, views, image, index, nindex, nthread);                                                       \
  }                                                                                                         \
  void or_backproject_##SUF(T* image, const T* views, const T* mumap, const T* psfs, int nx, int ny,        \
                            int nz, int px, int pz, int nview, T dy, const double* angles, int method,      \
                            const int* index, int nindex, int nthread, int conv_kind) {                     \
    auto P = make_problem<T>(nx, ny, nz, px, pz, nview, dy, mumap, psfs, angles, method, conv_kind);        \
    backproject(P, image, views, index, nindex, nthread);                                                   \
  }                                                                                                         \
  void or_mlem_##SUF(T* out, const T* x0, const T* y, const T* b, const T* mumap, const T* psfs, int nx,   \
                     int ny, int nz, int px, int pz, int nview, T dy, const double* angles, int method,      \
                     int niter, int nthread, int conv_kind) {                                               \
    auto P = make_problem<T>(nx, ny, nz, px, pz, nview, dy, mumap, psfs, angles, method, conv_kind);        \
    mlem(P, out, x0, y, b, niter, nthread);                                                                 \
  }                                                                                                         \
  void or_osem_##SUF(T* out, const T* x0, const T* y, const T* b, const T* mumap, const T* psfs, int nx,   \
                     int ny, int nz, int px, int pz, int nview, T dy, const double* angles, int method,      \
                     int nblocks, int niter, int nthread, int conv_kind) {                                  \
    auto P = make_problem<T>(nx, ny, nz, px, pz, nview, dy, mumap, psfs, angles, method, conv_kind);        \
    osem(P, out, x0, y, b, nblocks, niter, nthread);                                                        \
  }                                                                                                         \
  double or_grad_##SUF(T* gx, T* gmu, const T* image, const T* yobs, const T* mumap, const T* psfs, int nx, \
                       int ny, int nz, int px, int pz, int nview, T dy, const double* angles, int method,      \
                       const int* index, int nindex, int nthread, int conv_kind) {                            \
    auto P = make_problem<T>(nx, ny, nz, px, pz, nview, dy, mumap, psfs, angles, method, conv_kind);        \
    return grad(P, gx, gmu, image, yobs, index, nindex, nthread);                                           \
  }                                                                                                         \
  void or_atten_##SUF(T* atten, const T* mumap, int nx, int ny, int nz, T dy, double theta, int method) {   \
    RotWork<T> W(nx);                                                                                       \
    const long sxy = (long)nx * ny;                                                                         \
    std::vector<T> mr(sxy * nz), e((size_t)nx * nz);                                                       \
    for (int z = 0; z < nz; ++z) imrotate(mr.data() + sxy * z, mumap + sxy * z, theta, W, method, false);  \
    for (int y = 0; y < ny; ++y) {                                                                          \
      atten_plane(e.data(), mr.data(), nx, ny, nz, y, dy);                                                  \
      for (long z = 0; z < nz; ++z)                                                                         \
        for (long x = 0; x < nx; ++x) atten[x + nx * y + sxy * z] = e[x + nx * z];                          \
    }                                                                                                       \
  }                                                                                                         \
  void or_psf_gauss_##SUF(T* psf, int ny, int px, int pz, const double* fwhm_x, const double* fwhm_z) {     \
    psf_gauss(psf, ny, px, pz, fwhm_x, fwhm_z);                                                             \
  }
OR_PROJ(f32, float)
OR_PROJ(f64, double)

}  // extern "C"
